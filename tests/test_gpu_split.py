"""The split-precision form (rbc3d_set_precision(ctx, RBC3D_PREC_SPLIT)) against the
FP64 oracle on identical inputs.

The split form runs the PME grids / FFTs / gather in FP32, the Legendre
GEMMs of the global SH transforms as three-way split BF16 products with FP32
accumulation, and stores the wall self-matrix in FP32; the singular and
near-singular quadratures and GMRES stay FP64.  Its real-space pair sums are FP64 by
default; with rbc3d_set_split_pairs(ctx, 1) they run in FP32 (target-bin-relative
coordinates, FP32 accumulation), which the pair-sum tests here switch on.  Bar: every
output within 1e-6 relative (max-norm) of the FP64 oracle; each test prints the error it
measured."""
import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-6


@pytest.fixture(scope="module")
def R():
    import rbc3d_amd
    rbc3d_amd.lib()
    return rbc3d_amd


@pytest.fixture(scope="module")
def c1split(R, orc):
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    g = R.from_workload(w)
    g.set_precision(g.PREC_SPLIT)
    g.set_split_pairs(True)
    return g, orc.from_workload(w)


@pytest.mark.parametrize("c1,c2", [(1 / (4 * np.pi), 0.0), (0.0, -1 / (4 * np.pi)),
                                   (1 / (4 * np.pi), 1 / (4 * np.pi))])
def test_split_pair_sums_c1(c1split, c1, c2):
    g, o = c1split
    tl = o.tlist_rbc()
    vo, npairs = o.add_int_on_rbcs(c1, c2, tl)
    vg = g.add_int_on_rbcs(c1, c2, tl)
    ng, _ = g.pair_stats()
    err = relerr(vg, vo)
    print(f"split AddIntOnRbcs c1={c1:.3g} c2={c2:.3g}: rel err {err:.2e}, pairs {ng} vs {npairs}")
    # the FP32 cut-off test may only move pairs within rounding of rc; the discs'
    # diameter equals rc = 2, so antipodal nodes of one equator ring sit at rc to
    # within rounding (52 such pairs at C1, |r - rc| < 1e-6), and their Ewald
    # coefficients are erfc(xi rc) ~ 1e-8 of the near field
    assert abs(ng - npairs) <= max(2, 1e-4 * npairs)
    assert err < TOL


def test_split_pme_c1(c1split):
    g, o = c1split
    tl = o.tlist_rbc()
    c = 1 / (4 * np.pi)
    vo = o.pme(c, c, tl)
    vg = g.pme(c, c, tl)
    errs = [relerr(g.pme_grid(w), o.pme_grid(w)) for w in (0, 1, 2)]
    err = relerr(vg, vo)
    print(f"split PME: grids {errs[0]:.2e} {errs[1]:.2e} {errs[2]:.2e}, velocity {err:.2e}")
    assert max(errs) < TOL and err < TOL


@pytest.mark.parametrize("nlat_case", ["C1", "C2s"])
def test_split_glob_sph_trans(R, orc, nlat_case):
    from rbc3d_amd import configs
    if nlat_case == "C1":
        w = configs.make_workload("C1", seed=1)
    else:
        w = configs.make_workload("C2", seed=3, ncell=48, Nb=(48, 48, 48))
    g = R.from_workload(w)
    g.set_precision(g.PREC_SPLIT)
    o = orc.from_workload(w)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(o.cell_dof())
    vo = o.glob_sph_trans(c=u, direction=-1)
    e_syn = relerr(g.glob_sph_trans(c=u, direction=-1), vo)
    e_ana = relerr(g.glob_sph_trans(v=vo, direction=1), o.glob_sph_trans(v=vo, direction=1))
    print(f"split Glob_Sph_Trans ({nlat_case}): synthesis {e_syn:.2e}, analysis {e_ana:.2e}")
    assert e_syn < TOL and e_ana < TOL
    # the split products are far more accurate than one BF16 term (2^-9)
    assert e_syn < 1e-6 and e_ana < 1e-6


def test_split_cell_matvec_and_rhs_c1(c1split):
    g, o = c1split
    u = np.random.default_rng(3).standard_normal(o.cell_dof())
    e_mv = relerr(g.cell_matmult(u), o.cell_matmult(u))
    e_rhs = relerr(g.cell_rhs(), o.cell_rhs())
    e_vf = relerr(g.calc_velocity_field(o.tlist_rbc()), o.calc_velocity_field(o.tlist_rbc()))
    print(f"split C1: cell matvec {e_mv:.2e}, rhs {e_rhs:.2e}, CalcVelocityField {e_vf:.2e}")
    assert max(e_mv, e_rhs, e_vf) < TOL


def test_split_dense_subset(R, orc):
    """The dense 48-cell piece of the 20% packing: near-singular corrections active."""
    from rbc3d_amd import configs
    w = configs.make_workload("C2", seed=3, ncell=48, Nb=(48, 48, 48))
    g = R.from_workload(w)
    g.set_precision(g.PREC_SPLIT)
    g.set_split_pairs(True)
    o = orc.from_workload(w)
    tl = o.tlist_rbc()
    c = 1 / (4 * np.pi)
    vo, npairs = o.add_int_on_rbcs(c, -c, tl)
    e_pair = relerr(g.add_int_on_rbcs(c, -c, tl), vo)
    u = np.random.default_rng(2).standard_normal(o.cell_dof())
    e_mv = relerr(g.cell_matmult(u), o.cell_matmult(u))
    print(f"split dense 48 cells: AddIntOnRbcs {e_pair:.2e}, cell matvec {e_mv:.2e}")
    assert e_pair < TOL and e_mv < TOL


def test_split_default_pairs_are_fp64(R):
    """The split form's pair sums are FP64 unless FP32 pairs are asked for: AddIntOnRbcs
    (pair sums and FP64 quadratures, no PME) equals the plain form's bit for bit."""
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    g0, g1 = R.from_workload(w), R.from_workload(w)
    g1.set_precision(g1.PREC_SPLIT)
    tl = g0.tlist_rbc()
    c = 1 / (4 * np.pi)
    v0 = g0.add_int_on_rbcs(c, c, tl)
    assert np.array_equal(g1.add_int_on_rbcs(c, c, tl), v0)
    g1.set_split_pairs(True)
    assert not np.array_equal(g1.add_int_on_rbcs(c, c, tl), v0)


def test_split_switch_back_is_exact(R, orc):
    """Switching back to FP64 restores the plain form bit for bit."""
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    g = R.from_workload(w)
    u = np.random.default_rng(9).standard_normal(g.cell_dof())
    b0 = g.cell_matmult(u)
    g.set_precision(g.PREC_SPLIT)
    bs = g.cell_matmult(u)
    g.set_precision(g.PREC_FP64)
    b1 = g.cell_matmult(u)
    assert np.array_equal(b0, b1)
    assert not np.array_equal(b0, bs)


@pytest.fixture(scope="module")
def tube_split():
    import rbc3d_amd
    from oracle import oracle_py
    from rbc3d_amd import configs
    w = configs.make_workload("C3s", seed=2)
    g = rbc3d_amd.from_workload(w)
    g.set_precision(g.PREC_SPLIT)
    o = oracle_py.from_workload(w)
    g.prepare_sing_int_on_wall()
    o.prepare_wall()  # the oracle assembles the wall self-matrix only on request
    return w, g, o


def test_split_wall_operator(tube_split):
    w, g, o = tube_split
    f = np.random.default_rng(4).standard_normal(3 * w.wall[2]) * 0.1
    g.set_wall_density(f)
    o.set_wall_density(f)
    e_res = relerr(g.wall_residual_vel(), o.wall_residual_vel())
    u = np.random.default_rng(8).standard_normal(o.wall_dof())
    e_op = relerr(g.wall_matmult(u), o.wall_matmult(u))
    print(f"split wall: residual velocity {e_res:.2e}, no-slip operator {e_op:.2e}")
    assert e_res < TOL and e_op < TOL


def test_split_timestep_velocity(tube_split):
    """One tube timestep (Solve_RBC_Vel, NoSlipWall to 1e-8, Solve_RBC_Vel with the
    new traction) in the split form: the cell velocities match the FP64 oracle's."""
    w, g, o = tube_split
    f0 = np.zeros(3 * w.wall[2])
    g.set_wall_density(f0)
    o.set_wall_density(f0)
    g.solve_rbc_vel()
    o.solve_rbc_vel()
    ng, _, cg, _ = g.noslip_wall(1e-8, 300)
    no, _, co, _ = o.noslip_wall(1e-8, 300)
    vg, _, _ = g.solve_rbc_vel()
    vo, _, _ = o.solve_rbc_vel()
    e_v = relerr(vg, vo)
    e_f = relerr(g.wall_density(), o.wall_density())
    print(f"split tube timestep: GMRES {ng} vs {no} iterations, cell velocity {e_v:.2e}, "
          f"wall traction {e_f:.2e}")
    assert cg and co
    assert e_v < TOL


def test_chosen_form_timestep_velocity(tube_split):
    """The chosen form of the tube timestep: the split form with GMRES(150) against the
    FP64 oracle with GMRES(30), both to 1e-8: the cell velocities agree to the split
    form's tolerance, in fewer iterations."""
    w, g, o = tube_split
    f0 = np.zeros(3 * w.wall[2])
    g.set_wall_density(f0)
    o.set_wall_density(f0)
    g.solve_rbc_vel()
    o.solve_rbc_vel()
    g.set_gmres_restart(150)
    try:
        ng, _, cg, _ = g.noslip_wall(1e-8, 1000)
    finally:
        g.set_gmres_restart(30)
    no, _, co, _ = o.noslip_wall(1e-8, 1000)
    vg, _, _ = g.solve_rbc_vel()
    vo, _, _ = o.solve_rbc_vel()
    e_v = relerr(vg, vo)
    print(f"chosen tube timestep: GMRES(150) {ng} vs GMRES(30) {no} iterations, "
          f"cell velocity {e_v:.2e}")
    assert cg and co and ng < no
    assert e_v < TOL
