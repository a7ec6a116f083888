"""GPU parity of the wall path (AddIntOnWalls, the sparse wall self-matrix, the
no-slip operator and its GMRES solve) and of the cell-velocity solve, against the
oracle on a small periodic tube with cells."""
import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def tube():
    import rbc3d_amd
    from oracle import oracle_py
    from rbc3d_amd import configs
    w = configs.make_workload("C3s", seed=2)
    g = rbc3d_amd.from_workload(w)
    o = oracle_py.from_workload(w)
    rng = np.random.default_rng(4)
    nv = w.wall[2]
    f = rng.standard_normal(3 * nv) * 0.1
    g.set_wall_density(f)
    o.set_wall_density(f)
    return w, g, o


def test_wall_self_matrix(tube):
    w, g, o = tube
    import ctypes as C
    import rbc3d_amd
    g.prepare_sing_int_on_wall()
    o.prepare_wall()
    nnz = C.c_longlong()
    rbc3d_amd._call("rbc3d_wall_self_nnz", g.h, C.byref(nnz))
    rp, col, val = o.wall_csr()
    assert nnz.value == len(val)


def test_wall_only_spread_grid(tube):
    """The wall-only single-layer spread (every no-slip GMRES iteration): per-brick
    owner-computes over the fixed centroids, each point summing its triangles in
    ascending order with the reference's expressions, gives the oracle's grid bit for
    bit."""
    w, g, o = tube
    c = 1 / (4 * np.pi)
    g.pme_distrib_source(c, 0.0, cells=False, walls=True)
    L = __import__("oracle").oracle_py.L()
    L.orc_pme_distrib_source(o.h, c, 0.0, 0, 1)
    gg = g.pme_grid(0)
    og = o.pme_grid(0)[:len(gg)]
    assert np.abs(og).max() > 0
    assert np.array_equal(gg, og), np.abs(gg - og).max()


def test_add_int_on_walls_cell_targets(tube):
    w, g, o = tube
    tl = o.tlist_rbc()
    c = 1 / (4 * np.pi)
    assert relerr(g.add_int_on_walls(c, tl), o.add_int_on_walls(c, tl)) < TOL


def test_wall_operator_and_residual(tube):
    w, g, o = tube
    assert g.wall_dof() == o.wall_dof()
    u = np.random.default_rng(8).standard_normal(o.wall_dof())
    assert relerr(g.wall_matmult(u), o.wall_matmult(u)) < TOL
    # restore the traction the matvec overwrote (MyMatMult sets wall%f)
    f = np.random.default_rng(4).standard_normal(3 * w.wall[2]) * 0.1
    g.set_wall_density(f)
    o.set_wall_density(f)
    assert relerr(g.wall_residual_vel(), o.wall_residual_vel()) < TOL


def test_noslip_wall_gmres(tube):
    w, g, o = tube
    f = np.zeros(3 * w.wall[2])
    g.set_wall_density(f)
    o.set_wall_density(f)
    ng, rg, cg, vg = g.noslip_wall(1e-8, 60)
    no, ro, co, vo = o.noslip_wall(1e-8, 60)
    assert ng == no and cg == co
    assert abs(rg - ro) <= 1e-6 * ro
    assert relerr(g.wall_density(), o.wall_density()) < 1e-10
    # the update enforces no slip: residual wall velocity far below the background flow
    assert vg < 1e-3 and abs(vg - vo) <= 1e-8


def test_noslip_wall_gmres_restart(tube):
    """GMRES with a restart length other than PETSc's 30: 60 iterations at restart 20
    cross two explicit restarts; the oracle runs the same length.  (Long cycles are
    compared on the device alone, below: CGS without refinement amplifies the last-bit
    differences of the two operators with the cycle length, 4e-5 in the residual
    estimate after 45 iterations of one cycle.)"""
    w, g, o = tube
    f = np.zeros(3 * w.wall[2])
    g.set_wall_density(f)
    o.set_wall_density(f)
    g.set_gmres_restart(20)
    o.set_gmres_restart(20)
    try:
        ng, rg, cg, vg = g.noslip_wall(1e-8, 60)
        no, ro, co, vo = o.noslip_wall(1e-8, 60)
    finally:
        g.set_gmres_restart(30)
        o.set_gmres_restart(30)
    assert ng == no == 60 and cg == co
    assert abs(rg - ro) <= 1e-6 * ro
    assert relerr(g.wall_density(), o.wall_density()) < 1e-10


def test_gmres_restart_changes_only_the_restart_point(tube):
    """Restarted GMRES: a longer restart repeats the shorter one's first cycle exactly
    (same Krylov basis, same Givens estimates) and then departs from it; both solves
    reach the tolerance and leave the same no-slip residual."""
    import rbc3d_amd
    w, g, o = tube
    f = np.zeros(3 * w.wall[2])
    hist, vres = {}, {}
    for m in (30, 100):
        g.set_wall_density(f)
        g.set_gmres_restart(m)
        try:
            n, r, conv, vm = g.noslip_wall(1e-8, 1000)
        finally:
            g.set_gmres_restart(30)
        assert conv
        hist[m] = g.gmres_history()
        vres[m] = vm
        assert len(hist[m]) == n + 1
    h30, h100 = hist[30], hist[100]
    print(f"restart 30: {len(h30) - 1} iterations, restart 100: {len(h100) - 1}")
    assert np.array_equal(h30[:31], h100[:31])
    assert len(h100) < len(h30)
    assert h100[-1] <= 1e-8 * h100[0] and h30[-1] <= 1e-8 * h30[0]
    assert abs(vres[30] - vres[100]) < 1e-6
    g.set_wall_density(o.wall_density())  # the state the oracle holds, for the tests below
    with pytest.raises(rbc3d_amd.Rbc3dError):
        g.set_gmres_restart(256)
    with pytest.raises(rbc3d_amd.Rbc3dError):
        g.set_gmres_restart(0)


def test_cell_velocity_solve(tube):
    w, g, o = tube
    vg, ng, _ = g.solve_rbc_vel()
    vo, no, _ = o.solve_rbc_vel()
    assert ng == no == 0  # viscosity ratio 1: the right-hand side is the solution
    assert relerr(vg, vo) < TOL


def test_cell_velocity_gmres_lambda5():
    import rbc3d_amd
    from oracle import oracle_py
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    g = rbc3d_amd.from_workload(w)
    o = oracle_py.from_workload(w)
    vg, ng, rg = g.solve_rbc_vel()
    vo, no, ro = o.solve_rbc_vel()
    assert ng == no and ng > 0
    assert relerr(vg, vo) < 1e-10
