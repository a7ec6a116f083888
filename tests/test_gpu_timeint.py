"""GPU parity of the device time stepping (rbc3d_amd/csrc/timeint.hip, mech.hip) against
the oracle's restatement of TimeInt_Euler (oracle/timeint_oracle.py): deformed cells
relaxing over three Euler steps at viscosity ratio 5 (GMRES in every step), and one step
of a tube with cells, the background flow and NoSlipWall.

Bar: 1e-12 in the coordinates (cell size 1): the same arithmetic up to summation order,
with velocities from GMRES solves that take the same iterations."""
import numpy as np
import pytest

from oracle import oracle_py as O
from oracle import timeint_oracle as T
from rbc3d_amd import configs
from tests.test_oracle_timeint import NLAT, deformed_cells, ref_cell, workload

pytestmark = pytest.mark.gpu
TOL = 1e-12
ES, ED, EB = 1.0, 20.0, 0.03


def test_euler_steps_match_the_oracle():
    import rbc3d_amd as R
    rng = np.random.default_rng(7)
    rots = [configs.random_rotation(rng) for _ in range(2)]
    cells = deformed_cells([(2.5, 2.5, 2.5), (5.5, 5.6, 5.4)], seed=7, rots=rots)
    w = workload(cells, 5.0)
    ref = ref_cell()
    Ts, nsteps, eps = 0.02, 3, 0.05
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], ref_x(), [ES], [ED], [EB])
    t, ic, iw = g.timeint_euler(Ts, nsteps, eps)
    xg = g.rbc_coords()
    xo, its = T.timeint_euler(O.from_workload(w), cells, [ref], [ES], [ED], [EB], Ts, nsteps, eps,
                              w.Lb)
    disp = np.abs(xo - cells.x).max()
    err = np.abs(xg - xo).max()
    print(f"3 steps: moved {disp:.3e}, GPU vs oracle {err:.2e}; iterations {list(ic)} vs "
          f"{[i for i, _ in its]}", flush=True)
    assert disp > 1e-3
    assert err < TOL
    assert t == pytest.approx(nsteps * Ts, abs=1e-15)
    assert all(abs(a - b[0]) <= 1 for a, b in zip(ic, its))


def ref_x():
    from tests.test_mech_oracle import flat, grid, shape_fn
    from tests.test_oracle_timeint import AXES
    th, phi, _ = grid(NLAT)
    return flat(shape_fn(0, 0.0, axes=AXES)(*np.meshgrid(th, phi, indexing="ij")))


def test_pieces_match_the_oracle():
    """rbc3d_compute_rbc_vel, rbc3d_move_rbcs, rbc3d_rebox_rbcs (cells moved out of the
    box first), rbc3d_vol_constrain_rbcs and rbc3d_filter_rbcs one at a time."""
    import rbc3d_amd as R
    cells = deformed_cells([(2.5, 2.5, 2.5), (5.5, 5.6, 5.4)], seed=11)
    w = workload(cells, 5.0)
    ref = ref_cell()
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], ref_x(), [ES], [ED], [EB])
    v, ni, _ = g.compute_rbc_vel()
    o = O.from_workload(w)
    fs = [T.cell_force(x, NLAT, ref, ES, ED, EB)[1] for x in T.split_cells(cells.x, cells.nlat, 3)]
    o.set_density(1, np.concatenate(fs))
    vo, no, _ = o.solve_rbc_vel()
    e_v = np.abs(v - vo).max() / np.abs(vo).max()
    g.move_rbcs(0.5)
    xs = T.split_cells(cells.x, cells.nlat, 3)
    V = vo.reshape(3, -1)
    npt = 2 * NLAT * NLAT
    xs = [(x.reshape(3, -1) + 0.5 * V[:, c * npt:(c + 1) * npt]).ravel() for c, x in enumerate(xs)]
    e_m = np.abs(g.rbc_coords() - np.concatenate(xs)).max()
    g.rebox_rbcs()
    g.vol_constrain_rbcs(0.05)
    g.filter_rbcs()
    xs = [T.rebox(x, w.Lb) for x in xs]
    xs = [O.rbc_sph_project(T.vol_constrain(x, ref["vol"], 0.05), NLAT, NLAT) for x in xs]
    e_x = np.abs(g.rbc_coords() - np.concatenate(xs)).max()
    print(f"velocity {e_v:.2e} ({ni} vs {no} iterations), move {e_m:.2e}, "
          f"rebox + volume + filter {e_x:.2e}", flush=True)
    assert e_v < 1e-9 and e_m < TOL and e_x < TOL


def test_rebox_is_exact():
    import rbc3d_amd as R
    cells = deformed_cells([(2.5, 2.5, 2.5), (5.5, 5.6, 5.4)], seed=12)
    X = cells.x.reshape(2, 3, -1).copy()
    X[0] += np.array([8.0, -16.0, 24.0])[:, None]
    X[1] += np.array([0.0, 0.0, -8.0])[:, None]
    moved = X.ravel()
    import dataclasses
    g = R.from_workload(workload(dataclasses.replace(cells, x=moved), 5.0))
    g.rebox_rbcs()
    want = np.concatenate([T.rebox(x, (8.0, 8.0, 8.0)) for x in T.split_cells(moved, cells.nlat, 3)])
    assert np.array_equal(g.rbc_coords(), want)


def test_tube_step_with_noslip_wall():
    import rbc3d_amd as R
    w = configs.make_workload("C3s", seed=2)
    ref = ref_cell()
    Ts, eps = 0.02, 0.05
    # the discs relax toward the ellipsoid reference shape while the flow carries them
    refx = ref_x()
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], refx, [ES], [ED], [EB])
    t, ic, iw = g.timeint_euler(Ts, 1, eps, noslip_rtol=1e-7, noslip_maxit=200)
    o = O.from_workload(w)
    xo, its = T.timeint_euler(o, w.cells, [ref], [ES], [ED], [EB], Ts, 1, eps, w.Lb,
                              noslip=(1e-7, 200))
    err = np.abs(g.rbc_coords() - xo).max()
    e_f = np.abs(g.wall_density() - o.wall_density()).max() / np.abs(o.wall_density()).max()
    print(f"tube step: coordinates {err:.2e}, wall traction {e_f:.2e}, iterations "
          f"{ic[0]}/{iw[0]} vs {its[0]}", flush=True)
    assert err < TOL and e_f < 1e-9 and iw[0] == its[0][1]


def test_rk2_and_ab2_match_the_oracle():
    import rbc3d_amd as R
    rng = np.random.default_rng(3)
    rots = [configs.random_rotation(rng) for _ in range(2)]
    cells = deformed_cells([(2.5, 2.5, 2.5), (5.5, 5.6, 5.4)], seed=4, rots=rots)
    w = workload(cells, 5.0)
    ref = ref_cell()
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], ref_x(), [ES], [ED], [EB])
    t, ic, _ = g.timeint_rk2(2.0 / 3.0, 0.02, 2)
    xo, its = T.timeint_rk2(O.from_workload(w), cells, [ref], [ES], [ED], [EB], 2.0 / 3.0, 0.02, 2,
                            w.Lb)
    e_rk = np.abs(g.rbc_coords() - xo).max()
    assert t == pytest.approx(0.04, abs=1e-15)
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], ref_x(), [ES], [ED], [EB])
    t2, ic2, _ = g.timeint_ab2(0.02, 3)
    xa, its2 = T.timeint_ab2(O.from_workload(w), cells, [ref], [ES], [ED], [EB], 0.02, 3, w.Lb)
    e_ab = np.abs(g.rbc_coords() - xa).max()
    moved = np.abs(xa - cells.x).max()
    print(f"RK2 {e_rk:.2e} (iterations {list(ic)} vs {[i for i, _ in its]}), AB2 {e_ab:.2e} "
          f"(moved {moved:.3e})", flush=True)
    assert t2 == pytest.approx(0.04, abs=1e-15)  # the first (Euler) step does not advance time
    assert e_rk < TOL and e_ab < TOL and moved > 1e-3
    assert all(abs(a - b[0]) <= 1 for a, b in zip(ic, its))


def test_ab2_tube_with_noslip_wall():
    import rbc3d_amd as R
    w = configs.make_workload("C3s", seed=2)
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], ref_x(), [ES], [ED], [EB])
    _, ic, iw = g.timeint_ab2(0.02, 2, noslip_rtol=1e-7, noslip_maxit=200)
    o = O.from_workload(w)
    xo, its = T.timeint_ab2(o, w.cells, [ref_cell()], [ES], [ED], [EB], 0.02, 2, w.Lb,
                            noslip=(1e-7, 200))
    err = np.abs(g.rbc_coords() - xo).max()
    e_f = np.abs(g.wall_density() - o.wall_density()).max() / np.abs(o.wall_density()).max()
    print(f"AB2 tube: coordinates {err:.2e}, wall traction {e_f:.2e}, wall iterations "
          f"{list(iw)} vs {[b for _, b in its]}", flush=True)
    assert err < TOL and e_f < 1e-9 and list(iw) == [b for _, b in its]


def test_operators_after_a_cell_update_see_the_moved_cells():
    """A cell update (here ReboxRbcs bringing a cell back from one period away) leaves the
    surface sources of the old coordinates; the next operator call rebuilds them first,
    so AddIntOnRbcs and the cell matvec equal those of a context set up at the new
    coordinates."""
    import dataclasses
    import rbc3d_amd as R
    from tests.conftest import relerr
    cells = deformed_cells([(2.5, 2.5, 2.5), (5.5, 5.6, 5.4)], seed=13)
    w = workload(cells, 5.0)
    X = cells.x.reshape(2, 3, -1).copy()
    X[0] += np.array([8.0, 0.0, -8.0])[:, None]
    g1 = R.from_workload(dataclasses.replace(w, cells=dataclasses.replace(cells, x=X.ravel())))
    g1.rebox_rbcs()
    g2 = R.from_workload(w)
    tl = g2.tlist_rbc()
    c = 1 / (4 * np.pi)
    e_add = relerr(g1.add_int_on_rbcs(c, c, tl), g2.add_int_on_rbcs(c, c, tl))
    u = np.random.default_rng(3).standard_normal(g2.cell_dof())
    e_mv = relerr(g1.cell_matmult(u), g2.cell_matmult(u))
    print(f"after ReboxRbcs: AddIntOnRbcs {e_add:.2e}, cell matvec {e_mv:.2e}", flush=True)
    assert e_add < 1e-10 and e_mv < 1e-10


def test_maxit_does_not_stop_the_run():
    """A NoSlipWall that reaches its maxit is reported through its count and the step
    goes on (the reference prints the residual and continues)."""
    import rbc3d_amd as R
    w = configs.make_workload("C3s", seed=2)
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], ref_x(), [ES], [ED], [EB])
    t, ic, iw = g.timeint_euler(0.02, 2, 0.05, noslip_rtol=1e-14, noslip_maxit=3)
    print(f"maxit 3: wall iterations {list(iw)}, time {t}", flush=True)
    assert list(iw) == [3, 3] and t == pytest.approx(0.04)
    assert np.all(np.isfinite(g.rbc_coords()))
