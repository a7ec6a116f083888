"""Error paths of the C ABI, forced from the host so every kernel still runs to its end:
non-finite inputs (RBC3D_ENONFINITE), GMRES stopped by maxit (RBC3D_ENOCONV, with every
output written) and more than 32 neighbouring cells within rc of a target (RBC3D_ENBR;
the reference's NbrRbcList overflows silently, common/ModRbcSingInt.F90:319-362)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    import rbc3d_amd
    rbc3d_amd.lib()
    return rbc3d_amd


@pytest.fixture(scope="module")
def tube(R):
    from rbc3d_amd import configs
    w = configs.make_workload("C3s", seed=3)
    g = R.from_workload(w)
    g.prepare_sing_int_on_wall()
    yield w, g
    g.close()


def test_nonfinite_cell_coordinate_rejected(R):
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    g = R.System(**w.conf_kwargs())
    try:
        x = w.cells.x.copy()
        x[17] = np.nan
        cells = configs.Cells(w.cells.nlat0, w.cells.nlat, w.cells.celltype, x, w.cells.a3,
                              w.cells.detj, w.cells.f, w.cells.g)
        with pytest.raises(R.Rbc3dError) as e:
            g.set_cells(cells)
        assert e.value.code == 5
        # the context is still usable with finite input
        g.set_cells(w.cells)
        g.set_cell_density(1, w.cells.f)
    finally:
        g.close()


def test_nonfinite_target_rejected(R):
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    g = R.from_workload(w)
    try:
        tx, ti, act, A = g.tlist_rbc()
        tx = tx.copy()
        tx[5] = np.inf
        with pytest.raises(R.Rbc3dError) as e:
            g.add_int_on_rbcs(1.0, 0.0, (tx, ti, act, A))
        assert e.value.code == 5
    finally:
        g.close()


def test_nonfinite_density_ends_noslip_gmres(tube, R):
    """A NaN in the cells' force density reaches the no-slip right-hand side: GMRES stops
    on the non-finite residual and reports RBC3D_ENONFINITE instead of convergence."""
    w, g = tube
    f = w.cells.f.copy()
    f[3] = np.nan
    g.set_cell_density(1, f)
    try:
        g.set_wall_density(np.zeros(3 * w.wall[2]))
        g.solve_rbc_vel()
        with pytest.raises(R.Rbc3dError) as e:
            g.noslip_wall(1e-8, 100)
        assert e.value.code == 5
    finally:
        g.set_cell_density(1, w.cells.f)


def test_noslip_maxit_reports_not_converged(tube, R):
    """maxit before the tolerance: RBC3D_ENOCONV from the C ABI, every output written
    (a finite traction update and the iteration count), as the reference prints the
    residual and goes on (common/ModNoSlip.F90:120-126)."""
    w, g = tube
    nv = w.wall[2]
    g.set_wall_density(np.zeros(3 * nv))
    g.solve_rbc_vel()
    ni, rn, cv, vm = C.c_int(), C.c_double(), C.c_int(), C.c_double()
    code = R.lib().rbc3d_noslip_wall(g.h, 1e-8, 5, None, C.byref(ni), C.byref(rn), C.byref(cv),
                                     C.byref(vm))
    assert code == 6 and ni.value == 5 and cv.value == 0
    f = g.wall_density()
    assert np.isfinite(f).all() and np.abs(f).max() > 0 and np.isfinite(rn.value)
    # the Python binding reports it as not converged rather than raising
    g.set_wall_density(np.zeros(3 * nv))
    it, res, conv, _ = g.noslip_wall(1e-8, 5)
    assert it == 5 and not conv


def _crowded(R, orc):
    """48 spheres of radius 0.2 on a 0.5-spaced 4 x 4 x 3 lattice: targets on the central
    cells have all 47 others within rc = 2."""
    from rbc3d_amd import configs
    rng = np.random.default_rng(0)
    cs = np.array([(3.0 + 0.5 * i, 3.0 + 0.5 * j, 3.5 + 0.5 * k)
                   for i in range(4) for j in range(4) for k in range(3)])
    n = len(cs)
    cells = configs.build_cells(rng, ["sphere"] * n, [7] * n, [0.2] * n, cs, [1] * n)
    w = configs.Workload("crowded", (8.0, 8.0, 8.0), configs.alpha_ref_from_xi(2.0), 2.0,
                         (16, 16, 16), 6, (5.0,), (0.0, 0.0, 0.0), cells)
    return w


def test_more_than_32_neighbouring_cells(R, orc):
    w = _crowded(R, orc)
    g = R.from_workload(w)
    o = orc.from_workload(w)
    try:
        tl = o.tlist_rbc()
        with pytest.raises(R.Rbc3dError) as e:
            g.add_int_on_rbcs(1 / (4 * np.pi), 0.0, tl)
        assert e.value.code == 3
        with pytest.raises(RuntimeError, match="status 2"):
            o.add_int_on_rbcs(1 / (4 * np.pi), 0.0, tl)
        # the device and the context stay usable: the SH transforms of the cells still run
        nl = 7
        a, b = R.sh_anal_gau(nl, 2 * nl, w.cells.x[:3 * nl * 2 * nl], nt=3)
        assert np.isfinite(a).all()
    finally:
        g.close()
