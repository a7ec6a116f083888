"""GPU parity: the HIP path (C ABI) against the oracle on identical inputs.

Bars: integer/index work (cell lists, pair sets) bit-exact; FP64 velocities and
coefficients within 1e-12 relative (max-norm) of the oracle."""
import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def R():
    import rbc3d_amd
    rbc3d_amd.lib()
    return rbc3d_amd


@pytest.fixture(scope="module")
def c1w():
    from rbc3d_amd import configs
    return configs.make_workload("C1", seed=1)


@pytest.fixture(scope="module")
def c1pair(R, orc, c1w):
    g = R.from_workload(c1w)
    o = orc.from_workload(c1w)
    return g, o


def test_hashtable_bitexact(R, orc):
    rng = np.random.default_rng(5)
    for n, Nc, L in [(1000, (4, 4, 4), 8.0), (5000, (7, 5, 3), 15.0), (17, (3, 3, 3), 6.0)]:
        x = rng.uniform(-1.5 * L, 2.5 * L, size=(n, 3))
        # points exactly on bin boundaries and at the box edges
        x[:5] = np.array([[0.0, 0.0, 0.0], [L, L, L], [-L, 2 * L, 0.5 * L], [L / Nc[0], 0, 0],
                          [-L / Nc[0], -L / Nc[1], -L / Nc[2]]])
        xf = np.ascontiguousarray(x.reshape(-1, order="F"))
        iLbNc = np.array([1.0 / L * Nc[0], 1.0 / L * Nc[1], Nc[2] / L])
        hoc, nxt = R.hashtable_build(Nc, iLbNc, xf)
        ho = np.zeros_like(hoc)
        no = np.zeros(n, np.int32)
        orc.L().orc_hashtable_build(np.array(Nc, np.int32), iLbNc, n, xf, n, ho, no)
        assert np.array_equal(hoc, ho)
        assert np.array_equal(nxt, no)


@pytest.mark.parametrize("nlat", [13, 17])
def test_sh_transforms(R, orc, nlat):
    rng = np.random.default_rng(nlat)
    nlon = 2 * nlat
    nt = 3
    a = np.zeros((nlat, nlat, nt))
    b = np.zeros((nlat, nlat, nt))
    for k in range(nt):
        a[:, :, k] = np.triu(rng.standard_normal((nlat, nlat)))
        b[:, :, k] = np.triu(rng.standard_normal((nlat, nlat)))
        b[0, :, k] = 0
    af, bf = a.reshape(-1, order="F"), b.reshape(-1, order="F")
    g = R.sh_synth_gau(nlat, nlon, af, bf, nt=nt)
    go = np.zeros_like(g)
    orc.L().orc_sh_synth_gau(nlat, nlon, nt, go, nlat, nlon, af, bf, nlat, nlat)
    assert relerr(g, go) < TOL
    ge = R.sh_synth_equ(nlat, nlon, af, bf, nt=nt)
    geo = np.zeros_like(ge)
    orc.L().orc_sh_synth_equ(nlat, nlon, nt, geo, nlat, nlon, af, bf, nlat, nlat)
    assert relerr(ge, geo) < TOL
    a2, b2 = R.sh_anal_gau(nlat, nlon, g, nt=nt)
    ao = np.zeros_like(a2)
    bo = np.zeros_like(b2)
    orc.L().orc_sh_anal_gau(nlat, nlon, nt, g, nlat, nlon, ao, bo, nlat, nlat)
    mask = np.triu(np.ones((nlat, nlat), bool))[:, :, None].repeat(nt, 2).reshape(-1, order="F")
    assert relerr(a2[mask], ao[mask]) < TOL and relerr(b2[mask], bo[mask]) < TOL
    # band-limited round trip
    assert relerr(a2[mask], af[mask]) < 1e-12


@pytest.mark.parametrize("nlat", [13, 17])
def test_sh_gradient(R, orc, nlat):
    """ShGradGau on the GPU against the oracle (three fields at once)."""
    rng = np.random.default_rng(nlat + 7)
    nlon, nt = 2 * nlat, 3
    a = np.zeros((nlat, nlat, nt))
    b = np.zeros((nlat, nlat, nt))
    for k in range(nt):
        a[:, :, k] = np.triu(rng.standard_normal((nlat, nlat)))
        b[:, :, k] = np.triu(rng.standard_normal((nlat, nlat)))
    af = np.ascontiguousarray(a.reshape(-1, order="F"))
    bf = np.ascontiguousarray(b.reshape(-1, order="F"))
    vg, wg = R.sh_grad_gau(nlat, nlon, af, bf, nt=nt)
    vo = np.zeros(nlat * nlon * nt)
    wo = np.zeros(nlat * nlon * nt)
    orc.L().orc_sh_grad_gau(nlat, nlon, nt, vo, wo, nlat, nlon, af, bf, nlat, nlat)
    assert relerr(vg, vo) < TOL and relerr(wg, wo) < TOL


def test_surface_splines(c1pair):
    """The surface spline tables (values and the three derivative tables of x, a3,
    detJ, f detJ, g detJ) are the oracle's bit for bit: the device sums in the same order
    without FMA contraction from the same host tables.  The double-layer self patch
    subtracts spline positions ~0.01 apart at magnitude ~L, so a one-ulp table
    difference would move its velocities by ~1e-12."""
    g, o = c1pair
    nl = 13
    for which, nvar in [(0, 3), (1, 3), (2, 1), (3, 3), (4, 3)]:
        for irbc in range(2):
            ug = g.cell_spline(irbc, which, 2 * nl, 2 * nl, nvar)
            uo = o.cell_spline(irbc, which, 2 * nl, 2 * nl, nvar)
            for a, b in zip(ug, uo):
                assert np.array_equal(a, b), (which, irbc, np.abs(a - b).max())


@pytest.mark.parametrize("c1,c2", [(1 / (4 * np.pi), 0.0), (0.0, -1 / (4 * np.pi)),
                                   (1 / (4 * np.pi), 1 / (4 * np.pi))])
def test_add_int_on_rbcs_c1(c1pair, c1, c2):
    g, o = c1pair
    tl = o.tlist_rbc()
    vo, npairs = o.add_int_on_rbcs(c1, c2, tl)
    vg = g.add_int_on_rbcs(c1, c2, tl)
    ng, cs = g.pair_stats()
    assert ng == npairs
    assert cs == o.pair_checksum()
    assert relerr(vg, vo) < TOL


def test_pme_c1(c1pair):
    g, o = c1pair
    tl = o.tlist_rbc()
    c = 1 / (4 * np.pi)
    vo = o.pme(c, c, tl)
    vg = g.pme(c, c, tl)
    for which in (0, 1, 2):
        assert relerr(g.pme_grid(which), o.pme_grid(which)) < TOL, which
    assert relerr(vg, vo) < TOL


def test_calc_velocity_field_c1(c1pair):
    g, o = c1pair
    tl = o.tlist_rbc()
    assert relerr(g.calc_velocity_field(tl), o.calc_velocity_field(tl)) < TOL


def test_glob_sph_trans_and_matmult_c1(c1pair):
    g, o = c1pair
    rng = np.random.default_rng(11)
    u = rng.standard_normal(o.cell_dof())
    vg = g.glob_sph_trans(c=u, direction=-1)
    vo = o.glob_sph_trans(c=u, direction=-1)
    assert relerr(vg, vo) < TOL
    assert relerr(g.glob_sph_trans(v=vo, direction=1), o.glob_sph_trans(v=vo, direction=1)) < TOL
    bg = g.cell_matmult(u)
    bo = o.cell_matmult(u)
    assert relerr(bg, bo) < TOL


def test_cell_rhs_c1(c1pair):
    g, o = c1pair
    assert relerr(g.cell_rhs(), o.cell_rhs()) < TOL


def test_dense_subset(R, orc):
    """A 48-cell piece of the dense 20% packing: many neighbouring cells, near-singular
    corrections active; pair set bit-exact, velocities to 1e-12."""
    from rbc3d_amd import configs
    w = configs.make_workload("C2", seed=3, ncell=48, Nb=(48, 48, 48))
    g = R.from_workload(w)
    o = orc.from_workload(w)
    tl = o.tlist_rbc()
    c = 1 / (4 * np.pi)
    vo, npairs = o.add_int_on_rbcs(c, -c, tl)
    vg = g.add_int_on_rbcs(c, -c, tl)
    ng, cs = g.pair_stats()
    assert ng == npairs and cs == o.pair_checksum()
    assert relerr(vg, vo) < TOL
    u = np.random.default_rng(2).standard_normal(o.cell_dof())
    assert relerr(g.cell_matmult(u), o.cell_matmult(u)) < TOL


def test_dense_packing_source_filter(R, orc):
    """128 cells of the dense 20% packing in a 10.0 box: ten fine bins per dimension (the
    image path) holding 74 sources each on average, where the pair kernel's loader keeps
    only the sources within reach of each 64-target chunk.  Every target active: pair
    set bit-exact, velocities and the cell matvec to 1e-12."""
    from rbc3d_amd import configs
    w = configs.make_workload("C2", seed=4, ncell=128, Nb=(64, 64, 64))
    nf = 2 * int(w.Lb[0] // 2.0)  # fine bins per dimension at rc = 2
    assert nf >= 7 and w.cells.npoint >= 32 * nf ** 3  # image path, filter on
    g = R.from_workload(w)
    o = orc.from_workload(w)
    tl = o.tlist_rbc()
    c = 1 / (4 * np.pi)
    vo, npairs = o.add_int_on_rbcs(c, -c, tl)
    vg = g.add_int_on_rbcs(c, -c, tl)
    ng, cs = g.pair_stats()
    e_v = relerr(vg, vo)
    u = np.random.default_rng(5).standard_normal(o.cell_dof())
    e_m = relerr(g.cell_matmult(u), o.cell_matmult(u))
    print(f"dense 128-cell packing: {ng} pairs, AddIntOnRbcs {e_v:.2e}, matvec {e_m:.2e}",
          flush=True)
    assert ng == npairs and cs == o.pair_checksum()
    assert e_v < TOL and e_m < TOL


def test_geometry_reuse(R, orc):
    """Matvecs that reuse the geometry-only results (neighbour lists, projections) of
    the first one give exactly the values of matvecs that recompute them; after the
    cells are set again the reuse starts over, and the result matches the oracle."""
    from rbc3d_amd import configs
    w = configs.make_workload("C2", seed=3, ncell=48, Nb=(48, 48, 48))
    g = R.from_workload(w)
    rng = np.random.default_rng(21)
    dof = g.cell_dof()
    u1, u2 = rng.standard_normal(dof), rng.standard_normal(dof)
    g.set_geometry_reuse(True)
    g.cell_matmult(u1)
    b_reuse = g.cell_matmult(u2)
    assert g.near_stats()[1] > 0  # near-singular corrections are exercised
    g.set_geometry_reuse(False)
    b_full = g.cell_matmult(u2)
    assert np.array_equal(b_reuse, b_full)
    # move the cells: a stale reuse would keep the old neighbour lists and projections
    g.set_geometry_reuse(True)
    c = w.cells
    shift = np.zeros_like(c.x)
    off = 0
    for nl in c.nlat:
        npt = int(nl) * 2 * int(nl)
        shift[3 * off: 3 * off + npt] = 0.05  # x components of the cell block
        off += npt
    from dataclasses import replace
    c2 = replace(c, x=c.x + shift)
    g.set_cells(c2)
    g.set_cell_density(1, c2.f)
    g.set_cell_density(2, c2.g)
    b_moved = g.cell_matmult(u2)
    o = orc.from_workload(replace(w, cells=c2))
    assert relerr(b_moved, o.cell_matmult(u2)) < 2e-12
