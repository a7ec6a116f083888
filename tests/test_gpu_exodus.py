"""The wall path on one of the reference's own meshes, read from its Exodus file
(tests/golden/exodus/squeeze_D8_D6_L10.e: a tube narrowing from diameter 8 to 6 and back,
length 10, 587 vertices, 1,102 triangles) by rbc3d_read_wall_mesh and made periodic in z
(the box is the tube's length): three cells, two of them within rc of the wall, GPU
against the oracle for the wall self-matrix pattern, AddIntOnWalls at the cell nodes (the
real-space part, ~3e-5 at rc), the no-slip operator, the residual wall velocity and the
NoSlipWall solve.  Bar: 1e-12 relative (1e-10 for the
GMRES solution, same iteration count)."""
import os

import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12
MESH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "exodus",
                    "squeeze_D8_D6_L10.e")


@pytest.fixture(scope="module")
def squeeze():
    import rbc3d_amd
    from oracle import oracle_py
    from rbc3d_amd import configs
    x, e2v, nv, ne = rbc3d_amd.read_wall_mesh(MESH)
    X = x.reshape(3, nv) + np.array([5.0, 5.0, 5.0])[:, None]  # into the box [0, 10]^3
    wall = (X.ravel(), e2v, nv, ne)
    rng = np.random.default_rng(6)
    cells = configs.build_cells(rng, ["sphere"] * 3, [13] * 3, [0.8] * 3,
                                [(6.6, 5.0, 2.0), (5.1, 4.9, 5.0), (4.9, 3.5, 8.0)], [1, 1, 1])
    w = configs.Workload("squeeze", (10.0, 10.0, 10.0), configs.alpha_ref_from_xi(2.0), 2.0,
                         (32, 32, 32), 6, (1.0,), (0.0, 0.0, 1.0), cells, wall)
    g = rbc3d_amd.from_workload(w)
    o = oracle_py.from_workload(w)
    g.prepare_sing_int_on_wall()
    o.prepare_wall()
    return w, g, o


def test_self_matrix_and_cell_targets(squeeze):
    import ctypes as C
    import rbc3d_amd
    w, g, o = squeeze
    nnz = C.c_longlong()
    rbc3d_amd._call("rbc3d_wall_self_nnz", g.h, C.byref(nnz))
    assert nnz.value == len(o.wall_csr()[2])
    f = np.random.default_rng(4).standard_normal(3 * w.wall[2]) * 0.1
    g.set_wall_density(f)
    o.set_wall_density(f)
    tl = o.tlist_rbc()
    c = 1 / (4 * np.pi)
    vo = o.add_int_on_walls(c, tl)
    e = relerr(g.add_int_on_walls(c, tl), vo)
    print(f"squeeze mesh: self-matrix {nnz.value} entries, AddIntOnWalls at cells {e:.2e} "
          f"(max |v| {np.abs(vo).max():.2e})", flush=True)
    assert np.abs(vo).max() > 1e-6 and e < TOL


def test_operator_residual_and_noslip(squeeze):
    w, g, o = squeeze
    assert g.wall_dof() == o.wall_dof()
    u = np.random.default_rng(8).standard_normal(o.wall_dof())
    e_mv = relerr(g.wall_matmult(u), o.wall_matmult(u))
    f = np.random.default_rng(5).standard_normal(3 * w.wall[2]) * 0.1
    g.set_wall_density(f)
    o.set_wall_density(f)
    e_res = relerr(g.wall_residual_vel(), o.wall_residual_vel())
    f0 = np.zeros(3 * w.wall[2])
    g.set_wall_density(f0)
    o.set_wall_density(f0)
    ng, rg, cg, vg = g.noslip_wall(1e-8, 200)
    no, ro, co, vo = o.noslip_wall(1e-8, 200)
    e_f = relerr(g.wall_density(), o.wall_density())
    print(f"squeeze mesh: operator {e_mv:.2e}, residual {e_res:.2e}, NoSlipWall {ng} vs {no} "
          f"iterations, traction {e_f:.2e}, residual velocity after {vg:.2e}", flush=True)
    assert e_mv < TOL and e_res < TOL
    assert ng == no and cg and co and e_f < 1e-10
