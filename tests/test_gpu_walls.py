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
