"""GPU parity of the repulsion (rbc3d_amd/csrc/repulsion.hip) against the oracle's
restatement (oracle/repulsion_oracle.py): AddIntraCellRepulsionForce on thin cells of two
grids, InterCellRepulsion between cells a gap apart, over a flat wall with a rigid cell
(rigidsep), and between cells and the tube wall, and two Euler steps with the repulsion
on.  Bar: 1e-12 (the same arithmetic; the neighbour searches find the same nodes and
triangles)."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle_py as O
from oracle import repulsion_oracle as RP
from oracle import timeint_oracle as T
from rbc3d_amd import configs
from tests.test_mech_oracle import ellipsoid
from tests.test_oracle_repulsion import (EPS, cells_of, flat_wall, shape, sphere_fn, two_spheres,
                                         workload)

pytestmark = pytest.mark.gpu
TOL = 1e-12


def gpu(w):
    import rbc3d_amd as R
    return R.from_workload(w)


def test_intra_two_grids():
    rng = np.random.default_rng(5)
    xs13 = [shape(13, ellipsoid(1.0, 1.0, 0.08), (2.0, 2.0, 2.0), configs.random_rotation(rng)),
            shape(13, ellipsoid(1.1, 0.95, 0.7), (6.0, 2.0, 2.0))]
    xs17 = [shape(17, ellipsoid(1.0, 0.9, 0.1), (2.0, 6.0, 6.0), configs.random_rotation(rng))]
    c13, c17 = cells_of(xs13, 13), cells_of(xs17, 17)
    cells = configs.Cells(nlat0=np.concatenate([c13.nlat0, c17.nlat0]),
                          nlat=np.concatenate([c13.nlat, c17.nlat]),
                          celltype=np.concatenate([c13.celltype, c17.celltype]),
                          x=np.concatenate([c13.x, c17.x]), a3=np.concatenate([c13.a3, c17.a3]),
                          detj=np.concatenate([c13.detj, c17.detj]),
                          f=np.concatenate([c13.f, c17.f]), g=np.concatenate([c13.g, c17.g]))
    g = gpu(workload(cells))
    g.set_repulsion(True, 2.0)
    df, n = g.intra_cell_repulsion()
    want, nw = [], 0
    for x, nl in [(x, 13) for x in xs13] + [(x, 17) for x in xs17]:
        geo = O.rbc_geometry(x)
        d, h = RP.intra_force(x, geo["a3"], geo["area"], nl, 2.0)
        want.append(d)
        nw += h
    want = np.concatenate(want)
    err = np.abs(df - want).max() / np.abs(want).max()
    print(f"intra-cell force: {n} nodes (oracle {nw}), {err:.2e}", flush=True)
    assert n == nw > 0 and err < TOL


def compare_inter(w, xs, visc, thresh, rigidsep, wall=None):
    g = gpu(w)
    g.set_repulsion(True, 0.0, thresh, rigidsep)
    n, dmin = g.inter_cell_repulsion(EPS)
    xg = g.rbc_coords()
    o = O.from_workload(w)
    out, no, dmo = RP.inter_repulsion(o, xs, w.cells.nlat, w.cells.celltype, visc, EPS, w.Lb, thresh,
                                      rigidsep, wall=wall)
    xo = np.concatenate(out)
    moved = np.abs(xo - np.concatenate(xs)).max()
    err = np.abs(xg - xo).max()
    print(f"moved {n} nodes (oracle {no}) by up to {moved:.3e}; separation {dmin:.6f} vs "
          f"{dmo:.6f}; coordinates {err:.2e}", flush=True)
    assert n == no > 0 and moved > 0 and err < TOL and dmin == pytest.approx(dmo, rel=TOL)


def test_inter_cells():
    xs, _, _ = two_spheres(0.02)
    # and two ellipsoids 0.03 apart across the periodic boundary
    xs += [shape(25, ellipsoid(1.1, 0.95, 0.7), (7.0, 6.5, 6.0)),
           shape(25, ellipsoid(1.1, 0.95, 0.7), (1.23, 6.5, 6.0))]
    cells = cells_of(xs, 25)
    compare_inter(workload(cells), xs, (1.0,), 10.0, False)


def test_inter_wall_rigidsep():
    z0 = 1.0
    xs = [shape(17, sphere_fn(1.0), (4.3, 3.7, z0 + 1.03)),
          shape(17, sphere_fn(1.0), (1.2, 6.1, z0 + 1.01)),
          shape(17, sphere_fn(1.0), (7.9, 0.2, z0 + 1.02))]
    wall = flat_wall(z0=z0)
    cells = cells_of(xs, 17, celltype=[1, 2, 2])
    w = workload(cells, visc=(1.0, -1.0), wall=wall)
    compare_inter(w, xs, (1.0, -1.0), 0.5, True, wall=wall)


def test_inter_tube():
    w = configs.make_workload("C3s", seed=2)
    # discs pushed out toward the tube wall (radius 4) until some touch it
    xs = T.split_cells(w.cells.x, w.cells.nlat, 3)
    out = []
    for x in xs:
        X = x.reshape(3, -1)
        c = X.mean(1)
        r = np.hypot(c[0] - 4.5, c[1] - 4.5)
        s = (4.0 - 1.0 + 0.02 - r) / r
        X = X + np.array([(c[0] - 4.5) * s, (c[1] - 4.5) * s, 0.0])[:, None]
        out.append(X.ravel())
    cells = dataclasses.replace(w.cells, x=np.concatenate(out))
    w = dataclasses.replace(w, cells=cells)
    compare_inter(w, out, w.viscRat, 10.0, False, wall=w.wall)


def test_euler_steps_with_repulsion():
    from tests.test_gpu_timeint import ES, ED, EB, ref_x
    from tests.test_oracle_timeint import NLAT, deformed_cells, ref_cell
    import rbc3d_amd as R
    cells = deformed_cells([(2.5, 2.5, 2.5), (4.75, 2.5, 2.5)], seed=7)
    w = configs.Workload("deformed", (8.0, 8.0, 8.0), configs.alpha_ref_from_xi(2.0), 2.0,
                         (32, 32, 32), 6, (5.0,), (0.0, 0.0, 0.0), cells)
    Ts, eps, coef = 0.02, 0.1, 0.5
    g = R.from_workload(w)
    g.set_ref_cells([NLAT], ref_x(), [ES], [ED], [EB])
    g.set_repulsion(True, coef, 10.0, False)
    g.timeint_euler(Ts, 2, eps)
    xg = g.rbc_coords()
    xo, _ = T.timeint_euler(O.from_workload(w), cells, [ref_cell()], [ES], [ED], [EB], Ts, 2, eps,
                            w.Lb, repulsion=(coef, 10.0, False), viscRat=w.viscRat)
    x_norep, _ = T.timeint_euler(O.from_workload(w), cells, [ref_cell()], [ES], [ED], [EB], Ts, 2,
                                 eps, w.Lb)
    err = np.abs(xg - xo).max()
    effect = np.abs(xo - x_norep).max()
    print(f"two steps with repulsion: effect {effect:.3e}, GPU vs oracle {err:.2e}", flush=True)
    assert effect > 1e-4 and err < TOL


def test_compute_rbc_vel_with_intra_force():
    """Compute_Rbc_Vel with the intra-cell force on: thin cells at their reference shape,
    so the force density is the projected repulsion alone."""
    import rbc3d_amd as R
    rng = np.random.default_rng(9)
    pan = ellipsoid(1.0, 1.0, 0.08)
    xs = [shape(13, pan, (2.5, 2.5, 2.5), configs.random_rotation(rng)),
          shape(13, pan, (5.5, 5.0, 5.0), configs.random_rotation(rng))]
    cells = cells_of(xs, 13)
    w = configs.Workload("pancakes", (8.0, 8.0, 8.0), configs.alpha_ref_from_xi(2.0), 2.0,
                         (32, 32, 32), 6, (1.0,), (0.0, 0.0, 0.0), cells)
    ref = shape(13, pan)
    g = R.from_workload(w)
    g.set_ref_cells([13], ref, [1.0], [20.0], [0.03])
    g.set_repulsion(True, 2.0)
    v, ni, _ = g.compute_rbc_vel()
    o = O.from_workload(w)
    refg = O.rbc_geometry(ref)
    fs = [T.cell_force(x, 13, refg, 1.0, 20.0, 0.03, 2.0)[1] for x in xs]
    o.set_density(1, np.concatenate(fs))
    vo, no, _ = o.solve_rbc_vel()
    err = np.abs(v - vo).max() / np.abs(vo).max()
    print(f"velocity from the intra-cell force: |v| {np.abs(vo).max():.3e}, {ni} vs {no} "
          f"iterations, {err:.2e}", flush=True)
    assert np.abs(vo).max() > 1e-3 and ni == no and err < 1e-10
