"""Pins of the oracle's repulsion (oracle/repulsion_oracle.py, common/ModRepulsion.F90):

- AddIntraCellRepulsionForce on a flat oblate ellipsoid (a "pancake" of half-thickness c):
  the nodes of the first latitude ring face their mirror image across the thin body at
  distance 2 c cos(th1), so their force is ForceCoef exp(1 - 2 c cos(th1)/epsDist) a3 with
  epsDist = 2 sqrt(area)/nlat; a convex cell feels none; the force turns with the cell.
- InterCellRepulsion between two spheres a gap apart: nodes of one sphere closer than
  epsDist to the other move 0.5 (epsDist - d) along the line from the other centre (d the
  distance to that sphere), to the accuracy of the spline surface; the same across the
  periodic boundary; a node above a flat wall moves 0.5 (epsDist - h) straight up (exact);
  with rigidsep a rigid cell moves by the mean displacement of its displaced nodes.
"""
import math

import numpy as np
import pytest

from oracle import oracle_py as O
from oracle import repulsion_oracle as RP
from rbc3d_amd import configs
from tests.test_mech_oracle import ellipsoid, flat, grid

EPS = 0.05


def shape(nlat, fn, centre=(0.0, 0.0, 0.0), rot=None):
    th, phi, _ = grid(nlat)
    T, P = np.meshgrid(th, phi, indexing="ij")
    X = np.stack(fn(T, P), -1)
    if rot is not None:
        X = X @ rot.T
    return flat(X + np.asarray(centre))


def cells_of(xs, nlat, celltype=None):
    geos = [O.rbc_geometry(x) for x in xs]
    n = len(xs)
    z = np.zeros(sum(x.size for x in xs))
    return configs.Cells(nlat0=np.full(n, nlat, np.int32), nlat=np.full(n, nlat, np.int32),
                         celltype=np.ones(n, np.int32) if celltype is None else np.asarray(celltype, np.int32),
                         x=np.concatenate(xs), a3=np.concatenate([g["a3"] for g in geos]),
                         detj=np.concatenate([g["detj"] for g in geos]), f=z.copy(), g=z.copy())


def workload(cells, L=8.0, visc=(1.0,), wall=None):
    return configs.Workload("repulsion", (L, L, L), configs.alpha_ref_from_xi(2.0), 2.0,
                            (16, 16, 16), 4, tuple(visc), (0.0, 0.0, 0.0), cells, wall)


def intra(x, nlat, coef):
    g = O.rbc_geometry(x)
    return RP.intra_force(x, g["a3"], g["area"], nlat, coef), g


def test_intra_pancake_first_ring():
    nlat, c, coef = 13, 0.08, 2.0
    x = shape(nlat, ellipsoid(1.0, 1.0, c))
    (df, hits), g = intra(x, nlat, coef)
    th, _, _ = grid(nlat)
    eps = 2 * math.sqrt(g["area"]) / nlat
    DF = df.reshape(3, 2 * nlat, nlat).transpose(0, 2, 1)  # [d, ilat, ilon]
    A3 = g["a3"].reshape(3, 2 * nlat, nlat).transpose(0, 2, 1)
    for ilat in (0, nlat - 1):
        want = coef * math.exp(1 - 2 * c * math.cos(th[0]) / eps) * A3[:, ilat, :]
        err = np.abs(DF[:, ilat, :] - want).max() / np.abs(want).max()
        print(f"ring {ilat}: {err:.2e}")
        assert err < 1e-12
    assert hits >= 4 * nlat
    # the force points out of the body on both faces
    assert np.all(DF[2, 0, :] > 0) and np.all(DF[2, nlat - 1, :] < 0)


def test_intra_none_on_convex_cell():
    (df, hits), _ = intra(shape(13, ellipsoid(1.1, 0.95, 0.7)), 13, 2.0)
    assert hits == 0 and not df.any()


def test_intra_turns_with_the_cell():
    nlat = 13
    rot = configs.random_rotation(np.random.default_rng(3))
    (df0, h0), _ = intra(shape(nlat, ellipsoid(1.0, 1.0, 0.08)), nlat, 2.0)
    (df1, h1), _ = intra(shape(nlat, ellipsoid(1.0, 1.0, 0.08), (3.0, 1.0, 2.0), rot), nlat, 2.0)
    assert h0 == h1
    assert np.abs(rot @ df0.reshape(3, -1) - df1.reshape(3, -1)).max() < 1e-12


def sphere_fn(R):
    return lambda T, P: (R * np.sin(T) * np.cos(P), R * np.sin(T) * np.sin(P), R * np.cos(T))


def two_spheres(gap, shift=0.0, nlat=25):
    cA = np.array([2.0, 3.0, 3.0])
    cB = np.array([4.0 + gap + shift, 3.0, 3.0])
    xs = [shape(nlat, sphere_fn(1.0), cA), shape(nlat, sphere_fn(1.0), cB)]
    return xs, cA, cB


def run_inter(xs, nlat, visc=(1.0,), thresh=10.0, rigidsep=False, wall=None, celltype=None):
    cells = cells_of(xs, nlat, celltype)
    w = workload(cells, visc=visc, wall=wall)
    o = O.from_workload(w)
    return RP.inter_repulsion(o, xs, cells.nlat, cells.celltype, visc, EPS, w.Lb, thresh, rigidsep,
                              wall=wall)


def test_inter_two_spheres():
    xs, cA, cB = two_spheres(0.02)
    out, moved, dmin = run_inter(xs, 25)
    X = xs[0].reshape(3, -1)
    r = X - cB[:, None]
    dist = np.sqrt((r * r).sum(0))
    d = dist - 1.0
    want = np.where(d < EPS, 0.5 * (EPS - d) * r / dist, 0.0)
    sure = np.abs(d - EPS) > 1e-5
    err = np.abs((out[0].reshape(3, -1) - X) - want)[:, sure].max()
    n_want = int((d < EPS).sum())
    print(f"moved {moved} (sphere A: {n_want}), min separation {dmin:.6f}, error {err:.2e}")
    assert n_want > 0 and err < 1e-6
    assert abs(dmin - 0.02) < 1e-4
    # symmetric: sphere B moves the mirror way
    XB = xs[1].reshape(3, -1)
    dB = out[1].reshape(3, -1) - XB
    assert np.abs(dB[0].sum() + (out[0].reshape(3, -1) - X)[0].sum()) < 1e-6


def test_inter_across_the_periodic_boundary():
    xs, _, _ = two_spheres(0.02)
    ys, _, _ = two_spheres(0.02, shift=-8.0)
    a, _, _ = run_inter(xs, 25)
    b, _, _ = run_inter(ys, 25)
    assert np.abs(a[0] - b[0]).max() < 1e-12
    assert np.abs(a[1] - (b[1].reshape(3, -1) + np.array([8.0, 0, 0])[:, None]).ravel()).max() < 1e-12


def flat_wall(L=8.0, n=16, z0=1.0):
    """the plane z = z0 over the periodic square [0, L]^2: n x n squares, two triangles
    each; returns (x(nvert,3) Fortran-flat, e2v(nele,3) 1-based Fortran-flat, nvert, nele)"""
    g = np.linspace(0.0, L, n + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z0)], 1)
    idx = np.arange((n + 1) ** 2).reshape(n + 1, n + 1)
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = idx[i, j], idx[i + 1, j], idx[i + 1, j + 1], idx[i, j + 1]
            tris += [(a, b, c), (a, c, d)]
    e2v = np.asarray(tris) + 1
    return (np.asfortranarray(pts).ravel(order="F"), np.asfortranarray(e2v).ravel(order="F").astype(np.int32),
            len(pts), len(tris))


def test_inter_flat_wall():
    z0, h = 1.0, 0.03
    xs = [shape(17, sphere_fn(1.0), (4.3, 3.7, z0 + 1.0 + h))]
    wall = flat_wall(z0=z0)
    out, moved, dmin = run_inter(xs, 17, wall=wall)
    X = xs[0].reshape(3, -1)
    hgt = X[2] - z0
    want = np.zeros_like(X)
    want[2] = np.where(hgt < EPS, 0.5 * (EPS - hgt), 0.0)
    err = np.abs(out[0].reshape(3, -1) - X - want).max()
    print(f"wall: moved {moved}, min separation {dmin:.6f}, error {err:.2e}")
    assert moved == int((hgt < EPS).sum()) > 0 and err < 1e-14
    assert dmin == pytest.approx(hgt.min(), abs=1e-14)


def test_inter_rigidsep():
    z0, h = 1.0, 0.03
    xs = [shape(17, sphere_fn(1.0), (4.3, 3.7, z0 + 1.0 + h))]
    wall = flat_wall(z0=z0)
    X = xs[0].reshape(3, -1)
    hgt = X[2] - z0
    sel = hgt < EPS
    # a rigid cell (|viscRat| >= viscRatThresh) with rigidsep: the mean displacement
    out, _, _ = run_inter(xs, 17, visc=(-1.0,), thresh=0.5, rigidsep=True, wall=wall)
    shift = out[0].reshape(3, -1) - X
    assert np.abs(shift[2] - (0.5 * (EPS - hgt[sel])).mean()).max() < 1e-14
    assert np.abs(shift[:2]).max() < 1e-14
    # without rigidsep it does not move
    out, _, _ = run_inter(xs, 17, visc=(-1.0,), thresh=0.5, rigidsep=False, wall=wall)
    assert np.array_equal(out[0], xs[0])
