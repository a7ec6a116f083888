"""The repulsion of common/ModRepulsion.F90 restated for the oracle.

TEST INFRASTRUCTURE ONLY: the checker of the device repulsion; the product path never
imports this module.

  intra_force       AddIntraCellRepulsionForce (:126-266), one cell
  inter_repulsion   InterCellRepulsion (:270-402) with Closest_Neighbor_Cell (:480-546)
                    and Closest_Neighbor_Wall (:556-613)

The reference finds neighbours by walking 27 bins of its cell list (bins >= rc wide) and
keeps the minimum distance; these restatements take the minimum over every candidate
(first candidate in (cell, node) or triangle order on exact ties), which is the same
minimum wherever it can matter (within 2 epsDist of a node).  The projection onto the
neighbouring cell uses the x spline the oracle holds for that cell (the one built by its
last set_cells, as Closest_Neighbor_Cell uses rbc%spln_x from the step's start).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from oracle import oracle_py as O


def nint(t):
    """Fortran NINT: nearest integer, halves away from zero."""
    t = np.asarray(t, dtype=np.float64)
    tr = np.trunc(t)
    return np.where(np.abs(t - tr) == 0.5, tr + np.sign(t), np.rint(t))


def grid_angles(nlat):
    th = np.zeros(nlat)
    w = np.zeros(nlat)
    O.L().orc_gaqd(nlat, th, w)
    phi = np.arange(2 * nlat) * (2 * np.pi / (2 * nlat))
    return th, phi


def patch_radius(nlat):
    """RbcPolarPatch_Create (common/ModPolarPatch.F90:42): pi/sqrt(real(nlat)), real(8)
    under the reference's -freal-4-real-8 build."""
    return math.pi / math.sqrt(float(nlat))


def dist_on_sphere(th0, phi0, th1, phi1):
    """DistOnSphere (common/ModPolarPatch.F90:249-257)"""
    d = np.cos(th0 - th1) - np.sin(th0) * np.sin(th1) * (1. - np.cos(phi0 - phi1))
    return np.arccos(np.minimum(1., np.maximum(-1., d)))


def intra_force(x, a3, area, nlat, force_coef):
    """AddIntraCellRepulsionForce for one cell: x, a3 (nlat, 2 nlat, 3) flat; returns df in
    the same layout and the number of nodes with a force."""
    npt = 2 * nlat * nlat
    X = np.asarray(x, np.float64).reshape(3, npt)
    A = np.asarray(a3, np.float64).reshape(3, npt)
    th, phi = grid_angles(nlat)
    q = np.arange(npt)
    thq, phq = th[q % nlat], phi[q // nlat]
    eps = 2 * math.sqrt(area) / nlat
    epsRef = patch_radius(nlat)
    df = np.zeros((3, npt))
    hits = 0
    with np.errstate(invalid="ignore", divide="ignore"):
        for i in range(npt):
            xx0, xx1, xx2 = X[0, i] - X[0], X[1, i] - X[1], X[2, i] - X[2]
            rr = np.sqrt(xx0 * xx0 + xx1 * xx1 + xx2 * xx2)
            rrRef = dist_on_sphere(thq[i], phq[i], thq, phq)
            xxn = -(xx0 * A[0] + xx1 * A[1] + xx2 * A[2])
            # skip: rrRef < epsRef, rr > epsDist, xxn/rr < 0.707 (NaN compares false)
            ok = ~(rrRef < epsRef) & ~(rr > eps) & ~(xxn / rr < 0.707)
            if not ok.any():
                continue
            dist = rr[ok].min()
            if dist < eps:
                df[:, i] = force_coef * math.exp(1 - dist / eps) * A[:, i]
                hits += 1
    return df.ravel(), hits


def _min_dist_to_tri(xt, xe):
    """MinDistToTri (common/ModIntOnWalls.F90:480-577) through the oracle: distance and the
    closest point (1 - s - t) x1 + s x2 + t x3."""
    s0, t0 = C.c_double(), C.c_double()
    xcm = np.ascontiguousarray(xe.T.ravel())  # x(3,3) column-major: x[l + 3 d]
    rr = O.L().orc_min_dist_to_tri(np.ascontiguousarray(xt, np.float64), xcm, C.byref(s0),
                                   C.byref(t0))
    s, t = s0.value, t0.value
    return rr, (1 - s - t) * xe[0] + s * xe[1] + t * xe[2]


def inter_repulsion(o, xs, nlat, celltype, viscRat, eps, Lb, visc_thresh, rigidsep, wall=None):
    """InterCellRepulsion: xs per-cell coordinates (flat, (nlat, 2 nlat, 3)); `o` holds the
    cells' x splines (projection).  wall: None or (x(nvert,3) Fortran-flat, e2v(nele,3)
    1-based Fortran-flat, nvert, nele).  Returns the moved coordinates, the number of
    displaced nodes and the smallest separation below 2 eps (inf when none)."""
    Lb = np.asarray(Lb, np.float64)
    iLb = 1.0 / Lb
    ncell = len(xs)
    npts = [2 * int(n) ** 2 for n in nlat]
    Xs = [np.asarray(x, np.float64).reshape(3, -1) for x in xs]
    allX = np.concatenate(Xs, axis=1)
    owner = np.concatenate([np.full(n, c) for c, n in enumerate(npts)])
    local = np.concatenate([np.arange(n) for n in npts])
    if wall is not None:
        wx, e2v, nv, ne = wall
        WX = np.asarray(wx, np.float64).reshape(3, nv).T
        E = np.asarray(e2v).reshape(3, ne).T - 1
        tri = WX[E]  # (ne, 3 vertices, 3)
        cen = tri.mean(axis=1)
        wrad = np.sqrt(((tri - cen[:, None, :]) ** 2).sum(-1)).max()
    dxs = [np.zeros_like(X) for X in Xs]
    moved = 0
    dmin = math.inf
    for c in range(ncell):
        nl = int(nlat[c])
        th, phi = grid_angles(nl)
        for q in range(npts[c]):
            xi = Xs[c][:, q]
            # Closest_Neighbor_Cell
            other = owner != c
            xx = xi[:, None] - allX[:, other]
            xx = xx - nint(xx * iLb[:, None]) * Lb[:, None]
            rr = np.sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2])
            dist1, x1 = math.inf, None
            if rr.size:
                k = int(np.argmin(rr))
                dist1 = float(rr[k])
                if dist1 <= 2 * eps:
                    j = int(owner[other][k])
                    qj = int(local[other][k])
                    nlj = int(nlat[j])
                    thj, phj = grid_angles(nlj)
                    x0 = Xs[j][:, qj].copy()
                    d = xi - x0
                    d = d - nint(d * iLb) * Lb
                    xtar = x0 + d
                    _, _, x1 = o.find_projection(j, xtar, thj[qj % nlj], phj[qj // nlj])
                    e = xtar - x1
                    dist1 = math.sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])
            # Closest_Neighbor_Wall
            dist2, x2 = math.inf, None
            if wall is not None:
                dc = xi[None, :] - cen
                dc = dc - nint(dc * iLb) * Lb
                near = np.nonzero(np.sqrt((dc * dc).sum(1)) <= (eps + wrad) * (1 + 1e-9))[0]
                for e in near:
                    xe = tri[e]
                    d = xi - xe[0]
                    d = d - nint(d * iLb) * Lb
                    r, xp = _min_dist_to_tri(xe[0] + d, xe)
                    if r < dist2:
                        dist2, x2 = r, xp
            if dist1 < dist2:
                dist, xj = dist1, x1
            else:
                dist, xj = dist2, x2
            if dist <= 2 * eps:
                dmin = min(dmin, dist)
            if dist < eps:
                d = xi - xj
                d = d - nint(d * iLb) * Lb
                r = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
                dxs[c][:, q] = 0.5 * d * (eps - r) / r
                moved += 1
    out = []
    for c in range(ncell):
        X = Xs[c].copy()
        if abs(viscRat[int(celltype[c]) - 1]) < visc_thresh:
            X = X + dxs[c]
        elif rigidsep:
            sel = (dxs[c] ** 2).sum(0) > 0
            if sel.any():
                X = X + dxs[c][:, sel].sum(1)[:, None] / sel.sum()
        out.append(X.ravel())
    return out, moved, dmin
