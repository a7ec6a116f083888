"""TimeInt_Euler (common/ModTimeInt.F90:109-178) restated over the oracle's pieces.

TEST INFRASTRUCTURE ONLY: the checker of the device time stepping; the product path never
imports this module.

One step, in the reference's order:
  Compute_Rbc_Vel (:716-808)  RBC_ComputeGeometry per cell (orc_rbc_geometry), the surface
                              sources (orc_set_cells), f = -Shell_ResForce projected by
                              RBC_SphProject, Solve_RBC_Vel (orc_solve_rbc_vel; its
                              previous solution is the next initial guess)
  NoSlipWall (ModNoSlip:45)   when the oracle holds a wall
  rbc%x += Ts rbc%v (:131-136)
  ReboxRbcs (:1006-1020)
  4 x {VolConstrainRbcs (:938-963), InterCellRepulsion (common/ModRepulsion.F90:270-402,
  when the repulsion is on), FilterRbcs (:908-934)}

TimeInt_RK2 (:183-271) and TimeInt_AB2 (:275-341) use the same velocity evaluation
(Compute_Rbc_Vel, then NoSlipWall) and ReboxRbcs, without the volume constraint, the
repulsion and the filter.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from oracle import oracle_py as O
from oracle import repulsion_oracle as RP


def split_cells(v, nlat, ncomp):
    out, o = [], 0
    for nl in nlat:
        n = 2 * int(nl) * int(nl) * ncomp
        out.append(np.array(v[o:o + n], dtype=np.float64))
        o += n
    return out


def rebox(x, Lb):
    """ReboxRbcs of one cell x(nlat, nlon, 3) (flat, column-major)."""
    X = x.reshape(3, -1)
    for d in range(3):
        xc = 0.5 * (X[d].max() + X[d].min())
        X[d] = X[d] - np.floor(xc * (1.0 / Lb[d])) * Lb[d]
    return X.ravel()


def vol_constrain(x, ref_vol, eps_dist):
    """VolConstrainRbcs of one cell: fac = SIGN(MIN(ABS(fac), epsDist/20.), fac).  fac is a
    default REAL in the reference, double precision under its -freal-4-real-8 build
    (CMakeLists.txt:61, Makefile.in:49)."""
    g = O.rbc_geometry(x)
    fac = (ref_vol - g["vol"]) / g["area"]
    fac = float(np.copysign(min(abs(fac), eps_dist / 20.0), fac))
    return x + fac * g["a3"]


def cell_force(x, nlat0, ref, ES, ED, EB, force_coef=0.0):
    """rbc%f = -Shell_ResForce(rbc, rbcRef), AddIntraCellRepulsionForce (force_coef != 0),
    then RBC_SphProject(rbc, 3, rbc%f)."""
    g = O.rbc_geometry(x)
    f = O.shell_res_force(g, ref, ES, ED, EB)
    if force_coef == 0:
        return g, -O.rbc_sph_project(f, g["nlat"], nlat0)
    df, _ = RP.intra_force(x, g["a3"], g["area"], g["nlat"], force_coef)
    return g, O.rbc_sph_project(-f + df, g["nlat"], nlat0)


def timeint_euler(o, cells, refs, ES, ED, EB, Ts, nsteps, eps_dist, Lb, noslip=None,
                  repulsion=None, viscRat=None, wall=None):
    """nsteps Euler steps of the cells (a configs.Cells) on the oracle `o` (built with
    the same configuration); refs[k]: rbc_geometry of cell type k+1's reference cell.
    noslip: None, or (rtol, maxit) of NoSlipWall.  repulsion: None, or (ForceCoef,
    viscRatThresh, rigidsep), with viscRat (per type) and wall (the oracle's wall tuple,
    or None) for InterCellRepulsion.  Returns the final coordinates in the cells' layout
    and the per-step (cell, wall) GMRES iterations."""
    coef = repulsion[0] if repulsion is not None else 0.0
    nlat, nlat0, ct = cells.nlat, cells.nlat0, cells.celltype
    xs = split_cells(cells.x, nlat, 3)
    its = []
    for _ in range(nsteps):
        geos, fs = [], []
        for c, x in enumerate(xs):
            k = ct[c] - 1
            g, f = cell_force(x, int(nlat0[c]), refs[k], ES[k], ED[k], EB[k], coef)
            geos.append(g)
            fs.append(f)
        cur = dataclasses.replace(cells, x=np.concatenate(xs),
                                  a3=np.concatenate([g["a3"] for g in geos]),
                                  detj=np.concatenate([g["detj"] for g in geos]))
        o.set_cells(cur)
        o.set_density(1, np.concatenate(fs))
        v, ni, _ = o.solve_rbc_vel()
        nw = 0
        if noslip is not None:
            nw = o.noslip_wall(*noslip)[0]
        its.append((ni, nw))
        V = v.reshape(3, -1)
        off = 0
        for c, x in enumerate(xs):
            npt = 2 * int(nlat[c]) ** 2
            X = x.reshape(3, npt) + Ts * V[:, off:off + npt]
            xs[c] = rebox(X.ravel(), Lb)
            off += npt
        for _ in range(4):
            for c in range(len(xs)):
                xs[c] = vol_constrain(xs[c], refs[ct[c] - 1]["vol"], eps_dist)
            if repulsion is not None:
                xs, _, _ = RP.inter_repulsion(o, xs, nlat, ct, viscRat, eps_dist, Lb, repulsion[1],
                                              repulsion[2], wall=wall)
            for c in range(len(xs)):
                xs[c] = O.rbc_sph_project(xs[c], int(nlat[c]), int(nlat0[c]))
    return np.concatenate(xs), its


def velocity(o, cells, xs, refs, ES, ED, EB, noslip=None):
    """Compute_Rbc_Vel (+ NoSlipWall): the node velocities (3, np) of the cells at xs, and
    the (cell, wall) GMRES iterations."""
    nlat0, ct = cells.nlat0, cells.celltype
    geos, fs = [], []
    for c, x in enumerate(xs):
        k = ct[c] - 1
        g, f = cell_force(x, int(nlat0[c]), refs[k], ES[k], ED[k], EB[k])
        geos.append(g)
        fs.append(f)
    cur = dataclasses.replace(cells, x=np.concatenate(xs),
                              a3=np.concatenate([g["a3"] for g in geos]),
                              detj=np.concatenate([g["detj"] for g in geos]))
    o.set_cells(cur)
    o.set_density(1, np.concatenate(fs))
    v, ni, _ = o.solve_rbc_vel()
    nw = o.noslip_wall(*noslip)[0] if noslip is not None else 0
    return v.reshape(3, -1), (ni, nw)


def _add(xs, nlat, V, fn):
    out, off = [], 0
    for c, x in enumerate(xs):
        npt = 2 * int(nlat[c]) ** 2
        out.append(fn(x.reshape(3, npt), V[:, off:off + npt] if V is not None else None).ravel())
        off += npt
    return out


def timeint_rk2(o, cells, refs, ES, ED, EB, alpha, Ts, nsteps, Lb, noslip=None):
    """TimeInt_RK2(alpha): returns the final coordinates and the iterations of every
    evaluation."""
    c2, b2 = alpha, 1 / (2 * alpha)
    b1 = 1 - b2
    xs = split_cells(cells.x, cells.nlat, 3)
    its = []
    for _ in range(nsteps):
        k1, it = velocity(o, cells, xs, refs, ES, ED, EB, noslip)
        its.append(it)
        xs = _add(xs, cells.nlat, k1, lambda X, K: X + c2 * Ts * K)
        k2, it = velocity(o, cells, xs, refs, ES, ED, EB, noslip)
        its.append(it)
        xs = _add(xs, cells.nlat, k1, lambda X, K: X - c2 * Ts * K)
        off = 0
        out = []
        for c, x in enumerate(xs):
            npt = 2 * int(cells.nlat[c]) ** 2
            K1, K2 = k1[:, off:off + npt], k2[:, off:off + npt]
            out.append(rebox((x.reshape(3, npt) + Ts * (b1 * K1 + b2 * K2)).ravel(), Lb))
            off += npt
        xs = out
    return np.concatenate(xs), its


def timeint_ab2(o, cells, refs, ES, ED, EB, Ts, nsteps, Lb, noslip=None):
    """TimeInt_AB2: an Euler step from Compute_Rbc_Vel alone, then nsteps - 1 AB2 steps."""
    xs = split_cells(cells.x, cells.nlat, 3)
    v, it = velocity(o, cells, xs, refs, ES, ED, EB)
    its = [it]
    xs = _add(xs, cells.nlat, v, lambda X, V: X + Ts * V)
    for _ in range(1, nsteps):
        v1 = v
        v, it = velocity(o, cells, xs, refs, ES, ED, EB, noslip)
        its.append(it)
        off = 0
        out = []
        for c, x in enumerate(xs):
            npt = 2 * int(cells.nlat[c]) ** 2
            V, V1 = v[:, off:off + npt], v1[:, off:off + npt]
            out.append(rebox((x.reshape(3, npt) + Ts * (3 * V - V1) / 2).ravel(), Lb))
            off += npt
        xs = out
    return np.concatenate(xs), its
