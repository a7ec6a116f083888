"""ctypes binding of the CPU restatement (oracle/rbc3d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU-baseline
leg of bench.py, as the checker.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "librbc3d_oracle.so")

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcConf(C.Structure):
    _fields_ = [("Lb", C.c_double * 3), ("alpha", C.c_double), ("rc", C.c_double),
                ("Nb", C.c_int * 3), ("P", C.c_int), ("ntypes", C.c_int),
                ("viscRat", C.c_double * 8), ("vBkg", C.c_double * 3)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    if not os.path.exists(_LIB):
        build()
    L = C.CDLL(_LIB)
    vp = C.c_void_p
    sig = {
        "orc_create": (vp, [C.POINTER(OrcConf)]),
        "orc_destroy": (None, [vp]),
        "orc_set_threads": (C.c_int, [C.c_int]),
        "orc_set_cells": (C.c_int, [vp, C.c_int, ip, ip, ip, dp, dp, dp]),
        "orc_set_cell_density": (C.c_int, [vp, C.c_int, dp]),
        "orc_set_wall": (C.c_int, [vp, C.c_int, C.c_int, dp, ip]),
        "orc_set_wall_density": (C.c_int, [vp, dp]),
        "orc_set_gmres_restart": (C.c_int, [vp, C.c_int]),
        "orc_prepare_sing_int_on_wall": (C.c_int, [vp]),
        "orc_ntargets_rbc": (C.c_int, [vp]),
        "orc_ntargets_wall": (C.c_int, [vp]),
        "orc_tlist_rbc": (C.c_int, [vp, dp, ip, dp]),
        "orc_tlist_wall": (C.c_int, [vp, dp, ip, dp]),
        "orc_add_int_on_rbcs": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, dp, ip, ip, dp, dp,
                                          C.POINTER(C.c_longlong)]),
        "orc_add_int_on_walls": (C.c_int, [vp, C.c_double, C.c_int, dp, ip, ip, dp, dp]),
        "orc_pme_distrib_source": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int]),
        "orc_pme_transform": (C.c_int, [vp]),
        "orc_pme_add_interp_vel": (C.c_int, [vp, C.c_int, dp, ip, dp, dp]),
        "orc_calc_velocity_field": (C.c_int, [vp, C.c_int, dp, ip, ip, dp, dp]),
        "orc_glob_sph_trans": (C.c_int, [vp, dp, dp, C.c_int]),
        "orc_cell_dof": (C.c_int, [vp]),
        "orc_cell_matmult": (C.c_int, [vp, dp, dp]),
        "orc_cell_matmult_sample": (C.c_int, [vp, dp, dp, ip]),
        "orc_cell_rhs": (C.c_int, [vp, dp]),
        "orc_wall_dof": (C.c_int, [vp]),
        "orc_wall_matmult": (C.c_int, [vp, dp, dp]),
        "orc_wall_residual_vel": (C.c_int, [vp, dp]),
        "orc_ewald_sl": (None, [vp, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_ewald_dl": (None, [vp, C.c_double, C.POINTER(C.c_double)]),
        "orc_mask": (C.c_double, [vp, C.c_double]),
        "orc_bspline": (None, [C.c_double, C.c_int, C.POINTER(C.c_int), dp]),
        "orc_gauleg": (None, [C.c_double, C.c_double, C.c_int, dp, dp]),
        "orc_gauleg_sinh": (None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]),
        "orc_gaqd": (None, [C.c_int, dp, dp]),
        "orc_pbar_table": (None, [C.c_int, dp, C.c_int, dp]),
        "orc_hashtable_build": (C.c_int, [ip, dp, C.c_int, dp, C.c_int, ip, ip]),
        "orc_sh_anal_gau": (C.c_int, [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp,
                                      C.c_int, C.c_int]),
        "orc_sh_synth_gau": (C.c_int, [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp,
                                       C.c_int, C.c_int]),
        "orc_sh_synth_equ": (C.c_int, [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp,
                                       C.c_int, C.c_int]),
        "orc_fft_diff": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp]),
        "orc_sh_grad_gau": (C.c_int, [C.c_int, C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp,
                                      C.c_int, C.c_int]),
        "orc_min_dist_to_tri": (C.c_double, [dp, dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_tri_int_regular": (None, [vp, dp, dp, dp, dp]),
        "orc_tri_int_duffy": (None, [vp, dp, dp, dp, C.c_double, C.c_double, dp]),
        "orc_wall_self_nnz": (C.c_int, [vp]),
        "orc_wall_self_csr": (C.c_int, [vp, ip, ip, dp]),
        "orc_pme_grid": (C.c_int, [vp, C.c_int, dp]),
        "orc_fft3d_forward": (None, [C.c_int, C.c_int, C.c_int, dp, dp]),
        "orc_fft3d_inverse": (None, [C.c_int, C.c_int, C.c_int, dp, dp]),
        "orc_cell_spline": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, dp, dp]),
        "orc_pair_checksum": (C.c_ulonglong, [vp]),
        "orc_noslip_wall": (C.c_int, [vp, C.c_double, C.c_int, C.c_void_p, C.POINTER(C.c_int),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_get_wall_density": (C.c_int, [vp, dp]),
        "orc_solve_rbc_vel": (C.c_int, [vp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "orc_near_list": (C.c_int, [vp, C.c_int, dp, ip, C.c_int, ip, ip, ip, dp, C.c_int]),
        "orc_set_terms": (C.c_int, [vp, C.c_int]),
        "orc_spline_eval": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double, dp]),
        "orc_find_projection": (C.c_int, [vp, C.c_int, dp, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), dp]),
        "orc_spline_interp_raw": (None, [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, C.c_double,
                                         C.c_double, dp]),
        "orc_cell_mesh_size": (C.c_double, [vp, C.c_int]),
        "orc_rbc_geometry": (C.c_int, [C.c_int, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
        "orc_shell_res_force": (C.c_int, [C.c_int, dp, dp, dp, dp, dp, dp, dp, C.c_double, dp, dp,
                                          C.c_double, C.c_double, C.c_double, dp]),
        "orc_rbc_sph_project": (C.c_int, [C.c_int, C.c_int, C.c_int, dp]),
        "orc_set_deflation": (C.c_int, [vp, C.c_int]),
        "orc_sb_modes": (C.c_int, [vp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


_L = None


def L():
    global _L
    if _L is None:
        _L = lib()
    return _L


GEO_FIELDS = ("a1", "a2", "a3", "a", "a_rcp", "b", "detj")


def rbc_geometry(x):
    """RBC_ComputeGeometry of one cell: x(nlat, 2 nlat, 3) in the reference's
    (column-major) layout as a flat array; returns the fields (flat, column-major)
    and xc, area, vol, meshSize."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    npt = x.size // 3
    nlat = int(round((npt / 2) ** 0.5))
    assert 2 * nlat * nlat == npt
    out = {k: np.zeros(n * npt) for k, n in (("a1", 3), ("a2", 3), ("a3", 3), ("a", 4),
                                              ("a_rcp", 4), ("b", 4), ("detj", 1))}
    xc = np.zeros(3)
    area, vol, mesh = C.c_double(), C.c_double(), C.c_double()
    r = L().orc_rbc_geometry(nlat, x, out["a1"], out["a2"], out["a3"], out["a"], out["a_rcp"],
                             out["b"], out["detj"], xc, C.byref(area), C.byref(vol),
                             C.byref(mesh))
    assert r == 0
    out.update(xc=xc, area=area.value, vol=vol.value, mesh=mesh.value, nlat=nlat)
    return out


def shell_res_force(geo, ref, ES, ED, EB):
    """Shell_ResForce of one cell with geometry `geo` against the reference cell `ref`
    (both from rbc_geometry); f(nlat, nlon, 3) flat, Cartesian."""
    f = np.zeros_like(geo["a1"])
    r = L().orc_shell_res_force(geo["nlat"], geo["a1"], geo["a2"], geo["a3"], geo["a"],
                                geo["a_rcp"], geo["b"], geo["detj"], geo["area"], ref["a_rcp"],
                                ref["b"], ES, ED, EB, f)
    assert r == 0
    return f


def rbc_sph_project(u, nlat, nlat0):
    """RBC_SphProject of nvar fields u(nlat, 2 nlat, nvar) (flat): returns the projection."""
    u = np.array(u, dtype=np.float64).ravel()
    nvar = u.size // (2 * nlat * nlat)
    assert L().orc_rbc_sph_project(nlat, nlat0, nvar, u) == 0
    return u


def make_conf(Lb, alpha, rc, Nb, P, viscRat, vBkg=(0, 0, 0)):
    c = OrcConf()
    c.Lb[:] = list(Lb)
    c.alpha = alpha
    c.rc = rc
    c.Nb[:] = list(Nb)
    c.P = P
    c.ntypes = len(viscRat)
    for i, v in enumerate(viscRat):
        c.viscRat[i] = v
    c.vBkg[:] = list(vBkg)
    return c


class Oracle:
    """A restated reference system: conf + cells (+ one wall)."""

    def __init__(self, Lb, alpha, rc, Nb, P, viscRat, vBkg=(0, 0, 0)):
        self.conf = make_conf(Lb, alpha, rc, Nb, P, viscRat, vBkg)
        self.h = L().orc_create(C.byref(self.conf))

    def __del__(self):
        if getattr(self, "h", None):
            L().orc_destroy(self.h)
            self.h = None

    @staticmethod
    def threads(n=0):
        return L().orc_set_threads(n)

    def set_cells(self, cells):
        r = L().orc_set_cells(self.h, cells.ncell, np.ascontiguousarray(cells.nlat0, np.int32),
                              np.ascontiguousarray(cells.nlat, np.int32),
                              np.ascontiguousarray(cells.celltype, np.int32), cells.x, cells.a3,
                              cells.detj)
        assert r == 0
        self.np = L().orc_ntargets_rbc(self.h)
        self._nlat = np.asarray(cells.nlat, np.int32).copy()
        self._nlat0 = np.asarray(cells.nlat0, np.int32).copy()

    def set_density(self, which, dens):
        assert L().orc_set_cell_density(self.h, which, np.ascontiguousarray(dens)) == 0

    def set_wall(self, x, e2v, nvert, nele):
        assert L().orc_set_wall(self.h, nvert, nele, np.ascontiguousarray(x),
                                np.ascontiguousarray(e2v, np.int32)) == 0

    def set_wall_density(self, f):
        assert L().orc_set_wall_density(self.h, np.ascontiguousarray(f)) == 0

    def prepare_wall(self):
        assert L().orc_prepare_sing_int_on_wall(self.h) == 0

    def tlist_rbc(self):
        n = L().orc_ntargets_rbc(self.h)
        tx = np.zeros(3 * n)
        ti = np.zeros(3 * n, np.int32)
        Ac = np.zeros(n)
        L().orc_tlist_rbc(self.h, tx, ti, Ac)
        return tx, ti, np.ones(n, np.int32), Ac

    def tlist_wall(self):
        n = L().orc_ntargets_wall(self.h)
        tx = np.zeros(3 * n)
        ti = np.zeros(3 * n, np.int32)
        Ac = np.zeros(n)
        L().orc_tlist_wall(self.h, tx, ti, Ac)
        return tx, ti, np.ones(n, np.int32), Ac

    def add_int_on_rbcs(self, c1, c2, tl, v=None):
        tx, ti, act, Ac = tl
        n = len(Ac)
        v = np.zeros(3 * n) if v is None else v
        npairs = C.c_longlong(0)
        r = L().orc_add_int_on_rbcs(self.h, c1, c2, n, tx, ti, act, Ac, v, C.byref(npairs))
        if r:
            raise RuntimeError(f"orc_add_int_on_rbcs status {r}")
        return v, npairs.value

    def add_int_on_walls(self, c1, tl, v=None):
        tx, ti, act, Ac = tl
        n = len(Ac)
        v = np.zeros(3 * n) if v is None else v
        assert L().orc_add_int_on_walls(self.h, c1, n, tx, ti, act, Ac, v) == 0
        return v

    def pme(self, c1, c2, tl, cells=True, walls=True, v=None):
        tx, ti, act, Ac = tl
        n = len(Ac)
        v = np.zeros(3 * n) if v is None else v
        L().orc_pme_distrib_source(self.h, c1, c2, int(cells), int(walls))
        L().orc_pme_transform(self.h)
        L().orc_pme_add_interp_vel(self.h, n, tx, act, Ac, v)
        return v

    def calc_velocity_field(self, tl):
        tx, ti, act, Ac = tl
        n = len(Ac)
        v = np.zeros(3 * n)
        r = L().orc_calc_velocity_field(self.h, n, tx, ti, act, Ac, v)
        if r:
            raise RuntimeError(f"orc_calc_velocity_field status {r}")
        return v

    def cell_dof(self):
        return L().orc_cell_dof(self.h)

    def glob_sph_trans(self, v=None, c=None, direction=1):
        if direction == 1:
            c = np.zeros(self.cell_dof())
            L().orc_glob_sph_trans(self.h, np.ascontiguousarray(v), c, 1)
            return c
        v = np.zeros(3 * self.np)
        L().orc_glob_sph_trans(self.h, v, np.ascontiguousarray(c), -1)
        return v

    def cell_matmult(self, u):
        b = np.zeros(self.cell_dof())
        r = L().orc_cell_matmult(self.h, np.ascontiguousarray(u), b)
        if r:
            raise RuntimeError(f"orc_cell_matmult status {r}")
        return b

    def cell_matmult_sample(self, u, cells):
        """MyMatMult evaluated on the nodes of ``cells`` (0-based) only: b is valid on
        those cells' coefficient blocks."""
        act = np.zeros(len(self._nlat), np.int32)
        act[np.asarray(cells)] = 1
        b = np.zeros(self.cell_dof())
        r = L().orc_cell_matmult_sample(self.h, np.ascontiguousarray(u), b, act)
        if r:
            raise RuntimeError(f"orc_cell_matmult_sample status {r}")
        return b

    def cell_blocks(self, cells):
        """Indices of the Glob_Sph_Trans unknowns (3 nlat0^2 per cell) of ``cells``."""
        off = np.concatenate([[0], np.cumsum(3 * self._nlat0.astype(np.int64) ** 2)])
        return np.concatenate([np.arange(off[c], off[c + 1]) for c in cells])

    def cell_rhs(self):
        b = np.zeros(self.cell_dof())
        assert L().orc_cell_rhs(self.h, b) == 0
        return b

    def wall_dof(self):
        return L().orc_wall_dof(self.h)

    def wall_matmult(self, u):
        b = np.zeros(self.wall_dof())
        L().orc_wall_matmult(self.h, np.ascontiguousarray(u), b)
        return b

    def wall_residual_vel(self):
        v = np.zeros(3 * L().orc_ntargets_wall(self.h))
        L().orc_wall_residual_vel(self.h, v)
        return v

    def wall_csr(self):
        nnz = L().orc_wall_self_nnz(self.h)
        nv = L().orc_ntargets_wall(self.h)
        rp = np.zeros(3 * nv + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        L().orc_wall_self_csr(self.h, rp, col, val)
        return rp, col, val

    def pme_grid(self, which):
        N = int(np.prod(list(self.conf.Nb)))
        out = np.zeros((9 if which == 1 else 3) * N)
        L().orc_pme_grid(self.h, which, out)
        return out

    def cell_spline(self, irbc, which, m, n, nvar):
        u, u1, u2, u12 = (np.zeros(m * n * nvar) for _ in range(4))
        L().orc_cell_spline(self.h, irbc, which, u, u1, u2, u12)
        return u, u1, u2, u12

    def near_list(self, tl, i, cap=32):
        tx, ti, act, Ac = tl
        cells = np.zeros(cap, np.int32)
        il = np.zeros(cap, np.int32)
        jl = np.zeros(cap, np.int32)
        d = np.zeros(cap)
        n = L().orc_near_list(self.h, len(Ac), tx, ti, i, cells, il, jl, d, cap)
        return n, cells[:max(n, 0)], il[:max(n, 0)], jl[:max(n, 0)], d[:max(n, 0)]

    def set_gmres_restart(self, m):
        assert L().orc_set_gmres_restart(self.h, m) == 0

    def set_deflation(self, on):
        assert L().orc_set_deflation(self.h, int(bool(on))) == 0

    def sb_modes(self):
        m = np.zeros(18 * self.np)
        assert L().orc_sb_modes(self.h, m) == 0
        return m.reshape(6, 3, self.np)

    def noslip_wall(self, rtol=1e-8, maxit=60, finc0=None):
        ni = C.c_int()
        rn = C.c_double()
        vm = C.c_double()
        x0p = None
        if finc0 is not None:
            x0 = np.ascontiguousarray(finc0, dtype=np.float64)
            x0p = x0.ctypes.data_as(C.c_void_p)
        conv = L().orc_noslip_wall(self.h, rtol, maxit, x0p, C.byref(ni), C.byref(rn), C.byref(vm))
        return ni.value, rn.value, bool(conv), vm.value

    def wall_density(self):
        f = np.zeros(3 * L().orc_ntargets_wall(self.h))
        L().orc_get_wall_density(self.h, f)
        return f

    def solve_rbc_vel(self):
        if not hasattr(self, "_sol"):
            self._sol = np.zeros(self.cell_dof())
        v = np.zeros(3 * self.np)
        ni = C.c_int()
        rn = C.c_double()
        L().orc_solve_rbc_vel(self.h, self._sol, v, C.byref(ni), C.byref(rn))
        return v, ni.value, rn.value

    def pair_checksum(self):
        return L().orc_pair_checksum(self.h)

    TERM_PAIR, TERM_SELF, TERM_NEAR, TERM_LINEAR, TERM_ALL = 1, 2, 4, 8, 15

    def set_terms(self, terms):
        """Which parts of AddIntOnRbcs run (a test hook; all by default)."""
        assert L().orc_set_terms(self.h, terms) == 0

    def spline_eval(self, irbc, which, th, phi, nvar):
        out = np.zeros(nvar)
        assert L().orc_spline_eval(self.h, irbc, which, th, phi, out) == 0
        return out

    def find_projection(self, irbc, xtar, th0, phi0):
        t, p = C.c_double(th0), C.c_double(phi0)
        x0 = np.zeros(3)
        assert L().orc_find_projection(self.h, irbc, np.ascontiguousarray(xtar, np.float64),
                                       C.byref(t), C.byref(p), x0) == 0
        return t.value, p.value, x0

    def mesh_size(self, irbc):
        return L().orc_cell_mesh_size(self.h, irbc)

    def ewald_sl(self, r):
        A, B = C.c_double(), C.c_double()
        L().orc_ewald_sl(self.h, r, C.byref(A), C.byref(B))
        return A.value, B.value

    def ewald_dl(self, r):
        A = C.c_double()
        L().orc_ewald_dl(self.h, r, C.byref(A))
        return A.value

    def mask(self, x):
        return L().orc_mask(self.h, x)


def from_workload(w, threads=0):
    o = Oracle(**w.conf_kwargs())
    if threads:
        Oracle.threads(threads)
    o.set_cells(w.cells)
    o.set_density(1, w.cells.f)
    o.set_density(2, w.cells.g)
    if w.wall is not None:
        x, e2v, nv, ne = w.wall
        o.set_wall(x, e2v, nv, ne)
    return o
