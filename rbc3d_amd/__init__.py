"""Python host binding of librbc3d_gpu.so, the gfx950 implementation of RBC3D's
Ewald-split boundary-integral matvec (C ABI: include/rbc3d.h).

The names mirror the reference's Fortran module routines (common/ModIntOnRbcs.F90
AddIntOnRbcs, common/ModPME.F90 PME_*, common/ModSphpk.F90 Sh*, ...).  Arrays use
the reference's Fortran layouts, flattened in Fortran order.  There is no CPU
fallback: when the library is missing this module raises on first use.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbc3d_gpu.so")

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

STATUS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "neighbour-cell list overflow (>32)",
          4: "call out of order", 5: "non-finite input", 6: "GMRES did not converge", 7: "file missing, unreadable or malformed"}


class Rbc3dError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn}: status {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class Conf(C.Structure):
    _fields_ = [("Lb", C.c_double * 3), ("alpha", C.c_double), ("rc", C.c_double),
                ("Nb", C.c_int * 3), ("P", C.c_int), ("ntypes", C.c_int),
                ("viscRat", C.c_double * 8), ("vBkg", C.c_double * 3)]


_SIG = {
    "rbc3d_create": [C.POINTER(Conf), C.POINTER(C.c_void_p)],
    "rbc3d_destroy": [C.c_void_p],
    "rbc3d_set_cells": [C.c_void_p, C.c_int, ip, ip, ip, dp, dp, dp],
    "rbc3d_set_cell_density": [C.c_void_p, C.c_int, dp],
    "rbc3d_set_wall": [C.c_void_p, C.c_int, C.c_int, dp, ip],
    "rbc3d_set_wall_density": [C.c_void_p, dp],
    "rbc3d_prepare_sing_int_on_wall": [C.c_void_p],
    "rbc3d_wall_self_nnz": [C.c_void_p, C.POINTER(C.c_longlong)],
    "rbc3d_hashtable_build": [ip, dp, C.c_int, dp, C.c_int, ip, ip],
    "rbc3d_hashtable_build_noalloc": [ip, dp, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p],
    "rbc3d_add_int_on_rbcs": [C.c_void_p, C.c_double, C.c_double, C.c_int, dp, ip, ip, dp, dp],
    "rbc3d_add_int_on_walls": [C.c_void_p, C.c_double, C.c_int, dp, ip, ip, dp, dp],
    "rbc3d_pair_stats": [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_ulonglong)],
    "rbc3d_near_stats": [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
    "rbc3d_set_geometry_reuse": [C.c_void_p, C.c_int],
    "rbc3d_set_graph": [C.c_void_p, C.c_int],
    "rbc3d_set_deflation": [C.c_void_p, C.c_int],
    "rbc3d_sb_modes": [C.c_void_p, dp],
    "rbc3d_graph_ready": [C.c_void_p, C.POINTER(C.c_int)],
    "rbc3d_set_precision": [C.c_void_p, C.c_int],
    "rbc3d_set_gmres_restart": [C.c_void_p, C.c_int],
    "rbc3d_gmres_history": [C.c_void_p, dp, C.c_int, C.POINTER(C.c_int)],
    "rbc3d_mlp_create": [C.POINTER(C.c_void_p)],
    "rbc3d_mlp_destroy": [C.c_void_p],
    "rbc3d_mlp_set_stream": [C.c_void_p, C.c_void_p],
    "rbc3d_mlp_set_params": [C.c_void_p, fp],
    "rbc3d_mlp_get_params": [C.c_void_p, fp],
    "rbc3d_mlp_set_norm": [C.c_void_p, fp, fp, fp],
    "rbc3d_mlp_forward_dev": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "rbc3d_mlp_grad_dev": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.POINTER(C.c_double)],
    "rbc3d_mlp_adam_step_dev": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_float,
                                C.POINTER(C.c_double)],
    "rbc3d_wall_cell_velocity": [C.c_void_p, dp],
    "rbc3d_wall_vertex_normals": [C.c_void_p, dp],
    "rbc3d_noslip_wall_learned": [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int)],
    "rbc3d_pme_distrib_source": [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int],
    "rbc3d_pme_transform": [C.c_void_p],
    "rbc3d_pme_add_interp_vel": [C.c_void_p, C.c_int, dp, ip, dp, dp],
    "rbc3d_pme_grid": [C.c_void_p, C.c_int, dp],
    "rbc3d_sh_anal_gau": [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp, C.c_int, C.c_int],
    "rbc3d_sh_synth_gau": [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp, C.c_int, C.c_int],
    "rbc3d_sh_synth_equ": [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp, C.c_int, C.c_int],
    "rbc3d_sh_grad_gau": [C.c_int, C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, C.c_int,
                          C.c_int],
    "rbc3d_glob_sph_trans": [C.c_void_p, dp, dp, C.c_int],
    "rbc3d_rbc_compute_geometry": [C.c_int, ip, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp],
    "rbc3d_shell_res_force": [C.c_int, ip, ip, dp, dp, dp, dp, dp, dp, dp, dp, C.c_int, ip, dp, dp,
                              dp, dp, dp, dp],
    "rbc3d_rbc_sph_project": [C.c_int, ip, ip, C.c_int, dp],
    "rbc3d_cell_spline": [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, dp],
    "rbc3d_calc_velocity_field": [C.c_void_p, C.c_int, dp, ip, ip, dp, dp],
    "rbc3d_cell_dof": [C.c_void_p, C.POINTER(C.c_int)],
    "rbc3d_cell_matmult": [C.c_void_p, dp, dp],
    "rbc3d_cell_matmult_dev": [C.c_void_p, C.c_void_p, C.c_void_p],
    "rbc3d_cell_rhs": [C.c_void_p, dp],
    "rbc3d_wall_dof": [C.c_void_p, C.POINTER(C.c_int)],
    "rbc3d_wall_matmult": [C.c_void_p, dp, dp],
    "rbc3d_wall_residual_vel": [C.c_void_p, dp],
    "rbc3d_stream": [C.c_void_p, C.POINTER(C.c_void_p)],
    "rbc3d_sync": [C.c_void_p],
    "rbc3d_get_wall_density": [C.c_void_p, dp],
    "rbc3d_cell_rhs_dev": [C.c_void_p, C.c_void_p],
    "rbc3d_noslip_wall": [C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.POINTER(C.c_int),
                          C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "rbc3d_solve_rbc_vel": [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "rbc3d_set_ref_cells": [C.c_void_p, C.c_int, ip, dp, dp, dp, dp],
    "rbc3d_compute_rbc_vel": [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "rbc3d_move_rbcs": [C.c_void_p, C.c_double],
    "rbc3d_rebox_rbcs": [C.c_void_p],
    "rbc3d_vol_constrain_rbcs": [C.c_void_p, C.c_double],
    "rbc3d_filter_rbcs": [C.c_void_p],
    "rbc3d_get_rbc_coords": [C.c_void_p, dp],
    "rbc3d_set_repulsion": [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int],
    "rbc3d_set_split_pairs": [C.c_void_p, C.c_int],
    "rbc3d_timeint_rk2": [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                          C.POINTER(C.c_double), C.c_void_p, C.c_void_p],
    "rbc3d_timeint_ab2": [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int,
                          C.POINTER(C.c_double), C.c_void_p, C.c_void_p],
    "rbc3d_wall_mesh_info": [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "rbc3d_read_wall_mesh": [C.c_char_p, dp, ip],
    "rbc3d_restart_write": [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "rbc3d_restart_info_read": [C.c_char_p, C.c_void_p],
    "rbc3d_restart_read": [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "rbc3d_intra_cell_repulsion": [C.c_void_p, dp, C.POINTER(C.c_int)],
    "rbc3d_inter_cell_repulsion": [C.c_void_p, C.c_double, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double)],
    "rbc3d_timeint_euler": [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int,
                            C.POINTER(C.c_double), C.c_void_p, C.c_void_p],
    "rbc3d_timing": [C.c_void_p, C.c_int],
    "rbc3d_stage_time": [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                         C.POINTER(C.c_char_p)],
}

_lib = None


def lib():
    """Load librbc3d_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with rbc3d_amd/build.py "
                              "(there is no CPU fallback)")
        # The library does not need PyTorch.  A process that uses both must import
        # torch first: torch carries its own HIP runtime, the library then binds to
        # that one, and both see the device (the other order leaves torch without a
        # usable device).  bench.py, learned.py and the test session do so.
        L = C.CDLL(LIB_PATH)
        for name, args in _SIG.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        L.rbc3d_last_error.restype = C.c_char_p
        L.rbc3d_last_error.argtypes = []
        L.rbc3d_mlp_nparams.restype = C.c_int
        L.rbc3d_mlp_nparams.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIG) + ["rbc3d_last_error", "rbc3d_mlp_nparams"]


def _check(name, r):
    if r != 0:
        raise Rbc3dError(name, r, lib().rbc3d_last_error().decode())


def _call(name, *args):
    _check(name, getattr(lib(), name)(*args))


ENOCONV = 6


def _call_solve(name, *args):
    """A GMRES solve: RBC3D_ENOCONV (maxit reached; every output written) is reported
    through the converged flag rather than raised, as the reference prints the
    residual and goes on (common/ModNoSlip.F90:120-126)."""
    r = getattr(lib(), name)(*args)
    if r != ENOCONV:
        _check(name, r)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def hashtable_build(Nc, iLbNc, x, nPoint=None):
    """HashTable_Build (common/ModHashTable.F90:23-58): x is (ldx, 3) Fortran-flat."""
    x = _f64(x)
    ldx = len(x) // 3
    n = ldx if nPoint is None else nPoint
    Nc = _i32(Nc)
    hoc = np.zeros((Nc[0] + 2) * (Nc[1] + 2) * (Nc[2] + 2), np.int32)
    nxt = np.zeros(max(n, 1), np.int32)
    _call("rbc3d_hashtable_build", Nc, _f64(iLbNc), n, x, ldx, hoc, nxt)
    return hoc, nxt[:n]


def hashtable_build_noalloc(Nc, iLbNc, nPoint, d_x, ldx, d_hoc, d_next, stream=0):
    """HashTable_Build on device pointers without allocating (rbc3d_hashtable_build_noalloc):
    d_x (ldx, 3) FP64, d_hoc ((Nc1+2)(Nc2+2)(Nc3+2)) and d_next (nPoint) int32, all raw
    device addresses; queued on ``stream`` (a hipStream_t handle, 0 = null stream)."""
    _call("rbc3d_hashtable_build_noalloc", _i32(Nc), _f64(iLbNc), int(nPoint), d_x, int(ldx),
          d_hoc, d_next, stream)


def sh_anal_gau(nlat, nlon, g, nt=1, ida=None, jda=None):
    """ShAnalGau (common/ModSphpk.F90:74-105): g(nlat, nlon, nt) Fortran-flat -> a, b."""
    ida = ida or nlat
    jda = jda or nlat
    a = np.zeros(ida * jda * nt)
    b = np.zeros(ida * jda * nt)
    _call("rbc3d_sh_anal_gau", nlat, nlon, nt, _f64(g), nlat, nlon, a, b, ida, jda)
    return a, b


def sh_synth_gau(nlat, nlon, a, b, nt=1, ida=None, jda=None):
    ida = ida or nlat
    jda = jda or nlat
    g = np.zeros(nlat * nlon * nt)
    _call("rbc3d_sh_synth_gau", nlat, nlon, nt, g, nlat, nlon, _f64(a), _f64(b), ida, jda)
    return g


def sh_synth_equ(nlat, nlon, a, b, nt=1, ida=None, jda=None):
    ida = ida or nlat
    jda = jda or nlat
    g = np.zeros(nlat * nlon * nt)
    _call("rbc3d_sh_synth_equ", nlat, nlon, nt, g, nlat, nlon, _f64(a), _f64(b), ida, jda)
    return g


def sh_grad_gau(nlat, nlon, a, b, nt=1, ida=None, jda=None):
    """ShGradGau (common/ModSphpk.F90:377-417): coefficients a, b -> v = dg/dth, w = dg/dphi
    on the Gauss grid, (nlat, nlon, nt) Fortran-flat."""
    ida = ida or nlat
    jda = jda or nlat
    v = np.zeros(nlat * nlon * nt)
    w = np.zeros(nlat * nlon * nt)
    _call("rbc3d_sh_grad_gau", nlat, nlon, nt, v, w, nlat, nlon, _f64(a), _f64(b), ida, jda)
    return v, w


GEO_FIELDS = (("a1", 3), ("a2", 3), ("a3", 3), ("a", 4), ("a_rcp", 4), ("b", 4), ("detj", 1))


def rbc_compute_geometry(nlat, x):
    """RBC_ComputeGeometry (common/ModRbc.F90:404-509) of a batch of cells: nlat per cell,
    x the cells' (nlat, 2 nlat, 3) arrays concatenated (Fortran-flat).  Returns a dict of
    the concatenated fields (a1, a2, a3, a, a_rcp, b, detj) and per-cell xc (ncell, 3),
    area, vol, mesh."""
    nlat = np.ascontiguousarray(nlat, dtype=np.int32)
    ncell = len(nlat)
    npt = int(np.sum(2 * nlat.astype(np.int64) ** 2))
    out = {k: np.zeros(n * npt) for k, n in GEO_FIELDS}
    xc = np.zeros(3 * ncell)
    area, vol, mesh = np.zeros(ncell), np.zeros(ncell), np.zeros(ncell)
    _call("rbc3d_rbc_compute_geometry", ncell, nlat, _f64(x), out["a1"], out["a2"], out["a3"],
          out["a"], out["a_rcp"], out["b"], out["detj"], xc, area, vol, mesh)
    out.update(xc=xc.reshape(ncell, 3), area=area, vol=vol, mesh=mesh, nlat=nlat)
    return out


def shell_res_force(geo, celltype, ref_nlat, ref_a_rcp, ref_b, ES, ED, EB):
    """Shell_ResForce (common/ModRbc.F90:815-900) of a batch of cells with geometry `geo`
    (rbc_compute_geometry) against one reference cell per type (its nlat, a_rcp, b,
    concatenated over the types) and per-type moduli; returns f (Cartesian,
    concatenated (nlat, nlon, 3) arrays)."""
    nlat = np.ascontiguousarray(geo["nlat"], dtype=np.int32)
    f = np.zeros_like(geo["a1"])
    ref_nlat = np.ascontiguousarray(ref_nlat, dtype=np.int32)
    _call("rbc3d_shell_res_force", len(nlat), nlat, np.ascontiguousarray(celltype, dtype=np.int32),
          geo["a1"], geo["a2"], geo["a3"], geo["a"], geo["a_rcp"], geo["b"], geo["detj"],
          _f64(geo["area"]), len(ref_nlat), ref_nlat, _f64(ref_a_rcp), _f64(ref_b), _f64(ES),
          _f64(ED), _f64(EB), f)
    return f


def rbc_sph_project(nlat, nlat0, u, nvar):
    """RBC_SphProject (common/ModRbc.F90:659-686) of nvar fields per cell, concatenated."""
    u = np.array(u, dtype=np.float64).ravel()
    _call("rbc3d_rbc_sph_project", len(nlat), np.ascontiguousarray(nlat, dtype=np.int32),
          np.ascontiguousarray(nlat0, dtype=np.int32), int(nvar), u)
    return u


def read_wall_mesh(fn):
    """ReadWallMesh (common/ModIO.F90:547-605): an Exodus II Tri3 mesh.  Returns
    (x(nvert,3) Fortran-flat, e2v(nele,3) 1-based Fortran-flat, nvert, nele), the
    arguments of System.set_wall."""
    fb = os.fsencode(fn)
    nv, ne = C.c_int(), C.c_int()
    _call("rbc3d_wall_mesh_info", fb, C.byref(nv), C.byref(ne))
    x = np.zeros(3 * nv.value)
    e2v = np.zeros(3 * ne.value, np.int32)
    _call("rbc3d_read_wall_mesh", fb, x, e2v)
    return x, e2v, nv.value, ne.value


class RestartInfo(C.Structure):
    _fields_ = [("Lb", C.c_double * 3), ("lt", C.c_int), ("time", C.c_double),
                ("vBkg", C.c_double * 3), ("nrbc", C.c_int), ("nwall", C.c_int),
                ("npoint", C.c_longlong), ("nvert", C.c_longlong), ("nele", C.c_longlong)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def write_restart(fn, Lb, lt, time, vBkg, cells=None, starting_area=None, walls=()):
    """WriteRestart (common/ModIO.F90:742-786).  cells: dict with nlat0, nlon0, nlat, nlon,
    celltype and x (concatenated x(nlat,nlon,3) per cell); starting_area per cell;
    walls: sequence of (x(nvert,3), f(nvert,3), e2v(nele,3) 1-based, nvert, nele)."""
    info = RestartInfo()
    info.Lb[:] = [float(v) for v in Lb]
    info.lt = int(lt)
    info.time = float(time)
    info.vBkg[:] = [float(v) for v in vBkg]
    cell_args = [None] * 7
    if cells is not None and len(cells["nlat"]):
        info.nrbc = len(cells["nlat"])
        cell_args = [_i32(cells["nlat0"]), _i32(cells["nlon0"]), _i32(cells["nlat"]),
                     _i32(cells["nlon"]), _i32(cells["celltype"]), _f64(starting_area),
                     _f64(cells["x"])]
        info.npoint = int((cell_args[2].astype(np.int64) * cell_args[3]).sum())
        assert cell_args[6].size == 3 * info.npoint and cell_args[5].size == info.nrbc
    wall_args = [None] * 5
    if walls:
        info.nwall = len(walls)
        wall_args = [_i32([w[3] for w in walls]), _i32([w[4] for w in walls]),
                     _f64(np.concatenate([np.ravel(w[0]) for w in walls])),
                     _f64(np.concatenate([np.ravel(w[1]) for w in walls])),
                     _i32(np.concatenate([np.ravel(w[2]) for w in walls]))]
    keep = cell_args + wall_args
    _call("rbc3d_restart_write", os.fsencode(fn), C.byref(info), *[_ptr(a) for a in keep])


def read_restart(fn):
    """ReadRestart (common/ModIO.F90:792-921): a dict with Lb, lt, time, vBkg, the cells'
    nlat0, nlon0, nlat, nlon, celltype, starting_area, x and the walls as (x, f, e2v,
    nvert, nele) tuples."""
    fb = os.fsencode(fn)
    info = RestartInfo()
    _call("rbc3d_restart_info_read", fb, C.byref(info))
    nc, nw = info.nrbc, info.nwall
    ci = {k: np.zeros(nc, np.int32) for k in ("nlat0", "nlon0", "nlat", "nlon", "celltype")}
    area = np.zeros(nc)
    x = np.zeros(3 * info.npoint)
    nvert, nele = np.zeros(nw, np.int32), np.zeros(nw, np.int32)
    wx, wf = np.zeros(3 * info.nvert), np.zeros(3 * info.nvert)
    e2v = np.zeros(3 * info.nele, np.int32)
    _call("rbc3d_restart_read", fb, C.byref(info), _ptr(ci["nlat0"]), _ptr(ci["nlon0"]),
          _ptr(ci["nlat"]), _ptr(ci["nlon"]), _ptr(ci["celltype"]), _ptr(area), _ptr(x),
          _ptr(nvert), _ptr(nele), _ptr(wx), _ptr(wf), _ptr(e2v))
    walls, vo, eo = [], 0, 0
    for k in range(nw):
        nv, ne = int(nvert[k]), int(nele[k])
        walls.append((wx[3 * vo:3 * (vo + nv)], wf[3 * vo:3 * (vo + nv)], e2v[3 * eo:3 * (eo + ne)], nv, ne))
        vo += nv
        eo += ne
    return dict(Lb=tuple(info.Lb), lt=info.lt, time=info.time, vBkg=tuple(info.vBkg),
                starting_area=area, x=x, walls=walls, **ci)


class System:
    """The reference's module state (ModConf + ModData) on one GPU: parameters, cells,
    an optional wall, and the PME grids."""

    def __init__(self, Lb, alpha, rc, Nb, P, viscRat, vBkg=(0.0, 0.0, 0.0)):
        c = Conf()
        c.Lb[:] = list(Lb)
        c.alpha = alpha
        c.rc = rc
        c.Nb[:] = list(Nb)
        c.P = P
        c.ntypes = len(viscRat)
        for i, v in enumerate(viscRat):
            c.viscRat[i] = v
        c.vBkg[:] = list(vBkg)
        self.conf = c
        h = C.c_void_p()
        _call("rbc3d_create", C.byref(c), C.byref(h))
        self.h = h
        self.np = 0

    def close(self):
        if getattr(self, "h", None):
            lib().rbc3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- module data ---------------------------------------------------------
    def set_cells(self, cells):
        self.cells = cells  # the geometry last set (for callers that move cells)
        _call("rbc3d_set_cells", self.h, cells.ncell, _i32(cells.nlat0), _i32(cells.nlat),
              _i32(cells.celltype), _f64(cells.x), _f64(cells.a3), _f64(cells.detj))
        self.np = cells.npoint
        self.cells = cells

    def set_cell_density(self, which, dens):
        _call("rbc3d_set_cell_density", self.h, which, _f64(dens))

    def set_wall(self, x, e2v, nvert, nele):
        _call("rbc3d_set_wall", self.h, nvert, nele, _f64(x), _i32(e2v))
        self.nvert = nvert

    def set_wall_density(self, f):
        _call("rbc3d_set_wall_density", self.h, _f64(f))

    def prepare_sing_int_on_wall(self):
        _call("rbc3d_prepare_sing_int_on_wall", self.h)

    # --- target lists (TargetList_Create/Update, common/ModTargetList.F90) ----
    def tlist_rbc(self):
        cells = self.cells
        xs, idx, Ac = [], [], []
        Acoef = [1.0 + v if v > 0 else 1.0 for v in self.conf.viscRat[: self.conf.ntypes]]
        off = 0
        n = cells.npoint
        tx = np.zeros(3 * n)
        ti = np.zeros(3 * n, np.int32)
        A = np.zeros(n)
        p = 0
        for i in range(cells.ncell):
            nl = int(cells.nlat[i])
            nn = 2 * nl
            npt = nl * nn
            blk = cells.x[3 * off: 3 * (off + npt)].reshape(3, npt)
            tx[p:p + npt] = blk[0]
            tx[n + p:n + p + npt] = blk[1]
            tx[2 * n + p:2 * n + p + npt] = blk[2]
            q = np.arange(npt)
            ti[p:p + npt] = i + 1
            ti[n + p:n + p + npt] = q % nl + 1
            ti[2 * n + p:2 * n + p + npt] = q // nl + 1
            A[p:p + npt] = Acoef[cells.celltype[i] - 1]
            p += npt
            off += npt
        return tx, ti, np.ones(n, np.int32), A

    # --- hot path --------------------------------------------------------------
    def add_int_on_rbcs(self, c1, c2, tl, v=None):
        tx, ti, act, A = tl
        n = len(A)
        v = np.zeros(3 * n) if v is None else _f64(v)
        _call("rbc3d_add_int_on_rbcs", self.h, c1, c2, n, _f64(tx), _i32(ti), _i32(act), _f64(A), v)
        return v

    def pair_stats(self):
        n = C.c_longlong()
        s = C.c_ulonglong()
        _call("rbc3d_pair_stats", self.h, C.byref(n), C.byref(s))
        return n.value, s.value

    def set_geometry_reuse(self, enable):
        _call("rbc3d_set_geometry_reuse", self.h, int(bool(enable)))

    def set_deflation(self, on):
        """The reference's Deflate switch of the cell BIE (Pozrikidis' deflation of the
        solid-body and normal modes)."""
        _call("rbc3d_set_deflation", self.h, int(bool(on)))

    def sb_modes(self):
        m = np.zeros(18 * self.np)
        _call("rbc3d_sb_modes", self.h, m)
        return m.reshape(6, 3, self.np)

    def set_graph(self, on):
        """Replay the cell matvec from a recorded hipGraph at fixed geometry."""
        _call("rbc3d_set_graph", self.h, int(bool(on)))

    def graph_ready(self):
        r = C.c_int()
        _call("rbc3d_graph_ready", self.h, C.byref(r))
        return bool(r.value)

    PREC_FP64 = 0
    PREC_SPLIT = 1

    def set_precision(self, mode):
        """0: plain FP64 (default); 1: split form (FP32 PME, split-BF16 Legendre GEMMs,
        FP32 wall self-matrix; pair sums, quadratures and GMRES in FP64 -- FP32 pair sums
        only after set_split_pairs(True))."""
        _call("rbc3d_set_precision", self.h, int(mode))

    def set_split_pairs(self, fp32):
        """The split form's pair sums: FP64 (False, the default) or FP32 (True)."""
        _call("rbc3d_set_split_pairs", self.h, int(bool(fp32)))

    def set_gmres_restart(self, restart):
        """GMRES restart length of both solves (30, PETSc's default, unless set)."""
        _call("rbc3d_set_gmres_restart", self.h, int(restart))

    def near_stats(self):
        """(projections, near-singular corrections) of the last add_int_on_rbcs."""
        a = C.c_longlong()
        b = C.c_longlong()
        _call("rbc3d_near_stats", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def add_int_on_walls(self, c1, tl, v=None):
        tx, ti, act, A = tl
        n = len(A)
        v = np.zeros(3 * n) if v is None else _f64(v)
        _call("rbc3d_add_int_on_walls", self.h, c1, n, _f64(tx), _i32(ti), _i32(act), _f64(A), v)
        return v

    def pme_distrib_source(self, c1, c2, cells=True, walls=True):
        _call("rbc3d_pme_distrib_source", self.h, c1, c2, int(cells), int(walls))

    def pme_transform(self):
        _call("rbc3d_pme_transform", self.h)

    def pme_add_interp_vel(self, tl, v=None):
        tx, ti, act, A = tl
        n = len(A)
        v = np.zeros(3 * n) if v is None else _f64(v)
        _call("rbc3d_pme_add_interp_vel", self.h, n, _f64(tx), _i32(act), _f64(A), v)
        return v

    def pme(self, c1, c2, tl, cells=True, walls=True, v=None):
        self.pme_distrib_source(c1, c2, cells, walls)
        self.pme_transform()
        return self.pme_add_interp_vel(tl, v)

    def pme_grid(self, which):
        N = int(np.prod(list(self.conf.Nb)))
        out = np.zeros((9 if which == 1 else 3) * N)
        _call("rbc3d_pme_grid", self.h, which, out)
        return out

    def calc_velocity_field(self, tl):
        tx, ti, act, A = tl
        n = len(A)
        v = np.zeros(3 * n)
        _call("rbc3d_calc_velocity_field", self.h, n, _f64(tx), _i32(ti), _i32(act), _f64(A), v)
        return v

    def cell_dof(self):
        d = C.c_int()
        _call("rbc3d_cell_dof", self.h, C.byref(d))
        return d.value

    def glob_sph_trans(self, v=None, c=None, direction=1):
        if direction == 1:
            c = np.zeros(self.cell_dof())
            _call("rbc3d_glob_sph_trans", self.h, _f64(v), c, 1)
            return c
        v = np.zeros(3 * self.np)
        _call("rbc3d_glob_sph_trans", self.h, v, _f64(c), -1)
        return v

    def cell_matmult(self, u):
        b = np.zeros(self.cell_dof())
        _call("rbc3d_cell_matmult", self.h, _f64(u), b)
        return b

    def cell_matmult_dev(self, d_u, d_b):
        _call("rbc3d_cell_matmult_dev", self.h, C.c_void_p(d_u), C.c_void_p(d_b))

    def cell_rhs(self):
        b = np.zeros(self.cell_dof())
        _call("rbc3d_cell_rhs", self.h, b)
        return b

    def cell_spline(self, irbc, which, m, n, nvar):
        u, u1, u2, u12 = (np.zeros(m * n * nvar) for _ in range(4))
        _call("rbc3d_cell_spline", self.h, irbc, which, u, u1, u2, u12)
        return u, u1, u2, u12

    def wall_dof(self):
        d = C.c_int()
        _call("rbc3d_wall_dof", self.h, C.byref(d))
        return d.value

    def wall_matmult(self, u):
        b = np.zeros(self.wall_dof())
        _call("rbc3d_wall_matmult", self.h, _f64(u), b)
        return b

    def wall_residual_vel(self):
        v = np.zeros(3 * self.nvert)
        _call("rbc3d_wall_residual_vel", self.h, v)
        return v

    def wall_density(self):
        f = np.zeros(3 * self.nvert)
        _call("rbc3d_get_wall_density", self.h, f)
        return f

    def noslip_wall(self, rtol=1e-8, maxit=60, finc0=None):
        """NoSlipWall: returns (niter, residual estimate, converged, max |v| after)."""
        ni = C.c_int()
        rn = C.c_double()
        cv = C.c_int()
        vm = C.c_double()
        x0 = None
        if finc0 is not None:
            x0 = _f64(finc0)
            x0p = x0.ctypes.data_as(C.c_void_p)
        else:
            x0p = None
        _call_solve("rbc3d_noslip_wall", self.h, rtol, maxit, x0p, C.byref(ni), C.byref(rn), C.byref(cv),
              C.byref(vm))
        return ni.value, rn.value, bool(cv.value), vm.value

    def gmres_history(self):
        """Residual history of the last NoSlipWall solve (initial norm, then one
        estimate per iteration)."""
        n = C.c_int()
        _call("rbc3d_gmres_history", self.h, np.zeros(1), 0, C.byref(n))
        h = np.zeros(max(n.value, 1))
        _call("rbc3d_gmres_history", self.h, h, n.value, C.byref(n))
        return h[:n.value]

    def wall_cell_velocity(self):
        """The cells' velocity at the wall vertices, (nvert,3) flattened in Fortran order."""
        v = np.zeros(3 * self.nvert)
        _call("rbc3d_wall_cell_velocity", self.h, v)
        return v

    def wall_vertex_normals(self):
        n = np.zeros(3 * self.nvert)
        _call("rbc3d_wall_vertex_normals", self.h, n)
        return n

    def noslip_wall_learned(self, mlp, rtol=1e-8, maxit=60):
        """NoSlipWall warm-started by the network: (niter, residual, converged, max |v|
        after, fallback used)."""
        ni, rn, cv, vm, fb = C.c_int(), C.c_double(), C.c_int(), C.c_double(), C.c_int()
        _call_solve("rbc3d_noslip_wall_learned", self.h, mlp.h, rtol, maxit, C.byref(ni), C.byref(rn),
              C.byref(cv), C.byref(vm), C.byref(fb))
        return ni.value, rn.value, bool(cv.value), vm.value, bool(fb.value)

    def solve_rbc_vel(self):
        v = np.zeros(3 * self.np)
        ni = C.c_int()
        rn = C.c_double()
        _call("rbc3d_solve_rbc_vel", self.h, v.ctypes.data_as(C.c_void_p), C.byref(ni), C.byref(rn))
        return v, ni.value, rn.value

    # --- time stepping (common/ModTimeInt.F90) ----------------------------------
    def set_ref_cells(self, nlat, x, ES, ED, EB):
        """rbcRefs: one reference cell per type, x(nlat, 2 nlat, 3) concatenated."""
        _call("rbc3d_set_ref_cells", self.h, len(nlat), _i32(nlat), _f64(x), _f64(ES), _f64(ED),
              _f64(EB))

    def compute_rbc_vel(self):
        """Compute_Rbc_Vel: returns (v(3, np) flat, niter, residual)."""
        v = np.zeros(3 * self.np)
        ni = C.c_int()
        rn = C.c_double()
        _call_solve("rbc3d_compute_rbc_vel", self.h, v.ctypes.data_as(C.c_void_p), C.byref(ni),
                    C.byref(rn))
        return v, ni.value, rn.value

    def move_rbcs(self, Ts):
        _call("rbc3d_move_rbcs", self.h, Ts)

    def rebox_rbcs(self):
        _call("rbc3d_rebox_rbcs", self.h)

    def vol_constrain_rbcs(self, eps_dist):
        _call("rbc3d_vol_constrain_rbcs", self.h, eps_dist)

    def filter_rbcs(self):
        _call("rbc3d_filter_rbcs", self.h)

    def set_repulsion(self, on, force_coef=0.0, visc_rat_thresh=0.0, rigidsep=False):
        _call("rbc3d_set_repulsion", self.h, int(bool(on)), force_coef, visc_rat_thresh,
              int(bool(rigidsep)))

    def intra_cell_repulsion(self):
        """AddIntraCellRepulsionForce on the current coordinates: (df per-cell layout,
        number of nodes with a force)."""
        df = np.zeros(3 * self.np)
        n = C.c_int()
        _call("rbc3d_intra_cell_repulsion", self.h, df, C.byref(n))
        return df, n.value

    def inter_cell_repulsion(self, eps_dist):
        """InterCellRepulsion in place: (nodes moved, smallest separation below 2 eps)."""
        n = C.c_int()
        d = C.c_double()
        _call("rbc3d_inter_cell_repulsion", self.h, eps_dist, C.byref(n), C.byref(d))
        return n.value, d.value

    def rbc_coords(self):
        """The cells' coordinates in rbc3d_set_cells' per-cell layout."""
        x = np.zeros(3 * self.np)
        _call("rbc3d_get_rbc_coords", self.h, x)
        return x

    def timeint_euler(self, Ts, nsteps, eps_dist, noslip_rtol=1e-8, noslip_maxit=60, time=0.0):
        """TimeInt_Euler for nsteps: returns (time, cell iterations, wall iterations)."""
        t = C.c_double(time)
        ic = np.zeros(max(nsteps, 1), np.int32)
        iw = np.zeros(max(nsteps, 1), np.int32)
        _call("rbc3d_timeint_euler", self.h, Ts, nsteps, eps_dist, noslip_rtol, noslip_maxit,
              C.byref(t), ic.ctypes.data_as(C.c_void_p), iw.ctypes.data_as(C.c_void_p))
        return t.value, ic[:nsteps], iw[:nsteps]

    def timeint_rk2(self, alpha, Ts, nsteps, noslip_rtol=1e-8, noslip_maxit=60, time=0.0):
        """TimeInt_RK2(alpha): returns (time, cell iterations, wall iterations), two
        evaluations per step."""
        t = C.c_double(time)
        ic = np.zeros(max(2 * nsteps, 1), np.int32)
        iw = np.zeros(max(2 * nsteps, 1), np.int32)
        _call("rbc3d_timeint_rk2", self.h, alpha, Ts, nsteps, noslip_rtol, noslip_maxit,
              C.byref(t), ic.ctypes.data_as(C.c_void_p), iw.ctypes.data_as(C.c_void_p))
        return t.value, ic[:2 * nsteps], iw[:2 * nsteps]

    def timeint_ab2(self, Ts, nsteps, noslip_rtol=1e-8, noslip_maxit=60, time=0.0):
        """TimeInt_AB2 (nsteps position updates, the first an Euler step)."""
        t = C.c_double(time)
        ic = np.zeros(max(nsteps, 1), np.int32)
        iw = np.zeros(max(nsteps, 1), np.int32)
        _call("rbc3d_timeint_ab2", self.h, Ts, nsteps, noslip_rtol, noslip_maxit, C.byref(t),
              ic.ctypes.data_as(C.c_void_p), iw.ctypes.data_as(C.c_void_p))
        return t.value, ic[:nsteps], iw[:nsteps]

    def stream(self):
        s = C.c_void_p()
        _call("rbc3d_stream", self.h, C.byref(s))
        return s.value

    def sync(self):
        _call("rbc3d_sync", self.h)

    def timing(self, enable):
        _call("rbc3d_timing", self.h, int(enable))

    def stage_times(self):
        """{stage name: (total ms, launches)} from the HIP-event stage timer."""
        out = {}
        s = 0
        while True:
            ms = C.c_double()
            n = C.c_longlong()
            nm = C.c_char_p()
            if lib().rbc3d_stage_time(self.h, s, C.byref(ms), C.byref(n), C.byref(nm)) != 0:
                break
            out[nm.value.decode()] = (ms.value, n.value)
            s += 1
        return out


class Mlp:
    """The per-wall-node network of the learned NoSlipWall warm start (rbc3d_mlp_*).
    Device arrays are passed as raw pointers (e.g. torch tensor data_ptr())."""

    def __init__(self):
        self.h = C.c_void_p()
        _call("rbc3d_mlp_create", C.byref(self.h))
        self.nparams = lib().rbc3d_mlp_nparams()

    def close(self):
        if self.h:
            lib().rbc3d_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """Run the device-pointer entries on `stream` (a HIP stream handle, e.g.
        torch.cuda.current_stream().cuda_stream; 0 is the null stream)."""
        _call("rbc3d_mlp_set_stream", self.h, C.c_void_p(stream))

    def use_torch_stream(self):
        import torch
        self.set_stream(torch.cuda.current_stream().cuda_stream)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _call("rbc3d_mlp_set_params", self.h, p)

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        _call("rbc3d_mlp_get_params", self.h, p)
        return p

    def set_norm(self, mean_in, std_in, scale_out):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        _call("rbc3d_mlp_set_norm", self.h, f(mean_in), f(std_in), f(scale_out))

    def forward_dev(self, n, d_x, d_y):
        _call("rbc3d_mlp_forward_dev", self.h, n, C.c_void_p(d_x), C.c_void_p(d_y))

    def grad_dev(self, n, d_x, d_t, d_grad):
        loss = C.c_double()
        _call("rbc3d_mlp_grad_dev", self.h, n, C.c_void_p(d_x), C.c_void_p(d_t),
              C.c_void_p(d_grad), C.byref(loss))
        return loss.value

    def adam_step_dev(self, n, d_x, d_t, lr, want_loss=False):
        loss = C.c_double()
        _call("rbc3d_mlp_adam_step_dev", self.h, n, C.c_void_p(d_x), C.c_void_p(d_t),
              C.c_float(lr), C.byref(loss) if want_loss else None)
        return loss.value if want_loss else None


def from_workload(w):
    s = System(**w.conf_kwargs())
    s.set_cells(w.cells)
    s.set_cell_density(1, w.cells.f)
    s.set_cell_density(2, w.cells.g)
    if w.wall is not None:
        x, e2v, nv, ne = w.wall
        s.set_wall(x, e2v, nv, ne)
    return s
