"""The Fortran host path: a Fortran program calls the library through the
rbc3d_bind ISO_C_BINDING module (CalcVelocityField on the two-cell case) and its
velocities match the oracle's to 1e-12 with the identical pair set."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "fortran", "velocity_driver")


def _write_input(path, w, tl):
    c = w.cells
    tx, ti, act, A = tl
    ntypes = len(w.viscRat)
    visc = np.zeros(8)
    visc[:ntypes] = w.viscRat
    with open(path, "wb") as fh:
        np.array([c.ncell, c.npoint, len(A)], np.int32).tofile(fh)
        # rbc3d_conf component by component (unformatted stream I/O has no padding)
        np.array(list(w.Lb) + [w.alpha, w.rc], np.float64).tofile(fh)
        np.array(list(w.Nb) + [w.P, ntypes], np.int32).tofile(fh)
        np.array(list(visc) + list(w.vBkg), np.float64).tofile(fh)
        for a in (c.nlat0, c.nlat, c.celltype):
            np.asarray(a, np.int32).tofile(fh)
        for a in (c.x, c.a3, c.detj, c.f, c.g):
            np.asarray(a, np.float64).tofile(fh)
        np.asarray(tx, np.float64).tofile(fh)
        np.asarray(ti, np.int32).tofile(fh)
        np.asarray(act, np.int32).tofile(fh)
        np.asarray(A, np.float64).tofile(fh)


def test_fortran_host_calc_velocity_field(orc, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/fortran/velocity_driver is not built (run rbc3d_amd/build.py)")
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    o = orc.from_workload(w)
    tl = o.tlist_rbc()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, w, tl)
    print("fortran driver: C1 seed 1", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nT = len(tl[3])
    raw = np.fromfile(fout, np.uint8)
    v = raw[: 24 * nT].view(np.float64)
    npairs, csum = raw[24 * nT:].view(np.int64)
    vo = o.calc_velocity_field(tl)
    _, np_o = o.add_int_on_rbcs(0.0, 0.0, tl)
    assert int(npairs) == np_o
    assert int(np.uint64(csum)) == o.pair_checksum()
    assert relerr(v, vo) < 1e-12
