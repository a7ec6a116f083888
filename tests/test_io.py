"""The file formats of common/ModIO.F90 (rbc3d_amd/csrc/io.hip), host code, run on the CPU:

- ReadWallMesh (:547-605) on two of the reference's sample meshes: the coordinates and
  connectivity agree, value for value, with what scipy's independent netCDF reader gives;
  the tube of diameter 2 has every vertex at radius 1 and is a closed-up annulus (Euler
  characteristic 0).  Meshes scipy writes in the classic and 64-bit-offset encodings, with
  single-precision coordinates, read the same; a mesh that is not Tri3, has two element
  blocks or is missing is an RBC3D_EIO error.
- WriteRestart/ReadRestart (:742-921): a file written by the library has exactly the bytes
  of the record layout restated with Python's struct module; a Fortran program (amdflang)
  reads it with unformatted sequential input and writes it back byte for byte identical;
  the library reads back what it wrote; a truncated file and changed counts are errors,
  and so is a cell grid ReadRestart's dealiasing inference would not re-create.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

import rbc3d_amd as R

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = [os.path.join(HERE, "golden", "exodus", n) for n in ("new_cyl_D2_L13_33.e", "squeeze_D8_D6_L10.e")]


def scipy_mesh(fn):
    with netcdf_file(fn, "r", mmap=False) as nc:
        x = np.concatenate([np.asarray(nc.variables[v][:], np.float64) for v in ("coordx", "coordy", "coordz")])
        c = np.asarray(nc.variables["connect1"][:], np.int64)  # (nele, 3)
    return x, c


@pytest.mark.parametrize("fn", MESHES)
def test_exodus_matches_scipy(fn):
    x, e2v, nv, ne = R.read_wall_mesh(fn)
    xs, c = scipy_mesh(fn)
    assert nv * 3 == xs.size and ne == c.shape[0]
    assert np.array_equal(x, xs)
    assert np.array_equal(e2v.reshape(3, ne).T, c)  # e2v(nele,3) column-major = connect1


def test_exodus_tube_geometry():
    x, e2v, nv, ne = R.read_wall_mesh(MESHES[0])
    X = x.reshape(3, nv)
    assert np.abs(np.hypot(X[0], X[1]) - 1.0).max() < 1e-12
    assert X[2].min() == pytest.approx(-6.665) and X[2].max() == pytest.approx(6.665)
    E = e2v.reshape(3, ne).T
    edges = {tuple(sorted((int(t[a]), int(t[(a + 1) % 3])))) for t in E for a in range(3)}
    assert nv - len(edges) + ne == 0


def write_mesh(fn, version, npe=3, blocks=1, dtype="f"):
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]], np.float64)
    conn = np.array([[1, 2, 3], [2, 4, 3]], np.int32)
    with netcdf_file(fn, "w", version=version) as nc:
        nc.title = "Tri3 test mesh"
        nc.createDimension("time_step", None)
        nc.createDimension("num_nodes", 4)
        nc.createDimension("num_elem", 2)
        nc.createDimension("num_el_blk", blocks)
        nc.createDimension("num_el_in_blk1", 2 if blocks == 1 else 1)
        nc.createDimension("num_nod_per_el1", npe)
        t = nc.createVariable("time_whole", "d", ("time_step",))
        t[0] = 0.0
        for k, v in enumerate(("coordx", "coordy", "coordz")):
            var = nc.createVariable(v, dtype, ("num_nodes",))
            var[:] = x[:, k]
        cv = nc.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        cv.elem_type = "TRI3"
        if npe == 3:
            cv[:] = conn if blocks == 1 else conn[:1]
        else:
            cv[:] = np.c_[conn, conn[:, :1]] if blocks == 1 else np.c_[conn, conn[:, :1]][:1]
    return x, conn


@pytest.mark.parametrize("version", [1, 2])
@pytest.mark.parametrize("dtype", ["f", "d"])
def test_exodus_encodings(tmp_path, version, dtype):
    fn = str(tmp_path / "m.e")
    x, conn = write_mesh(fn, version, dtype=dtype)
    xr, e2v, nv, ne = R.read_wall_mesh(fn)
    assert (nv, ne) == (4, 2)
    assert np.array_equal(xr.reshape(3, 4).T, x.astype(dtype).astype(np.float64))
    assert np.array_equal(e2v.reshape(3, 2).T, conn)


def test_exodus_errors(tmp_path):
    fn = str(tmp_path / "quad.e")
    write_mesh(fn, 2, npe=4)
    with pytest.raises(R.Rbc3dError, match="Tri3") as e:
        R.read_wall_mesh(fn)
    assert e.value.code == 7
    write_mesh(fn, 2, blocks=2)
    with pytest.raises(R.Rbc3dError, match="element block"):
        R.read_wall_mesh(fn)
    with pytest.raises(R.Rbc3dError, match="does not exist"):
        R.read_wall_mesh(str(tmp_path / "missing.e"))
    (tmp_path / "junk.e").write_bytes(b"not a mesh at all")
    with pytest.raises(R.Rbc3dError, match="netCDF"):
        R.read_wall_mesh(str(tmp_path / "junk.e"))


def sample_state(seed=0):
    rng = np.random.default_rng(seed)
    # grids ReadRestart re-creates: nlat0 + 2 (nlat/nlat0 < 1.5) or k nlat0
    nlat = np.array([7, 8, 12], np.int32)
    cells = dict(nlat0=np.array([5, 6, 4], np.int32), nlon0=np.array([10, 12, 8], np.int32),
                 nlat=nlat, nlon=2 * nlat, celltype=np.array([1, 2, 1], np.int32),
                 x=rng.standard_normal(int((3 * nlat * 2 * nlat).sum())))
    area = rng.uniform(10, 14, 3)
    walls = []
    for nv, ne in ((6, 7), (4, 3)):
        walls.append((rng.standard_normal(3 * nv), rng.standard_normal(3 * nv),
                      rng.integers(1, nv + 1, 3 * ne).astype(np.int32), nv, ne))
    return cells, area, walls


def expected_bytes(Lb, lt, time, vBkg, cells, area, walls):
    """the restart records restated: each payload framed by its length (little-endian
    4-byte integer) before and after; 'i' 4-byte integers, 'd' 8-byte reals"""
    out = bytearray()

    def rec(fmt, *vals):
        body = struct.pack("<" + fmt, *vals)
        out.extend(struct.pack("<i", len(body)) + body + struct.pack("<i", len(body)))

    rec("3d", *Lb)
    rec("i", lt)
    rec("d", time)
    rec("3d", *vBkg)
    rec("i", len(cells["nlat"]))
    o = 0
    for i in range(len(cells["nlat"])):
        n = 3 * int(cells["nlat"][i]) * int(cells["nlon"][i])
        rec("2i", int(cells["nlat0"][i]), int(cells["nlon0"][i]))
        rec("2i", int(cells["nlat"][i]), int(cells["nlon"][i]))
        rec("i", int(cells["celltype"][i]))
        rec("d", float(area[i]))
        rec(f"{n}d", *cells["x"][o:o + n])
        o += n
    rec("i", len(walls))
    for wx, wf, e2v, nv, ne in walls:
        rec("2i", nv, ne)
        rec(f"{3 * nv}d", *wx)
        rec(f"{3 * nv}d", *wf)
        rec(f"{3 * ne}i", *[int(v) for v in e2v])
    return bytes(out)


def test_restart_layout_and_roundtrip(tmp_path):
    cells, area, walls = sample_state()
    Lb, lt, time, vBkg = (8.0, 9.0, 10.5), 1234, 12.25, (0.0, 0.1, 1.0)
    fn = str(tmp_path / "restart.dat")
    R.write_restart(fn, Lb, lt, time, vBkg, cells, area, walls)
    data = open(fn, "rb").read()
    assert data == expected_bytes(Lb, lt, time, vBkg, cells, area, walls)
    r = R.read_restart(fn)
    assert r["Lb"] == Lb and r["lt"] == lt and r["time"] == time and r["vBkg"] == vBkg
    for k in ("nlat0", "nlon0", "nlat", "nlon", "celltype", "x"):
        assert np.array_equal(r[k], cells[k]), k
    assert np.array_equal(r["starting_area"], area)
    for a, b in zip(r["walls"], walls):
        assert all(np.array_equal(u, v) for u, v in zip(a[:3], b[:3])) and a[3:] == b[3:]


def test_restart_through_fortran(tmp_path):
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    if not shutil.which(fc):
        pytest.fail(f"no Fortran compiler at {fc}")
    exe = str(tmp_path / "restart_copy")
    subprocess.run([fc, "-O1", os.path.join(HERE, "fortran", "restart_copy.f90"), "-o", exe], check=True)
    cells, area, walls = sample_state(1)
    fn, fn2 = str(tmp_path / "a.dat"), str(tmp_path / "b.dat")
    R.write_restart(fn, (8.0, 8.0, 16.0), 7, 0.5, (0.0, 0.0, 1.0), cells, area, walls)
    out = subprocess.run([exe, fn, fn2], check=True, capture_output=True, text=True).stdout
    assert "lt=7 nrbc=3 nwall=2" in out
    assert open(fn2, "rb").read() == open(fn, "rb").read()
    # and no cells, no walls
    R.write_restart(fn, (8.0, 8.0, 16.0), 0, 0.0, (0.0, 0.0, 0.0))
    subprocess.run([exe, fn, fn2], check=True, capture_output=True)
    assert open(fn2, "rb").read() == open(fn, "rb").read()
    r = R.read_restart(fn2)
    assert r["x"].size == 0 and r["walls"] == []


@pytest.mark.parametrize("nlat0,nlat,ok", [(13, 15, True), (6, 12, True), (4, 12, True),
                                             (13, 13, False), (6, 10, False), (5, 16, False)])
def test_restart_grid_inference(tmp_path, nlat0, nlat, ok):
    """ReadRestart's dealiasing inference: 100 (nlat = nlat0 + 2) below a ratio of 1.5,
    else the integer quotient; a grid RBC_Create would not re-create is refused, as the
    reference stops on it (a cell with nlat = nlat0 among them)."""
    cells = dict(nlat0=np.array([nlat0], np.int32), nlon0=np.array([2 * nlat0], np.int32),
                 nlat=np.array([nlat], np.int32), nlon=np.array([2 * nlat], np.int32),
                 celltype=np.array([1], np.int32), x=np.zeros(3 * 2 * nlat * nlat))
    fn = str(tmp_path / "g.dat")
    R.write_restart(fn, (8.0, 8.0, 8.0), 0, 0.0, (0, 0, 0), cells, np.array([12.0]))
    if ok:
        assert R.read_restart(fn)["nlat"][0] == nlat
    else:
        with pytest.raises(R.Rbc3dError, match="invalid array dimension"):
            R.read_restart(fn)


def test_restart_errors(tmp_path):
    cells, area, walls = sample_state(2)
    fn = str(tmp_path / "r.dat")
    R.write_restart(fn, (8.0, 8.0, 8.0), 1, 1.0, (0, 0, 0), cells, area, walls)
    data = open(fn, "rb").read()
    open(fn, "wb").write(data[:-10])
    with pytest.raises(R.Rbc3dError, match="truncated") as e:
        R.read_restart(fn)
    assert e.value.code == 7
    with pytest.raises(R.Rbc3dError, match="cannot read"):
        R.read_restart(str(tmp_path / "none.dat"))
