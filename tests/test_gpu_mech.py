"""GPU parity of the per-cell geometry and membrane mechanics (rbc3d_amd/csrc/mech.hip)
against the oracle (oracle/mech_oracle.c): RBC_ComputeGeometry, Shell_ResForce and
RBC_SphProject on a batch of deformed cells of two grid types in one launch, and on 512
cells at nlat 17 (the batched-cell workload's grid).  Bar: 1e-12 max-norm relative per
output (the same arithmetic up to summation order)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests.conftest import relerr
from tests.test_mech_oracle import flat, grid, shape_fn

pytestmark = pytest.mark.gpu
TOL = 1e-12


def batch(nlats, seed):
    xs = []
    rng = np.random.default_rng(seed)
    for k, nl in enumerate(nlats):
        th, phi, _ = grid(nl)
        T, P = np.meshgrid(th, phi, indexing="ij")
        X = shape_fn(seed + k, 0.1, lmax=4)(T, P) + rng.uniform(-5, 5, 3)
        xs.append(flat(X))
    return xs


def split(v, nlats, k):
    out, o = [], 0
    for nl in nlats:
        n = 2 * nl * nl * k
        out.append(v[o:o + n])
        o += n
    return out


def run_case(R, nlats, types, seed, moduli):
    xs = batch(nlats, seed)
    # one reference cell per type (its grid is that of its cells)
    ref_nl = [13, 17]
    refs = []
    for k, nl in enumerate(ref_nl):
        th, phi, _ = grid(nl)
        T, P = np.meshgrid(th, phi, indexing="ij")
        refs.append(O.rbc_geometry(flat(shape_fn(100 + k, 0.03, lmax=3)(T, P))))
    geo = R.rbc_compute_geometry(nlats, np.concatenate(xs))
    ES, ED, EB = (np.array([m[i] for m in moduli]) for i in range(3))
    f = R.shell_res_force(geo, types, ref_nl, np.concatenate([r["a_rcp"] for r in refs]),
                          np.concatenate([r["b"] for r in refs]), ES, ED, EB)
    errs = {}
    for name, k in (("a1", 3), ("a2", 3), ("a3", 3), ("a", 4), ("a_rcp", 4), ("b", 4), ("detj", 1)):
        errs[name] = 0.0
        for c, part in enumerate(split(geo[name], nlats, k)):
            og = O.rbc_geometry(xs[c])
            errs[name] = max(errs[name], relerr(part, og[name]))
    e_cell = 0.0
    e_f = 0.0
    for c, (x, part) in enumerate(zip(xs, split(f, nlats, 3))):
        og = O.rbc_geometry(x)
        e_cell = max(e_cell, abs(geo["area"][c] / og["area"] - 1), abs(geo["vol"][c] / og["vol"] - 1),
                     np.abs(geo["xc"][c] - og["xc"]).max() / np.abs(og["xc"]).max(),
                     abs(geo["mesh"][c] / og["mesh"] - 1))
        t = types[c] - 1
        fo = O.shell_res_force(og, refs[t], ES[t], ED[t], EB[t])
        e_f = max(e_f, relerr(part, fo))
    return errs, e_cell, e_f


def test_geometry_and_force_two_grids():
    import rbc3d_amd as R
    nlats = [13, 17, 13, 17, 17, 13]
    types = np.array([1, 2, 1, 2, 2, 1], np.int32)
    errs, e_cell, e_f = run_case(R, nlats, types, 3, [(1.0, 50.0, 0.03), (2.0, 20.0, 0.1)])
    print(f"geometry {errs}, per-cell sums {e_cell:.2e}, force {e_f:.2e}", flush=True)
    assert max(errs.values()) < TOL and e_cell < TOL and e_f < TOL


def test_geometry_and_force_512_cells():
    import rbc3d_amd as R
    nlats = [17] * 512
    types = np.array([2] * 512, np.int32)
    xs = batch(nlats, 11)
    geo = R.rbc_compute_geometry(nlats, np.concatenate(xs))
    th, phi, _ = grid(17)
    T, P = np.meshgrid(th, phi, indexing="ij")
    ref = O.rbc_geometry(flat(shape_fn(101, 0.03, lmax=3)(T, P)))
    ref13 = O.rbc_geometry(flat(shape_fn(100, 0.03, lmax=3)(*np.meshgrid(*grid(13)[:2], indexing="ij"))))
    f = R.shell_res_force(geo, types, [13, 17], np.concatenate([ref13["a_rcp"], ref["a_rcp"]]),
                          np.concatenate([ref13["b"], ref["b"]]), [1.0, 2.0], [50.0, 20.0],
                          [0.03, 0.1])
    fs = split(f, nlats, 3)
    e = 0.0
    for c in range(0, 512, 37):  # a sample of the cells against the oracle
        og = O.rbc_geometry(xs[c])
        e = max(e, relerr(split(geo["b"], nlats, 4)[c], og["b"]),
                relerr(fs[c], O.shell_res_force(og, ref, 2.0, 20.0, 0.1)))
    print(f"512 cells: sampled b and force {e:.2e}", flush=True)
    assert e < TOL


def test_sph_project_two_grids():
    import rbc3d_amd as R
    nlats = [13, 17, 13]
    nlat0 = [13, 11, 9]
    rng = np.random.default_rng(2)
    u = rng.standard_normal(sum(2 * n * n * 3 for n in nlats))
    pg = R.rbc_sph_project(nlats, nlat0, u, 3)
    e = 0.0
    for part, ou, nl, n0 in zip(split(pg, nlats, 3), split(u, nlats, 3), nlats, nlat0):
        e = max(e, relerr(part, O.rbc_sph_project(ou, nl, n0)))
    print(f"RBC_SphProject: {e:.2e}", flush=True)
    assert e < TOL


def test_geometry_rejects_non_finite():
    import rbc3d_amd as R
    x = np.concatenate(batch([13], 1))
    x[7] = np.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        R.rbc_compute_geometry([13], x)
