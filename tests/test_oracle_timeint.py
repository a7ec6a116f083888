"""Pins of the oracle's time stepping (oracle/timeint_oracle.py, TimeInt_Euler in
common/ModTimeInt.F90:109-178):

- Compute_Rbc_Vel + the Euler update dissipate the membrane energy: the rate
  dE/dt = sum(f_shell . v dA) (f_shell the energy's gradient, Shell_ResForce) is negative,
  because the velocity is the Stokes response to the force density -f_shell; a flipped
  sign of the force, of the velocity or of the update makes it positive.  Over several
  full steps (volume constraint and filter included) the energy of a relaxing cell falls.
- a cell of its reference shape, rotated and moved, is at rest (frame indifference).
- ReboxRbcs moves a cell shifted by whole periods back, by whole periods.
- VolConstrainRbcs moves the volume toward the reference volume, to second order, and
  by exactly epsDist/20 along the normals when the limit applies.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle_py as O
from oracle import timeint_oracle as T
from rbc3d_amd import configs
from tests.test_mech_oracle import flat, grid, shape_fn, skalak_energy, unflat

NLAT = 13
AXES = (1.1, 0.95, 0.7)


def ref_cell(nlat=NLAT):
    th, phi, _ = grid(nlat)
    P = np.meshgrid(th, phi, indexing="ij")
    return O.rbc_geometry(flat(shape_fn(0, 0.0, axes=AXES)(*P)))


def deformed_cells(centres, seed, amp=0.08, nlat=NLAT, rots=None):
    th, phi, _ = grid(nlat)
    P = np.meshgrid(th, phi, indexing="ij")
    xs = []
    for k, c in enumerate(centres):
        X = shape_fn(seed + k, amp, axes=AXES)(*P)
        if rots is not None:
            X = X @ rots[k].T
        xs.append(flat(X + np.asarray(c)))
    geos = [O.rbc_geometry(x) for x in xs]
    n = len(xs)
    z = np.zeros(sum(x.size for x in xs))
    return configs.Cells(nlat0=np.full(n, nlat, np.int32), nlat=np.full(n, nlat, np.int32),
                         celltype=np.ones(n, np.int32), x=np.concatenate(xs),
                         a3=np.concatenate([g["a3"] for g in geos]),
                         detj=np.concatenate([g["detj"] for g in geos]), f=z.copy(), g=z.copy(),
                         centres=np.asarray(centres))


def workload(cells, visc, L=8.0):
    return configs.Workload("deformed", (L, L, L), configs.alpha_ref_from_xi(2.0), 2.0,
                            (32, 32, 32), 6, (visc,), (0.0, 0.0, 0.0), cells)


def energy(x, ref, ES, ED):
    return skalak_energy(O.rbc_geometry(x), ref, ES, ED, grid(NLAT)[2])


@pytest.mark.parametrize("visc", [1.0, 5.0])
def test_velocity_dissipates_membrane_energy(visc):
    ES, ED = 1.0, 20.0
    ref = ref_cell()
    cells = deformed_cells([(4.0, 4.0, 4.0)], seed=3)
    o = O.from_workload(workload(cells, visc))
    g, f = T.cell_force(cells.x, NLAT, ref, ES, ED, 0.0)
    o.set_density(1, f)
    v, _, _ = o.solve_rbc_vel()
    vx = v  # one cell: v(3, np) is its x(nlat, nlon, 3) layout
    dA = unflat(g["detj"], NLAT) * grid(NLAT)[2][:, None]
    f_shell = -unflat(f, NLAT, 3)
    rate = np.sum(np.sum(f_shell * unflat(vx, NLAT, 3), -1) * dA)
    h = 1e-4
    dE = (energy(cells.x + h * vx, ref, ES, ED) - energy(cells.x - h * vx, ref, ES, ED)) / (2 * h)
    print(f"viscRat {visc}: dE/dt {dE:.6e}, sum f.v dA {rate:.6e}")
    assert rate < 0 and dE < 0
    # f is the projected force: the rate matches the energy's derivative to the
    # projection's truncation (modes beyond the grid)
    assert abs(rate - dE) < 2e-2 * abs(dE)


def test_full_steps_relax_the_cell():
    ES, ED, EB = 1.0, 20.0, 0.0
    ref = ref_cell()
    cells = deformed_cells([(4.0, 4.0, 4.0)], seed=5)
    w = workload(cells, 5.0)
    o = O.from_workload(w)
    e = [energy(cells.x, ref, ES, ED)]
    x = cells.x
    for _ in range(4):
        x, its = T.timeint_euler(o, dataclasses.replace(cells, x=x), [ref], [ES], [ED], [EB],
                                 0.02, 1, 0.05, w.Lb)
        e.append(energy(x, ref, ES, ED))
    print("energy", e)
    assert all(b < a for a, b in zip(e, e[1:]))


def test_reference_shape_is_at_rest():
    ES, ED, EB = 1.0, 20.0, 0.03
    ref = ref_cell()
    rng = np.random.default_rng(4)
    rots = [configs.random_rotation(rng) for _ in range(2)]
    cells = deformed_cells([(2.5, 2.5, 2.5), (5.5, 5.6, 5.4)], seed=0, amp=0.0, rots=rots)
    w = workload(cells, 5.0)
    x, _ = T.timeint_euler(O.from_workload(w), cells, [ref], [ES], [ED], [EB], 0.05, 1, 0.05, w.Lb)
    xf = np.concatenate([O.rbc_sph_project(xc, NLAT, NLAT) for xc in T.split_cells(cells.x, cells.nlat, 3)])
    d = np.abs(x - xf).max()
    print(f"reference shapes after one step: moved {d:.2e} beyond the filter")
    assert d < 1e-10


def test_rebox():
    cells = deformed_cells([(4.0, 4.0, 4.0)], seed=1)
    L = (8.0, 8.0, 8.0)
    X = cells.x.reshape(3, -1)
    for shift in ((8.0, -16.0, 24.0), (0.0, 0.0, -8.0)):
        moved = (X + np.asarray(shift)[:, None]).ravel()
        back = T.rebox(moved.copy(), L).reshape(3, -1)
        # moved back by whole periods, the centre of the coordinate range in [0, L)
        assert np.abs(back - X).max() < 1e-13
        c = 0.5 * (back.max(1) + back.min(1))
        assert np.all(c >= 0) and np.all(c < 8.0)


def test_vol_constrain():
    ref = ref_cell()
    cells = deformed_cells([(4.0, 4.0, 4.0)], seed=2, amp=0.0)
    g0 = O.rbc_geometry(cells.x)
    x = cells.x + 0.004 * g0["a3"]  # an inflated reference shape
    v0 = O.rbc_geometry(x)["vol"] - ref["vol"]
    v1 = O.rbc_geometry(T.vol_constrain(x, ref["vol"], 1.0))["vol"] - ref["vol"]
    print(f"volume error {v0:.3e} -> {v1:.3e}")
    # first order removed: what is left is the curvature term, ~ H fac ~ 4e-3 of it
    assert v0 > 0 and abs(v1) < 1e-2 * abs(v0)
    # the limit: |fac| = epsDist/20, along -a3 for an inflated cell
    g = O.rbc_geometry(x)
    eps = 0.01
    d = (T.vol_constrain(x, ref["vol"], eps) - x).reshape(3, -1)
    assert np.allclose(d, -(eps / 20) * g["a3"].reshape(3, -1), rtol=0, atol=1e-14)


def trajectories(method, hs, T):
    """final coordinates of a relaxing deformed cell (viscosity ratio 1, EB 0) after time T
    with steps hs"""
    ES, ED, EB = 1.0, 2.0, 0.0
    ref = ref_cell()
    cells = deformed_cells([(4.0, 4.0, 4.0)], seed=5, amp=0.12)
    w = workload(cells, 1.0)
    out = []
    for h in hs:
        o = O.from_workload(w)
        n = int(round(T / h))
        if method == "rk2":
            x, _ = T_.timeint_rk2(o, cells, [ref], [ES], [ED], [EB], 2.0 / 3.0, h, n, w.Lb)
        else:
            x, _ = T_.timeint_ab2(o, cells, [ref], [ES], [ED], [EB], h, n, w.Lb)
        out.append(x)
    return out


T_ = T


@pytest.mark.parametrize("method", ["rk2", "ab2"])
def test_second_order_integrators(method):
    """RK2 (Ralston, alpha = 2/3) and AB2 converge at second order: halving the step
    quarters the difference between successive solutions (Euler would halve it)."""
    x1, x2, x3 = trajectories(method, [0.01, 0.005, 0.0025], 0.02)
    e1, e2 = np.abs(x1 - x2).max(), np.abs(x2 - x3).max()
    print(f"{method}: successive differences {e1:.3e}, {e2:.3e}, ratio {e1 / e2:.2f}")
    assert 3.5 < e1 / e2 < 4.6
