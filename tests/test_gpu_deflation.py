"""GPU parity of the deflated cell BIE (common/ModVelSolver.F90:144-449) against the
oracle: the solid-body modes, and the deflated Solve_RBC_Vel at viscosity ratio 10 (same
GMRES iteration count, velocities to 1e-10 as the undeflated solve's parity test) and
for rigid cells (viscRat -1: the recovered rigid motion and the stored potential)."""
import dataclasses

import numpy as np
import pytest

from tests.conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("visc", [10.0, -1.0])
def test_deflated_solve(visc):
    import rbc3d_amd
    from oracle import oracle_py
    from rbc3d_amd import configs
    w = dataclasses.replace(configs.make_workload("C1", seed=1), viscRat=(visc,))
    g = rbc3d_amd.from_workload(w)
    o = oracle_py.from_workload(w)
    try:
        g.set_deflation(True)
        o.set_deflation(True)
        vg, ng, _ = g.solve_rbc_vel()
        vo, no, _ = o.solve_rbc_vel()
        e_m = relerr(g.sb_modes(), o.sb_modes())
        e_v = relerr(vg, vo)
        print(f"viscRat {visc}: {ng} vs {no} iterations, modes {e_m:.2e}, velocity {e_v:.2e}",
              flush=True)
        assert ng == no and ng > 0
        assert e_m < 1e-12 and e_v < 1e-10
        # a second solve from the first one's solution (the double-layer density the
        # solve left, rigid cells' stored potential included) also agrees
        vg2, ng2, _ = g.solve_rbc_vel()
        vo2, no2, _ = o.solve_rbc_vel()
        assert ng2 == no2 and relerr(vg2, vo2) < 1e-10
    finally:
        g.close()
