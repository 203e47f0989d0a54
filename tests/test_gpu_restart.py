"""Checkpoint and resume of a device-resident run through the reference's
restart record rg####c.d (gfm_amd/restart.py; SAV3F C/Sav3f.f90:21-88,
read_restart_data C/Initsv.f90:225-316).

The restart record does not hold every array the iteration carries (the
reaction rates, the wall-function viscosity on faces, velocities at
non-face positions are re-derived), so a resumed run is not the
uninterrupted one -- in the reference either.  What is pinned: the GPU run
resumed from a GPU checkpoint equals the oracle resumed from the same
record, to the tolerance of the other parity tests."""
import numpy as np
import pytest

import oracle
from gfm_amd import case as cs
from gfm_amd import restart as rs
from gfm_amd.solver import CombustionSpace
from helpers import field_errors

pytestmark = pytest.mark.gpu


def test_gpu_resume_matches_oracle_resume(tmp_path):
    base = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    base.ctrl.sweep_order = 1
    g0 = base.clone()
    with CombustionSpace(g0) as s:
        s.sphase(6)
        s.download()
    path = tmp_path / "rg0001c.d"
    rs.save_rg(g0, path)
    fresh = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    fresh.ctrl.sweep_order = 1
    fresh.ctrl.itr_gas = 0
    rs.load_rg(fresh, path)
    assert fresh.ctrl.itr_gas == 6
    o, g = fresh.clone(), fresh.clone()
    oracle.initsv_tail(o)
    h_o = oracle.sphase(o, 2)
    with CombustionSpace(g) as s:
        h_g = s.sphase(2)
        s.download()
    errs = field_errors(o, g)
    print("resumed at 6, 2 iterations:", h_o, h_g, errs, flush=True)
    np.testing.assert_allclose(h_g, h_o, rtol=1e-9)
    assert max(errs.values()) < 1e-8
    assert g.ctrl.itr_gas == o.ctrl.itr_gas == 8
