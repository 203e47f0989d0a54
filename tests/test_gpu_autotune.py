"""The launch-shape autotuner (gfm_hip_autotune): the iterations it runs
while trying the shapes, and the iterations after it with the chosen ones
(graph replay), give fields identical bit for bit to the default shapes'
on the same case, in FP64 and in mixed precision."""
import numpy as np
import pytest

from gfm_amd import case as cs
from gfm_amd.solver import CombustionSpace

pytestmark = pytest.mark.gpu

FIELDS = ("p", "t", "dnst", "theta", "tmu", "gama", "gf", "ug")


@pytest.mark.parametrize("precision", [0, 1])
def test_autotuned_shapes_are_bit_identical(precision):
    base = cs.furnace_large(nx=48, ny=24, nz=16, seed=3)
    base.ctrl.sweep_order = 1
    base.ctrl.precision = precision
    a, b = base.clone(), base.clone()
    with CombustionSpace(a) as s:
        ha = list(s.sphase(12)) + list(s.sphase(5))
        s.download()
    with CombustionSpace(b) as s:
        report = s.autotune(2)
        print(report, flush=True)
        assert b.ctrl.itr_gas == 12
        hb = [b.ctrl.res_mass]
        hb += list(s.sphase(5))
        s.download()
    assert len(report.splitlines()) > 10
    assert ha[11] == hb[0] and ha[12:] == hb[1:]
    for f in FIELDS:
        np.testing.assert_array_equal(a.arrays[f], b.arrays[f], err_msg=f)
