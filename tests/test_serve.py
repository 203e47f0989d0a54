"""Batched serving (gfm_amd/serve.py): the sweep is reproducible and drawn
from the stated sizes and operating ranges (CPU); a small sweep completes
in a pool of worker processes sharing the GPU (GPU)."""
import pytest

from gfm_amd import serve


def test_sweep_is_seeded_and_in_range():
    a, b = serve.sweep(24, 5), serve.sweep(24, 5)
    assert a == b
    assert serve.sweep(24, 6) != a
    for j in a:
        assert tuple(j["grid"]) in serve.SIZES
        assert 20.0 <= j["fuel_u"] <= 40.0 and 900.0 <= j["air_t"] <= 1300.0


def test_block_corr_needs_a_tolerance():
    assert serve.main(["--block-corr"]) == 2


def test_stall_criterion_and_lowest_common_level():
    assert not serve.stalled([1.0, 0.5, 0.2, 0.05, 0.01, 0.005])
    assert serve.stalled([0.5, 0.2, 0.004, 0.005, 0.006, 0.005])  # no progress in the last third
    assert serve.stalled([0.5, 0.2, 0.1, 0.06, 0.03, 0.011])      # ends above 1e-2
    mk = lambda q, lv: dict(index=q, first_below={f"{t:g}": ({"iterations": 5, "seconds": 1.0} if t >= lv else None)
                                                     for t in serve.LEVELS})
    level, times = serve.lowest_common_level([mk(0, 1e-4), mk(1, 1e-3), mk(2, 5e-5)])
    assert level == 1e-3 and set(times) == {0, 1, 2}
    assert serve.lowest_common_level([mk(0, 1.0)]) == (None, {})


@pytest.mark.gpu
@pytest.mark.parametrize("tol,bc,learned", [(0.0, False, False), (0.01, True, False), (0.01, False, True)])
def test_small_sweep_completes(tol, bc, learned):
    cases = [dict(index=q, grid=(64, 32, 24), fuel_u=25.0 + 5 * q, air_t=1000.0, seed=10 + q) for q in range(3)]
    summary, recs = serve.serve(cases, workers=2, cap=4, precision="mixed", log=lambda m: print(m, flush=True),
                                pcorr_tol=tol, block_corr=bc, learned=learned)
    assert summary["failed"] == 0
    assert summary["learned_guess"] == learned
    assert summary["block_correction"] == bc
    assert all(r["iterations"] == 4 for r in recs)
    assert summary["cell_updates"] == 3 * 4 * 64 * 32 * 24
