"""The staged entry points gfm_hip_adlbl3 / gfm_hip_calc_residual on host
systems (C/Adlbl3.f90:10-97, :239-271) against the oracle.

The system is the manufactured diagonally dominant one of the oracle
tests, on the box the reference relaxes for scalars, with the furnace's
cell codes (walls and blocked cells pinned).  After a few sweeps the GPU
and the oracle agree to rounding (same plane order, same per-plane
operator; the line solves differ only in elimination order, PCR against
the reference's TDMA); after many sweeps both reach the manufactured
solution.
"""
import numpy as np
import pytest

import oracle
from gfm_amd import case as cs
from gfm_amd.solver import CombustionSpace, GfmError
from helpers import manufactured

pytestmark = pytest.mark.gpu


def _setup(order, sweeps, nx=20, precision=0, ny=14, nz=12):
    c = cs.furnace_small(nx=nx, ny=ny, nz=nz, seed=5)
    c.ctrl.sweep_order = order
    c.ctrl.precision = precision
    c.ctrl.ntimesg[3] = sweeps
    box = (4, c.ctrl.mp - 2, 4, c.ctrl.np - 2, 4, c.ctrl.lp - 2)
    return c, box


@pytest.mark.parametrize("order,nx,ny,nz", [(1, 20, 14, 12), (0, 20, 14, 12), (1, 72, 14, 12), (1, 34, 50, 12),
                                            (1, 34, 80, 12), (1, 20, 96, 64), (1, 34, 100, 12), (1, 34, 140, 12),
                                            (1, 20, 160, 12), (1, 20, 180, 12), (1, 20, 220, 12)])
def test_adlbl3_matches_oracle(order, nx, ny, nz):
    """Three sweeps from xt = 0 in the box: every value within 1e-12 of the
    oracle's (relative to the largest).  nx=72 (36 planes per colour) runs
    the tiled line kernel.  Lines of up to 160 rows take the partition
    solve with 2, 4, 6, 8, 10 rows per lane (ny = 14, 50, 80, 100, 140 and
    160 -- the y-lines of the 384x160x96 furnace; the 96x64 plane of the
    192x96x64 furnace: 6 and 4); ny = 180, 220 the PCR solve with 6 and 8
    rows per lane."""
    c, box = _setup(order, 3, nx, ny=ny, nz=nz)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=11)
    xo = xt.copy(order="F")
    oracle.adlbl3(c, 4, ap, as_, bs, xo, box)
    with CombustionSpace(c, initsv=False) as s:
        xg = xt.copy(order="F")
        s.adlbl3(4, ap, as_, bs, xg, box)
    err = np.max(np.abs(xg - xo)) / np.max(np.abs(xo))
    print("adlbl3 order", order, "nx", nx, "ny", ny, "nz", nz, "max rel diff", err, flush=True)
    assert err < 1e-12
    assert np.max(np.abs(xg[sl] - xs[sl])) > 1e-6  # not yet converged: the comparison is of the iterates


@pytest.mark.parametrize("nx,ny,nz", [(20, 14, 12), (72, 14, 12), (20, 96, 64), (20, 160, 24)])
def test_adlbl3_fp32_matches_oracle(nx, ny, nz):
    """Mixed-precision context: coefficients rounded to FP32 and FP32 line
    sweeps (nx=72: tiles of 32 planes of one colour and a partial tile;
    ny=160: the partition solve with 10 rows per lane).
    After three sweeps every value within 2e-5 of the FP64 oracle's
    (relative to the largest); after 60 the manufactured solution is met
    to 2e-5 (the FP32 rounding floor of the line solves)."""
    c, box = _setup(1, 3, nx, precision=1, ny=ny, nz=nz)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=14)
    xo = xt.copy(order="F")
    oracle.adlbl3(c, 4, ap, as_, bs, xo, box)
    with CombustionSpace(c, initsv=False) as s:
        xg = xt.copy(order="F")
        s.adlbl3(4, ap, as_, bs, xg, box)
        err = np.max(np.abs(xg - xo)) / np.max(np.abs(xo))
        c.ctrl.ntimesg[3] = 60
        xg = xt.copy(order="F")
        s.adlbl3(4, ap, as_, bs, xg, box)
    print("adlbl3 fp32 nx", nx, "3 sweeps rel diff", err, flush=True)
    assert 1e-9 < err < 2e-5  # FP32 arithmetic really ran, within its tolerance
    opn = c.arrays["ibcell"][sl] <= 0
    sol = np.max(np.abs(xg[sl][opn] - xs[sl][opn])) / np.max(np.abs(xs[sl][opn]))
    print("adlbl3 fp32 nx", nx, "60 sweeps solution error", sol, flush=True)
    assert sol < 2e-5


def test_fp32_sweeps_keep_fp64_values_of_pinned_rows():
    """In a mixed-precision context the FP32 working copy is written back
    only where the line kernel does not pin the row: rows with a_p = 0
    inside the box keep FP64 values that FP32 cannot hold."""
    c, box = _setup(1, 3, 20, precision=1)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=15)
    rng = np.random.default_rng(3)
    ii = rng.integers(0, xt[sl].shape[0], 12)
    jj = rng.integers(0, xt[sl].shape[1], 12)
    kk = rng.integers(0, xt[sl].shape[2], 12)
    pos = (box[0] - 1 + 2 * ii, box[2] - 1 + 2 * jj, box[4] - 1 + 2 * kk)
    ap[pos] = 0.0
    vals = 1.0 / 3.0 + 1e-9 * np.arange(12)
    xt[pos] = vals
    assert np.all(vals.astype(np.float32).astype(np.float64) != vals)
    with CombustionSpace(c, initsv=False) as s:
        s.adlbl3(4, ap, as_, bs, xt, box)
    np.testing.assert_array_equal(xt[pos], vals)
    opn = (c.arrays["ibcell"][sl] <= 0) & (ap[sl] > 0)
    assert np.max(np.abs(xt[sl][opn])) > 0  # the open rows were relaxed


@pytest.mark.parametrize("order", [1, 0])
def test_adlbl3_reaches_solution(order):
    c, box = _setup(order, 60)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=12)
    with CombustionSpace(c, initsv=False) as s:
        r0 = s.calc_residual(ap, as_, bs, xt, box)
        s.adlbl3(4, ap, as_, bs, xt, box)
        r1 = s.calc_residual(ap, as_, bs, xt, box)
    opn = c.arrays["ibcell"][sl] <= 0
    assert np.max(np.abs(xt[sl][opn] - xs[sl][opn])) < 1e-10
    assert np.all(xt[sl][~opn] == 0.0)  # pinned rows keep their value
    assert r1 < 1e-10 * r0


def test_calc_residual_matches_oracle():
    c, box = _setup(1, 1)
    rng = np.random.default_rng(4)
    shp = c.shape
    ap = np.asfortranarray(rng.uniform(1, 2, shp))
    as_ = np.asfortranarray(rng.uniform(0, 1, shp + (3, 2)))
    bs = np.asfortranarray(rng.uniform(-1, 1, shp))
    xt = np.asfortranarray(rng.uniform(-1, 1, shp))
    ro = oracle.calc_residual(c, ap, as_, bs, xt, box)
    with CombustionSpace(c, initsv=False) as s:
        rg = s.calc_residual(ap, as_, bs, xt, box)
    assert rg == pytest.approx(ro, rel=1e-12)


@pytest.mark.parametrize("bad", [(0, 0), (1, 1), (2, 2)])
def test_box_outside_lattice_is_rejected(bad):
    """A box whose two-step neighbours leave the lattice, or whose ends have
    different parity, is refused before anything runs."""
    c, box = _setup(1, 1)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=13)
    b = list(box)
    axis, kind = bad
    if kind == 0:
        b[2 * axis] = 2                     # begin below 3
    elif kind == 1:
        b[2 * axis + 1] = c.shape[axis] - 1  # end above M-2
    else:
        b[2 * axis + 1] -= 1                # ends of different parity
    with CombustionSpace(c, initsv=False) as s:
        with pytest.raises(GfmError, match="box"):
            s.adlbl3(4, ap, as_, bs, xt, tuple(b))
        with pytest.raises(GfmError, match="box"):
            s.calc_residual(ap, as_, bs, xt, tuple(b))
