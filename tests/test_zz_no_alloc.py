"""gfm_hip_calc_residual_noalloc (calc_residual allocating nothing, on a
system already in device memory) against gfm_hip_calc_residual (the
context entry point, host arrays), through the C-ABI.  Two calls in a row,
the second on a smaller lattice, so the static partial-sum array is reused
with fewer rows."""
import ctypes as ct

import numpy as np
import pytest

from gfm_amd import case as cs
from gfm_amd.solver import CombustionSpace, GfmError, lib
from helpers import manufactured

pytestmark = pytest.mark.gpu


def _bind():
    L = lib()
    P, I = ct.c_void_p, ct.c_int
    L.gfm_hip_calc_residual_noalloc.argtypes = [I, I, I, I, P, P, P, P, P] + [I] * 6 + [P, P]
    L.gfm_hip_calc_residual_noalloc.restype = I
    return L


def _system(nx, ny, nz, seed):
    c = cs.furnace_small(nx=nx, ny=ny, nz=nz, seed=seed)
    mp, np_, lp = c.shape
    box = (4, mp - 2, 4, np_ - 2, 4, lp - 2)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed)
    # a state that is not the solution: the residual is of order one
    rng = np.random.default_rng(seed + 1)
    xt = np.asfortranarray(xs * (1.0 + 0.1 * rng.uniform(-1, 1, xs.shape)))
    return c, box, ap, as_, bs, xt


class _Hip:
    """Device buffers through the HIP runtime libgfm_hip.so links against
    (loaded with it; torch bundles a runtime of its own whose pointers
    this one does not know)."""

    def __init__(self):
        lib()
        self.L = ct.CDLL("libamdhip64.so.7")
        self.bufs = []
        assert self.L.hipSetDevice(0) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ct.c_void_p()
        assert self.L.hipMalloc(ct.byref(p), ct.c_size_t(a.nbytes)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ct.c_void_p), ct.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype=dtype)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(ct.c_void_p), p, ct.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(c, box, ap, as_, bs, xt):
    L = _bind()
    h = _Hip()
    try:
        flat = lambda a: a.ravel(order="F")
        ib = h.put(flat(c.arrays["ibcell"]).astype(np.int16))
        dap, das, dbs, dxt = (h.put(flat(a).astype(np.float64)) for a in (ap, as_, bs, xt))
        out = h.put(np.array([-1.0]))
        mp, np_, lp = c.shape
        rc = L.gfm_hip_calc_residual_noalloc(mp, np_, lp, c.ctrl.ncells, ib, dap, das, dbs, dxt, *box, out, None)
        assert rc == 0, L.gfm_hip_last_error()
        return float(h.get(out, 1, np.float64)[0])
    finally:
        h.free()


def test_noalloc_matches_context_entry_point():
    # the second system is smaller (fewer (J,K) rows than the first)
    for dims, seed in (((24, 16, 12), 5), ((12, 10, 8), 6)):
        c, box, ap, as_, bs, xt = _system(*dims, seed)
        with CombustionSpace(c, initsv=False) as s:
            want = s.calc_residual(ap, as_, bs, xt, box)
        got = _noalloc(c, box, ap, as_, bs, xt)
        print(dims, "context", want, "no-alloc", got, flush=True)
        assert want > 1e-6
        assert got == pytest.approx(want, rel=1e-13)


def test_noalloc_sub_box():
    """A box inside the lattice (not the whole p' box) on a third, larger
    system after the smaller ones."""
    c, box, ap, as_, bs, xt = _system(32, 20, 14, 7)
    mp, np_, lp = c.shape
    sub = (6, mp - 6, 8, np_ - 4, 4, lp - 6)
    with CombustionSpace(c, initsv=False) as s:
        want = s.calc_residual(ap, as_, bs, xt, sub)
    got = _noalloc(c, sub, ap, as_, bs, xt)
    assert got == pytest.approx(want, rel=1e-13)
