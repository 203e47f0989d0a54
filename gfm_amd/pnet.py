"""Host mirror of include/gfm_pnet.h: the U-Net that supplies the initial
guess of the pressure-correction (p') solve, trained and run on the GPU by
the library's BF16 matrix-core kernels.  No CPU fallback: every call goes
through libgfm_hip.so."""
import ctypes as ct

import numpy as np

from .solver import GfmError, lib as _lib

DEFAULT_CHANNELS = (32, 32, 64, 64)


def _bind():
    L = _lib()
    if getattr(L, "_pnet_bound", False):
        return L
    P = ct.c_void_p
    F = ct.POINTER(ct.c_float)
    L.gfm_pnet_create.argtypes = [ct.c_int, ct.POINTER(ct.c_int), ct.c_int, ct.c_longlong, ct.c_uint64, ct.POINTER(P)]
    L.gfm_pnet_destroy.argtypes = [P]
    L.gfm_pnet_param_count.argtypes = [P, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t)]
    L.gfm_pnet_get_params.argtypes = [P, F, ct.c_size_t]
    L.gfm_pnet_set_params.argtypes = [P, F, ct.c_size_t]
    L.gfm_pnet_forward.argtypes = [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, F, F]
    L.gfm_pnet_loss_grad.argtypes = [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, F, F, ct.POINTER(ct.c_double), F]
    L.gfm_pnet_adam_step.argtypes = [P, ct.c_float, ct.c_float, ct.c_float, ct.c_float]
    L.gfm_pnet_train.argtypes = [P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, F, F, ct.c_int, ct.c_int, ct.c_float, ct.c_uint64,
                                 ct.POINTER(ct.c_double)]
    L._pnet_bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise GfmError(f"{what} failed (code {rc}): {_lib().gfm_hip_last_error().decode(errors='replace')}")


def _fp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


class PNet:
    """Network with buffers for up to max_batch grids of max_cells cells."""

    def __init__(self, channels=DEFAULT_CHANNELS, max_batch=1, max_cells=4096, seed=0, device=0):
        L = _bind()
        self.channels = tuple(int(c) for c in channels)
        ch = (ct.c_int * 4)(*self.channels)
        self._h = ct.c_void_p()
        _check(L.gfm_pnet_create(int(device), ch, int(max_batch), int(max_cells), int(seed), ct.byref(self._h)),
               "gfm_pnet_create")
        nt, ne = ct.c_size_t(), ct.c_size_t()
        _check(L.gfm_pnet_param_count(self._h, ct.byref(nt), ct.byref(ne)), "gfm_pnet_param_count")
        self.n_params, self.n_effective = nt.value, ne.value

    @property
    def handle(self):
        return self._h

    def get_params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        _check(_lib().gfm_pnet_get_params(self._h, _fp(p), p.size), "gfm_pnet_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib().gfm_pnet_set_params(self._h, _fp(p), p.size), "gfm_pnet_set_params")

    def forward(self, x):
        """x [B,D,H,W,2] -> y [B,D,H,W]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        B, D, H, W, _ = x.shape
        y = np.zeros((B, D, H, W), dtype=np.float32)
        _check(_lib().gfm_pnet_forward(self._h, B, D, H, W, _fp(x), _fp(y)), "gfm_pnet_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        B, D, H, W, _ = x.shape
        loss = ct.c_double()
        g = np.zeros(self.n_params, dtype=np.float32)
        _check(_lib().gfm_pnet_loss_grad(self._h, B, D, H, W, _fp(x), _fp(t), ct.byref(loss), _fp(g)), "gfm_pnet_loss_grad")
        return loss.value, g

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(_lib().gfm_pnet_adam_step(self._h, lr, beta1, beta2, eps), "gfm_pnet_adam_step")

    def train(self, x, t, epochs, batch, lr=1e-3, seed=0):
        """x [N,D,H,W,2], t [N,D,H,W] -> mean loss per epoch."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        N, D, H, W, _ = x.shape
        out = np.zeros(epochs)
        _check(_lib().gfm_pnet_train(self._h, N, D, H, W, _fp(x), _fp(t), int(epochs), int(batch), float(lr), int(seed),
                                     out.ctypes.data_as(ct.POINTER(ct.c_double))), "gfm_pnet_train")
        return out

    def close(self):
        if self._h:
            _lib().gfm_pnet_destroy(self._h)
            self._h = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def forward_flops(channels, d, h, w, batch=1):
    """Multiply-add FLOPs (2 per MAC) of one forward pass: the ten 3x3x3
    convolutions at their padded channel widths (what the matrix cores
    execute) and the 1x1x1 output layer."""
    c = channels
    v = [batch * (d >> q) * (h >> q) * (w >> q) for q in range(4)]
    convs = [(0, 32, c[0]), (0, c[0], c[0]), (1, c[0], c[1]), (1, c[1], c[1]), (2, c[1], c[2]), (2, c[2], c[2]),
             (3, c[2], c[3]), (2, c[3] + c[2], c[2]), (1, c[2] + c[1], c[1]), (0, c[1] + c[0], c[0])]
    return sum(2 * 27 * ci * co * v[lv] for lv, ci, co in convs) + 2 * c[0] * v[0]


def _main():
    import argparse
    import time
    ap = argparse.ArgumentParser(prog="python -m gfm_amd.pnet",
                                 description="time the forward pass of the p' guess network on one grid")
    ap.add_argument("--grid", type=int, nargs=3, default=(192, 96, 64), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    nx, ny, nz = a.grid
    rng = np.random.default_rng(0)
    x = np.zeros((1, nz, ny, nx, 2), dtype=np.float32)
    x[..., 0] = rng.standard_normal((1, nz, ny, nx))
    x[..., 1] = 1.0
    with PNet(max_batch=1, max_cells=nx * ny * nz) as net:
        net.forward(x)
        t0 = time.perf_counter()
        for _ in range(a.reps):
            net.forward(x)
        dt = (time.perf_counter() - t0) / a.reps
    fl = forward_flops(DEFAULT_CHANNELS, nz, ny, nx)
    print(f"forward {nx}x{ny}x{nz}: {dt * 1e3:.3f} ms per pass including host copies, {fl / 1e9:.1f} GFLOP, "
          f"{fl / dt / 1e12:.1f} TFLOP/s")


if __name__ == "__main__":
    _main()
