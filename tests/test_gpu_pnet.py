"""The p' guess network's BF16 matrix-core kernels (forward, input and
weight gradients, Adam) against the CPU restatement oracle/pnet_oracle.py,
which rounds to BF16 at the same points, so the two differ only by
accumulation order (FP32 on the GPU, FP64 on the CPU) and the rare BF16
rounding it flips.  Tolerances: forward output within 5e-3 relative L2,
loss within 1e-3, each layer's weight and bias gradient within 1e-2
relative L2.  Each test runs every voxel tile of the convolution
(GFM_PCONV_TILE: 4, 2 or 1 tiles of 16 voxels per wavefront, or one tile
with the steps split over four wavefronts) as well as the choice made by
grid size, which on these small grids is the split one."""
import numpy as np
import pytest

import pnet_oracle as po
from gfm_amd.pnet import PNet

pytestmark = pytest.mark.gpu

CH = (32, 32, 64, 64)


def _batch(rng, B, D, H, W):
    x = np.zeros((B, D, H, W, 2), dtype=np.float32)
    x[..., 0] = rng.standard_normal((B, D, H, W))
    x[..., 1] = rng.uniform(size=(B, D, H, W)) > 0.15
    t = rng.standard_normal((B, D, H, W)).astype(np.float32)
    return x, t


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


TILES = ["auto", "4", "2", "1", "1x4"]


def _forward_case(shape, seed):
    rng = np.random.default_rng(seed)
    p = po.init_params(CH, seed=seed + 1).astype(np.float32)
    x, _ = _batch(rng, *shape)
    yo, _ = po.forward(p.astype(np.float64), CH, x.astype(np.float64), bf16=True)
    return p, x, yo


@pytest.mark.parametrize("w", [24, 32])
def test_forward_matches_oracle(monkeypatch, w):
    """W = 24: no tile width divides a row, every tap is gathered; W = 32:
    the x-taps of tiles 2, 1 and 1x4 come from lane shifts of one row."""
    p, x, yo = _forward_case((2, 8, 16, w), 0)
    for tile in TILES:
        monkeypatch.setenv("GFM_PCONV_TILE", tile)
        with PNet(CH, max_batch=2, max_cells=8 * 16 * w) as net:
            net.set_params(p)
            assert np.array_equal(net.get_params(), p)
            y = net.forward(x)
        err = _rel(y, yo)
        print(f"tile {tile}: forward rel err {err:.2e}", flush=True)
        assert err < 5e-3, tile


def test_forward_matches_oracle_large_grid():
    """A 64x64x64 grid, where the choice by grid size takes 4 tiles per
    wavefront on the finest level and the split tile on the coarser ones."""
    p, x, yo = _forward_case((1, 64, 64, 64), 7)
    with PNet(CH, max_batch=1, max_cells=64 ** 3) as net:
        net.set_params(p)
        y = net.forward(x)
    err = _rel(y, yo)
    print("forward 64^3 rel err", err, flush=True)
    assert err < 5e-3


def test_unknown_tile_is_rejected(monkeypatch):
    monkeypatch.setenv("GFM_PCONV_TILE", "3")
    x = np.zeros((1, 8, 8, 8, 2), dtype=np.float32)
    with PNet(CH, max_batch=1, max_cells=512) as net:
        with pytest.raises(RuntimeError, match="GFM_PCONV_TILE"):
            net.forward(x)


@pytest.mark.parametrize("w", [16, 64])
def test_gradients_match_oracle(monkeypatch, w):
    """W = 16: lane-shifted x-taps for tiles 1 and 1x4 only; W = 64: for every
    tile, the 96- and 128-channel input gradients included."""
    rng = np.random.default_rng(2)
    p = po.init_params(CH, seed=3).astype(np.float32)
    x, t = _batch(rng, 2, 8, 8, w)
    lo, go = po.loss_grad(p.astype(np.float64), CH, x.astype(np.float64), t.astype(np.float64), bf16=True)
    layers, off_out, n = po.layout(CH)
    for tile in TILES:
        monkeypatch.setenv("GFM_PCONV_TILE", tile)
        with PNet(CH, max_batch=2, max_cells=8 * 8 * w) as net:
            net.set_params(p)
            loss, g = net.loss_grad(x, t)
        print(f"tile {tile}: loss", loss, lo, flush=True)
        assert loss == pytest.approx(lo, rel=1e-3)
        for l, L in enumerate(layers):
            ci = L["ca"] + L["cb"]
            sw = slice(L["ow"], L["ow"] + 27 * L["co"] * ci)
            sb = slice(L["ob"], L["ob"] + L["co"])
            ew, eb = _rel(g[sw], go[sw]), _rel(g[sb], go[sb])
            print(f"  layer {l}: weight grad rel err {ew:.2e}, bias {eb:.2e}", flush=True)
            assert ew < 1e-2 and eb < 1e-2, (tile, l)
        so = slice(off_out, n)
        assert _rel(g[so], go[so]) < 1e-2, tile


def test_training_reduces_loss():
    """Learn a fixed smoothing of the input (a 7-point average applied
    twice): the mean loss of the last epoch is below a quarter of the
    first's."""
    rng = np.random.default_rng(4)
    N, D, H, W = 16, 8, 16, 16
    x = np.zeros((N, D, H, W, 2), dtype=np.float32)
    x[..., 0] = rng.standard_normal((N, D, H, W))
    x[..., 1] = 1.0
    s = x[..., 0].astype(np.float64)
    for _ in range(2):
        sp = np.pad(s, ((0, 0), (1, 1), (1, 1), (1, 1)))
        s = (sp[:, 1:-1, 1:-1, 1:-1] + sp[:, :-2, 1:-1, 1:-1] + sp[:, 2:, 1:-1, 1:-1] + sp[:, 1:-1, :-2, 1:-1]
             + sp[:, 1:-1, 2:, 1:-1] + sp[:, 1:-1, 1:-1, :-2] + sp[:, 1:-1, 1:-1, 2:]) / 7.0
    t = (s / s.std()).astype(np.float32)
    with PNet(CH, max_batch=4, max_cells=D * H * W, seed=5) as net:
        hist = net.train(x, t, epochs=12, batch=4, lr=2e-3, seed=6)
    print("epoch losses", hist, flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.25 * hist[0]
