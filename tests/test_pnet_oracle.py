"""Pins the CPU restatement of the p' guess network (oracle/pnet_oracle.py)
before it is used to check the GPU kernels: the convolution against a
brute-force loop, its input and weight gradients as exact adjoints, and the
whole backward pass against central finite differences of the FP64
network."""
import numpy as np
import pytest

import pnet_oracle as po

CH = (32, 32, 32, 32)


def test_conv3_matches_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 3, 4, 5, 2))
    W = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    got = po.conv3(x, W, b)
    ref = np.zeros((1, 3, 4, 5, 3))
    for z in range(3):
        for y in range(4):
            for xx in range(5):
                for t in range(27):
                    dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                    zz, yy, x2 = z + dz, y + dy, xx + dx
                    if 0 <= zz < 3 and 0 <= yy < 4 and 0 <= x2 < 5:
                        ref[0, z, y, xx] += W[t] @ x[0, zz, yy, x2]
                ref[0, z, y, xx] += b
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13)


def test_conv3_gradients_are_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 3, 5, 3))
    W = rng.standard_normal((27, 4, 3))
    g = rng.standard_normal((2, 4, 3, 5, 4))
    lhs = np.sum(po.conv3(x, W, None) * g)
    assert lhs == pytest.approx(np.sum(x * po.conv3_input_grad(g, W)), rel=1e-12)
    assert lhs == pytest.approx(np.sum(W * po.conv3_weight_grad(g, x)), rel=1e-12)


def test_pool_up_adjoint():
    rng = np.random.default_rng(2)
    f = rng.standard_normal((1, 4, 4, 6, 3))
    c = rng.standard_normal((1, 2, 2, 3, 3))
    assert np.sum(po.up(c) * f) == pytest.approx(np.sum(c * po.unup(f)), rel=1e-12)
    assert np.sum(po.pool(f) * c) == pytest.approx(np.sum(f * po.up(c) * 0.125), rel=1e-12)


def test_round_bf16():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.140625, 1e-3])
    r = po.round_bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie -> even
    assert r[2] == 1.0 + 2 ** -7  # above the tie -> up
    assert r[3] == -3.140625      # representable
    assert abs(r[4] - 1e-3) <= 2 ** -9 * 1e-3


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = po.init_params(CH, seed=4)
    x = np.zeros((1, 8, 8, 8, 2))
    x[..., 0] = rng.standard_normal((1, 8, 8, 8))
    x[..., 1] = (rng.uniform(size=(1, 8, 8, 8)) > 0.1)
    t = rng.standard_normal((1, 8, 8, 8))
    loss, g = po.loss_grad(p, CH, x, t, bf16=False)
    layers, off_out, n = po.layout(CH)
    idx = []
    for l, L in enumerate(layers):
        ci = L["ca"] + L["cb"]
        creal = 2 if l == 0 else ci
        for _ in range(3):
            tp, co, c = rng.integers(27), rng.integers(L["co"]), rng.integers(creal)
            idx.append(L["ow"] + (tp * L["co"] + co) * ci + c)
        idx.append(L["ob"] + rng.integers(L["co"]))
    idx += [off_out, off_out + CH[0]]
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (po.loss_grad(pp, CH, x, t, bf16=False)[0] - po.loss_grad(pm, CH, x, t, bf16=False)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=2e-5, abs=1e-9), i
