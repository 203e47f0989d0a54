"""CPU restatement (numpy, FP64) of the learned p' initial-guess network of
include/gfm_pnet.h: forward pass, loss and hand-written backward pass.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the GPU kernels
(gfm_amd/csrc/gfm_pnet.hip); nothing in the product imports it.

The reference has no such network (it starts the p' solve from zero,
C/Gslv3v.f90:157-161); this file restates the network defined in
include/gfm_pnet.h.  With bf16=True every tensor the GPU stores in BF16 is
rounded here at the same point (round to nearest even from FP32), so the
two differ only by accumulation order; with bf16=False it is the exact FP64
network, whose gradients tests/test_pnet_oracle.py checks by finite
differences.
"""
import numpy as np

NTAP = 27


def round_bf16(a):
    """Round to BF16 (nearest even, from FP32) and return as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(ch):
    """Conv layers (ca, cb, co, up_a) and offsets in the flat parameter vector."""
    c = ch
    spec = [(32, 0, c[0], 0), (c[0], 0, c[0], 0), (c[0], 0, c[1], 0), (c[1], 0, c[1], 0), (c[1], 0, c[2], 0),
            (c[2], 0, c[2], 0), (c[2], 0, c[3], 0), (c[3], c[2], c[2], 1), (c[2], c[1], c[1], 1), (c[1], c[0], c[0], 1)]
    layers, off = [], 0
    for ca, cb, co, up in spec:
        ow = off
        off += NTAP * co * (ca + cb)
        ob = off
        off += co
        layers.append(dict(ca=ca, cb=cb, co=co, up=up, ow=ow, ob=ob))
    return layers, off, off + c[0] + 1


def unpack(params, ch):
    layers, off_out, n = layout(ch)
    assert params.size == n
    Ws, bs = [], []
    for L in layers:
        ci = L["ca"] + L["cb"]
        Ws.append(params[L["ow"]:L["ow"] + NTAP * L["co"] * ci].reshape(NTAP, L["co"], ci).astype(np.float64))
        bs.append(params[L["ob"]:L["ob"] + L["co"]].astype(np.float64))
    wout = params[off_out:off_out + ch[0]].astype(np.float64)
    bout = float(params[off_out + ch[0]])
    return layers, Ws, bs, wout, bout


def _taps():
    for t in range(NTAP):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def conv3(x, W, b):
    """x [B,D,H,W,Ci], W [27,Co,Ci] -> [B,D,H,W,Co] (zero padding), + b."""
    B, D, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, D, H, Wd, W.shape[1]))
    for t, dz, dy, dx in _taps():
        out += xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + Wd, :] @ W[t].T
    return out + (b if b is not None else 0.0)


def conv3_input_grad(g, W):
    """Gradient of conv3 with respect to its input: g [B,D,H,W,Co] -> [.., Ci]."""
    B, D, H, Wd, _ = g.shape
    gp = np.pad(g, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, D, H, Wd, W.shape[2]))
    for t, dz, dy, dx in _taps():
        out += gp[:, 1 - dz:1 - dz + D, 1 - dy:1 - dy + H, 1 - dx:1 - dx + Wd, :] @ W[t]
    return out


def conv3_weight_grad(g, x):
    """dW[t][co][ci] = sum_v g[v][co] x[v+off(t)][ci]."""
    B, D, H, Wd, Ci = x.shape
    Co = g.shape[-1]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros((NTAP, Co, Ci))
    g2 = g.reshape(-1, Co)
    for t, dz, dy, dx in _taps():
        xs = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + Wd, :].reshape(-1, Ci)
        dW[t] = g2.T @ xs
    return dW


def pool(x):
    B, D, H, W, C = x.shape
    return x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def unup(g):
    """Gradient of up(): sum over the 8 children."""
    B, D, H, W, C = g.shape
    return g.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def forward(params, ch, x, bf16=True):
    """x [B,D,H,W,2] -> (y [B,D,H,W], cache)."""
    R = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers, Ws, bs, wout, bout = unpack(params, ch)
    if bf16:
        Ws = [round_bf16(W) for W in Ws]
    relu = lambda a: np.maximum(a, 0.0)
    B, D, H, W, _ = x.shape
    x0 = np.zeros((B, D, H, W, 32))
    x0[..., :2] = R(x)
    a = {}
    a["x0"] = x0
    a["a0"] = R(relu(conv3(x0, Ws[0], bs[0])))
    a["s0"] = R(relu(conv3(a["a0"], Ws[1], bs[1])))
    a["p1"] = R(pool(a["s0"]))
    a["a1"] = R(relu(conv3(a["p1"], Ws[2], bs[2])))
    a["s1"] = R(relu(conv3(a["a1"], Ws[3], bs[3])))
    a["p2"] = R(pool(a["s1"]))
    a["a2"] = R(relu(conv3(a["p2"], Ws[4], bs[4])))
    a["s2"] = R(relu(conv3(a["a2"], Ws[5], bs[5])))
    a["p3"] = R(pool(a["s2"]))
    a["s3"] = R(relu(conv3(a["p3"], Ws[6], bs[6])))
    a["c2"] = np.concatenate([up(a["s3"]), a["s2"]], axis=-1)
    a["d2"] = R(relu(conv3(a["c2"], Ws[7], bs[7])))
    a["c1"] = np.concatenate([up(a["d2"]), a["s1"]], axis=-1)
    a["d1"] = R(relu(conv3(a["c1"], Ws[8], bs[8])))
    a["c0"] = np.concatenate([up(a["d1"]), a["s0"]], axis=-1)
    a["d0"] = R(relu(conv3(a["c0"], Ws[9], bs[9])))
    y = a["d0"] @ wout + bout
    return y, dict(a=a, Ws=Ws, layers=layers, wout=wout)


def loss_grad(params, ch, x, t, bf16=True):
    """Mean squared error over open cells (x[...,1] > 0) and its gradient
    with respect to the flat parameter vector."""
    R = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, cache = forward(params, ch, x, bf16)
    a, Ws, layers, wout = cache["a"], cache["Ws"], cache["layers"], cache["wout"]
    m = (x[..., 1] > 0).astype(np.float64)
    n = m.sum()
    r = m * (y - t)
    loss = float((r * r).sum() / n)
    dy = 2.0 * r / n
    if bf16:
        dy = dy.astype(np.float32).astype(np.float64)
    grads = np.zeros(params.size)
    _, off_out, _ = layout(ch)
    grads[off_out:off_out + ch[0]] = np.einsum("bzyx,bzyxc->c", dy, a["d0"])
    grads[off_out + ch[0]] = dy.sum()

    def put(l, g, xin):
        L = layers[l]
        ci = L["ca"] + L["cb"]
        grads[L["ow"]:L["ow"] + NTAP * L["co"] * ci] = conv3_weight_grad(g, xin).ravel()
        grads[L["ob"]:L["ob"] + L["co"]] = g.reshape(-1, L["co"]).sum(axis=0)

    mask = lambda g, act: g * (act > 0)
    g = R(dy[..., None] * wout)
    g = mask(g, a["d0"])
    put(9, g, a["c0"])
    gi = R(conv3_input_grad(g, Ws[9]))
    gs0 = gi[..., ch[1]:]
    g = R(unup(gi[..., :ch[1]]))
    g = mask(g, a["d1"])
    put(8, g, a["c1"])
    gi = R(conv3_input_grad(g, Ws[8]))
    gs1 = gi[..., ch[2]:]
    g = R(unup(gi[..., :ch[2]]))
    g = mask(g, a["d2"])
    put(7, g, a["c2"])
    gi = R(conv3_input_grad(g, Ws[7]))
    gs2 = gi[..., ch[3]:]
    g = R(unup(gi[..., :ch[3]]))
    g = mask(g, a["s3"])
    put(6, g, a["p3"])
    gp3 = R(conv3_input_grad(g, Ws[6]))
    gs2 = R(gs2 + up(gp3) * 0.125)
    g = mask(gs2, a["s2"])
    put(5, g, a["a2"])
    g = mask(R(conv3_input_grad(g, Ws[5])), a["a2"])
    put(4, g, a["p2"])
    gp2 = R(conv3_input_grad(g, Ws[4]))
    gs1 = R(gs1 + up(gp2) * 0.125)
    g = mask(gs1, a["s1"])
    put(3, g, a["a1"])
    g = mask(R(conv3_input_grad(g, Ws[3])), a["a1"])
    put(2, g, a["p1"])
    gp1 = R(conv3_input_grad(g, Ws[2]))
    gs0 = R(gs0 + up(gp1) * 0.125)
    g = mask(gs0, a["s0"])
    put(1, g, a["a0"])
    g = mask(R(conv3_input_grad(g, Ws[1])), a["a0"])
    put(0, g, a["x0"])
    return loss, grads


def init_params(ch, seed):
    """Random parameters in the flat layout (layer 0's padding weights 0)."""
    rng = np.random.default_rng(seed)
    layers, off_out, n = layout(ch)
    p = np.zeros(n)
    for l, L in enumerate(layers):
        ci = L["ca"] + L["cb"]
        creal = 2 if l == 0 else ci
        W = rng.standard_normal((NTAP, L["co"], ci)) * np.sqrt(2.0 / (NTAP * creal))
        if l == 0:
            W[..., 2:] = 0.0
        p[L["ow"]:L["ow"] + W.size] = W.ravel()
        p[L["ob"]:L["ob"] + L["co"]] = 0.1 * rng.standard_normal(L["co"])
    p[off_out:off_out + ch[0]] = rng.standard_normal(ch[0]) / np.sqrt(ch[0])
    p[off_out + ch[0]] = 0.05
    return p
