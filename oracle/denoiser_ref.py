"""FP64 restatement of the learned denoiser (lighthouse2_amd/denoiser.py).

TEST INFRASTRUCTURE: imported only by tests/ as the checker of the gfx950
denoiser kernels. float64 numpy; `quant` rounds a tensor to bfloat16 (round
to nearest even) at every point where the GPU stores bf16, so the GPU and
this reference differ only by fp32 accumulation order. The backward pass is
checked against central finite differences in tests/test_denoiser_ref.py.
"""
import numpy as np

LEVELS, CH, CIN, COUT = 5, 32, 16, 16


def bf16(a):
    a = np.asarray(a, np.float64).astype(np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _taps(x):
    """(H, W, C) -> (H, W, 9, C): the 3x3 neighbourhood, zero padded, tap = 3 dy + dx."""
    H, W, C = x.shape
    p = np.zeros((H + 2, W + 2, C), x.dtype)
    p[1:-1, 1:-1] = x
    return np.stack([p[dy:dy + H, dx:dx + W] for dy in range(3) for dx in range(3)], 2)


def conv(x, w, b=None):
    """x (H, W, Cin), w [9][Cout][Cin] -> (H, W, Cout)."""
    y = np.einsum("hwtc,toc->hwo", _taps(x), w)
    return y if b is None else y + b


def conv_dx(gy, w):
    """Gradient of conv with respect to x."""
    H, W, _ = gy.shape
    Cin = w.shape[2]
    gx = np.zeros((H + 2, W + 2, Cin))
    contrib = np.einsum("hwo,toc->hwtc", gy, w)
    for t in range(9):
        dy, dx = t // 3, t % 3
        gx[dy:dy + H, dx:dx + W] += contrib[:, :, t]
    return gx[1:-1, 1:-1]


def conv_dw(x, gy):
    """Gradients of conv with respect to w [9][Cout][Cin] and b [Cout]."""
    return np.einsum("hwtc,hwo->toc", _taps(x), gy), gy.sum((0, 1))


def pool2(x):
    H, W, C = x.shape
    return x.reshape(H // 2, 2, W // 2, 2, C).mean((1, 3))


def pool2_bwd(gy):
    return np.repeat(np.repeat(gy, 2, 0), 2, 1) * 0.25


def upcat(coarse, skip):
    return np.concatenate([np.repeat(np.repeat(coarse, 2, 0), 2, 1), skip], 2)


def upcat_bwd(gcat, C0):
    H, W, _ = gcat.shape
    g0 = gcat[..., :C0].reshape(H // 2, 2, W // 2, 2, C0).sum((1, 3))
    return g0, gcat[..., C0:]


def relmse(out, target):
    o, t = out[..., :3], target[..., :3]
    den = t * t + 0.01
    n = o.size
    loss = ((o - t) ** 2 / den).sum() / n
    g = np.zeros_like(out)
    g[..., :3] = 2 * (o - t) / den / n
    return loss, g


def layer_names():
    names = []
    for lvl in range(LEVELS):
        names += [f"e1_{lvl}", f"e2_{lvl}"]
    for lvl in range(LEVELS - 2, -1, -1):
        names += [f"d1_{lvl}", f"d2_{lvl}"]
    return names + ["out"]


def forward(params, x, quant=None):
    """params: {name: (w, b)}; x (H, W, 16). Returns (out, activations)."""
    q = quant or (lambda a: a)
    relu = lambda a: np.maximum(a, 0.0)
    A = {"in_0": q(x)}
    for lvl in range(LEVELS):
        A[f"a_{lvl}"] = q(relu(conv(A[f"in_{lvl}"], *params[f"e1_{lvl}"])))
        A[f"b_{lvl}"] = q(relu(conv(A[f"a_{lvl}"], *params[f"e2_{lvl}"])))
        if lvl < LEVELS - 1:
            A[f"in_{lvl + 1}"] = q(pool2(A[f"b_{lvl}"]))
    prev = A[f"b_{LEVELS - 1}"]
    for lvl in range(LEVELS - 2, -1, -1):
        A[f"cat_{lvl}"] = upcat(prev, A[f"b_{lvl}"])
        A[f"c_{lvl}"] = q(relu(conv(A[f"cat_{lvl}"], *params[f"d1_{lvl}"])))
        A[f"d_{lvl}"] = q(relu(conv(A[f"c_{lvl}"], *params[f"d2_{lvl}"])))
        prev = A[f"d_{lvl}"]
    out = conv(A["d_0"], *params["out"])
    return out, A


def backward(params, A, gout, quant=None):
    """Parameter gradients {name: (gw, gb)} given dLoss/dout."""
    q = quant or (lambda a: a)
    G = {}
    grads = {}
    gout = q(gout)
    grads["out"] = conv_dw(A["d_0"], gout)
    G["d_0"] = q(conv_dx(gout, params["out"][0]))
    for lvl in range(LEVELS - 1):
        g = G[f"d_{lvl}"] * (A[f"d_{lvl}"] > 0)
        grads[f"d2_{lvl}"] = conv_dw(A[f"c_{lvl}"], g)
        G[f"c_{lvl}"] = q(conv_dx(g, params[f"d2_{lvl}"][0]))
        g = G[f"c_{lvl}"] * (A[f"c_{lvl}"] > 0)
        grads[f"d1_{lvl}"] = conv_dw(A[f"cat_{lvl}"], g)
        gcat = q(conv_dx(g, params[f"d1_{lvl}"][0]))
        gc, gs = upcat_bwd(gcat, CH)
        G[f"d_{lvl + 1}" if lvl + 1 < LEVELS - 1 else f"b_{LEVELS - 1}"] = q(gc)
        G[f"b_{lvl}"] = gs
    for lvl in range(LEVELS - 1, -1, -1):
        if lvl < LEVELS - 1:
            G[f"b_{lvl}"] = q(G[f"b_{lvl}"] + pool2_bwd(G[f"in_{lvl + 1}"]))
        g = G[f"b_{lvl}"] * (A[f"b_{lvl}"] > 0)
        grads[f"e2_{lvl}"] = conv_dw(A[f"a_{lvl}"], g)
        G[f"a_{lvl}"] = q(conv_dx(g, params[f"e2_{lvl}"][0]))
        g = G[f"a_{lvl}"] * (A[f"a_{lvl}"] > 0)
        grads[f"e1_{lvl}"] = conv_dw(A[f"in_{lvl}"], g)
        if lvl > 0:
            G[f"in_{lvl}"] = q(conv_dx(g, params[f"e1_{lvl}"][0]))
    return grads


def loss_and_grads(params, x, target, quant=None):
    out, A = forward(params, x, quant)
    loss, gout = relmse(out, target)
    return loss, backward(params, A, gout, quant), out
