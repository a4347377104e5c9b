"""The FP64 denoiser reference (oracle/denoiser_ref.py) pinned by mathematics:
adjoint identities of every linear layer, closed forms of the convolution
taps, and central finite differences of the whole network. CPU only."""
import numpy as np

import denoiser_ref as R

rng = np.random.default_rng(0)


def test_conv_taps_closed_form():
    x = rng.standard_normal((6, 7, 3))
    w = np.zeros((9, 3, 3))
    w[4] = np.eye(3)                                   # centre tap: identity
    assert np.allclose(R.conv(x, w), x)
    w = np.zeros((9, 3, 3))
    w[5] = np.eye(3)                                   # tap (dy=1, dx=2): y[h, w] = x[h, w + 1]
    y = R.conv(x, w)
    assert np.allclose(y[:, :-1], x[:, 1:]) and np.allclose(y[:, -1], 0)
    w = np.zeros((9, 3, 3))
    w[1] = np.eye(3)                                   # tap (0, 1): y[h, w] = x[h - 1, w]
    y = R.conv(x, w)
    assert np.allclose(y[1:], x[:-1]) and np.allclose(y[0], 0)


def test_adjoints_of_linear_layers():
    x = rng.standard_normal((8, 10, 5))
    w = rng.standard_normal((9, 4, 5))
    gy = rng.standard_normal((8, 10, 4))
    lhs = (R.conv(x, w) * gy).sum()
    assert np.isclose(lhs, (x * R.conv_dx(gy, w)).sum())
    assert np.isclose(lhs, (w * R.conv_dw(x, gy)[0]).sum())
    assert np.allclose(R.conv_dw(x, gy)[1], gy.sum((0, 1)))
    g2 = rng.standard_normal((4, 5, 5))
    assert np.isclose((R.pool2(x) * g2).sum(), (x * R.pool2_bwd(g2)).sum())
    c = rng.standard_normal((4, 5, 3))
    gc = rng.standard_normal((8, 10, 8))
    lhs = (R.upcat(c, x) * gc).sum()
    d0, d1 = R.upcat_bwd(gc, 3)
    assert np.isclose(lhs, (c * d0).sum() + (x * d1).sum())


def test_network_gradients_match_finite_differences():
    H, W = 16, 16
    params = {}
    for name in R.layer_names():
        ci = R.CIN if name == "e1_0" else (2 * R.CH if name.startswith("d1") else R.CH)
        co = R.COUT if name == "out" else R.CH
        params[name] = (rng.standard_normal((9, co, ci)) * np.sqrt(2.0 / (9 * ci)), rng.standard_normal(co) * 0.1)
    x = rng.uniform(0, 1, (H, W, R.CIN))
    t = rng.uniform(0.1, 1, (H, W, 4))
    loss, grads, _ = R.loss_and_grads(params, x, t)
    pick = np.random.default_rng(1)
    checked = 0
    for name in ("e1_0", "e2_2", "e1_4", "d1_3", "d2_0", "out"):
        w, b = params[name]
        for _ in range(3):
            idx = tuple(pick.integers(0, s) for s in w.shape)
            if name == "out":
                idx = (idx[0], pick.integers(0, 3), idx[2])     # channels that reach the loss
            h = 1e-6
            w[idx] += h
            lp = R.loss_and_grads(params, x, t)[0]
            w[idx] -= 2 * h
            lm = R.loss_and_grads(params, x, t)[0]
            w[idx] += h
            fd = (lp - lm) / (2 * h)
            an = grads[name][0][idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-9, (name, idx, fd, an)
            checked += 1
        j = int(pick.integers(0, 3 if name == "out" else len(b)))
        b[j] += 1e-6
        lp = R.loss_and_grads(params, x, t)[0]
        b[j] -= 2e-6
        lm = R.loss_and_grads(params, x, t)[0]
        b[j] += 1e-6
        assert abs((lp - lm) / 2e-6 - grads[name][1][j]) <= 1e-6 * max(1.0, abs(grads[name][1][j])) + 1e-9
    assert checked == 18 and loss > 0


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0], np.float64)
    r = R.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == -2.5 and r[4] == 0.0
