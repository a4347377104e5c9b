"""Learned denoiser on the gfx950 matrix cores against the
FP64 reference (oracle/denoiser_ref.py) with bf16 rounding emulated where
the kernels store bf16. Tolerances: fp32-accumulated outputs within 1e-5 of
the output scale; bf16-stored outputs within 1 bf16 ulp (2^-7 relative);
whole-network gradients within 3e-2 relative L2 per layer (bf16 rounding
flips of intermediate activations propagate)."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import denoiser as D
import denoiser_ref as R

pytestmark = pytest.mark.gpu


def dev_bf16(torch, a):
    return torch.from_numpy(D.bf16_bits(a).reshape(-1).copy()).cuda()


def host_f32(t):
    return t.cpu().numpy().astype(np.float32)


@pytest.mark.parametrize("cin,cout", [(16, 32), (32, 32), (64, 32), (32, 16), (16, 16)])
def test_conv_forward_dx_dw_match_fp64(torch_cuda, cin, cout):
    torch = torch_cuda
    L = lh.library()
    rng = np.random.default_rng(cin * 7 + cout)
    H, W = 20, 70                                   # ragged: W not a multiple of the 64-pixel tile
    x = R.bf16(rng.standard_normal((H, W, cin)))
    w = R.bf16(rng.standard_normal((9, cout, cin)) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    gy = R.bf16(rng.standard_normal((H, W, cout)))
    dx_, dw_ = dev_bf16(torch, x), dev_bf16(torch, w)
    db_ = torch.from_numpy(b).cuda()
    sh = torch.cuda.current_stream().cuda_stream
    y32 = torch.zeros(H * W * cout, dtype=torch.float32, device="cuda")
    assert L.lh2hip_dn_conv3x3(dx_.data_ptr(), cin, dw_.data_ptr(), db_.data_ptr(), y32.data_ptr(), cout, H, W, 2, sh) == 0
    yb = torch.zeros(H * W * cout, dtype=torch.int16, device="cuda")
    assert L.lh2hip_dn_conv3x3(dx_.data_ptr(), cin, dw_.data_ptr(), db_.data_ptr(), yb.data_ptr(), cout, H, W, 1, sh) == 0
    torch.cuda.synchronize()
    ref = R.conv(x, w, b)
    got = host_f32(y32).reshape(H, W, cout)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    gotb = D.bf16_to_f32(yb.cpu().numpy()).reshape(H, W, cout)
    refb = R.bf16(np.maximum(ref, 0))
    assert np.all(np.abs(gotb - refb) <= np.abs(refb) * 2 ** -7 + 1e-30)
    # input gradient through the flipped weight layout written by the Adam kernel
    wf = torch.from_numpy(w.astype(np.float32).reshape(-1).copy()).cuda()
    z = torch.zeros_like(wf)
    wfwd = torch.zeros(wf.numel(), dtype=torch.int16, device="cuda")
    wbwd = torch.zeros(wf.numel(), dtype=torch.int16, device="cuda")
    assert L.lh2hip_dn_adam(wf.data_ptr(), z.data_ptr(), z.data_ptr(), None, wf.numel(), 0.0, 0.9, 0.999, 1e-8, 1,
                            wfwd.data_ptr(), wbwd.data_ptr(), cout, cin, sh) == 0
    if cin <= 64:
        gyb = dev_bf16(torch, gy)
        gx = torch.zeros(H * W * cin, dtype=torch.float32, device="cuda")
        assert L.lh2hip_dn_conv3x3(gyb.data_ptr(), cout, wbwd.data_ptr(), None, gx.data_ptr(), cin, H, W, 2, sh) == 0
        torch.cuda.synchronize()
        rgx = R.conv_dx(gy, w)
        assert np.max(np.abs(host_f32(gx).reshape(H, W, cin) - rgx)) <= 1e-5 * np.max(np.abs(rgx))
    ws = torch.zeros(int(L.lh2hip_dn_dw_workspace_bytes(cin, cout)) // 4, dtype=torch.float32, device="cuda")
    gw = torch.zeros(9 * cout * cin, dtype=torch.float32, device="cuda")
    gb = torch.zeros(cout, dtype=torch.float32, device="cuda")
    gyb = dev_bf16(torch, gy)
    assert L.lh2hip_dn_conv3x3_dw(dx_.data_ptr(), cin, gyb.data_ptr(), cout, H, W, gw.data_ptr(), gb.data_ptr(), ws.data_ptr(), sh) == 0
    torch.cuda.synchronize()
    rgw, rgb = R.conv_dw(x, gy)
    assert np.max(np.abs(host_f32(gw).reshape(9, cout, cin) - rgw)) <= 1e-5 * np.max(np.abs(rgw))
    assert np.max(np.abs(host_f32(gb) - rgb)) <= 1e-5 * np.max(np.abs(rgb))


def test_network_forward_and_gradients_match_fp64(torch_cuda):
    torch = torch_cuda
    H, W = 32, 48
    params = D.init_params(seed=3)
    rng = np.random.default_rng(4)
    for k, (w, b) in params.items():               # non-zero biases exercise the bias paths
        params[k] = (w, (rng.standard_normal(len(b)) * 0.05).astype(np.float32) * (0 if k == "out" else 1))
    net = D.UNetDenoiser(H, W, params=params)
    x = np.zeros((H, W, 16))
    x[..., :10] = rng.uniform(0, 1, (H, W, 10))
    t = rng.uniform(0.05, 1.0, (H, W, 4)).astype(np.float32)
    net.set_input_packed(dev_bf16(torch, x))
    out = net.forward()
    loss = net.loss_backward(torch.from_numpy(t).cuda())
    torch.cuda.synchronize()
    rp = {k: (R.bf16(w.astype(np.float64)), b.astype(np.float64)) for k, (w, b) in params.items()}
    rloss, rgrads, rout = R.loss_and_grads(rp, x, t.astype(np.float64), quant=R.bf16)
    o = out.cpu().numpy()[..., :3]
    assert np.max(np.abs(o - rout[..., :3])) <= 2e-2 * np.max(np.abs(rout[..., :3]))
    assert abs(float(loss.item()) - rloss) <= 1e-2 * rloss
    g = net.grads_numpy()
    for name, (rgw, rgb) in rgrads.items():
        gw = g[name][0].reshape(rgw.shape)
        err = np.linalg.norm(gw - rgw) / max(np.linalg.norm(rgw), 1e-30)
        assert err < 3e-2, (name, err)


def test_training_reduces_relmse(torch_cuda):
    """Denoise a noisy synthetic image (smooth radiance x albedo, features
    clean): 150 Adam steps must cut the relMSE loss by more than half."""
    torch = torch_cuda
    H, W = 64, 64
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:H, 0:W] / 64.0
    clean = np.stack([0.5 + 0.4 * np.sin(4 * xx), 0.5 + 0.4 * np.cos(3 * yy), 0.5 + 0.3 * np.sin(2 * (xx + yy)), np.ones_like(xx)], -1)
    noisy = clean * rng.gamma(2.0, 0.5, (H, W, 1))
    albedo = np.concatenate([clean[..., :3], np.ones((H, W, 1))], -1)
    normal = np.zeros((H, W, 4))
    depth = np.ones((H, W))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    net = D.UNetDenoiser(H, W, seed=1)
    net.set_input(dev(noisy), dev(albedo), dev(normal), dev(depth))
    tgt = dev(clean)
    losses = []
    for i in range(150):
        net.forward()
        losses.append(float(net.loss_backward(tgt).item()))
        net.step(lr=2e-3)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])


def test_conv_of_upsampled_concatenation_matches_fp64(torch_cuda):
    """The decoder's first convolution read straight from the coarse level and
    the skip connection equals the convolution of their concatenation."""
    torch = torch_cuda
    L = lh.library()
    rng = np.random.default_rng(21)
    H, W = 22, 70                                   # even, ragged against the 4 x 64 tile
    coarse = R.bf16(rng.standard_normal((H // 2, W // 2, 32)))
    skip = R.bf16(rng.standard_normal((H, W, 32)))
    w = R.bf16(rng.standard_normal((9, 32, 64)) * 0.2)
    b = rng.standard_normal(32).astype(np.float32)
    sh = torch.cuda.current_stream().cuda_stream
    dc, ds, dw_ = dev_bf16(torch, coarse), dev_bf16(torch, skip), dev_bf16(torch, w)
    db_ = torch.from_numpy(b).cuda()
    y = torch.zeros(H * W * 32, dtype=torch.float32, device="cuda")
    assert L.lh2hip_dn_conv3x3_cat(dc.data_ptr(), ds.data_ptr(), dw_.data_ptr(), db_.data_ptr(), y.data_ptr(), 32, H, W, 2, sh) == 0
    cat = torch.zeros(H * W * 64, dtype=torch.int16, device="cuda")
    assert L.lh2hip_dn_upcat(dc.data_ptr(), 32, ds.data_ptr(), 32, H, W, cat.data_ptr(), sh) == 0
    y2 = torch.zeros_like(y)
    assert L.lh2hip_dn_conv3x3(cat.data_ptr(), 64, dw_.data_ptr(), db_.data_ptr(), y2.data_ptr(), 32, H, W, 2, sh) == 0
    torch.cuda.synchronize()
    ref = R.conv(R.upcat(coarse, skip), w, b)
    got = host_f32(y).reshape(H, W, 32)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    assert np.array_equal(host_f32(y), host_f32(y2))
    assert L.lh2hip_dn_conv3x3_cat(dc.data_ptr(), ds.data_ptr(), dw_.data_ptr(), None, y.data_ptr(), 32, H + 1, W, 2, sh) != 0


def _conv_rows(x, w, b, rows=256):
    """FP64 conv3x3 (as denoiser_ref.conv) computed in row blocks through matmul."""
    H, W, C = x.shape
    wm = w.transpose(0, 2, 1).reshape(9 * C, -1)
    p = np.zeros((H + 2, W + 2, C))
    p[1:-1, 1:-1] = x
    out = np.empty((H, W, wm.shape[1]))
    for r in range(0, H, rows):
        h = min(rows, H - r)
        t = np.stack([p[r + dy:r + dy + h, dx:dx + W] for dy in range(3) for dx in range(3)], 2)
        out[r:r + h] = (t.reshape(h * W, 9 * C) @ wm).reshape(h, W, -1)
    return out + b


@pytest.mark.parametrize("cin,cout,cat", [(16, 32, False), (32, 32, False), (32, 16, False), (32, 64, False), (64, 32, True)])
def test_tall_image_walks_several_bands_per_block(torch_cuda, cin, cout, cat):
    """An image tall enough that each block walks several 4-row bands down its
    strip (rows fetched ahead into a ring), with ragged last strip and band."""
    torch = torch_cuda
    L = lh.library()
    rng = np.random.default_rng(cin + cout)
    H, W = 4098, 70
    w = R.bf16(rng.standard_normal((9, cout, cin)) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    sh = torch.cuda.current_stream().cuda_stream
    dw_, db_ = dev_bf16(torch, w), torch.from_numpy(b).cuda()
    y = torch.zeros(H * W * cout, dtype=torch.float32, device="cuda")
    if cat:
        coarse = R.bf16(rng.standard_normal((H // 2, W // 2, 32)))
        skip = R.bf16(rng.standard_normal((H, W, 32)))
        dc, ds = dev_bf16(torch, coarse), dev_bf16(torch, skip)
        assert L.lh2hip_dn_conv3x3_cat(dc.data_ptr(), ds.data_ptr(), dw_.data_ptr(), db_.data_ptr(), y.data_ptr(), cout, H, W, 2, sh) == 0
        x = R.upcat(coarse, skip)
    else:
        x = R.bf16(rng.standard_normal((H, W, cin)))
        dx_ = dev_bf16(torch, x)
        assert L.lh2hip_dn_conv3x3(dx_.data_ptr(), cin, dw_.data_ptr(), db_.data_ptr(), y.data_ptr(), cout, H, W, 2, sh) == 0
    torch.cuda.synchronize()
    ref = _conv_rows(x, w, b)
    got = host_f32(y).reshape(H, W, cout)
    err = np.abs(got - ref)
    assert err.max() <= 1e-5 * np.abs(ref).max(), np.unravel_index(err.argmax(), err.shape)


@pytest.mark.parametrize("cin,H", [(16, 22), (32, 22), (16, 2050), (32, 2050)])
def test_conv_with_fused_pool_equals_conv_then_pool(torch_cuda, cin, H):
    """The encoder convolution that also writes the 2x2 average of its output
    gives exactly dn_conv3x3 (relu, bf16) followed by dn_pool2, with 16-pixel
    tiles (22 rows: a small image) and 64-pixel tiles (2050 rows)."""
    torch = torch_cuda
    L = lh.library()
    rng = np.random.default_rng(40 + cin)
    W = 70                                          # the last 4-row tile has only 2 rows; ragged strip
    x = R.bf16(rng.standard_normal((H, W, cin)))
    w = R.bf16(rng.standard_normal((9, 32, cin)) * 0.3)
    b = (rng.standard_normal(32) * 0.1).astype(np.float32)
    sh = torch.cuda.current_stream().cuda_stream
    dx_, dw_, db_ = dev_bf16(torch, x), dev_bf16(torch, w), torch.from_numpy(b).cuda()
    y = torch.zeros(H * W * 32, dtype=torch.int16, device="cuda")
    yp = torch.zeros(H * W * 8, dtype=torch.int16, device="cuda")
    assert L.lh2hip_dn_conv3x3_pool(dx_.data_ptr(), cin, dw_.data_ptr(), db_.data_ptr(), y.data_ptr(), 32, H, W, yp.data_ptr(), sh) == 0
    y2 = torch.zeros_like(y)
    yp2 = torch.zeros_like(yp)
    assert L.lh2hip_dn_conv3x3(dx_.data_ptr(), cin, dw_.data_ptr(), db_.data_ptr(), y2.data_ptr(), 32, H, W, 1, sh) == 0
    assert L.lh2hip_dn_pool2(y2.data_ptr(), 32, H, W, yp2.data_ptr(), sh) == 0
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), y2.cpu().numpy())
    assert np.array_equal(yp.cpu().numpy(), yp2.cpu().numpy())
    if H > 100:
        return                                      # the unfused pair is checked against FP64 at this height elsewhere
    ref = R.pool2(R.bf16(np.maximum(R.conv(x, w, b), 0)))
    got = D.bf16_to_f32(yp.cpu().numpy()).reshape(H // 2, W // 2, 32)
    assert np.all(np.abs(got - R.bf16(ref)) <= np.abs(R.bf16(ref)) * 2 ** -7 + 1e-30)
    assert L.lh2hip_dn_conv3x3_pool(dx_.data_ptr(), cin, dw_.data_ptr(), None, y.data_ptr(), 32, H - 1, W, yp.data_ptr(), sh) != 0


@pytest.mark.parametrize("w,h", [(70, 40), (1920, 1080)])
def test_fused_display_equals_unfused(torch_cuda, w, h):
    """Inference with the output convolution writing the display image
    (lh2hip_dn_conv3x3_display) gives the image of set_input_frame + forward
    + output_frame bit for bit: a frame inside a padded image with 16-pixel
    tiles (70 x 40 in 48 x 80) and with 64-pixel, 8-row tiles (1080p in
    1088 x 1920)."""
    torch = torch_cuda
    rng = np.random.default_rng(w + h)
    Hp, Wp = (h + 15) // 16 * 16, (w + 15) // 16 * 16
    net = D.UNetDenoiser(Hp, Wp, seed=5)
    acc = torch.from_numpy(rng.uniform(0, 2, (2, h, w, 4)).astype(np.float32)).cuda()
    feat = rng.integers(0, 2 ** 32, (h, w, 4), dtype=np.uint64).astype(np.uint32)
    feat[..., 2] = rng.uniform(0.1, 50, (h, w)).astype(np.float32).view(np.uint32)   # hit distance
    feat[0, 0, 2] = np.float32(np.inf).view(np.uint32)                                 # a miss
    f = torch.from_numpy(feat.view(np.int32)).cuda()
    ref = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    net.set_input_frame(acc, f, 4, w, h)
    net.forward()
    net.output_frame(f, w, h, ref)
    got = torch.full((h, w, 4), -1.0, dtype=torch.float32, device="cuda")
    net.denoise_frame(acc, f, 4, w, h, got)
    torch.cuda.synchronize()
    r, g = ref.cpu().numpy(), got.cpu().numpy()
    assert np.isfinite(r).all() and (r[..., :3] > 0).any()
    assert np.array_equal(r.view(np.uint32), g.view(np.uint32))


def test_elementwise_kernels_vector_and_scalar_paths_agree(torch_cuda):
    """relu_bwd, pool2, pool2_bwd (= and +=), upcat and upcat_bwd on 16-byte
    aligned tensors (8 channels per thread) and on tensors one element off
    alignment (the per-element kernels): identical bits."""
    torch = torch_cuda
    L = lh.library()
    rng = np.random.default_rng(77)
    H, W, C = 22, 70, 32
    sh = torch.cuda.current_stream().cuda_stream

    def dev(a, off):
        """bf16 bits of a on the device starting `off` elements into a buffer."""
        bits = torch.from_numpy(D.bf16_bits(a).reshape(-1).copy())
        buf = torch.zeros(bits.numel() + 8, dtype=torch.int16, device="cuda")
        buf[off:off + bits.numel()] = bits.cuda()
        return buf[off:off + bits.numel()]

    def run(off):
        x = dev(rng_state["x"], off)
        g = dev(rng_state["g"], off)
        co = dev(rng_state["co"], off)
        out = {}
        y = dev(np.zeros((H, W, C), np.float32), off)
        assert L.lh2hip_dn_relu_bwd(x.data_ptr(), g.data_ptr(), y.data_ptr(), H * W * C, sh) == 0
        out["relu"] = y
        p = dev(np.zeros((H // 2, W // 2, C), np.float32), off)
        assert L.lh2hip_dn_pool2(x.data_ptr(), C, H, W, p.data_ptr(), sh) == 0
        out["pool"] = p
        dxs = dev(np.zeros((H, W, C), np.float32), off)
        assert L.lh2hip_dn_pool2_bwd(co.data_ptr(), C, H, W, dxs.data_ptr(), 0, sh) == 0
        out["pool_bwd"] = dxs.clone()
        assert L.lh2hip_dn_pool2_bwd(co.data_ptr(), C, H, W, dxs.data_ptr(), 1, sh) == 0
        out["pool_bwd_acc"] = dxs
        cat = dev(np.zeros((H, W, 2 * C), np.float32), off)
        assert L.lh2hip_dn_upcat(co.data_ptr(), C, x.data_ptr(), C, H, W, cat.data_ptr(), sh) == 0
        out["upcat"] = cat
        dco = dev(np.zeros((H // 2, W // 2, C), np.float32), off)
        dsk = dev(np.zeros((H, W, C), np.float32), off)
        assert L.lh2hip_dn_upcat_bwd(cat.data_ptr(), C, C, H, W, dco.data_ptr(), dsk.data_ptr(), sh) == 0
        out["upcat_bwd_coarse"], out["upcat_bwd_skip"] = dco, dsk
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in out.items()}

    rng_state = {"x": R.bf16(rng.standard_normal((H, W, C))), "g": R.bf16(rng.standard_normal((H, W, C))),
                 "co": R.bf16(rng.standard_normal((H // 2, W // 2, C)))}
    vec, sca = run(0), run(1)
    for k in vec:
        assert np.array_equal(vec[k], sca[k]), k
    # the scalar path itself against the definitions
    x, g = rng_state["x"], rng_state["g"]
    assert np.array_equal(D.bf16_to_f32(vec["relu"]).reshape(H, W, C), np.where(x > 0, g, 0).astype(np.float32))
    assert np.array_equal(D.bf16_to_f32(vec["upcat_bwd_skip"]).reshape(H, W, C), x)


@pytest.mark.parametrize("cin,cout", [(16, 32), (32, 32), (32, 16)])
def test_large_tiles_equal_fp64_and_small_tiles(torch_cuda, cin, cout):
    """1930 x 202 (ragged in both directions) runs the 8 x 64 tiles with the
    weights fetched two taps ahead (32 input channels). Checked against FP64
    dot products on sampled pixels, and bit for bit against crops of the same
    input run through the 4 x 16 tiles (same tap and channel order of the
    products, so the same fp32 sums)."""
    torch = torch_cuda
    L = lh.library()
    rng = np.random.default_rng(cin + 3 * cout)
    H, W = 202, 1930
    x = R.bf16(rng.standard_normal((H, W, cin)))
    w = R.bf16(rng.standard_normal((9, cout, cin)) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    sh = torch.cuda.current_stream().cuda_stream
    dw_, db_ = dev_bf16(torch, w), torch.from_numpy(b).cuda()

    def run(xa, flags):
        h, wd = xa.shape[:2]
        xd = dev_bf16(torch, xa)
        y = torch.zeros(h * wd * cout, dtype=torch.float32 if flags & 2 else torch.int16, device="cuda")
        assert L.lh2hip_dn_conv3x3(xd.data_ptr(), cin, dw_.data_ptr(), db_.data_ptr(), y.data_ptr(), cout, h, wd, flags, sh) == 0
        torch.cuda.synchronize()
        y = y.cpu().numpy()
        return (y if flags & 2 else D.bf16_to_f32(y)).reshape(h, wd, cout)

    full = run(x, 2)                                     # fp32 out, no ReLU
    py, px = rng.integers(0, H, 3000), rng.integers(0, W, 3000)
    xp = np.zeros((H + 2, W + 2, cin))
    xp[1:-1, 1:-1] = x
    ref = np.zeros((3000, cout)) + b
    for t in range(9):
        ref += xp[py + t // 3, px + t % 3] @ w[t].T
    assert np.max(np.abs(full[py, px] - ref)) <= 1e-5 * np.max(np.abs(ref))
    fullb = run(x, 1)                                    # bf16 out with ReLU
    for (r0, r1, c0, c1) in ((0, 40, 0, 60), (150, 202, 1868, 1930), (91, 130, 700, 760)):
        a0, a1, b0, b1 = max(r0 - 1, 0), min(r1 + 1, H), max(c0 - 1, 0), min(c1 + 1, W)
        crop = x[a0:a1, b0:b1]
        for flags, whole in ((2, full), (1, fullb)):
            part = run(crop, flags)
            assert np.array_equal(part[r0 - a0:r1 - a0, c0 - b0:c1 - b0].view(np.uint32),
                                  np.ascontiguousarray(whole[r0:r1, c0:c1]).view(np.uint32))
