"""The allocation-free weight gradient (lh2hip_dn_conv3x3_dw_noalloc, no
counterpart in the reference) against lh2hip_dn_conv3x3_dw with a caller
workspace, through the C ABI: two calls in a row, the second on a smaller
input, each bit-identical to the workspace entry point (same block
partition and reduction order) and within 1e-5 of the FP64 reference; the
shapes beyond the built-in array are rejected."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import denoiser as D
import denoiser_ref as R

pytestmark = pytest.mark.gpu


def _dev_bf16(torch, a):
    return torch.from_numpy(D.bf16_bits(a).reshape(-1).copy()).cuda()


def _grads(torch, L, cin, cout, H, W, seed):
    rng = np.random.default_rng(seed)
    x = R.bf16(rng.standard_normal((H, W, cin)))
    gy = R.bf16(rng.standard_normal((H, W, cout)))
    dx, dgy = _dev_bf16(torch, x), _dev_bf16(torch, gy)
    sh = torch.cuda.current_stream().cuda_stream
    ws = torch.zeros(int(L.lh2hip_dn_dw_workspace_bytes(cin, cout)) // 4, dtype=torch.float32, device="cuda")
    gw = torch.full((9 * cout * cin,), 7.0, dtype=torch.float32, device="cuda")
    gb = torch.full((cout,), 7.0, dtype=torch.float32, device="cuda")
    assert L.lh2hip_dn_conv3x3_dw(dx.data_ptr(), cin, dgy.data_ptr(), cout, H, W, gw.data_ptr(), gb.data_ptr(), ws.data_ptr(), sh) == 0
    nw = torch.full_like(gw, -3.0)
    nb = torch.full_like(gb, -3.0)
    assert L.lh2hip_dn_conv3x3_dw_noalloc(dx.data_ptr(), cin, dgy.data_ptr(), cout, H, W, nw.data_ptr(), nb.data_ptr(), sh) == 0
    torch.cuda.synchronize()
    return x, gy, gw.cpu().numpy(), gb.cpu().numpy(), nw.cpu().numpy(), nb.cpu().numpy()


def test_noalloc_weight_gradient_equals_workspace_entry_point(torch_cuda):
    torch = torch_cuda
    L = lh.library()
    assert L.lh2hip_dn_dw_scratch_bytes() >= max(L.lh2hip_dn_dw_workspace_bytes(ci, co)
                                                 for ci, co in [(16, 32), (32, 32), (64, 32), (32, 16), (16, 16)])
    # first call: the largest channel shape on a ragged 200-pixel-wide image; second: a smaller one
    for cin, cout, H, W, seed in [(64, 32, 96, 200, 11), (32, 16, 20, 70, 12)]:
        print(f"noalloc dw: Cin {cin} Cout {cout} {H}x{W} seed {seed}", flush=True)
        x, gy, gw, gb, nw, nb = _grads(torch, L, cin, cout, H, W, seed)
        assert np.array_equal(nw.view(np.uint32), gw.view(np.uint32)), "dw differs from the workspace entry point"
        assert np.array_equal(nb.view(np.uint32), gb.view(np.uint32)), "db differs from the workspace entry point"
        rgw, rgb = R.conv_dw(x, gy)
        assert np.max(np.abs(nw.reshape(9, cout, cin) - rgw)) <= 1e-5 * np.max(np.abs(rgw))
        assert np.max(np.abs(nb - rgb)) <= 1e-5 * np.max(np.abs(rgb))


def test_noalloc_rejects_unsupported_shapes(torch_cuda):
    torch = torch_cuda
    L = lh.library()
    x = torch.zeros(64 * 64 * 48, dtype=torch.int16, device="cuda")
    dw = torch.zeros(9 * 48 * 48, dtype=torch.float32, device="cuda")
    sh = torch.cuda.current_stream().cuda_stream
    # 48 channels: no kernel for the shape
    assert L.lh2hip_dn_conv3x3_dw_noalloc(x.data_ptr(), 48, x.data_ptr(), 32, 8, 8, dw.data_ptr(), None, sh) != 0
    # wider than 512 column segments of 64 pixels: more blocks than the array holds partials for (checked on the host)
    assert L.lh2hip_dn_conv3x3_dw_noalloc(x.data_ptr(), 16, x.data_ptr(), 16, 1, 512 * 64 + 1, dw.data_ptr(), None, sh) != 0
    assert L.lh2hip_dn_conv3x3_dw_noalloc(None, 16, x.data_ptr(), 16, 8, 8, dw.data_ptr(), None, sh) != 0
    torch.cuda.synchronize()
