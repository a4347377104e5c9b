"""Learned denoiser: a 5-level convolutional U-Net (3x3 kernels, 32 channels)
trained and run with the gfx950 kernels of librendercore_hip.so (bf16 NHWC
activations and weights, fp32 accumulation on the matrix cores). No
counterpart in the reference, whose denoisers are SVGF and OptiX's.

Network (channels): input 16 = [radiance 3, albedo 3, normal 3, depth 1, 0 x 6]
  encoder level l = 0..3: a = relu(conv(in)), b = relu(conv(a)), in_{l+1} = avgpool2(b)
  bottleneck (level 4):   a = relu(conv(in)), b = relu(conv(a))
  decoder level l = 3..0: cat = [upsample2(prev), b_l], c = relu(conv(cat)), d = relu(conv(c))
                          (the forward reads cat from its two halves; the backward
                          pass materialises it for the weight gradient)
  output: conv(d_0) -> 16 channels fp32, the first 3 are the denoised radiance.
Loss: relMSE over the 3 output channels. Optimiser: Adam on fp32 master weights.
Torch tensors serve as device buffers only; every operation is a kernel of
the library (lh2hip_dn_*).
"""
import ctypes

import numpy as np

from .core import library, _check

LEVELS = 5
CH = 32
CIN = 16
COUT = 16


def layer_specs():
    """(name, Cin, Cout) of the 19 convolutions, in forward order."""
    spec = []
    for lvl in range(LEVELS):
        spec.append((f"e1_{lvl}", CIN if lvl == 0 else CH, CH))
        spec.append((f"e2_{lvl}", CH, CH))
    for lvl in range(LEVELS - 2, -1, -1):
        spec.append((f"d1_{lvl}", 2 * CH, CH))
        spec.append((f"d2_{lvl}", CH, CH))
    spec.append(("out", CH, COUT))
    return spec


def init_params(seed=0):
    """He-normal weights [9][Cout][Cin] and zero biases (float32, numpy)."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, ci, co in layer_specs():
        w = rng.standard_normal((9, co, ci)) * np.sqrt(2.0 / (9 * ci))
        if name == "e1_0":
            w[:, :, 10:] = 0.0                    # padding channels of the input
        if name == "out":
            w[:, 3:, :] = 0.0                     # only 3 output channels are used
        out[name] = (w.astype(np.float32), np.zeros(co, np.float32))
    return out


class UNetDenoiser:
    def __init__(self, H, W, seed=0, params=None, device="cuda"):
        import torch
        if H % 16 or W % 16:
            raise ValueError("image size must be a multiple of 16 (4 pooling steps)")
        self.torch, self.H, self.W = torch, H, W
        self.stream = torch.cuda.current_stream()
        self.sh = self.stream.cuda_stream
        self.t = 0
        bf = dict(dtype=torch.int16, device=device)
        f32 = dict(dtype=torch.float32, device=device)
        self.p = {}
        params = params or init_params(seed)
        for name, ci, co in layer_specs():
            w, b = params[name]
            d = {"ci": ci, "co": co,
                 "w": torch.from_numpy(np.ascontiguousarray(w)).to(device), "b": torch.from_numpy(np.ascontiguousarray(b)).to(device)}
            for k in ("w", "b"):
                d["m" + k] = torch.zeros_like(d[k])
                d["v" + k] = torch.zeros_like(d[k])
                d["g" + k] = torch.zeros_like(d[k])
            d["wf"] = torch.zeros(9 * co * ci, **bf)
            d["wb"] = torch.zeros(9 * co * ci, **bf)
            self.p[name] = d
        self._cast()
        self.ws = torch.zeros(int(library().lh2hip_dn_dw_workspace_bytes(2 * CH, CH)) // 4 + 64, **f32)
        # activations per level (kept for the backward pass) and their gradients
        self.a, self.g = {}, {}
        for lvl in range(LEVELS):
            h, w = H >> lvl, W >> lvl
            self.a[f"in_{lvl}"] = torch.zeros(h * w * (CIN if lvl == 0 else CH), **bf)
            for k in ("a", "b"):
                self.a[f"{k}_{lvl}"] = torch.zeros(h * w * CH, **bf)
                self.g[f"{k}_{lvl}"] = torch.zeros(h * w * CH, **bf)
            self.g[f"in_{lvl}"] = torch.zeros(h * w * CH, **bf)
            if lvl < LEVELS - 1:
                for k in ("c", "d"):
                    self.a[f"{k}_{lvl}"] = torch.zeros(h * w * CH, **bf)
                    self.g[f"{k}_{lvl}"] = torch.zeros(h * w * CH, **bf)
                self.a[f"cat_{lvl}"] = torch.zeros(h * w * 2 * CH, **bf)
                self.g[f"cat_{lvl}"] = torch.zeros(h * w * 2 * CH, **bf)
        self.out = torch.zeros(H * W * COUT, **f32)
        self.dout = torch.zeros(H * W * COUT, **bf)
        self.tmp = torch.zeros(H * W * 2 * CH, **bf)
        self.loss = torch.zeros(1, **f32)

    # ---- helpers --------------------------------------------------------
    def _cast(self):
        L = library()
        for d in self.p.values():
            _check(L.lh2hip_dn_adam(d["w"].data_ptr(), d["mw"].data_ptr(), d["vw"].data_ptr(), None, d["w"].numel(),
                                    0.0, 0.9, 0.999, 1e-8, 1, d["wf"].data_ptr(), d["wb"].data_ptr(), d["co"], d["ci"], self.sh))

    def _conv(self, x, name, y, h, w, flags):
        d = self.p[name]
        _check(library().lh2hip_dn_conv3x3(x.data_ptr(), d["ci"], d["wf"].data_ptr(), d["b"].data_ptr(), y.data_ptr(), d["co"],
                                           h, w, flags, self.sh))

    def _conv_dx(self, gy, name, gx, h, w):
        d = self.p[name]
        _check(library().lh2hip_dn_conv3x3(gy.data_ptr(), d["co"], d["wb"].data_ptr(), None, gx.data_ptr(), d["ci"], h, w, 0, self.sh))

    def _conv_dw(self, x, name, gy, h, w):
        d = self.p[name]
        _check(library().lh2hip_dn_conv3x3_dw(x.data_ptr(), d["ci"], gy.data_ptr(), d["co"], h, w, d["gw"].data_ptr(),
                                              d["gb"].data_ptr(), self.ws.data_ptr(), self.sh))

    def _relu_bwd(self, y, gy, out):
        _check(library().lh2hip_dn_relu_bwd(y.data_ptr(), gy.data_ptr(), out.data_ptr(), y.numel(), self.sh))

    # ---- public API -----------------------------------------------------
    def set_input(self, radiance, albedo, normal, depth):
        """Device float32 tensors (H, W, 4) x 3 and (H, W)."""
        _check(library().lh2hip_dn_pack_input(radiance.data_ptr(), albedo.data_ptr(), normal.data_ptr(), depth.data_ptr(),
                                              self.H, self.W, self.a["in_0"].data_ptr(), self.sh))

    def set_input_packed(self, packed_bf16):
        self.a["in_0"].copy_(packed_bf16.reshape(-1))

    def forward(self):
        """Runs the network on the current input; returns the fp32 output (H, W, 16)."""
        self._conv(self.a["in_0"], "e1_0", self.a["a_0"], self.H, self.W, 1)
        self._forward_from_a0()
        self._conv(self.a["d_0"], "out", self.out, self.H, self.W, 2)
        return self.out.view(self.H, self.W, COUT)

    def _forward_from_a0(self):
        """Every layer after e1_0 and before the output convolution."""
        L = library()
        H, W = self.H, self.W
        for lvl in range(LEVELS):
            h, w = H >> lvl, W >> lvl
            if lvl > 0:
                self._conv(self.a[f"in_{lvl}"], f"e1_{lvl}", self.a[f"a_{lvl}"], h, w, 1)
            if lvl < LEVELS - 1:                   # b_l and its 2x2 average in_{l+1} from one kernel
                d = self.p[f"e2_{lvl}"]
                _check(L.lh2hip_dn_conv3x3_pool(self.a[f"a_{lvl}"].data_ptr(), CH, d["wf"].data_ptr(), d["b"].data_ptr(),
                                                self.a[f"b_{lvl}"].data_ptr(), CH, h, w, self.a[f"in_{lvl + 1}"].data_ptr(), self.sh))
            else:
                self._conv(self.a[f"a_{lvl}"], f"e2_{lvl}", self.a[f"b_{lvl}"], h, w, 1)
        prev = self.a[f"b_{LEVELS - 1}"]
        for lvl in range(LEVELS - 2, -1, -1):
            h, w = H >> lvl, W >> lvl
            d = self.p[f"d1_{lvl}"]
            _check(L.lh2hip_dn_conv3x3_cat(prev.data_ptr(), self.a[f"b_{lvl}"].data_ptr(), d["wf"].data_ptr(), d["b"].data_ptr(),
                                           self.a[f"c_{lvl}"].data_ptr(), CH, h, w, 1, self.sh))
            self._conv(self.a[f"c_{lvl}"], f"d2_{lvl}", self.a[f"d_{lvl}"], h, w, 1)
            prev = self.a[f"d_{lvl}"]

    def loss_backward(self, target):
        """relMSE against target (device float32 (H, W, 4)); fills every
        parameter gradient. Returns the loss as a device scalar."""
        L = library()
        H, W = self.H, self.W
        self.loss.zero_()
        _check(L.lh2hip_dn_relmse(self.out.data_ptr(), COUT, target.data_ptr(), H, W, self.dout.data_ptr(), self.loss.data_ptr(), self.sh))
        g, a, tmp = self.g, self.a, self.tmp
        self._conv_dw(a["d_0"], "out", self.dout, H, W)
        self._conv_dx(self.dout, "out", g["d_0"], H, W)
        for lvl in range(0, LEVELS - 1):
            h, w = H >> lvl, W >> lvl
            self._relu_bwd(a[f"d_{lvl}"], g[f"d_{lvl}"], tmp[:h * w * CH])
            self._conv_dw(a[f"c_{lvl}"], f"d2_{lvl}", tmp, h, w)
            self._conv_dx(tmp, f"d2_{lvl}", g[f"c_{lvl}"], h, w)
            self._relu_bwd(a[f"c_{lvl}"], g[f"c_{lvl}"], tmp[:h * w * CH])
            prev = a[f"d_{lvl + 1}"] if lvl + 1 < LEVELS - 1 else a[f"b_{LEVELS - 1}"]
            _check(L.lh2hip_dn_upcat(prev.data_ptr(), CH, a[f"b_{lvl}"].data_ptr(), CH, h, w, a[f"cat_{lvl}"].data_ptr(), self.sh))
            self._conv_dw(a[f"cat_{lvl}"], f"d1_{lvl}", tmp, h, w)
            self._conv_dx(tmp, f"d1_{lvl}", g[f"cat_{lvl}"], h, w)
            coarse = g[f"d_{lvl + 1}"] if lvl + 1 < LEVELS - 1 else g[f"b_{LEVELS - 1}"]
            _check(L.lh2hip_dn_upcat_bwd(g[f"cat_{lvl}"].data_ptr(), CH, CH, h, w, coarse.data_ptr(), g[f"b_{lvl}"].data_ptr(), self.sh))
        for lvl in range(LEVELS - 1, -1, -1):
            h, w = H >> lvl, W >> lvl
            if lvl < LEVELS - 1:
                _check(L.lh2hip_dn_pool2_bwd(g[f"in_{lvl + 1}"].data_ptr(), CH, h, w, g[f"b_{lvl}"].data_ptr(), 1, self.sh))
            self._relu_bwd(a[f"b_{lvl}"], g[f"b_{lvl}"], tmp[:h * w * CH])
            self._conv_dw(a[f"a_{lvl}"], f"e2_{lvl}", tmp, h, w)
            self._conv_dx(tmp, f"e2_{lvl}", g[f"a_{lvl}"], h, w)
            self._relu_bwd(a[f"a_{lvl}"], g[f"a_{lvl}"], tmp[:h * w * CH])
            self._conv_dw(a[f"in_{lvl}"], f"e1_{lvl}", tmp, h, w)
            if lvl > 0:
                self._conv_dx(tmp, f"e1_{lvl}", g[f"in_{lvl}"], h, w)
        return self.loss

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step on every parameter (also refreshes the bf16 copies)."""
        L = library()
        self.t += 1
        for d in self.p.values():
            _check(L.lh2hip_dn_adam(d["w"].data_ptr(), d["mw"].data_ptr(), d["vw"].data_ptr(), d["gw"].data_ptr(), d["w"].numel(),
                                    lr, b1, b2, eps, self.t, d["wf"].data_ptr(), d["wb"].data_ptr(), d["co"], d["ci"], self.sh))
            _check(L.lh2hip_dn_adam(d["b"].data_ptr(), d["mb"].data_ptr(), d["vb"].data_ptr(), d["gb"].data_ptr(), d["b"].numel(),
                                    lr, b1, b2, eps, self.t, None, None, 0, 0, self.sh))

    def weights_blob(self):
        """The network's inference weights as one blob (the job server's
        lh2hip_server_set_denoiser layout): per layer in forward order, the
        bf16 forward weights [9][Cout][Cin] then the fp32 bias [Cout]."""
        self.torch.cuda.synchronize()
        parts = []
        for name, ci, co in layer_specs():
            d = self.p[name]
            parts.append(d["wf"].cpu().numpy().astype(np.int16).tobytes())
            parts.append(d["b"].cpu().numpy().astype(np.float32).tobytes())
        blob = b"".join(parts)
        assert len(blob) == library().lh2hip_dn_weights_bytes()
        return blob

    def set_input_frame(self, acc32, features, spp, w, h):
        """Input from a filter-estimator frame on the device (acc32 float4
        [2 h w], features uint4 [h w]; the frame's w x h inside this network's
        H x W): the same kernel the job server runs."""
        _check(library().lh2hip_dn_frame_input(acc32.data_ptr(), features.data_ptr(), 1.0 / spp, w, h, self.H, self.W,
                                               self.a["in_0"].data_ptr(), self.sh))

    def denoise_frame(self, acc32, features, spp, w, h, image):
        """Inference on a filter-estimator frame, the output convolution writing
        the display image itself (no fp32 output tensor): the image equals
        set_input_frame + forward + output_frame's."""
        L, H, W = library(), self.H, self.W
        self.set_input_frame(acc32, features, spp, w, h)
        self._conv(self.a["in_0"], "e1_0", self.a["a_0"], H, W, 1)
        self._forward_from_a0()
        d = self.p["out"]
        _check(L.lh2hip_dn_conv3x3_display(self.a["d_0"].data_ptr(), d["wf"].data_ptr(), d["b"].data_ptr(), features.data_ptr(), w, h,
                                           image.data_ptr(), H, W, self.sh))

    def output_frame(self, features, w, h, image):
        """Display image (w x h float4, device) from the last forward pass, as the job server writes it."""
        _check(library().lh2hip_dn_frame_output(self.out.data_ptr(), self.H, self.W, features.data_ptr(), w, h, image.data_ptr(), self.sh))

    def params_numpy(self):
        return {k: (d["w"].cpu().numpy(), d["b"].cpu().numpy()) for k, d in self.p.items()}

    def grads_numpy(self):
        return {k: (d["gw"].cpu().numpy(), d["gb"].cpu().numpy()) for k, d in self.p.items()}

    def flops_forward(self):
        """Multiply-add FLOPs of one forward pass (2 per MAC)."""
        f = 0
        for name, ci, co in layer_specs():
            lvl = 0 if name == "out" else int(name.split("_")[1])
            f += 2 * 9 * ci * co * (self.H >> lvl) * (self.W >> lvl)
        return f


def bf16_bits(a):
    """float32 numpy -> bf16 bits (round to nearest even), as int16."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r.view(np.int16)


def bf16_to_f32(bits):
    return (np.asarray(bits).astype(np.uint16).astype(np.uint32) << 16).view(np.float32)
