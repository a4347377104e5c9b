/* denoise.h - kernels of the learned denoiser (a 5-level convolutional
 * U-Net, 3x3 kernels, 32 channels): bf16 activations and weights in NHWC,
 * fp32 accumulation on the gfx950 matrix cores. No counterpart in the
 * reference (its denoisers are SVGF and OptiX's). */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace lh {

typedef uint16_t bf16_t;   /* bfloat16 bits */

enum { DN_RELU = 1, DN_OUT_F32 = 2 };

/* y[H][W][Cout] = act(conv3x3(x[H][W][Cin], w) + bias), zero padding.
 * w: [9][Cout][Cin] bf16 (tap = 3 * dy + dx); Cin, Cout multiples of 16,
 * Cout <= 64. flags: DN_RELU, DN_OUT_F32 (y is fp32 instead of bf16). */
hipError_t dn_conv3x3(const bf16_t* x, int Cin, const bf16_t* w, const float* bias, void* y, int Cout, int H, int W, int flags, hipStream_t s);
/* y = relu(conv3x3(x) + bias) (Cin 16 or 32, Cout 32, bf16) and, from the same
 * tile, ypool = dn_pool2(y) ([H/2][W/2][32]; H, W even). */
hipError_t dn_conv3x3_pool(const bf16_t* x, int Cin, const bf16_t* w, const float* bias, bf16_t* y, int Cout, int H, int W,
	bf16_t* ypool, hipStream_t s);
/* The same convolution (Cin 64, Cout 32) of the concatenation
 * [nearest 2x upsample of coarse[H/2][W/2][32], skip[H][W][32]], read from
 * the two tensors (H, W even); equal to dn_upcat followed by dn_conv3x3. */
hipError_t dn_conv3x3_cat(const bf16_t* coarse, const bf16_t* skip, const bf16_t* w, const float* bias, void* y, int Cout, int H, int W,
	int flags, hipStream_t s);
/* image = the display image of conv3x3(x) + bias (32 -> 16 channels) for a
 * w x h frame inside the Hp x Wp network image: what dn_conv3x3 (fp32 out)
 * followed by dn_frame_output computes, without the fp32 output tensor. */
hipError_t dn_conv3x3_display(const bf16_t* x, const bf16_t* wts, const float* bias, const uint4* features, int w, int h, float4* image,
	int Hp, int Wp, hipStream_t s);
/* Weight gradient: dw[9][Cout][Cin] = sum_p x[p + tap] dy[p], db[Cout] = sum_p dy[p]
 * (both overwritten). ws: workspace of dn_dw_workspace_bytes(Cin, Cout). */
size_t dn_dw_workspace_bytes(int Cin, int Cout);
hipError_t dn_conv3x3_dw(const bf16_t* x, int Cin, const bf16_t* dy, int Cout, int H, int W, float* dw, float* db, void* ws, hipStream_t s);
/* dn_conv3x3_dw with the workspace in a file-scope device array of
 * dn_dw_scratch_bytes() (no allocation; calls overlapping on one device race). */
size_t dn_dw_scratch_bytes();
hipError_t dn_conv3x3_dw_noalloc(const bf16_t* x, int Cin, const bf16_t* dy, int Cout, int H, int W, float* dw, float* db, hipStream_t s);
hipError_t dn_relu_bwd(const bf16_t* y, const bf16_t* dy, bf16_t* dyPre, size_t n, hipStream_t s);
/* 2x2 average pool (H, W even) and its gradient (dx = or += dy / 4). */
hipError_t dn_pool2(const bf16_t* x, int C, int H, int W, bf16_t* y, hipStream_t s);
hipError_t dn_pool2_bwd(const bf16_t* dy, int C, int H, int W, bf16_t* dx, int accumulate, hipStream_t s);
/* out[H][W][C0 + C1] = (nearest 2x upsample of coarse[H/2][W/2][C0], skip[H][W][C1]); backward:
 * dcoarse = 2x2 sum of dcat[..][0:C0] (overwrite), dskip (=) dcat[..][C0:]. */
hipError_t dn_upcat(const bf16_t* coarse, int C0, const bf16_t* skip, int C1, int H, int W, bf16_t* out, hipStream_t s);
hipError_t dn_upcat_bwd(const bf16_t* dcat, int C0, int C1, int H, int W, bf16_t* dcoarse, bf16_t* dskip, hipStream_t s);
/* Network input: [radiance.xyz, albedo.xyz, normal.xyz, depth, 0 x 6] as bf16 (16 channels). */
hipError_t dn_pack_input(const float4* radiance, const float4* albedo, const float4* normal, const float* depth, int H, int W, bf16_t* out, hipStream_t s);
/* relMSE loss over the first 3 channels of out (fp32, Cout channels):
 * loss += sum (o - t)^2 / (t^2 + 0.01) / (3 H W); dout (bf16, Cout channels) = its gradient. */
hipError_t dn_relmse(const float* out, int Cout, const float4* target, int H, int W, bf16_t* dout, float* loss, hipStream_t s);
/* Adam step on fp32 parameters; for a conv weight [9][Cout][Cin] (Cout > 0)
 * also writes the bf16 forward layout and the flipped backward layout
 * wbwd[8 - tap][Cin][Cout] used by the input-gradient convolution. */
hipError_t dn_adam(float* p, float* m, float* v, const float* g, size_t n, float lr, float b1, float b2, float eps, int t,
	bf16_t* wfwd, bf16_t* wbwd, int Cout, int Cin, hipStream_t s);

/* Network input from a filter-estimator frame: radiance = (direct + indirect)
 * sums * invSpp, demodulated by the G-buffer albedo and range-compressed,
 * log(1 + radiance / max(albedo, 0.02)); albedo, normal and 0.1 log(1 + depth)
 * (depth clamped to 1e4, misses at 1e4) from the packed features; zero in the
 * padding beyond w x h. acc32: float4 [2 w h]; features: uint4 [w h]. */
hipError_t dn_frame_input(const float4* acc32, const uint4* features, float invSpp, int w, int h, int Hp, int Wp, bf16_t* out, hipStream_t s);
/* Denoised image from the network output: radiance = (exp(o) - 1) *
 * max(albedo, 0.02) per channel (o clamped at 0), written as the filter
 * core's display value radiance^(1/4) (the sqrt of the sqrt SVGF's last pass
 * writes), w = 0. */
hipError_t dn_frame_output(const float* out, int Hp, int Wp, const uint4* features, int w, int h, float4* image, hipStream_t s);

} // namespace lh
