/* unet.h - inference of the learned denoiser on the host side of the core:
 * the 5-level U-Net forward pass composed from the denoiser kernels, for the
 * job server. The same network is trained and run from Python with the same
 * kernels (lighthouse2_amd/denoiser.py). No counterpart in the reference. */
#pragma once
#include "rendercore.h"
#include "denoise.h"

namespace lh {

class UNetForward
{
public:
	static const int LEVELS = 5, CH = 32, CIN = 16, COUT = 16, LAYERS = 2 * LEVELS + 2 * (LEVELS - 1) + 1;
	/* (Cin, Cout) of layer i in forward order: e1_0 e2_0 .. e1_4 e2_4, d1_3 d2_3 .. d1_0 d2_0, out */
	static void LayerShape(int i, int& cin, int& cout);
	/* The weight blob: per layer in forward order, bf16 weights [9][Cout][Cin]
	 * (tap = 3 dy + dx) then fp32 bias [Cout]. */
	static size_t BlobBytes();
	void SetWeights(const void* blob, size_t bytes, hipStream_t s);
	bool Ready() const { return ready; }
	/* Input buffer [Hp][Wp][16] bf16 for a Hp x Wp image (multiples of 16). */
	bf16_t* Input(int Hp, int Wp);
	/* Runs the network on the input; returns the fp32 output [Hp][Wp][16]. */
	const float* Forward(hipStream_t s);
	/* The display image [h w] of a filter-core frame (acc32 float4 [2 h w],
	 * features uint4 [h w]), the output convolution writing it directly; equal
	 * to Input + dn_frame_input + Forward + dn_frame_output. */
	void DenoiseFrame(const float4* acc32, const uint4* features, float invSpp, int w, int h, float4* image, hipStream_t s);

private:
	void Middle(hipStream_t s);   /* every layer after the first and before the output convolution */
	const bf16_t* Wt(int i) const { return (const bf16_t*)((const char*)weights.p + wOff[i]); }
	const float* Bi(int i) const { return (const float*)((const char*)weights.p + bOff[i]); }
	bool ready = false;
	int H = 0, W = 0;
	DevBuf weights;
	size_t wOff[LAYERS] = {}, bOff[LAYERS] = {};
	DevBuf in[LEVELS], a[LEVELS], b[LEVELS], c[LEVELS], d[LEVELS], out;
};

} // namespace lh
