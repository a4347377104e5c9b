/* svgf.h - spatiotemporal variance-guided filter (SVGF) of RenderCore_HIP:
 * prepareFilterKernel, applyFilterKernel x 3 and finalizeNoTAAKernel of
 * lib/CUDA/shared_kernel_code/finalize_shared.h:169-484,587-600, driven as
 * RenderCore_Optix7Filter's FinalizeRender (rendercore.cpp:897-935). */
#pragma once
#include "rendercore.h"

namespace lh {

/* Camera of the previous frame, prepared on the host (prepareFilter,
 * finalize_shared.h:296-313). */
struct SvgfCam { float4 prevPos, prevE, prevRight, prevUp; };
SvgfCam svgf_camera(const lighthouse2::ViewPyramid& prevView, int h);

/* All per-pixel state of the filter between frames. */
class Svgf
{
public:
	/* The buffers are cleared on stream s, ahead of the frames drawn on it;
	 * without a stream the constructor returns once they are clear. */
	Svgf(int w, int h, hipStream_t s = nullptr);
	int w, h;
	/* One frame. accumulator: device float4 [2 * w * h] (direct, then indirect
	 * radiance sums); features: device uint4 [w * h] {albedo HDRtoRGB32,
	 * PackNormal2 normal, depth bits, specular << 4 | matid << 6} (the history
	 * count in bits 0..3 of w is kept by the filter); worldPos: device float4
	 * {position, packed normal | specular bits}; deltaDepth: device float4
	 * (z, w = depth derivatives). worldPos and deltaDepth are read in place by
	 * the frame's kernels (stream-ordered on s), not copied. out: device float4
	 * [w * h], the display image. */
	/* Clears every buffer, history included: the next Frame is a first frame. */
	void Reset(hipStream_t s);
	void Frame(const float4* accumulator, const uint4* features, const float4* worldPos, const float4* deltaDepth,
		const lighthouse2::ViewPyramid& prevView, float j0, float j1, float prevj0, float prevj1, int samplesTaken,
		float clampDirect, float clampIndirect, int camIsStationary, float4* out, hipStream_t s);
	/* taa: after the last pass, TAApass and unsharpenTAA produce the display
	 * image (finalize_shared.h:490-575); otherwise finalizeNoTAA. The TAA
	 * history (prevPixels) is kept only while taa is on. */
	int taa = 0;
	DevBuf features, worldPos, prevWorldPos, motion, moments, prevMoments, shading, filteredIn, filteredOut, lastPass;
	DevBuf prevPixels, taaOut;
};

} // namespace lh
