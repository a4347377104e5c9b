/* shade.h - wavefront path-tracing kernels (generate / shade / connect /
 * resolve / finalize) of RenderCore_HIP. */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "waves.h"

namespace lh {

/* Device material, the 128-byte CUDAMaterial of the reference
 * (lib/RenderCore_PrimeRef/core_settings.h:87-136): half diffuse/transmittance,
 * flags, 16 8-bit Disney parameters + eta, six texture-map descriptors. */
struct GpuMaterial
{
	uint32_t diffuse_rg, diffuse_b_medium_r, medium_gb, flags;
	uint4 params;
	uint4 t0, t1, n0, n1, s, r;
};

enum MaterialFlags
{
	MF_ISDIELECTRIC = 1, MF_DIFFUSEMAPISHDR = 2, MF_HASDIFFUSEMAP = 4, MF_HASNORMALMAP = 8, MF_HASSPECULARITYMAP = 16,
	MF_HASROUGHNESSMAP = 32, MF_ISANISOTROPIC = 64, MF_HAS2NDNORMALMAP = 128, MF_HAS2NDDIFFUSEMAP = 512,
	MF_HASSMOOTHNORMALS = 2048, MF_HASALPHA = 4096
};

/* One job's share of a pass that several render jobs share (the job
 * server's batches): its paths hold path slots [slotBase, slotBase + n) of
 * the pass, slot slotBase being path jobBase of the job's frame; the rest is
 * what the job's own ShadeParams would hold. */
struct JobSeg
{
	uint32_t slotBase, jobBase;
	int w, h, pass, maxDepth;
	float spreadAngle;
	uint32_t shift;               /* optix7 blue-noise shift */
	int blueSlot;
	uint32_t R0add;               /* reference estimator: added to the pass-wide R0 (samplesTaken * 7907) */
	uint4* gFeatures;
	float4* gWorldPos;
	float4* gDeltaDepth;
	float4 camPos, camP1, camRight, camUp;
	uint32_t R0w[8];              /* optix7 / filter estimators: shade seed per wave (index pathLength) */
};
static const int JOB_SEGS_MAX = 16;

/* Per-wave constants of the shade kernel (replaces the reference's
 * __constant__ globals, lib/RenderCore_PrimeRef/kernels/.cuda.cu:20-45, so
 * several jobs can be in flight). */
struct ShadeParams
{
	const ShadeInst* inst;
	const GpuMaterial* mats;
	const float4* triLights;      /* CoreLightTri, 6 float4 each */
	const float4* pointLights;    /* CorePointLight, 2 float4 */
	const float4* spotLights;     /* CoreSpotLight, 3 float4 */
	const float4* dirLights;      /* CoreDirectionalLight, 2 float4 */
	int nTri, nPoint, nSpot, nDir;
	int w, h;
	uint32_t jobBase;             /* global job index of path slot 0 of this pass */
	int pathLength, maxDepth;
	uint32_t R0;
	int pass;
	float eps;
	int probePixel;
	int* probe;                   /* {inst, prim, t bits} */
	/* texel arrays (SyncStorageType) and the ray cone spread for texture LOD */
	const uint32_t* argb32;
	const float4* argb128;
	const uint32_t* nrm32;
	float spreadAngle;
	/* skydome: float4 texels, size, rows of the world-to-sky 3x3 */
	const float4* sky;
	int skyW, skyH;
	float4 w2s0, w2s1, w2s2;
	/* optix7 estimator: blue-noise tables and shift, firefly clamp, NEE on/off */
	const uint32_t* blueNoise;
	uint32_t shift;
	float clampValue;
	int connections;
	int bsdf;                     /* optix7 estimator: 0 Lambert (sharedBSDFs/lambert.h), 1 Disney (sharedBSDFs/disney.h) */
	/* filter estimator (RenderCore_Optix7Filter): G-buffer of the first diffuse
	 * vertex (null: filter off), blue-noise slot, path slots per pass (indirect
	 * radiance goes to slot + passPaths), camera for the depth derivatives */
	uint4* gFeatures;
	float4* gWorldPos;
	float4* gDeltaDepth;
	int blueSlot;
	uint32_t passPaths;
	float4 camPos, camP1, camRight, camUp;
	/* several jobs in one pass: nseg > 0 segments (device table, ascending
	 * slotBase); the path buffers then hold pass slots instead of path indices */
	const JobSeg* segs;
	int nseg;
};

/* Path-state buffers of one pass (SoA). */
struct PathBuffers
{
	float4* rays[2];       /* Ray4 {O, tmin; D, tmax} */
	float4* state[2];      /* throughput.xyz, flags bits */
	uint32_t* pathId[2];   /* global job index */
	float4* hits;
	float4* shadowRays;
	float4* shadowPot;     /* contribution.xyz, path slot bits */
	uint32_t* occluded;
	float4* pathAcc;       /* per path slot, summed in path order */
	uint32_t* counters;    /* [0] extension count in, [1] extension out, [2] shadow, [3] trace work counter, [4] second trace counter */
	float2* aux[2];        /* optix7 estimator: postponed bsdf pdf, packed normal of the last vertex */
};

/* Eye rays of paths jobBase .. jobBase + n - 1 of a frame into pass slots
 * slotBase .. slotBase + n - 1 (total: paths of the whole pass, 0 = n; the
 * launch with slotBase 0 resets the wave counters to it). storeSlot: the path
 * buffers hold the pass slot (passes shared by several jobs), else the path index. */
hipError_t launch_generate(const CamParams& c, uint32_t R0, uint32_t jobBase, int n, float eps, const PathBuffers& b, int out, hipStream_t s,
	uint32_t slotBase = 0, uint32_t total = 0, bool storeSlot = false);
hipError_t launch_shade(const ShadeParams& p, const PathBuffers& b, int in, int maxCount, hipStream_t s);
/* Adds each unoccluded connection to its path slot; countW is added to w
 * (optix7's connections count 1 there, rendercore_optix7/optix/.optix.cu:153). */
hipError_t launch_connect(const PathBuffers& b, int maxCount, hipStream_t s, float countW = 0.0f);
/* fresh: write the sums instead of adding them to the accumulators (the
 * first pass of a frame after a restart; every pixel gets >= 1 sample). */
hipError_t launch_resolve(const float4* pathAcc, int pixels, int samples, float4* acc32, double* acc64, hipStream_t s, bool fresh = false);
hipError_t launch_finalize(const float4* acc32, int pixels, float scale, float4* image, hipStream_t s);
hipError_t launch_set_u32(uint32_t* p, int n, uint32_t v, hipStream_t s);
/* optix7 estimator (lib/rendercore_optix7): eye rays with blue-noise pixel and
 * lens samples (optix/.optix.cu:85-129) and the MIS shade kernel
 * (kernels/pathtracer.h:54-238) with the Lambert BSDF (sharedBSDFs/lambert.h). */
hipError_t launch_generate7(const CamParams& c, const ShadeParams& p, uint32_t jobBase, int n, float eps, const PathBuffers& b, int out, hipStream_t s,
	uint32_t slotBase = 0, uint32_t total = 0, bool storeSlot = false);
hipError_t launch_shade7(const ShadeParams& p, const PathBuffers& b, int in, int maxCount, hipStream_t s);
/* Blue-noise tables in the layout of Heitz et al.'s sampler: [0, 65536)
 * sequence values dimension + 256 * index, [65536, 196608) scrambling tile
 * and [196608, 327680) ranking tile, 8 entries per texel of a 128 x 128 tile,
 * plus 64 zero entries of padding (the ranking reads of dimensions above 7
 * at the last texel run past the tables). Filled by a seeded generator. */
void blue_noise_tables(uint32_t* out, uint32_t seed);
static const int BLUE_NOISE_WORDS = 65536 * 5 + 64;
static const uint32_t BLUE_NOISE_SEED = 0x2545f491u;
/* MAXPATHLENGTH of rendercore_optix7 (core_settings.h:25) */
static const int MAXPATHLENGTH_OPTIX7 = 3;
/* filter estimator (lib/RenderCore_Optix7Filter): pinhole pixel-corner eye
 * rays with the filter on, blue-noise jittered thin-lens rays with it off
 * (optix/.optix.cu:66-129), and the shade kernel that also writes the SVGF
 * G-buffer and splits direct and indirect radiance (kernels/pathtracer.h:76-306). */
hipError_t launch_generateF(const CamParams& c, const ShadeParams& p, uint32_t jobBase, int n, float eps, const PathBuffers& b, int out, hipStream_t s,
	uint32_t slotBase = 0, uint32_t total = 0, bool storeSlot = false);
hipError_t launch_shadeF(const ShadeParams& p, const PathBuffers& b, int in, int maxCount, hipStream_t s);
/* MAXPATHLENGTH of RenderCore_Optix7Filter (core_settings.h:25) */
static const int MAXPATHLENGTH_FILTER = 7;
/* counters[0] = counters[1] (next wave size, also logged), counters[1] = counters[2] = 0.
 * snapshot (if given) receives counters[0..3] as they were before; live (if
 * given, may be pinned host memory) the next wave size. */
hipError_t launch_advance(uint32_t* counters, uint32_t* logSlot, hipStream_t s, uint32_t* snapshot = nullptr, uint32_t* live = nullptr);
hipError_t launch_copy_u32(uint32_t* dst, const uint32_t* src, hipStream_t s);

} // namespace lh
