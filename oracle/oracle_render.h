/* oracle_render.h - CPU oracle of the reference path-tracing estimator.
 * TEST INFRASTRUCTURE (see oracle_math.h). */
#pragma once
#include "oracle_trace.h"
#include "../include/lh2hip_types.h"

namespace orc {

using lh2hip::CoreTri;
using lh2hip::CoreMaterial;
using lh2hip::CoreLightTri;
using lh2hip::CorePointLight;
using lh2hip::CoreSpotLight;
using lh2hip::CoreDirectionalLight;
using lh2hip::ViewPyramid;

/* Internal material, as RenderCore::SetMaterials builds CUDAMaterial
 * (lib/RenderCore_PrimeRef/rendercore.cpp:360-402; core_settings.h:87-136). */
struct OMat
{
	uint16_t diffuse[3], transm[3];
	uint flags;
	uint params[4];
	/* texture maps t0 (diffuse), t1 (detail diffuse), n0, n1 (normal maps), s, r
	 * (specularity, roughness): x = width | height << 16, y = half uscale |
	 * half vscale << 16, z = half uoffs | half voffs << 16, w = first texel. */
	uint maps[6][4];
};

/* What SetMaterials needs of a texture (CoreTexDesc, host variant). */
struct TexInfo { uint width, height, flags, firstPixel; int storage; };

enum { S_SPECULAR = 1 };
enum { M_ISDIELECTRIC = 1, M_DIFFUSEMAPISHDR = 2, M_HASDIFFUSEMAP = 4, M_HASNORMALMAP = 8, M_HASSPECULARITYMAP = 16,
	M_HASROUGHNESSMAP = 32, M_HAS2NDNORMALMAP = 128, M_HAS2NDDIFFUSEMAP = 512, M_HASSMOOTHNORMALS = 2048, M_HASALPHA = 4096 };

struct RenderScene : Scene
{
	std::vector<std::vector<CoreTri>> tris;
	std::vector<OMat> mats;
	std::vector<CoreLightTri> triLights;
	std::vector<CorePointLight> pointLights;
	std::vector<CoreSpotLight> spotLights;
	std::vector<CoreDirectionalLight> dirLights;
	/* textures: continuous texel arrays per storage kind (SyncStorageType) */
	std::vector<TexInfo> tex;
	std::vector<uint> argb32, nrm32;
	std::vector<f4> argb128;
	/* skydome: float4 texels (w = 0) and the world-to-sky transform */
	std::vector<f4> sky;
	int skyW = 0, skyH = 0;
	float worldToSky[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
	/* blue-noise tables of the optix7 estimator: 5 * 65536 + 16 words */
	std::vector<uint> blueNoise;
};

OMat convert_material(const CoreMaterial& m, const std::vector<TexInfo>& tex);

/* Texel lookups of the shading code (exposed for the tests). */
f4 fetch_texel(const RenderScene& s, float tu, float tv, int o, int w, int h, int storage);
f4 fetch_texel_trilinear(const RenderScene& s, float lambda, float tu, float tv, int o, int w, int h);
f3 sample_sky(const RenderScene& s, f3 D);

struct RenderParams
{
	ViewPyramid view;   /* spreadAngle sets the texture LOD cone */
	int w, h, spp, maxDepth;
	float eps;
};

/* One frame (spp samples per pixel) of the PrimeRef estimator; per-path
 * contributions are summed in fp32 per path (pathAcc: w*h*spp float4). */
void render_frame(const RenderScene& s, const RenderParams& p, uint R0gen, int pass, float* pathAcc, int threads);

struct Render7Params
{
	ViewPyramid view;
	int w, h, spp, pass;
	float eps, clampValue;
	uint shift;
	int bsdf;   /* 0 Lambert, 1 Disney */
};

/* One frame of the optix7 estimator (Lambert BSDF). camSeed is camRNGseed at
 * the frame's start; the per-wave shade seeds are RandomUInt(camSeed) + k *
 * 91771. pathAcc: w*h*spp float4 (xyz radiance, w unoccluded connections).
 * Returns the number of shade waves the frame ran. */
int render_frame7(const RenderScene& s, const Render7Params& p, uint camSeed, float* pathAcc, int threads);

struct RenderFParams
{
	ViewPyramid view;   /* aperture and distortion as the core uses them */
	int w, h, spp, pass, blueSlot, bsdf;
	float eps, clampValue;
	float j0, j1;       /* TAA jitter in pixels (0, 0: none) */
};
/* One frame of the filter estimator. pathAcc: 2 * w*h*spp float4, direct
 * radiance (w: depth at the first hit plus unoccluded connections) then
 * indirect; features (w*h uint4), worldPos and deltaDepth (w*h float4) are
 * the G-buffer, written for sample 0 only; null features: filter off (no
 * G-buffer, jittered thin-lens eye rays, everything in the direct sum).
 * Returns the number of shade waves the frame ran. */
int render_frameF(const RenderScene& s, const RenderFParams& p, uint camSeed, float* pathAcc, uint* features, float* worldPos,
	float* deltaDepth, int threads);
void blue_noise_default(uint* out, uint seed);
void blue_noise_sample4(const RenderScene& s, int x, int y, int sampleIndex, int dim, float* out4);
uint normal_pack(const float* n3);
f3 sample_sky_small(const RenderScene& s, f3 D);
void disney_probe_eval(const float* color3, const uint* params4, const float* transm3, const float* N3, const float* T3,
	const float* wo3, const float* wi3, float* out4);
void disney_probe_sample(const float* color3, const uint* params4, const float* transm3, const float* N3, const float* T3,
	const float* wo3, const float* r3, float* out8);
void ggx_probe(int which, const float* v3, const float* m3, float ax, float ay, float r0, float r1, float* out3);
void normal_unpack(uint p, float* n3);

/* Microbenchmark rays. */
void primary_ray(const ViewPyramid& v, int w, int h, uint R0, int job, float tmin, float* ray8);
bool bounce_ray(const RenderScene& s, const float* rayIn8, const Hit& h, uint pathId, uint R0, float tmin, float* rayOut8);

} // namespace orc
