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
};

enum { S_SPECULAR = 1 };
enum { M_ISDIELECTRIC = 1, M_HASDIFFUSEMAP = 4, M_HASNORMALMAP = 8, M_HASSPECULARITYMAP = 16,
	M_HASROUGHNESSMAP = 32, M_HAS2NDNORMALMAP = 128, M_HAS2NDDIFFUSEMAP = 512, M_HASSMOOTHNORMALS = 2048, M_HASALPHA = 4096 };

struct RenderScene : Scene
{
	std::vector<std::vector<CoreTri>> tris;
	std::vector<OMat> mats;
	std::vector<CoreLightTri> triLights;
	std::vector<CorePointLight> pointLights;
	std::vector<CoreSpotLight> spotLights;
	std::vector<CoreDirectionalLight> dirLights;
};

OMat convert_material(const CoreMaterial& m);

struct RenderParams
{
	ViewPyramid view;
	int w, h, spp, maxDepth;
	float eps;
};

/* One frame (spp samples per pixel) of the PrimeRef estimator; per-path
 * contributions are summed in fp32 per path (pathAcc: w*h*spp float4). */
void render_frame(const RenderScene& s, const RenderParams& p, uint R0gen, int pass, float* pathAcc, int threads);

/* Microbenchmark rays. */
void primary_ray(const ViewPyramid& v, int w, int h, uint R0, int job, float tmin, float* ray8);
bool bounce_ray(const RenderScene& s, const float* rayIn8, const Hit& h, uint pathId, uint R0, float tmin, float* rayOut8);

} // namespace orc
