/* waves.h - ray-wave kernels (primary generation, diffuse bounce). */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lh {

/* Camera for ray generation: ViewPyramid fields (common_classes.h:452-475) + size. */
struct CamParams
{
	float pos[3], p1[3], p2[3], p3[3];
	float aperture, spreadAngle, distortion;
	int w, h;
};

/* Per-instance shading record: inverse transform rows, the mesh's fat
 * triangles (CoreTri, 13 float4) and vertex soup (3 float4 per triangle). */
struct ShadeInst
{
	float4 inv0, inv1, inv2;
	const float4* coreTris;
	const float4* verts;
	int mesh, pad[3];
};

hipError_t launch_primary(const CamParams& c, uint32_t R0, float tmin, float4* rays, uint32_t* pathId, hipStream_t s);
hipError_t launch_bounce(const ShadeInst* inst, const float4* raysIn, const float4* hits, const uint32_t* pathIn, int n,
	uint32_t R0, float tmin, float4* raysOut, uint32_t* pathOut, uint32_t* counter, hipStream_t s);

} // namespace lh
