/* lh2hip_types.h - data formats shared by the caller and RenderCore_HIP.
 *
 * Every struct here is byte-compatible with the Lighthouse 2 type of the same
 * name, so a RenderSystem compiled against the Lighthouse 2 headers can hand
 * its arrays straight to this core. Layouts are restated, field by field, from:
 *   vector placeholders  lib/RenderSystem/common_types.h:59-68 (alignment 8 for
 *                        2-vectors, 4 for 3-vectors, 16 for 4-vectors)
 *   CoreTri              lib/RenderSystem/common_classes.h:57-97   (208 bytes)
 *   CoreMaterial         lib/RenderSystem/common_classes.h:219-328
 *   CoreTexDesc          lib/RenderSystem/common_classes.h:330-359
 *   Core*Light           lib/RenderSystem/common_classes.h:365-446
 *   ViewPyramid          lib/RenderSystem/common_classes.h:452-475
 *   mat4                 lib/RenderSystem/common_types.h:460-465   (row-major, translation in cell[3,7,11])
 *   CoreStats            lib/RenderSystem/core_api_base.h:30-64
 *   Ray4 / hit record    lib/RenderCore_PrimeRef/core_settings.h:151-153,
 *                        lib/rendercore_optix7/optix/.optix.cu:174-184
 * The header is plain C (usable from ctypes, cgo or C++); C++ callers get the
 * same names inside namespace lh2hip.
 */
#ifndef LH2HIP_TYPES_H
#define LH2HIP_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
#define LH2_ALIGN(x) alignas(x)
namespace lh2hip {
#else
#define LH2_ALIGN(x) _Alignas(x)
#endif

typedef struct { LH2_ALIGN(8) float x; float y; } lh2_float2;
typedef struct { float x, y, z; } lh2_float3;
typedef struct { LH2_ALIGN(16) float x; float y, z, w; } lh2_float4;
typedef struct { LH2_ALIGN(8) int32_t x; int32_t y; } lh2_int2;
typedef struct { LH2_ALIGN(8) uint32_t x; uint32_t y; } lh2_uint2;
typedef struct { LH2_ALIGN(16) int32_t x; int32_t y, z, w; } lh2_int4;

/* Convergence mode for Render (common_classes.h:40-44). */
enum { LH2_CONVERGE = 0, LH2_RESTART = 1 };

/* Fat shading triangle, 13 float4 = 208 bytes (common_classes.h:57-97). */
typedef struct
{
	float u0, u1, u2; int32_t ltriIdx;
	float v0, v1, v2; uint32_t material;
	lh2_float3 vN0; float Nx;
	lh2_float3 vN1; float Ny;
	lh2_float3 vN2; float Nz;
	lh2_float3 T; float area;
	lh2_float3 B; float invArea;
	lh2_float3 alpha; float LOD;
	lh2_float3 vertex0; float dummy0;
	lh2_float3 vertex1; float dummy1;
	lh2_float3 vertex2; float dummy2;
	float u1_0, u1_1, u1_2; float dummy3;
	float v1_0, v1_1, v1_2; float dummy4;
} CoreTri;

/* Material description as sent by RenderSystem (common_classes.h:219-328). */
typedef struct
{
	lh2_float3 value;            /* default if no map; 1e-32 = unset */
	int32_t textureID;           /* -1 = no map */
	float scale;
	lh2_float2 uvscale, uvoffset;
	lh2_uint2 size;
} CoreVec3Value;

typedef struct
{
	float value;                 /* default if no map; 1e32 = unset */
	int32_t textureID;
	int32_t component;
	float scale;
	lh2_float2 uvscale, uvoffset;
	lh2_uint2 size;
} CoreScalarValue;

typedef struct
{
	CoreVec3Value color, detailColor, normals, detailNormals;
	uint32_t flags;              /* 1 = smooth normals, 2 = alpha */
	CoreVec3Value absorption;
	CoreScalarValue metallic, subsurface, specular, roughness, specularTint, anisotropic,
		sheen, sheenTint, clearcoat, clearcoatGloss, transmission, eta;
	CoreScalarValue reflection, refraction, ior;
	uint8_t pbrtMaterialType;    /* MaterialType : 8 bit-field */
	CoreScalarValue urough, vrough;
	CoreVec3Value Ks, eta_rgb;
	CoreScalarValue sigma;
	uint8_t thin;
	CoreScalarValue specTrans, diffTrans;
	CoreVec3Value scatterDistance;
	CoreScalarValue flatness;
	CoreVec3Value Kr, opacity;
} CoreMaterial;

/* Texel storage kinds (common_classes.h:330-335). */
enum { LH2_ARGB32 = 0, LH2_ARGB128 = 1, LH2_NRM32 = 2 };

/* Host-side texture descriptor (common_classes.h:336-359, host variant). */
typedef struct
{
	union { lh2_float4* fdata; uint32_t* idata; };
	uint32_t width, height, flags, pixelCount, firstPixel, MIPlevels;
	int32_t storage;
} CoreTexDesc;

/* Emissive triangle (common_classes.h:365-387), 96 bytes. */
typedef struct
{
	lh2_float3 centre; float energy;
	lh2_float3 N; float area;
	lh2_float3 radiance; int32_t dummy2;
	lh2_float3 vertex0; int32_t triIdx;
	lh2_float3 vertex1; int32_t instIdx;
	lh2_float3 vertex2; int32_t dummy1;
} CoreLightTri;

typedef struct { lh2_float3 position; float energy; lh2_float3 radiance; int32_t dummy; } CorePointLight;
typedef struct { lh2_float3 position; float cosInner; lh2_float3 radiance; float cosOuter; lh2_float3 direction; int32_t dummy; } CoreSpotLight;
typedef struct { lh2_float3 direction; float energy; lh2_float3 radiance; int32_t dummy; } CoreDirectionalLight;

/* Camera view (common_classes.h:452-475). */
typedef struct
{
	lh2_float3 pos, p1, p2, p3;
	float aperture, spreadAngle, imagePlane, focalDistance, distortion;
} ViewPyramid;

/* Row-major 4x4 matrix (common_types.h:460-465). */
typedef struct { float cell[16]; } lh2_mat4;

/* Per-frame statistics (core_api_base.h:30-64). */
typedef struct
{
	char* deviceName;
	uint32_t SMcount, ccMajor, ccMinor, VRAM;
	uint32_t argb32TexelCount, argb128TexelCount, nrm32TexelCount;
	float bvhBuildTime;
	uint32_t totalRays, totalExtensionRays, totalShadowRays;
	float renderTime, frameOverhead;
	uint32_t primaryRayCount; float traceTime0;
	uint32_t bounce1RayCount; float traceTime1;
	uint32_t deepRayCount; float traceTimeX;
	float shadowTraceTime, shadeTime, filterTime;
	int32_t probedInstid, probedTriid; float probedDist;
	lh2_float3 probedWorldPos;
} CoreStats;

/* Ray as queried by the traversal: origin + tmin, direction + tmax
 * (Prime RTP_BUFFER_FORMAT_RAY_ORIGIN_TMIN_DIRECTION_TMAX). 32 bytes. */
typedef struct { lh2_float4 O4, D4; } Ray4;

/* Closest-hit record, 16 bytes (.optix.cu:174-184 packing):
 *   x = bits(u16(65535*u) | u16(65535*v) << 16), y = bits(instance id),
 *   z = bits(primitive id, 0xffffffff on a miss), w = t. */
typedef lh2_float4 HitRecord;

#ifdef __cplusplus
static_assert(sizeof(CoreTri) == 208, "CoreTri must be 208 bytes");
static_assert(sizeof(CoreVec3Value) == 48 && sizeof(CoreScalarValue) == 40, "material value layout");
static_assert(sizeof(CoreMaterial) == 1344, "CoreMaterial layout");
static_assert(sizeof(CoreLightTri) == 96 && sizeof(CorePointLight) == 32, "light layout");
static_assert(sizeof(CoreSpotLight) == 48 && sizeof(CoreDirectionalLight) == 32, "light layout");
static_assert(sizeof(ViewPyramid) == 68 && sizeof(lh2_mat4) == 64, "view layout");
static_assert(sizeof(Ray4) == 32 && sizeof(HitRecord) == 16, "ray layout");
} /* namespace lh2hip */
#endif

#endif /* LH2HIP_TYPES_H */
