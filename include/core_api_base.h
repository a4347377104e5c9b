/* core_api_base.h - the Lighthouse 2 render-core plugin interface, as
 * implemented by RenderCore_HIP.
 *
 * This restates the virtual interface of lib/RenderSystem/core_api_base.h:81-119
 * with the same method order (the vtable layout is the binary contract between
 * RenderSystem and a core loaded with dlopen + dlsym("CreateCore"),
 * lib/RenderSystem/core_api_base.cpp:97-130) and byte-compatible argument types
 * from lh2hip_types.h. A Lighthouse 2 RenderSystem built against its own headers
 * can load librendercore_hip.so unchanged.
 */
#pragma once

#include "lh2hip_types.h"

namespace lighthouse2 {

using lh2hip::CoreTri;
using lh2hip::CoreMaterial;
using lh2hip::CoreTexDesc;
using lh2hip::CoreLightTri;
using lh2hip::CorePointLight;
using lh2hip::CoreSpotLight;
using lh2hip::CoreDirectionalLight;
using lh2hip::ViewPyramid;
using lh2hip::CoreStats;
typedef lh2hip::lh2_int2 int2;
typedef lh2hip::lh2_float3 float3;
typedef lh2hip::lh2_float4 float4;
typedef lh2hip::lh2_mat4 mat4;

enum Convergence { Converge = 0, Restart = 1 };

/* Render target handle (lib/platform/system.h:220-238). RenderCore_HIP is
 * headless: it only reads width and height; ID 0 means "no GL texture". The
 * image is read back with lh2hip_read_image. */
class GLTexture
{
public:
	uint32_t ID = 0;
	uint32_t width = 0, height = 0;
};

class CoreAPI_Base
{
public:
	virtual CoreStats GetCoreStats() const = 0;
	virtual void Init() = 0;
	virtual void SetProbePos( const int2 pos ) = 0;
	virtual void SetTarget( GLTexture* target, const uint32_t spp ) = 0;
	virtual void Setting( const char* name, float value ) = 0;
	virtual void Render( const ViewPyramid& view, const Convergence converge, bool async ) = 0;
	virtual void WaitForRender() = 0;
	virtual void Shutdown() = 0;
	virtual void SetTextures( const CoreTexDesc* tex, const int textureCount ) = 0;
	virtual void SetMaterials( CoreMaterial* mat, const int materialCount ) = 0;
	virtual void SetLights( const CoreLightTri* triLights, const int triLightCount,
		const CorePointLight* pointLights, const int pointLightCount,
		const CoreSpotLight* spotLights, const int spotLightCount,
		const CoreDirectionalLight* directionalLights, const int directionalLightCount ) = 0;
	virtual void SetSkyData( const float3* pixels, const uint32_t width, const uint32_t height, const mat4& worldToLight ) = 0;
	virtual void SetGeometry( const int meshIdx, const float4* vertexData, const int vertexCount, const int triangleCount, const CoreTri* triangles ) = 0;
	virtual void SetInstance( const int instanceIdx, const int modelIdx, const mat4& transform ) = 0;
	virtual void FinalizeInstances() = 0;
};

} // namespace lighthouse2

/* Plugin entry point (lib/RenderCore_PrimeRef/core_api.cpp:18-22). */
extern "C" __attribute__((visibility("default"))) lighthouse2::CoreAPI_Base* CreateCore();
