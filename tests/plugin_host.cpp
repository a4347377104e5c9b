/* plugin_host.cpp - a render-system stand-in that loads librendercore_hip.so
 * the way Lighthouse 2 loads a core (CoreAPI_Base::CreateCoreAPI,
 * lib/RenderSystem/core_api_base.cpp:97-130: dlopen with RTLD_NOW |
 * RTLD_GLOBAL, dlsym("CreateCore"), Init) and drives it only through the
 * CoreAPI_Base vtable, in RenderSystem::SynchronizeSceneData's order
 * (lib/RenderSystem/rendersystem.cpp:46-211): materials, geometry per mesh,
 * lights, instances + truncation, FinalizeInstances, then SetTarget, Setting
 * and Render. Test infrastructure.
 *
 * usage: plugin_host <library> <scene file> <output file>
 * scene file (little endian): int32 w, h, spp, maxDepth; float epsilon;
 * uint32 seed; ViewPyramid; int32 nMat, CoreMaterial[nMat]; int32 nMesh, per
 * mesh int32 triCount, float4[3 triCount], CoreTri[triCount]; int32 nTriLight,
 * CoreLightTri[n]; int32 nInst, per instance int32 mesh, float[16].
 * output: float4 image [w h], double accumulator [w h 4], CoreStats. */
#include "core_api_base.h"
#include <dlfcn.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace lighthouse2;

template <typename T> static void rd(FILE* f, T* p, size_t n = 1)
{
	if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "scene file truncated\n"); exit(2); }
}

int main(int argc, char** argv)
{
	if (argc != 4) { fprintf(stderr, "usage: plugin_host <library> <scene> <output>\n"); return 2; }
	void* lib = dlopen(argv[1], RTLD_NOW | RTLD_GLOBAL);
	if (!lib) { fprintf(stderr, "dlopen: %s\n", dlerror()); return 3; }
	typedef CoreAPI_Base* (*CreateFn)();
	CreateFn create = (CreateFn)dlsym(lib, "CreateCore");
	typedef int (*ReadImageFn)(void*, lh2hip::lh2_float4*);
	typedef int (*ReadAccFn)(void*, double*);
	ReadImageFn readImage = (ReadImageFn)dlsym(lib, "lh2hip_plugin_read_image");
	ReadAccFn readAcc = (ReadAccFn)dlsym(lib, "lh2hip_plugin_read_accumulator_f64");
	if (!create || !readImage || !readAcc) { fprintf(stderr, "missing symbol\n"); return 3; }
	FILE* f = fopen(argv[2], "rb");
	if (!f) { perror("scene"); return 2; }
	int w, h, spp, maxDepth;
	float eps;
	uint32_t seed;
	ViewPyramid view;
	rd(f, &w), rd(f, &h), rd(f, &spp), rd(f, &maxDepth), rd(f, &eps), rd(f, &seed), rd(f, &view);
	int nMat;
	rd(f, &nMat);
	std::vector<CoreMaterial> mats(nMat);
	rd(f, mats.data(), mats.size());
	int nMesh;
	rd(f, &nMesh);
	std::vector<std::vector<float4>> verts(nMesh);
	std::vector<std::vector<CoreTri>> tris(nMesh);
	for (int m = 0; m < nMesh; m++)
	{
		int n;
		rd(f, &n);
		verts[m].resize(3 * (size_t)n), tris[m].resize(n);
		rd(f, verts[m].data(), verts[m].size()), rd(f, tris[m].data(), tris[m].size());
	}
	int nLight;
	rd(f, &nLight);
	std::vector<CoreLightTri> lights(nLight);
	rd(f, lights.data(), lights.size());
	int nInst;
	rd(f, &nInst);
	std::vector<int> instMesh(nInst);
	std::vector<mat4> instT(nInst);
	for (int i = 0; i < nInst; i++) rd(f, &instMesh[i]), rd(f, instT[i].cell, 16);
	fclose(f);

	CoreAPI_Base* core = create();
	core->Init();
	core->SetMaterials(mats.data(), nMat);
	for (int m = 0; m < nMesh; m++) core->SetGeometry(m, verts[m].data(), 3 * (int)tris[m].size(), (int)tris[m].size(), tris[m].data());
	core->SetLights(lights.data(), nLight, nullptr, 0, nullptr, 0, nullptr, 0);
	for (int i = 0; i < nInst; i++) core->SetInstance(i, instMesh[i], instT[i]);
	mat4 identity = {};
	for (int i = 0; i < 4; i++) identity.cell[5 * i] = 1.0f;
	core->SetInstance(nInst, -1, identity);
	core->FinalizeInstances();
	core->Setting("maxDepth", (float)maxDepth);
	core->Setting("epsilon", eps);
	core->Setting("seed", (float)seed);
	GLTexture target;
	target.ID = 0, target.width = (uint32_t)w, target.height = (uint32_t)h;
	core->SetTarget(&target, (uint32_t)spp);
	core->Render(view, Restart, false);
	core->WaitForRender();
	const CoreStats stats = core->GetCoreStats();
	std::vector<lh2hip::lh2_float4> image((size_t)w * h);
	std::vector<double> acc((size_t)w * h * 4);
	if (readImage(core, image.data()) != 0 || readAcc(core, acc.data()) != 0) { fprintf(stderr, "readback failed\n"); return 4; }
	core->Shutdown();
	FILE* o = fopen(argv[3], "wb");
	if (!o) { perror("output"); return 2; }
	fwrite(image.data(), sizeof(image[0]), image.size(), o);
	fwrite(acc.data(), sizeof(double), acc.size(), o);
	fwrite(&stats, sizeof(stats), 1, o);
	fclose(o);
	printf("rendered %dx%d, %u primary rays, render time %.3f ms\n", w, h, stats.primaryRayCount, stats.renderTime * 1e3);
	return 0;
}
