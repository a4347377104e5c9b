/* rendercore.cpp - RenderCore_HIP host core.
 *
 * Host orchestration of lib/RenderCore_PrimeRef/rendercore.cpp (Init :100-142,
 * SetTarget :148-208, SetGeometry :214-222, SetInstance :228-242,
 * FinalizeInstances :248-277, SetMaterials :360-402, SetLights :408-460,
 * Setting :484-502, Render :535-632) on HIP: BVHs are built on the GPU per
 * SetGeometry, the TLAS is rebuilt when instances change, and a frame is
 * enqueued without host round trips: every wave's ray count stays in device
 * memory and the traversal and shade kernels read it there (the reference
 * copies the counters to the host after every bounce, rendercore.cpp:592).
 */
#include "rendercore.h"
#include "svgf.h"
#include "../../include/lh2hip.h"
#include <chrono>
#include <cstring>
#include <cmath>
#include <algorithm>

namespace lh {

using lighthouse2::CoreTri;
using lighthouse2::CoreMaterial;
using lighthouse2::CoreTexDesc;
using lighthouse2::CoreLightTri;
using lighthouse2::CorePointLight;
using lighthouse2::CoreSpotLight;
using lighthouse2::CoreDirectionalLight;
using lighthouse2::ViewPyramid;
using lighthouse2::GLTexture;
using lighthouse2::Convergence;
using lighthouse2::Restart;
using lighthouse2::Converge;
using lighthouse2::int2;
using lighthouse2::mat4;

static inline void hipCheck(hipError_t e, const char* what)
{
	if (e == hipErrorOutOfMemory) throw CoreError(LH2HIP_ERR_OOM, std::string(what) + ": out of device memory");
	if (e != hipSuccess) throw CoreError(LH2HIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::release()
{
	if (p) (void)hipFree(p);
	p = nullptr, bytes = 0;
}

void DevBuf::ensure(size_t n)
{
	if (n <= bytes && p) return;
	release();
	if (n == 0) return;
	hipCheck(hipMalloc(&p, n), "hipMalloc");
	bytes = n;
}

void DevBuf::upload(const void* src, size_t n, hipStream_t s)
{
	ensure(n);
	if (n) hipCheck(hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
}

/* Round-to-nearest-even float -> half (the reference stores material colours
 * as CUDA half, core_settings.h:112). */
static uint16_t to_half(float f)
{
	uint32_t x;
	memcpy(&x, &f, 4);
	const uint32_t sign = (x >> 16) & 0x8000u;
	x &= 0x7fffffffu;
	if (x >= 0x7f800000u) return (uint16_t)(sign | (x > 0x7f800000u ? 0x7e00u : 0x7c00u));
	if (x >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
	if (x < 0x38800000u)
	{
		if (x < 0x33000000u) return (uint16_t)sign;
		const uint32_t e = x >> 23, m = (x & 0x7fffffu) | 0x800000u, shift = 126u - e;
		uint32_t h = m >> shift;
		const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
		if (rem > half || (rem == half && (h & 1u))) h++;
		return (uint16_t)(sign | h);
	}
	uint32_t h = (x >> 13) - (112u << 10);
	const uint32_t rem = x & 0x1fffu;
	if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
	return (uint16_t)(sign | h);
}

/* TOCHAR (rendercore.cpp:358), saturated to one byte. */
static inline uint32_t to_byte(float a)
{
	const float s = a * 255.0f;
	if (!(s > 0.0f)) return 0;
	if (s >= 255.0f) return 255;
	return (uint32_t)s;
}

static uint32_t xor32_host(uint32_t& s) { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; } /* RandomUInt(seed), system.cpp:51 */

void Wavefront::resizeTarget(int pixels, hipStream_t s)
{
	if (pixels <= pixelCap) return;
	pixelCap = pixels + (pixels >> 4);
	/* two halves: the filter estimator keeps direct and indirect radiance apart */
	acc32B.ensure(sizeof(float4) * 2 * pixelCap);
	acc64B.ensure(sizeof(double) * 8 * pixelCap);
	gFeaturesB.ensure(16 * (size_t)pixelCap), gWorldPosB.ensure(16 * (size_t)pixelCap), gDeltaDepthB.ensure(16 * (size_t)pixelCap);
	/* the shade kernel writes the G-buffer only where a path reaches a diffuse
	 * surface; the filter still reads the other pixels, so they start at 0 */
	hipCheck(hipMemsetAsync(gFeaturesB.p, 0, gFeaturesB.bytes, s), "memset");
	hipCheck(hipMemsetAsync(gWorldPosB.p, 0, gWorldPosB.bytes, s), "memset");
	hipCheck(hipMemsetAsync(gDeltaDepthB.p, 0, gDeltaDepthB.bytes, s), "memset");
	imageB.ensure(sizeof(float4) * pixelCap);
	ensureHost();
	hipCheck(hipMemsetAsync(imageB.p, 0, imageB.bytes, s), "memset");
}

void Wavefront::ensureHost()
{
	if (hostW) return;
	void* p = nullptr;
	hipCheck(hipHostMalloc(&p, sizeof(uint32_t) * HOST_WORDS, hipHostMallocDefault), "hipHostMalloc");
	hostW = (uint32_t*)p;
	memset(hostW, 0, sizeof(uint32_t) * HOST_WORDS);
	memset(hostW + 4, 0xff, 3 * sizeof(uint32_t));   /* no probe hit yet */
}

void Wavefront::clearAccumulators(hipStream_t s)
{
	if (acc32B.p) hipCheck(hipMemsetAsync(acc32B.p, 0, acc32B.bytes, s), "memset");
	if (acc64B.p) hipCheck(hipMemsetAsync(acc64B.p, 0, acc64B.bytes, s), "memset");
}

void Wavefront::resizePaths(size_t n, hipStream_t s)
{
	if (n <= pathCap) return;
	pathCap = n;
	for (int i = 0; i < 2; i++)
		rays[i].ensure(32 * n), state[i].ensure(16 * n), pathIdB[i].ensure(4 * n), auxB[i].ensure(8 * n);
	hitsB.ensure(16 * n), shadowRaysB.ensure(32 * n), shadowPotB.ensure(16 * n), occludedB.ensure(4 * n), pathAccB.ensure(32 * n);
	countersB.ensure(64);
	ensureHost();                                 /* one log slot per wave, maxDepth <= 1024 */
	hipCheck(hipMemsetAsync(countersB.p, 0, 64, s), "memset");
}

void Wavefront::release()
{
	for (int i = 0; i < 2; i++) rays[i].release(), state[i].release(), pathIdB[i].release(), auxB[i].release();
	hitsB.release(), shadowRaysB.release(), shadowPotB.release(), occludedB.release(), pathAccB.release();
	countersB.release(), acc32B.release(), acc64B.release(), imageB.release();
	if (hostW) (void)hipHostFree(hostW), hostW = nullptr;
	gFeaturesB.release(), gWorldPosB.release(), gDeltaDepthB.release();
	pathCap = 0, pixelCap = 0;
}

PathBuffers Wavefront::buffers()
{
	PathBuffers b;
	for (int i = 0; i < 2; i++)
		b.rays[i] = rays[i].as<float4>(), b.state[i] = state[i].as<float4>(), b.pathId[i] = pathIdB[i].as<uint32_t>(), b.aux[i] = auxB[i].as<float2>();
	b.hits = hitsB.as<float4>(), b.shadowRays = shadowRaysB.as<float4>(), b.shadowPot = shadowPotB.as<float4>();
	b.occluded = occludedB.as<uint32_t>(), b.pathAcc = pathAccB.as<float4>(), b.counters = countersB.as<uint32_t>();
	return b;
}

RenderCore::RenderCore() {}
RenderCore::~RenderCore()
{
	try { Shutdown(); }
	catch (...) {}   /* a server still holds the core: its owner destroys it first */
}

void RenderCore::Init()
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw CoreError(LH2HIP_ERR_MISSING, "no HIP device");
	int dev = 0;
	hipCheck(hipGetDevice(&dev), "hipGetDevice");
	hipDeviceProp_t prop;
	hipCheck(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
	deviceCUs = prop.multiProcessorCount;
	stats.SMcount = (uint32_t)prop.multiProcessorCount;
	stats.ccMajor = (uint32_t)prop.major, stats.ccMinor = (uint32_t)prop.minor;
	stats.VRAM = (uint32_t)(prop.totalGlobalMem >> 20);
	static char name[256];
	strncpy(name, prop.name, 255);
	stats.deviceName = name;
	stats.probedTriid = -1;
	hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
	for (int i = 0; i < 8; i++) hipCheck(hipEventCreate(&ev[i]), "hipEventCreate");
	persistentBlocks = deviceCUs * 8;
	counters.ensure(2 * TRACE_QUEUE_WORDS * sizeof(uint32_t));   /* ray feeds: direct queries, render waves */
	hipCheck(hipMemsetAsync(counters.p, 0, counters.bytes, stream), "memset");
	int clockKHz = 0;
	if (hipDeviceGetAttribute(&clockKHz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && clockKHz > 0) wallClockNs = 1e6 / clockKHz;
	overflow.ensure(trace_overflow_bytes(persistentBlocks));
	hipCheck(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking), "hipStreamCreate");
	for (int i = 0; i < 2 + 2 * SIDE_TIMED; i++)
		hipCheck(i < 2 ? hipEventCreateWithFlags(&sideEv[i], hipEventDisableTiming) : hipEventCreate(&sideEv[i]), "hipEventCreate");
	sideCounters.ensure((4 + TRACE_QUEUE_WORDS) * sizeof(uint32_t));
	overflow2.ensure(trace_overflow_bytes(persistentBlocks));
}

void RenderCore::Shutdown()
{
	if (!stream) return;
	if (serverCount.load() > 0) throw CoreError(LH2HIP_ERR_STATE, "Shutdown: destroy the job servers of this core first");
	(void)hipStreamSynchronize(stream);
	(void)hipStreamSynchronize(stream2);
	meshes.clear();
	tlasNodes.release(), instRecs.release(), shadeInsts.release(), overflow.release(), counters.release();
	overflow2.release(), sideCounters.release();
	materials.release(), triLights.release(), pointLights.release(), spotLights.release(), dirLights.release();
	wf.release();
	svgf.reset();
	for (int i = 0; i < 8; i++) (void)hipEventDestroy(ev[i]);
	for (int i = 0; i < 2 + 2 * SIDE_TIMED; i++) (void)hipEventDestroy(sideEv[i]);
	(void)hipStreamDestroy(stream2);
	(void)hipStreamDestroy(stream);
	stream = stream2 = nullptr;
}

void RenderCore::requireNoJobs(const char* what) const
{
	if (pendingJobs.load() > 0)
		throw CoreError(LH2HIP_ERR_STATE, std::string(what) + ": a job server has jobs pending on this scene (wait for them first)");
}

void RenderCore::SetProbePos(const int2 pos) { probePos = pos; }

void RenderCore::SetTarget(GLTexture* target, const uint32_t spp)
{
	if (!target) throw CoreError(LH2HIP_ERR_INVALID, "SetTarget: null target");
	SetTargetSize((int)target->width, (int)target->height, (int)spp);
}

void RenderCore::SetTargetSize(int w, int h, int spp)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	if (w <= 0 || h <= 0 || spp <= 0 || spp > 65535 || (int64_t)w * h > (1 << 28))
		throw CoreError(LH2HIP_ERR_INVALID, "SetTarget: bad size");
	/* path indices (eye-ray jitter, RNG seeds, path slots) are 32-bit */
	if ((uint64_t)w * (uint64_t)h * (uint64_t)spp > ((uint64_t)1 << 32))
		throw CoreError(LH2HIP_ERR_INVALID, "SetTarget: width * height * spp must not exceed 2^32");
	scrw = w, scrh = h, scrspp = spp;
	wf.resizeTarget(w * h, stream);
	wf.clearAccumulators(stream);
	samplesTaken = 0;
}

void RenderCore::Setting(const char* name, float value)
{
	if (!name) throw CoreError(LH2HIP_ERR_INVALID, "Setting: null name");
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	if (!strcmp(name, "nodeFormat") || !strcmp(name, "refit")) requireNoJobs("Setting");
	if (!strcmp(name, "epsilon")) geometryEpsilon = value;
	else if (!strcmp(name, "clampValue")) clampValue = value;
	else if (!strcmp(name, "maxDepth"))
	{
		if (!(value >= 1 && value <= 1024)) throw CoreError(LH2HIP_ERR_INVALID, "Setting maxDepth: 1..1024");
		maxDepth = (int)value;
	}
	else if (!strcmp(name, "refit"))
	{	/* not in the reference (its OptiX BLAS is rebuilt on every SetGeometry) */
		if (value != 0 && value != 1) throw CoreError(LH2HIP_ERR_INVALID, "Setting refit: 0 (rebuild) or 1 (refit)");
		refitSetting = value != 0;
	}
	else if (!strcmp(name, "nodeFormat"))
	{
		if (value != 0 && value != 1) throw CoreError(LH2HIP_ERR_INVALID, "Setting nodeFormat: 0 (fp32) or 1 (q8)");
		nodeFormat = (int)value;
		instancesDirty = true;
	}
	else if (!strcmp(name, "traceLoop"))
	{
		if (value != 0 && value != 1 && value != 2) throw CoreError(LH2HIP_ERR_INVALID, "Setting traceLoop: 0 (merged), 1 (while-while) or 2 (greedy)");
		traceLoop = (int)value;
		scene.loop = traceLoop;
	}
	else if (!strcmp(name, "traceRefillLanes"))
	{
		if (!(value >= 1 && value <= 64)) throw CoreError(LH2HIP_ERR_INVALID, "Setting traceRefillLanes: 1..64");
		refillLanes = (int)value;
		scene.refillLanes = refillLanes;
	}
	else if (!strcmp(name, "traceChunk"))
	{
		if (!(value >= 1 && value <= 1024)) throw CoreError(LH2HIP_ERR_INVALID, "Setting traceChunk: 1..1024");
		chunk = (int)value;
		scene.chunk = chunk;
	}
	else if (!strcmp(name, "traceTriBias"))
	{
		if (!(value >= 1 && value <= 256)) throw CoreError(LH2HIP_ERR_INVALID, "Setting traceTriBias: 1..256");
		triBias = (int)value;
		scene.triBias = triBias;
	}
	else if (!strcmp(name, "shadowLane"))
	{
		if (value != 0 && value != 1) throw CoreError(LH2HIP_ERR_INVALID, "Setting shadowLane: 0 or 1");
		shadowLane = value != 0;
	}
	else if (!strcmp(name, "seed")) seedSetting = camSeed = (uint32_t)value; /* replaces camRNGseed's 0x12345678 */
	else if (!strcmp(name, "estimator"))
	{	/* 0: RenderCore_PrimeRef's shadeKernel, 1: rendercore_optix7's (Lambert BSDF) */
		if (value != 0 && value != 1 && value != 2) throw CoreError(LH2HIP_ERR_INVALID, "Setting estimator: 0 (reference), 1 (optix7) or 2 (optix7 filter)");
		estimator = (int)value;
	}
	else if (!strcmp(name, "filter")) filterOn = value != 0;          /* RenderCore_Optix7Filter/rendercore.cpp:676 */
	else if (!strcmp(name, "clampDirect")) clampDirect = value;      /* :674 */
	else if (!strcmp(name, "clampIndirect")) clampIndirect = value;  /* :675 */
	else if (!strcmp(name, "TAA"))
	{	/* :677 */
		if (value != 0 && value != 1) throw CoreError(LH2HIP_ERR_INVALID, "Setting TAA: 0 or 1");
		taaOn = value != 0;
	}
	else if (!strcmp(name, "bsdf"))
	{	/* optix7 estimator's BSDF: the compile-time choice of kernels/bsdf.h:7-22 */
		if (value != 0 && value != 1) throw CoreError(LH2HIP_ERR_INVALID, "Setting bsdf: 0 (Lambert) or 1 (Disney)");
		bsdfSetting = (int)value;
	}
	else if (!strcmp(name, "traceMinRaysPerLane"))
	{
		if (!(value >= 0 && value <= 64)) throw CoreError(LH2HIP_ERR_INVALID, "Setting traceMinRaysPerLane: 0..64");
		minRaysPerLane = value;
	}
	else if (!strcmp(name, "traceBlocksPerCU"))
	{
		if (!(value >= 1 && value <= 64)) throw CoreError(LH2HIP_ERR_INVALID, "Setting traceBlocksPerCU: 1..64");
		hipCheck(hipStreamSynchronize(stream), "sync");
		hipCheck(hipStreamSynchronize(stream2), "sync");
		persistentBlocks = deviceCUs * (int)value;
		overflow.ensure(trace_overflow_bytes(persistentBlocks));
		overflow2.ensure(trace_overflow_bytes(persistentBlocks));
	}
	/* unknown keys are ignored, as in the reference cores */
}

/* BLAS build, entirely on the device: primitive boxes, binned SAH level by
 * level, then the layout passes (depth-first f32 nodes, 4-wide q8 nodes, leaf
 * triangle order) and the leaf-triangle gather. The host reads back only the
 * per-level node counts, the node counts of both formats and the root box. */
void RenderCore::buildMesh(MeshData& m)
{
	const auto t0 = std::chrono::high_resolution_clock::now();
	const int n = m.triCount;
	DevBuf lo, hi;
	lo.ensure(sizeof(float4) * n), hi.ensure(sizeof(float4) * n);
	hipCheck(tri_bounds(m.verts.as<float4>(), n, lo.as<float4>(), hi.as<float4>(), stream), "tri_bounds");
	DeviceBuild b;
	hipCheck(gpu_build(lo.as<float4>(), hi.as<float4>(), n, 4, 255, b, stream), "gpu_build");
	lo.release(), hi.release();
	m.dBuildNodes.release(), m.dBuildPrims.release();
	m.dBuildNodes.p = b.nodes, m.dBuildNodes.bytes = sizeof(BuildNode) * (size_t)(2 * n);
	m.dBuildPrims.p = b.prims, m.dBuildPrims.bytes = sizeof(uint32_t) * (size_t)n;
	m.buildNodeCount = b.nodeCount;
	m.levelStart = b.levelStart;
	m.buildDepth = b.depth;
	const size_t interior = std::max(1, (b.nodeCount - 1) / 2);
	m.nodesF32.ensure(sizeof(NodeF32) * interior);
	m.nodesQ8.ensure(sizeof(NodeQ8) * interior);
	m.dF32Kids.ensure(sizeof(int32_t) * 2 * interior);
	m.dQ8Map.ensure(sizeof(int32_t) * 5 * interior);
	m.dOrder.ensure(sizeof(uint32_t) * n);
	m.dLast.ensure((size_t)n);
	LayoutOut L;
	L.f32 = m.nodesF32.as<NodeF32>(), L.f32Kids = m.dF32Kids.as<int32_t>();
	L.q8 = m.nodesQ8.as<NodeQ8>(), L.q8Map = m.dQ8Map.as<int32_t>();
	L.order = m.dOrder.as<uint32_t>(), L.last = m.dLast.as<uint8_t>();
	int counts[3];
	hipCheck(layout_device(m.dBuildNodes.as<BuildNode>(), m.dBuildPrims.as<uint32_t>(), b.nodeCount, b.levelStart, nullptr, L, counts, stream), "bvh layout");
	m.tris.ensure(sizeof(float4) * 3 * (size_t)n);
	hipCheck(gather_tris(m.verts.as<float4>(), m.dOrder.as<uint32_t>(), m.dLast.as<uint8_t>(), n, m.tris.as<float4>(), stream), "gather_tris");
	BuildNode root;
	hipCheck(hipMemcpyAsync(&root, m.dBuildNodes.p, sizeof(BuildNode), hipMemcpyDeviceToHost, stream), "D2H root");
	hipCheck(hipStreamSynchronize(stream), "sync");
	const bool rootLeaf = root.count > 0;
	m.nodeCountF32 = (size_t)counts[0], m.nodeCountQ8 = (size_t)counts[1];
	m.rootF32 = rootLeaf ? REF_LEAF : 0u, m.rootQ8 = rootLeaf ? REF_LEAF : 0u;
	m.depthF32 = b.depth, m.stackF32 = b.depth + 1;
	m.depthQ8 = counts[2], m.stackQ8 = 3 * counts[2] + 1;
	m.leafCount = (size_t)(b.nodeCount + 1) / 2;
	for (int c = 0; c < 3; c++) m.lo[c] = root.lo[c], m.hi[c] = root.hi[c];
	const auto t1 = std::chrono::high_resolution_clock::now();
	m.buildMs = std::chrono::duration<float, std::milli>(t1 - t0).count();
	stats.bvhBuildTime = m.buildMs * 0.001f;
}

/* Refit keeps the build tree's topology and recomputes its boxes from the new
 * vertices on the device: one launch per tree level from the deepest up (a
 * leaf box is the min / max of its triangles' vertices, as tri_bounds computes
 * them, an interior box the union of its two children), then the f32 and q8
 * nodes are rewritten from the refitted boxes through the layout maps and the
 * leaf triangles gathered again. */
void RenderCore::refitMesh(MeshData& m)
{
	const auto t0 = std::chrono::high_resolution_clock::now();
	for (int d = (int)m.levelStart.size() - 2; d >= 0; d--)
		hipCheck(refit_level(m.dBuildNodes.as<BuildNode>(), m.dBuildPrims.as<uint32_t>(), m.verts.as<float4>(),
			m.levelStart[d], m.levelStart[d + 1] - m.levelStart[d], stream), "refit_level");
	hipCheck(refit_layout(m.dBuildNodes.as<BuildNode>(), m.nodesF32.as<NodeF32>(), m.dF32Kids.as<int32_t>(), (int)m.nodeCountF32,
		m.nodesQ8.as<NodeQ8>(), m.dQ8Map.as<int32_t>(), (int)m.nodeCountQ8, stream), "refit_layout");
	hipCheck(gather_tris(m.verts.as<float4>(), m.dOrder.as<uint32_t>(), m.dLast.as<uint8_t>(), m.triCount, m.tris.as<float4>(), stream), "gather_tris");
	BuildNode root;
	hipCheck(hipMemcpyAsync(&root, m.dBuildNodes.p, sizeof(BuildNode), hipMemcpyDeviceToHost, stream), "D2H root");
	hipCheck(hipStreamSynchronize(stream), "sync");
	for (int c = 0; c < 3; c++) m.lo[c] = root.lo[c], m.hi[c] = root.hi[c];
	const auto t1 = std::chrono::high_resolution_clock::now();
	m.buildMs = std::chrono::duration<float, std::milli>(t1 - t0).count();
	stats.bvhBuildTime = m.buildMs * 0.001f;
}

void RenderCore::SetGeometry(const int meshIdx, const lighthouse2::float4* vertexData, const int vertexCount, const int triangleCount, const CoreTri* triangles)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("SetGeometry");
	if (meshIdx < 0 || meshIdx > (int)meshes.size()) throw CoreError(LH2HIP_ERR_INVALID, "SetGeometry: meshes must be added in sequential order");
	if (triangleCount < 0 || vertexCount != 3 * triangleCount || (triangleCount > 0 && !vertexData))
		throw CoreError(LH2HIP_ERR_INVALID, "SetGeometry: vertexCount must be 3 * triangleCount");
	if (triangleCount >= (1 << 29)) throw CoreError(LH2HIP_ERR_INVALID, "SetGeometry: too many triangles");
	if (meshIdx == (int)meshes.size()) meshes.emplace_back();
	MeshData& m = meshes[meshIdx];
	/* animation (rendercore_optix7/core_mesh.cpp:34-61 rebuilds): with the refit setting and an
	 * unchanged triangle count the existing tree is refitted */
	const bool refit = refitSetting && triangleCount > 0 && m.triCount == triangleCount && m.buildNodeCount > 0;
	m.triCount = triangleCount;
	m.verts.upload(vertexData, sizeof(float4) * 3 * (size_t)triangleCount, stream);
	m.maxMaterial = -1;
	if (triangles)
	{
		m.coreTris.upload(triangles, sizeof(CoreTri) * (size_t)triangleCount, stream);
		for (int i = 0; i < triangleCount; i++) m.maxMaterial = std::max(m.maxMaterial, (int)std::min<uint32_t>(triangles[i].material, 0x7fffffffu));
	}
	else m.coreTris.release();
	if (refit) refitMesh(m);
	else if (triangleCount > 0) buildMesh(m);
	else
	{
		m.dBuildNodes.release(), m.dBuildPrims.release(), m.dF32Kids.release(), m.dQ8Map.release(), m.dOrder.release(), m.dLast.release();
		m.buildNodeCount = 0, m.levelStart.clear(), m.nodeCountF32 = m.nodeCountQ8 = 0;
	}
	instancesDirty = true;
}

void RenderCore::SetInstance(const int instanceIdx, const int modelIdx, const mat4& transform)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("SetInstance");
	if (modelIdx == -1)
	{
		if ((int)instances.size() > instanceIdx && instanceIdx >= 0) instances.resize(instanceIdx);
		instancesDirty = true;
		return;
	}
	if (instanceIdx < 0 || instanceIdx > (int)instances.size()) throw CoreError(LH2HIP_ERR_INVALID, "SetInstance: instances must be added in sequential order");
	if (modelIdx < 0 || modelIdx >= (int)meshes.size()) throw CoreError(LH2HIP_ERR_INVALID, "SetInstance: unknown mesh");
	if (instanceIdx == (int)instances.size()) instances.emplace_back();
	instances[instanceIdx].mesh = modelIdx;
	memcpy(instances[instanceIdx].T, transform.cell, 64);
	instancesDirty = true;
}

/* mat4::Inverted (common_types.h:559-606): MESA cofactors; singular -> identity. */
static void invert(const float* c, float* out)
{
	float inv[16];
	inv[0] = c[5] * c[10] * c[15] - c[5] * c[11] * c[14] - c[9] * c[6] * c[15] + c[9] * c[7] * c[14] + c[13] * c[6] * c[11] - c[13] * c[7] * c[10];
	inv[1] = -c[1] * c[10] * c[15] + c[1] * c[11] * c[14] + c[9] * c[2] * c[15] - c[9] * c[3] * c[14] - c[13] * c[2] * c[11] + c[13] * c[3] * c[10];
	inv[2] = c[1] * c[6] * c[15] - c[1] * c[7] * c[14] - c[5] * c[2] * c[15] + c[5] * c[3] * c[14] + c[13] * c[2] * c[7] - c[13] * c[3] * c[6];
	inv[3] = -c[1] * c[6] * c[11] + c[1] * c[7] * c[10] + c[5] * c[2] * c[11] - c[5] * c[3] * c[10] - c[9] * c[2] * c[7] + c[9] * c[3] * c[6];
	inv[4] = -c[4] * c[10] * c[15] + c[4] * c[11] * c[14] + c[8] * c[6] * c[15] - c[8] * c[7] * c[14] - c[12] * c[6] * c[11] + c[12] * c[7] * c[10];
	inv[5] = c[0] * c[10] * c[15] - c[0] * c[11] * c[14] - c[8] * c[2] * c[15] + c[8] * c[3] * c[14] + c[12] * c[2] * c[11] - c[12] * c[3] * c[10];
	inv[6] = -c[0] * c[6] * c[15] + c[0] * c[7] * c[14] + c[4] * c[2] * c[15] - c[4] * c[3] * c[14] - c[12] * c[2] * c[7] + c[12] * c[3] * c[6];
	inv[7] = c[0] * c[6] * c[11] - c[0] * c[7] * c[10] - c[4] * c[2] * c[11] + c[4] * c[3] * c[10] + c[8] * c[2] * c[7] - c[8] * c[3] * c[6];
	inv[8] = c[4] * c[9] * c[15] - c[4] * c[11] * c[13] - c[8] * c[5] * c[15] + c[8] * c[7] * c[13] + c[12] * c[5] * c[11] - c[12] * c[7] * c[9];
	inv[9] = -c[0] * c[9] * c[15] + c[0] * c[11] * c[13] + c[8] * c[1] * c[15] - c[8] * c[3] * c[13] - c[12] * c[1] * c[11] + c[12] * c[3] * c[9];
	inv[10] = c[0] * c[5] * c[15] - c[0] * c[7] * c[13] - c[4] * c[1] * c[15] + c[4] * c[3] * c[13] + c[12] * c[1] * c[7] - c[12] * c[3] * c[5];
	inv[11] = -c[0] * c[5] * c[11] + c[0] * c[7] * c[9] + c[4] * c[1] * c[11] - c[4] * c[3] * c[9] - c[8] * c[1] * c[7] + c[8] * c[3] * c[5];
	inv[12] = -c[4] * c[9] * c[14] + c[4] * c[10] * c[13] + c[8] * c[5] * c[14] - c[8] * c[6] * c[13] - c[12] * c[5] * c[10] + c[12] * c[6] * c[9];
	inv[13] = c[0] * c[9] * c[14] - c[0] * c[10] * c[13] - c[8] * c[1] * c[14] + c[8] * c[2] * c[13] + c[12] * c[1] * c[10] - c[12] * c[2] * c[9];
	inv[14] = -c[0] * c[5] * c[14] + c[0] * c[6] * c[13] + c[4] * c[1] * c[14] - c[4] * c[2] * c[13] - c[12] * c[1] * c[6] + c[12] * c[2] * c[5];
	inv[15] = c[0] * c[5] * c[10] - c[0] * c[6] * c[9] - c[4] * c[1] * c[10] + c[4] * c[2] * c[9] + c[8] * c[1] * c[6] - c[8] * c[2] * c[5];
	const float det = c[0] * inv[0] + c[1] * inv[4] + c[2] * inv[8] + c[3] * inv[12];
	for (int i = 0; i < 16; i++) out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
	if (det != 0)
	{
		const float invdet = 1.0f / det;
		for (int i = 0; i < 16; i++) out[i] = inv[i] * invdet;
	}
}

void RenderCore::FinalizeInstances()
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("FinalizeInstances");
	instancesDirty = true;
	updateTopLevel();
}

/* TLAS over instance world boxes: 8 corners of the mesh box through the
 * instance transform (an identity instance keeps the mesh box), built with the
 * GPU builder, one instance per leaf. */
void RenderCore::updateTopLevel()
{
	if (!instancesDirty) return;
	const int ni = (int)instances.size();
	std::vector<InstanceRec> recs(ni);
	std::vector<ShadeInst> sh(ni);
	std::vector<float4> blo, bhi;
	std::vector<int> live;
	int maxStack = 0;
	for (int i = 0; i < ni; i++)
	{
		const InstanceData& I = instances[i];
		const MeshData& m = meshes[I.mesh];
		float inv[16];
		invert(I.T, inv);
		bool identity = true;
		for (int k = 0; k < 16; k++) if (I.T[k] != ((k % 5 == 0) ? 1.0f : 0.0f)) identity = false;
		InstanceRec& r = recs[i];
		r.inv0 = make_float4(inv[0], inv[1], inv[2], inv[3]);
		r.inv1 = make_float4(inv[4], inv[5], inv[6], inv[7]);
		r.inv2 = make_float4(inv[8], inv[9], inv[10], inv[11]);
		r.nodes = nodeFormat == FMT_Q8 ? m.nodesQ8.p : m.nodesF32.p;
		r.tris = m.tris.as<float4>();
		r.root = nodeFormat == FMT_Q8 ? m.rootQ8 : m.rootF32;
		r.flags = identity ? 1u : 0u;
		ShadeInst& s = sh[i];
		s.inv0 = r.inv0, s.inv1 = r.inv1, s.inv2 = r.inv2;
		s.coreTris = m.coreTris.as<float4>();
		s.verts = m.verts.as<float4>();
		s.mesh = I.mesh;
		if (m.triCount == 0) continue;
		maxStack = std::max(maxStack, nodeFormat == FMT_Q8 ? m.stackQ8 : m.stackF32);
		float lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
		if (identity) for (int c = 0; c < 3; c++) lo[c] = m.lo[c], hi[c] = m.hi[c];
		else
			for (int c = 0; c < 8; c++)
			{
				const float p[3] = { (c & 1) ? m.hi[0] : m.lo[0], (c & 2) ? m.hi[1] : m.lo[1], (c & 4) ? m.hi[2] : m.lo[2] };
				for (int a = 0; a < 3; a++)
				{
					const float w = I.T[a * 4] * p[0] + I.T[a * 4 + 1] * p[1] + I.T[a * 4 + 2] * p[2] + I.T[a * 4 + 3];
					lo[a] = std::min(lo[a], w), hi[a] = std::max(hi[a], w);
				}
			}
		blo.push_back(make_float4(lo[0], lo[1], lo[2], 0)), bhi.push_back(make_float4(hi[0], hi[1], hi[2], 0));
		live.push_back(i);
	}
	instRecs.upload(recs.data(), recs.size() * sizeof(InstanceRec), stream);
	shadeInsts.upload(sh.data(), sh.size() * sizeof(ShadeInst), stream);
	scene = TraceScene();
	scene.format = nodeFormat;
	scene.loop = traceLoop;
	scene.triBias = triBias;
	scene.refillLanes = refillLanes;
	scene.chunk = chunk;
	scene.inst = instRecs.as<InstanceRec>();
	sceneReady = !live.empty();
	if (live.size() == 1 && recs[live[0]].flags == 1u && live[0] == 0)
	{	/* one identity instance: traverse its BLAS directly */
		scene.twoLevel = 0;
		scene.nodes = recs[0].nodes, scene.tris = recs[0].tris, scene.root = recs[0].root;
		scene.maxStack = maxStack;
	}
	else if (!live.empty())
	{
		const int n = (int)live.size();
		DevBuf dlo, dhi;
		dlo.upload(blo.data(), n * sizeof(float4), stream);
		dhi.upload(bhi.data(), n * sizeof(float4), stream);
		DevBuf liveMap;
		liveMap.upload(live.data(), n * sizeof(int), stream);
		DeviceBuild b;
		hipCheck(gpu_build(dlo.as<float4>(), dhi.as<float4>(), n, 1, 1, b, stream), "tlas build");
		/* one instance per leaf: leaf refs point straight at the instance id */
		tlasNodes.ensure(sizeof(NodeF32) * std::max(1, (b.nodeCount - 1) / 2));
		LayoutOut L;
		L.f32 = tlasNodes.as<NodeF32>();
		const hipError_t e = layout_device(b.nodes, b.prims, b.nodeCount, b.levelStart, liveMap.as<uint32_t>(), L, nullptr, stream);
		const int depth = b.depth;
		free_build(b);
		hipCheck(e, "tlas layout");
		scene.twoLevel = 1;
		scene.tlas = tlasNodes.as<NodeF32>();
		scene.tlasRoot = n == 1 ? (REF_LEAF | (uint32_t)live[0]) : 0u;
		scene.maxStack = maxStack + depth + 1;
	}
	if (scene.maxStack > TRACE_MAX_STACK) throw CoreError(LH2HIP_ERR_INVALID, "BVH too deep for the traversal stack");
	hipCheck(hipStreamSynchronize(stream), "sync");
	instancesDirty = false;
}

/* Persistent grid for a wave of `count` rays: the full grid, or fewer blocks
 * (not below one per CU) so that each lane gets at least minRaysPerLane rays. */
int RenderCore::blocksFor(int count) const
{
	if (minRaysPerLane <= 0.0f) return persistentBlocks;
	const int want = (int)((double)count / (TRACE_BLOCK * (double)minRaysPerLane)) + 1;
	return std::max(std::min(deviceCUs, persistentBlocks), std::min(persistentBlocks, want));
}

const TraceScene& RenderCore::Scene()
{
	updateTopLevel();
	if (!sceneReady) throw CoreError(LH2HIP_ERR_STATE, "no scene: call SetGeometry, SetInstance and FinalizeInstances first");
	return scene;
}

void RenderCore::TraceClosest(const void* rays, void* hits, int count, hipStream_t s)
{
	if (count < 0 || (count > 0 && (!rays || !hits))) throw CoreError(LH2HIP_ERR_INVALID, "trace: bad arguments");
	const TraceScene& sc = Scene();
	hipCheck(launch_trace_dev(sc, MODE_CLOSEST, (const float4*)rays, hits, count, nullptr, counters.as<uint32_t>(), overflow.as<uint32_t>(), blocksFor(count), s ? s : 0), "trace");
}

void RenderCore::TraceClosestCounted(const void* rays, void* hits, int count, hipStream_t s, uint64_t* out2)
{
	if (count < 0 || (count > 0 && (!rays || !hits))) throw CoreError(LH2HIP_ERR_INVALID, "trace: bad arguments");
	const TraceScene& sc = Scene();
	DevBuf st;
	st.ensure(sizeof(TraceStats));
	hipCheck(hipMemsetAsync(st.p, 0, sizeof(TraceStats), s), "memset");
	hipCheck(hipMemsetAsync((char*)st.p + offsetof(TraceStats, tStart), 0xff, 2 * sizeof(unsigned long long), s), "memset");
	hipCheck(launch_trace_dev(sc, MODE_CLOSEST, (const float4*)rays, hits, count, nullptr, counters.as<uint32_t>(),
		overflow.as<uint32_t>(), blocksFor(count), s, st.as<TraceStats>()), "trace");
	TraceStats h;
	hipCheck(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s), "D2H");
	hipCheck(hipStreamSynchronize(s), "sync");
	out2[0] += h.nodes, out2[1] += h.tris;
	out2[2] += h.nodeIters, out2[3] += h.nodeLanes, out2[4] += h.leafIters, out2[5] += h.leafLanes;
	out2[6] += h.loopIters, out2[7] += h.loopLanes, out2[8] += h.tailIters, out2[9] += h.tailLanes;
	if (h.tEnd > h.tStart && h.tDrained >= h.tStart && h.tDrained != ~0ull)
		out2[10] += (uint64_t)((h.tDrained - h.tStart) * wallClockNs), out2[11] += (uint64_t)((h.tEnd - h.tStart) * wallClockNs);
}

void RenderCore::TraceAny(const void* rays, uint32_t* mask, int count, hipStream_t s)
{
	if (count < 0 || (count > 0 && (!rays || !mask))) throw CoreError(LH2HIP_ERR_INVALID, "trace: bad arguments");
	const TraceScene& sc = Scene();
	hipCheck(hipMemsetAsync(mask, 0, sizeof(uint32_t) * ((count + 31) / 32), s), "memset");
	hipCheck(launch_trace_dev(sc, MODE_ANY_MASK, (const float4*)rays, mask, count, nullptr, counters.as<uint32_t>(), overflow.as<uint32_t>(), blocksFor(count), s ? s : 0), "trace");
}

void RenderCore::Bounce(const void* raysIn, const void* hits, const uint32_t* pathIn, int count, uint32_t R0, float tmin,
	void* raysOut, uint32_t* pathOut, uint32_t* countOut, hipStream_t s)
{
	Scene();
	hipCheck(launch_bounce(shadeInsts.as<ShadeInst>(), (const float4*)raysIn, (const float4*)hits, pathIn, count, R0, tmin,
		(float4*)raysOut, pathOut, countOut, s), "bounce");
}

void RenderCore::BvhInfo(int mesh, int64_t* o) const
{
	if (mesh < 0 || mesh >= (int)meshes.size()) throw CoreError(LH2HIP_ERR_INVALID, "bvh_info: unknown mesh");
	const MeshData& m = meshes[mesh];
	o[0] = (int64_t)m.buildNodeCount, o[1] = (int64_t)m.leafCount, o[2] = m.buildDepth, o[3] = (int64_t)m.nodeCountF32;
	o[4] = (int64_t)m.nodeCountQ8, o[5] = m.depthQ8, o[6] = (int64_t)(m.buildMs * 1000.0f), o[7] = m.triCount;
}

int RenderCore::BvhDownload(int mesh, void* nodesOut, int maxNodes, uint32_t* primsOut, int maxPrims) const
{
	if (mesh < 0 || mesh >= (int)meshes.size()) throw CoreError(LH2HIP_ERR_INVALID, "bvh_download: unknown mesh");
	const MeshData& m = meshes[mesh];
	if (nodesOut && maxNodes >= m.buildNodeCount && m.buildNodeCount > 0)
		hipCheck(hipMemcpyAsync(nodesOut, m.dBuildNodes.p, (size_t)m.buildNodeCount * sizeof(BuildNode), hipMemcpyDeviceToHost, stream), "D2H");
	if (primsOut && maxPrims >= m.triCount && m.buildNodeCount > 0)
		hipCheck(hipMemcpyAsync(primsOut, m.dBuildPrims.p, (size_t)m.triCount * 4, hipMemcpyDeviceToHost, stream), "D2H");
	hipCheck(hipStreamSynchronize(stream), "sync");
	return m.buildNodeCount;
}

void RenderCore::BvhLayoutDownload(int mesh, void* f32Out, int maxF32, void* q8Out, int maxQ8, void* trisOut, int maxTris) const
{
	if (mesh < 0 || mesh >= (int)meshes.size()) throw CoreError(LH2HIP_ERR_INVALID, "bvh_layout_download: unknown mesh");
	const MeshData& m = meshes[mesh];
	if ((f32Out && maxF32 < (int)m.nodeCountF32) || (q8Out && maxQ8 < (int)m.nodeCountQ8) || (trisOut && maxTris < m.triCount))
		throw CoreError(LH2HIP_ERR_INVALID, "bvh_layout_download: output array too small");
	if (f32Out && m.nodeCountF32) hipCheck(hipMemcpyAsync(f32Out, m.nodesF32.p, m.nodeCountF32 * sizeof(NodeF32), hipMemcpyDeviceToHost, stream), "D2H");
	if (q8Out && m.nodeCountQ8) hipCheck(hipMemcpyAsync(q8Out, m.nodesQ8.p, m.nodeCountQ8 * sizeof(NodeQ8), hipMemcpyDeviceToHost, stream), "D2H");
	if (trisOut && m.triCount) hipCheck(hipMemcpyAsync(trisOut, m.tris.p, (size_t)m.triCount * 3 * sizeof(float4), hipMemcpyDeviceToHost, stream), "D2H");
	hipCheck(hipStreamSynchronize(stream), "sync");
}

/* HostTexture::PixelsNeeded (host_texture.cpp:117-122). */
static uint64_t mip_texels(uint32_t w, uint32_t h, int levels)
{
	uint64_t n = 0;
	for (int i = 0; i < levels; i++) n += (uint64_t)w * h, w >>= 1, h >>= 1;
	return n;
}

void RenderCore::SetMaterials(CoreMaterial* mat, const int materialCount)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("SetMaterials");
	if (materialCount < 0 || (materialCount > 0 && !mat)) throw CoreError(LH2HIP_ERR_INVALID, "SetMaterials: bad arguments");
	std::vector<GpuMaterial> g(std::max(1, materialCount));
	memset(g.data(), 0, g.size() * sizeof(GpuMaterial));
	for (int i = 0; i < materialCount; i++)
	{
		const CoreMaterial& m = mat[i];
		GpuMaterial& d = g[i];
		const uint32_t dr = to_half(m.color.value.x), dg = to_half(m.color.value.y), db = to_half(m.color.value.z);
		const uint32_t tr = to_half(1.0f - m.absorption.value.x), tg = to_half(1.0f - m.absorption.value.y), tb = to_half(1.0f - m.absorption.value.z);
		d.diffuse_rg = dr | (dg << 16);
		d.diffuse_b_medium_r = db | (tr << 16);
		d.medium_gb = tg | (tb << 16);
		d.params.x = to_byte(m.metallic.value) + (to_byte(m.subsurface.value) << 8) + (to_byte(m.specular.value) << 16) + (to_byte(m.roughness.value) << 24);
		d.params.y = to_byte(m.specularTint.value) + (to_byte(m.anisotropic.value) << 8) + (to_byte(m.sheen.value) << 16) + (to_byte(m.sheenTint.value) << 24);
		d.params.z = to_byte(m.clearcoat.value) + (to_byte(m.clearcoatGloss.value) << 8) + (to_byte(m.transmission.value) << 16);
		memcpy(&d.params.w, &m.eta.value, 4);
		/* CUDAMaterial::Map (rendercore.h:82-89): width | height << 16, half
		 * uv scale, half uv offset, first texel */
		const auto map = [&](uint4& q, int id, const decltype(m.color.uvscale)& sc, const decltype(m.color.uvoffset)& of, bool mipmapped, int storage = LH2_ARGB32) {
			if (id < 0) return;
			if (id >= (int)textures.size()) throw CoreError(LH2HIP_ERR_STATE, "SetMaterials: texture id beyond SetTextures (set textures first)");
			const TexInfo& t = textures[id];
			/* the kernels read texels (x, y) < (w, h) of level 0, and for a
			 * trilinear map of all 5 levels: reject maps they would read past */
			const uint64_t need = mipmapped ? mip_texels(t.width, t.height, 5) : (uint64_t)t.width * t.height;
			if (t.width == 0 || t.height == 0 || t.width > 32767 || t.height > 32767 || (mipmapped && ((t.width >> 4) == 0 || (t.height >> 4) == 0)))
				throw CoreError(LH2HIP_ERR_INVALID, "SetMaterials: texture size unsupported (1..32767; a diffuse map needs at least 16x16 for 5 MIP levels)");
			if (t.pixelCount < need) throw CoreError(LH2HIP_ERR_INVALID, "SetMaterials: texture holds fewer texels than its size (and MIP levels) need");
			if (t.storage != storage) throw CoreError(LH2HIP_ERR_INVALID, "SetMaterials: normal maps need NRM32 texels, other maps ARGB32");
			q.x = (t.width & 0xffffu) | ((t.height & 0xffffu) << 16);
			q.y = (uint32_t)to_half(sc.x) | ((uint32_t)to_half(sc.y) << 16);
			q.z = (uint32_t)to_half(of.x) | ((uint32_t)to_half(of.y) << 16);
			q.w = t.firstPixel;
		};
		map(d.t0, m.color.textureID, m.color.uvscale, m.color.uvoffset, true);
		map(d.t1, m.detailColor.textureID, m.detailColor.uvscale, m.detailColor.uvoffset, false);
		map(d.n0, m.normals.textureID, m.normals.uvscale, m.normals.uvoffset, false, LH2_NRM32);
		map(d.n1, m.detailNormals.textureID, m.detailNormals.uvscale, m.detailNormals.uvoffset, false, LH2_NRM32);
		map(d.s, m.specular.textureID, m.specular.uvscale, m.specular.uvoffset, false);
		map(d.r, m.roughness.textureID, m.roughness.uvscale, m.roughness.uvoffset, false);
		const bool hdr = m.color.textureID >= 0 && (textures[m.color.textureID].flags & 8u);
		d.flags = (m.eta.value < 1 ? MF_ISDIELECTRIC : 0) + (hdr ? MF_DIFFUSEMAPISHDR : 0) + (m.color.textureID != -1 ? MF_HASDIFFUSEMAP : 0) +
			(m.normals.textureID != -1 ? MF_HASNORMALMAP : 0) + (m.specular.textureID != -1 ? MF_HASSPECULARITYMAP : 0) +
			(m.roughness.textureID != -1 ? MF_HASROUGHNESSMAP : 0) + (m.detailNormals.textureID != -1 ? MF_HAS2NDNORMALMAP : 0) +
			(m.detailColor.textureID != -1 ? MF_HAS2NDDIFFUSEMAP : 0) + ((m.flags & 1) ? MF_HASSMOOTHNORMALS : 0) + ((m.flags & 2) ? MF_HASALPHA : 0);
	}
	materials.upload(g.data(), g.size() * sizeof(GpuMaterial), stream);
	this->materialCount = materialCount;
}

void RenderCore::SetLights(const CoreLightTri* tl, const int nt, const CorePointLight* pl, const int np,
	const CoreSpotLight* sl, const int ns, const CoreDirectionalLight* dl, const int nd)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("SetLights");
	if (nt < 0 || np < 0 || ns < 0 || nd < 0) throw CoreError(LH2HIP_ERR_INVALID, "SetLights: negative count");
	if (nt) triLights.upload(tl, nt * sizeof(CoreLightTri), stream);
	if (np) pointLights.upload(pl, np * sizeof(CorePointLight), stream);
	if (ns) spotLights.upload(sl, ns * sizeof(CoreSpotLight), stream);
	if (nd) dirLights.upload(dl, nd * sizeof(CoreDirectionalLight), stream);
	lightCounts[0] = nt, lightCounts[1] = np, lightCounts[2] = ns, lightCounts[3] = nd;
}

void RenderCore::SetTextures(const CoreTexDesc* tex, const int textureCount)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("SetTextures");
	if (textureCount < 0 || (textureCount > 0 && !tex)) throw CoreError(LH2HIP_ERR_INVALID, "SetTextures: bad arguments");
	/* SyncStorageType (rendercore.cpp:310-355): one continuous array per storage
	 * kind, textures in order, firstPixel = start in that array. ARGB128 texels
	 * are copied whole (float4 each). */
	std::vector<uint32_t> a32, n32;
	std::vector<float4> a128;
	textures.assign(textureCount, TexInfo());
	for (int i = 0; i < textureCount; i++)
	{
		const CoreTexDesc& t = tex[i];
		if (t.pixelCount > 0 && !t.idata) throw CoreError(LH2HIP_ERR_INVALID, "SetTextures: texture without texels");
		if (t.storage < LH2_ARGB32 || t.storage > LH2_NRM32) throw CoreError(LH2HIP_ERR_INVALID, "SetTextures: unknown texel storage");
		TexInfo& d = textures[i];
		d.width = t.width, d.height = t.height, d.flags = t.flags, d.storage = t.storage, d.pixelCount = t.pixelCount;
		if (t.storage == LH2_ARGB128)
		{
			d.firstPixel = (uint32_t)a128.size();
			const float4* src = (const float4*)t.fdata;
			a128.insert(a128.end(), src, src + t.pixelCount);
		}
		else
		{
			std::vector<uint32_t>& dst = t.storage == LH2_NRM32 ? n32 : a32;
			d.firstPixel = (uint32_t)dst.size();
			dst.insert(dst.end(), t.idata, t.idata + t.pixelCount);
		}
	}
	a32.resize(std::max<size_t>(a32.size(), 16)), n32.resize(std::max<size_t>(n32.size(), 16)), a128.resize(std::max<size_t>(a128.size(), 16));
	texARGB32.upload(a32.data(), a32.size() * 4, stream);
	texNRM32.upload(n32.data(), n32.size() * 4, stream);
	texARGB128.upload(a128.data(), a128.size() * 16, stream);
	hipCheck(hipStreamSynchronize(stream), "sync");
	stats.argb32TexelCount = 0, stats.argb128TexelCount = 0, stats.nrm32TexelCount = 0;
	for (int i = 0; i < textureCount; i++)
	{
		if (tex[i].storage == LH2_ARGB32) stats.argb32TexelCount += tex[i].pixelCount;
		else if (tex[i].storage == LH2_ARGB128) stats.argb128TexelCount += tex[i].pixelCount;
		else stats.nrm32TexelCount += tex[i].pixelCount;
	}
}

void RenderCore::SetSkyData(const lighthouse2::float3* pixels, const uint32_t width, const uint32_t height, const mat4& worldToLight)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("SetSkyData");
	/* rendercore.cpp:466-478: float3 texels stored as float4 (w = 0), followed
	 * by the 64x64-box downsampled dome the optix7 estimator samples after a
	 * diffuse bounce (rendercore_optix7/rendercore.cpp:716-740; float sums, rows
	 * outer, columns inner, then times 1/4096). */
	const size_t n = (size_t)width * height;
	if (n > 0 && !pixels) throw CoreError(LH2HIP_ERR_INVALID, "SetSkyData: no pixels");
	const uint32_t sw = width >> 6, sh = height >> 6;
	std::vector<float4> px(std::max<size_t>(n + (size_t)sw * sh, 1), make_float4(0, 0, 0, 0));
	for (size_t i = 0; i < n; i++) px[i] = make_float4(pixels[i].x, pixels[i].y, pixels[i].z, 0.0f);
	for (uint32_t y = 0; y < sh; y++) for (uint32_t x = 0; x < sw; x++)
	{
		float t0 = 0, t1 = 0, t2 = 0, t3 = 0;
		const float4* tile = px.data() + (size_t)x * 64 + (size_t)y * 64 * width;
		for (int v = 0; v < 64; v++) for (int u = 0; u < 64; u++)
		{
			const float4 q = tile[u + (size_t)v * width];
			t0 += q.x, t1 += q.y, t2 += q.z, t3 += q.w;
		}
		const float k = 1.0f / (64 * 64);
		px[n + x + (size_t)y * sw] = make_float4(t0 * k, t1 * k, t2 * k, t3 * k);
	}
	skyPixels.upload(px.data(), px.size() * sizeof(float4), stream);
	hipCheck(hipStreamSynchronize(stream), "sync");
	skyW = (int)width, skyH = (int)height;
	for (int r = 0; r < 3; r++) for (int c = 0; c < 4; c++) worldToSky[4 * r + c] = worldToLight.cell[4 * r + c];
}

void RenderCore::SetBlueNoise(const uint8_t* sequence, const uint8_t* scrambling, const uint8_t* ranking)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);
	requireNoJobs("SetBlueNoise");
	if (!sequence || !scrambling || !ranking) throw CoreError(LH2HIP_ERR_INVALID, "SetBlueNoise: null table");
	std::vector<uint32_t> bn(BLUE_NOISE_WORDS, 0u);
	for (int i = 0; i < 65536; i++) bn[i] = sequence[i];
	for (int i = 0; i < 128 * 128 * 8; i++) bn[65536 + i] = scrambling[i], bn[3 * 65536 + i] = ranking[i];
	blueNoise.upload(bn.data(), bn.size() * 4, stream);
	hipCheck(hipStreamSynchronize(stream), "sync");
}

void RenderCore::checkRenderable()
{
	for (const InstanceData& I : instances)
	{
		const MeshData& m = meshes[I.mesh];
		if (m.triCount == 0) continue;
		if (m.maxMaterial < 0) throw CoreError(LH2HIP_ERR_STATE, "Render: a mesh was set without CoreTri data (traversal-only geometry)");
		if (m.maxMaterial >= materialCount) throw CoreError(LH2HIP_ERR_STATE, "Render: a triangle references a material beyond SetMaterials");
	}
}

/* Shade and camera constants of one Render call (the reference's staged
 * __constant__ globals, lib/RenderCore_PrimeRef/rendercore.cpp:535-584). */
FrameSetup RenderCore::frameParams(const ViewPyramid& view, int w, int h, int spp, int depth, uint32_t R0gen, int samplesTakenAtStart,
	float jitterX, float jitterY, int estimatorOverride)
{
	const int estimator = estimatorOverride >= 0 ? estimatorOverride : this->estimator;
	FrameSetup f;
	memcpy(f.cam.pos, &view.pos, 12), memcpy(f.cam.p1, &view.p1, 12), memcpy(f.cam.p2, &view.p2, 12), memcpy(f.cam.p3, &view.p3, 12);
	f.cam.aperture = view.aperture, f.cam.spreadAngle = view.spreadAngle, f.cam.distortion = view.distortion, f.cam.w = w, f.cam.h = h;
	ShadeParams& P = f.P;
	memset(&P, 0, sizeof(P));
	P.inst = shadeInsts.as<ShadeInst>();
	P.mats = materials.as<GpuMaterial>();
	P.triLights = triLights.as<float4>(), P.pointLights = pointLights.as<float4>();
	P.spotLights = spotLights.as<float4>(), P.dirLights = dirLights.as<float4>();
	P.nTri = lightCounts[0], P.nPoint = lightCounts[1], P.nSpot = lightCounts[2], P.nDir = lightCounts[3];
	P.w = w, P.h = h, P.maxDepth = depth, P.pass = samplesTakenAtStart, P.eps = geometryEpsilon;
	P.probePixel = -1;
	P.argb32 = texARGB32.as<uint32_t>(), P.argb128 = texARGB128.as<float4>(), P.nrm32 = texNRM32.as<uint32_t>();
	P.spreadAngle = view.spreadAngle;
	P.sky = skyPixels.as<float4>(), P.skyW = skyW, P.skyH = skyH;
	P.w2s0 = make_float4(worldToSky[0], worldToSky[1], worldToSky[2], 0.0f);
	P.w2s1 = make_float4(worldToSky[4], worldToSky[5], worldToSky[6], 0.0f);
	P.w2s2 = make_float4(worldToSky[8], worldToSky[9], worldToSky[10], 0.0f);
	P.clampValue = clampValue;
	P.bsdf = bsdfSetting;
	P.connections = (P.nTri + P.nPoint + P.nSpot + P.nDir) > 0;   /* noDirectLightsInScene, rendercore_optix7/rendercore.cpp:708 */
	f.estimator = estimator;
	if (estimator == ESTIMATOR_OPTIX7 || estimator == ESTIMATOR_FILTER)
	{
		std::lock_guard<std::mutex> lockNoise(blueNoiseMutex);   /* job servers call this from their submitting threads */
		if (!blueNoise.p)
		{
			std::vector<uint32_t> bn(BLUE_NOISE_WORDS);
			blue_noise_tables(bn.data(), BLUE_NOISE_SEED);
			blueNoise.upload(bn.data(), bn.size() * 4, stream);
			hipCheck(hipStreamSynchronize(stream), "sync");
		}
		P.blueNoise = blueNoise.as<uint32_t>();
		P.shift = shiftSeed;
		f.maxDepth = estimator == ESTIMATOR_OPTIX7 ? MAXPATHLENGTH_OPTIX7 : MAXPATHLENGTH_FILTER;
	}
	if (estimator == ESTIMATOR_FILTER)
	{	/* rendercore.cpp:780-797: with the filter on, aperture 0 and no distortion */
		P.blueSlot = blueSlot;
		if (filterOn && w == scrw && h == scrh)
		{
			P.gFeatures = wf.gFeatures(), P.gWorldPos = wf.gWorldPos(), P.gDeltaDepth = wf.gDeltaDepth();
			f.cam.aperture = 0;
		}
		f.cam.distortion = 0;
		/* TAA jitter (rendercore.cpp:781-786): p1..p3 move by (j0 / w) right + (j1 / h) up.
		 * Eye rays span the unjittered right/up from the jittered p1 (camP1); the
		 * shade kernel's depth derivatives use the jittered corners throughout. */
		lighthouse2::float3 p1 = view.p1, p2 = view.p2, p3 = view.p3;
		if (jitterX != 0 || jitterY != 0)
		{
			const float rx = view.p2.x - view.p1.x, ry = view.p2.y - view.p1.y, rz = view.p2.z - view.p1.z;
			const float ux = view.p3.x - view.p1.x, uy = view.p3.y - view.p1.y, uz = view.p3.z - view.p1.z;
			const float sx = jitterX / (float)w, sy = jitterY / (float)h;
			const float jx = sx * rx + sy * ux, jy = sx * ry + sy * uy, jz = sx * rz + sy * uz;
			p1.x += jx, p1.y += jy, p1.z += jz, p2.x += jx, p2.y += jy, p2.z += jz, p3.x += jx, p3.y += jy, p3.z += jz;
		}
		P.camPos = make_float4(view.pos.x, view.pos.y, view.pos.z, 0);
		P.camP1 = make_float4(p1.x, p1.y, p1.z, 0);
		P.camRight = make_float4(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z, 0);
		P.camUp = make_float4(p3.x - p1.x, p3.y - p1.y, p3.z - p1.z, 0);
	}
	f.R0gen = R0gen, f.samplesTaken = samplesTakenAtStart, f.spp = spp, f.pixels = w * h;
	if (estimator == ESTIMATOR_REFERENCE) f.maxDepth = depth;
	return f;
}

/* Samples [s0, s1) of a frame: the wavefront loop of RenderCore::Render
 * (rendercore.cpp:564-611) over passes of at most passCap paths, each pass
 * resolved into acc32/acc64 in sample order. The image depends on neither
 * the pass size nor on how the samples are split between calls: a path's
 * random numbers are keyed by its index within the frame. */
int RenderCore::runSamples(const FrameSetup& f, int s0, int s1, size_t passCap, Wavefront& w, float4* acc32, double* acc64,
	uint32_t* queue, uint32_t* ovf, hipStream_t st, hipEvent_t* evs, float* timing, const SideLane* side, bool fresh)
{
	const int pixels = f.pixels;
	const int passSamples = std::max(1, std::min(s1 - s0, (int)(std::min(passCap, Wavefront::PASS_CAP) / (size_t)pixels)));
	w.resizePaths((size_t)passSamples * pixels, st);
	int deepest = 0, sideTimed = 0;
	for (int a = s0; a < s1; a += passSamples)
	{
		PassPart part = { &f, a, std::min(s1, a + passSamples), acc32, acc64, fresh && a == s0 };
		deepest = std::max(deepest, runPass(&part, 1, w, queue, ovf, st, evs, timing, side, nullptr, nullptr, &sideTimed));
	}
	if (side && side->timing && timing && sideTimed > 0)
	{	/* shadow-lane durations, read once the lane has finished */
		hipCheck(hipEventSynchronize(side->timing[2 * sideTimed - 1]), "sync");
		for (int i = 0; i < sideTimed; i++)
		{
			float c = 0;
			if (hipEventElapsedTime(&c, side->timing[2 * i], side->timing[2 * i + 1]) == hipSuccess) timing[4] += c;
		}
	}
	return deepest;
}

/* One pass of the wavefront loop over the paths of one or several jobs:
 * generate -> (extend -> shade -> connect) per wave -> resolve. With several
 * parts the jobs' paths share every launch: part i holds pass slots
 * [slotBase_i, slotBase_i + (s1 - s0) * pixels_i), the path buffers carry
 * pass slots, and the shade kernels look each path's job up in a segment
 * table (segsHost staged, copied to segsDev on the stream; every pass reads
 * a wave count back before the next pass stages its table). All parts use
 * one estimator. A job's paths compute exactly what they compute in a pass
 * of their own: every random number is keyed by the path's index within its
 * job's frame, and each path slot is written by its own path only. */
int RenderCore::runPass(const PassPart* parts, int np, Wavefront& w, uint32_t* queue, uint32_t* ovf, hipStream_t st, hipEvent_t* evs,
	float* timing, const SideLane* side, const JobSeg* segsDev, JobSeg* segsHost, int* sideTimedIo)
{
	if (np < 1 || np > JOB_SEGS_MAX) throw CoreError(LH2HIP_ERR_INVALID, "runPass: 1..16 parts");
	const bool multi = np > 1;
	if (multi && (!segsDev || !segsHost)) throw CoreError(LH2HIP_ERR_INVALID, "runPass: several parts need a segment table");
	const FrameSetup& f0 = *parts[0].f;
	const TraceScene& sc = scene;
	const bool o7 = f0.estimator == ESTIMATOR_OPTIX7, fl = f0.estimator == ESTIMATOR_FILTER, mis = o7 || fl;
	uint32_t slotBase[JOB_SEGS_MAX];
	uint32_t n = 0;
	int maxDepthAll = 0;
	for (int i = 0; i < np; i++)
	{
		const FrameSetup& f = *parts[i].f;
		if (f.estimator != f0.estimator) throw CoreError(LH2HIP_ERR_INVALID, "runPass: parts of one estimator only");
		slotBase[i] = n;
		n += (uint32_t)(parts[i].s1 - parts[i].s0) * (uint32_t)f.pixels;
		maxDepthAll = std::max(maxDepthAll, f.maxDepth);
	}
	if (n > Wavefront::PASS_CAP) throw CoreError(LH2HIP_ERR_INVALID, "runPass: more than 2^24 paths");
	w.resizePaths(n, st);
	PathBuffers B = w.buffers();
	/* optix7 cores: R0 of shade wave k is RandomUInt(camRNGseed) + k * 91771,
	 * the seed advancing once per wave the frame runs (rendercore.cpp:881-905) */
	uint32_t R0wave[JOB_SEGS_MAX][MAXPATHLENGTH_FILTER + 1] = {};
	if (mis)
		for (int i = 0; i < np; i++)
		{
			uint32_t cs = parts[i].f->camSeed;
			for (int k = 1; k <= parts[i].f->maxDepth && k <= MAXPATHLENGTH_FILTER; k++) R0wave[i][k] = xor32_host(cs) + (uint32_t)k * 91771u;
		}
	ShadeParams P = f0.P;
	P.jobBase = (uint32_t)parts[0].s0 * (uint32_t)f0.pixels;
	P.passPaths = n;
	P.segs = nullptr, P.nseg = 0;
	if (multi)
	{
		for (int i = 0; i < np; i++)
		{
			const FrameSetup& f = *parts[i].f;
			JobSeg& S = segsHost[i];
			memset(&S, 0, sizeof(S));
			S.slotBase = slotBase[i], S.jobBase = (uint32_t)parts[i].s0 * (uint32_t)f.pixels;
			S.w = f.P.w, S.h = f.P.h, S.pass = f.P.pass, S.maxDepth = f.P.maxDepth, S.spreadAngle = f.P.spreadAngle;
			S.shift = f.P.shift, S.blueSlot = f.P.blueSlot, S.R0add = (uint32_t)f.samplesTaken * 7907u;
			S.gFeatures = f.P.gFeatures, S.gWorldPos = f.P.gWorldPos, S.gDeltaDepth = f.P.gDeltaDepth;
			S.camPos = f.P.camPos, S.camP1 = f.P.camP1, S.camRight = f.P.camRight, S.camUp = f.P.camUp;
			for (int k = 0; k < 8; k++) S.R0w[k] = R0wave[i][k];
		}
		hipCheck(hipMemcpyAsync((void*)segsDev, segsHost, sizeof(JobSeg) * np, hipMemcpyHostToDevice, st), "H2D");
		P.segs = segsDev, P.nseg = np, P.probe = nullptr;
	}
	for (int i = 0; i < np; i++)
	{
		const FrameSetup& f = *parts[i].f;
		const int k = parts[i].s1 - parts[i].s0;
		const uint32_t jb = (uint32_t)parts[i].s0 * (uint32_t)f.pixels, cnt = (uint32_t)k * (uint32_t)f.pixels;
		ShadeParams Pi = f.P;
		Pi.jobBase = jb, Pi.passPaths = n;
		if (o7) hipCheck(launch_generate7(f.cam, Pi, jb, (int)cnt, geometryEpsilon, B, 0, st, slotBase[i], n, multi), "generate");
		else if (fl) hipCheck(launch_generateF(f.cam, Pi, jb, (int)cnt, geometryEpsilon, B, 0, st, slotBase[i], n, multi), "generate");
		else hipCheck(launch_generate(f.cam, f.R0gen, jb, (int)cnt, geometryEpsilon, B, 0, st, slotBase[i], n, multi), "generate");
	}
	int deepest = 0;
	bool sidePending = false;
	int sideTimed = sideTimedIo ? *sideTimedIo : 0;
	/* the main stream waits for the shadow lane before anything that writes
	 * what the lane reads (shade: shadow rays, potentials, path slots) or
	 * reads what it writes (resolve: path slots) */
	const auto joinSide = [&]() {
		if (sidePending) hipCheck(hipStreamWaitEvent(st, side->done, 0), "wait"), sidePending = false;
	};
	int in = 0;
	for (int pathLength = 1; pathLength <= maxDepthAll; pathLength++)
	{
		if (evs) hipCheck(hipEventRecord(evs[0], st), "event");
		hipCheck(launch_trace_dev(sc, MODE_CLOSEST, B.rays[in], B.hits, (int)n, B.counters + 0, queue, ovf, persistentBlocks, st), "extend");
		if (evs) hipCheck(hipEventRecord(evs[1], st), "event");
		P.pathLength = pathLength;
		joinSide();
		if (mis)
		{
			P.R0 = R0wave[0][std::min(pathLength, MAXPATHLENGTH_FILTER)];
			if (o7) hipCheck(launch_shade7(P, B, in, (int)n, st), "shade");
			else hipCheck(launch_shadeF(P, B, in, (int)n, st), "shade");
		}
		else
		{	/* several parts: each job adds its samplesTaken * 7907 (JobSeg::R0add) */
			P.R0 = multi ? (uint32_t)pathLength * 91771u : (uint32_t)f0.samplesTaken * 7907u + (uint32_t)pathLength * 91771u;
			hipCheck(launch_shade(P, B, in, (int)n, st), "shade");
		}
		if (evs) hipCheck(hipEventRecord(evs[2], st), "event");
		/* the reference estimator connects from every vertex but the last;
		 * optix7's from every vertex (its NEE precedes the depth cut) */
		uint32_t* const logSlot = w.countLog() + pathLength;
		bool advanced = false;
		if ((mis || pathLength < maxDepthAll) && side)
		{	/* the advance (which clears the shadow count) snapshots the counters
			 * for the side stream, which then traces and connects */
			hipCheck(launch_advance(B.counters, logSlot, st, side->counters, w.hostLive()), "advance");
			advanced = true;
			hipCheck(hipEventRecord(side->shaded, st), "event");
			hipCheck(hipStreamWaitEvent(side->st, side->shaded, 0), "wait");
			const bool timed = side->timing && sideTimed < SIDE_TIMED;
			if (timed) hipCheck(hipEventRecord(side->timing[2 * sideTimed], side->st), "event");
			PathBuffers Bs = B;
			Bs.counters = side->counters;
			hipCheck(launch_trace_dev(sc, MODE_ANY_FLAG, B.shadowRays, B.occluded, (int)n, side->counters + 2, side->queue, side->ovf, persistentBlocks,
				side->st), "connect");
			hipCheck(launch_connect(Bs, (int)n, side->st, mis ? 1.0f : 0.0f), "connect");
			if (timed) hipCheck(hipEventRecord(side->timing[2 * sideTimed + 1], side->st), "event"), sideTimed++;
			hipCheck(hipEventRecord(side->done, side->st), "event");
			sidePending = true;
		}
		else if (mis || pathLength < maxDepthAll)
		{
			hipCheck(launch_trace_dev(sc, MODE_ANY_FLAG, B.shadowRays, B.occluded, (int)n, B.counters + 2, queue, ovf, persistentBlocks, st), "connect");
			hipCheck(launch_connect(B, (int)n, st, mis ? 1.0f : 0.0f), "connect");
		}
		if (evs) hipCheck(hipEventRecord(evs[3], st), "event");
		if (!advanced) hipCheck(launch_advance(B.counters, logSlot, st, nullptr, w.hostLive()), "advance");
		in ^= 1;
		deepest = std::max(deepest, pathLength);
		/* stop early once every path has terminated (checked every 4 waves;
		 * after every wave for optix7, whose loop reads the count back each time) */
		if (mis || (pathLength & 3) == 0 || pathLength == maxDepthAll)
		{
			hipCheck(hipStreamSynchronize(st), "sync");
			const uint32_t live = *(volatile uint32_t*)w.hostLive();   /* written by the advance */
			if (evs && timing)
			{
				float a1, b2, c;
				(void)hipEventElapsedTime(&a1, evs[0], evs[1]), (void)hipEventElapsedTime(&b2, evs[1], evs[2]), (void)hipEventElapsedTime(&c, evs[2], evs[3]);
				timing[std::min(pathLength - 1, 2)] += a1, timing[3] += b2;
				if (!side) timing[4] += c;
			}
			if (live == 0) break;
		}
	}
	joinSide();
	for (int i = 0; i < np; i++)
	{
		const FrameSetup& f = *parts[i].f;
		const int k = parts[i].s1 - parts[i].s0;
		hipCheck(launch_resolve(B.pathAcc + slotBase[i], f.pixels, k, parts[i].acc32, parts[i].acc64, st, parts[i].fresh), "resolve");
		if (f.P.gFeatures)
			hipCheck(launch_resolve(B.pathAcc + n + slotBase[i], f.pixels, k, parts[i].acc32 + f.pixels, parts[i].acc64 + 4 * (size_t)f.pixels, st,
				parts[i].fresh), "resolve");
	}
	if (sideTimedIo) *sideTimedIo = sideTimed;
	return deepest;
}

void RenderCore::Render(const ViewPyramid& view, const Convergence converge, bool async)
{
	std::unique_lock<std::shared_mutex> lockScene(sceneMutex);   /* a server's pass reads the same scene buffers */
	if (scrw == 0) throw CoreError(LH2HIP_ERR_STATE, "Render: SetTarget first");
	updateTopLevel();
	if (!sceneReady) return; /* no geometry yet: nothing to render (rendercore_optix7/rendercore.cpp:821) */
	checkRenderable();
	const auto t0 = std::chrono::high_resolution_clock::now();
	bool fresh = false;   /* the frame's first pass writes the accumulators (no clear needed) */
	if (converge == Restart || firstConvergingFrame)
	{
		fresh = true;
		samplesTaken = 0;
		firstConvergingFrame = true;
		if (estimator != ESTIMATOR_FILTER) camSeed = seedSetting;   /* the filter core keeps its seed (RenderCore_Optix7Filter/rendercore.cpp:749) */
	}
	if (converge == Converge) firstConvergingFrame = false;
	uint32_t R0gen = 0;
	if (estimator == ESTIMATOR_OPTIX7) xor32_host(shiftSeed);   /* once per frame, rendercore_optix7/rendercore.cpp:855 */
	else if (estimator == ESTIMATOR_REFERENCE) R0gen = xor32_host(camSeed);   /* RenderCore_PrimeRef/rendercore.cpp:556 */
	if (estimator == ESTIMATOR_FILTER)
	{	/* jitter the view for TAA (RenderCore_Optix7Filter/rendercore.cpp:754-758) */
		static const float haltonX[4] = { 0.3f, 0.7f, 0.2f, 0.8f }, haltonY[4] = { 0.2f, 0.8f, 0.7f, 0.3f };
		j0 = taaOn && filterOn ? haltonX[frameCycle] : 0.0f;
		j1 = taaOn && filterOn ? haltonY[frameCycle] : 0.0f;
		frameCycle = (frameCycle + 1) & 3;
	}
	FrameSetup f = frameParams(view, scrw, scrh, scrspp, maxDepth, R0gen, samplesTaken, estimator == ESTIMATOR_FILTER ? j0 : 0.0f,
		estimator == ESTIMATOR_FILTER ? j1 : 0.0f);
	f.camSeed = camSeed;
	f.P.probePixel = probePos.x + scrw * probePos.y;
	f.P.probe = wf.probe();
	float timing[5] = { 0, 0, 0, 0, 0 };
	SideLane lane;
	lane.st = stream2, lane.counters = sideCounters.as<uint32_t>(), lane.queue = sideCounters.as<uint32_t>() + 4, lane.ovf = overflow2.as<uint32_t>();
	lane.shaded = sideEv[0], lane.done = sideEv[1], lane.timing = sideEv + 2;
	const int waves = runSamples(f, 0, scrspp, Wavefront::PASS_CAP, wf, wf.acc32(), wf.acc64(), counters.as<uint32_t>() + TRACE_QUEUE_WORDS,
		overflow.as<uint32_t>(), stream, ev, timing, shadowLane ? &lane : nullptr, fresh);
	if (estimator != ESTIMATOR_REFERENCE) for (int k = 0; k < waves; k++) xor32_host(camSeed);
	const int pixels = scrw * scrh;
	samplesTaken += scrspp;
	bool filtered = false;
	if (estimator == ESTIMATOR_FILTER && f.P.gFeatures)
	{	/* FinalizeRender (RenderCore_Optix7Filter/rendercore.cpp:897-935) */
		blueSlot = (blueSlot + 1) & 255;
		if (!svgf || svgf->w != scrw || svgf->h != scrh) svgf.reset(new Svgf(scrw, scrh, stream));
		svgf->taa = taaOn;
		hipCheck(hipEventRecord(ev[4], stream), "event");
		svgf->Frame(wf.acc32(), wf.gFeatures(), wf.gWorldPos(), wf.gDeltaDepth(), prevView, j0, j1, prevj0, prevj1, samplesTaken, clampDirect,
			clampIndirect, samplesTaken == scrspp ? 0 : 1, wf.image(), stream);
		prevj0 = j0, prevj1 = j1;
		hipCheck(hipEventRecord(ev[5], stream), "event");
		filtered = true;
	}
	else
	{
		if (estimator == ESTIMATOR_FILTER) blueSlot = (blueSlot + 1) & 255;
		hipCheck(launch_finalize(wf.acc32(), pixels, 1.0f / (float)samplesTaken, wf.image(), stream), "finalize");
	}
	prevView = view;
	if (!async)
	{
		hipCheck(hipStreamSynchronize(stream), "sync");
		int probe[3];
		memcpy(probe, (const void*)wf.probe(), 12);   /* pinned: written by the shade kernel */
		stats.probedInstid = probe[0], stats.probedTriid = probe[1];
		memcpy(&stats.probedDist, &probe[2], 4);
		float ms = 0;
		if (filtered && hipEventElapsedTime(&ms, ev[4], ev[5]) == hipSuccess) stats.filterTime = ms * 1e-3f;
		/* wave sizes of the (last) pass: the advance after wave k logged the size of wave k + 1 */
		uint32_t logged[MAXPATHLENGTH_FILTER + 2] = {};
		const int nlog = std::min(waves, MAXPATHLENGTH_FILTER + 1);
		if (nlog > 1) memcpy(logged, (const void*)wf.countLog(), sizeof(uint32_t) * (nlog + 1));   /* pinned: written by the advances */
		stats.bounce1RayCount = nlog > 1 ? logged[1] : 0;
		stats.deepRayCount = 0;
		for (int k = 2; k < nlog; k++) stats.deepRayCount += logged[k];
	}
	const auto t1 = std::chrono::high_resolution_clock::now();
	stats.renderTime = std::chrono::duration<float>(t1 - t0).count();
	stats.traceTime0 = timing[0] * 1e-3f, stats.traceTime1 = timing[1] * 1e-3f, stats.traceTimeX = timing[2] * 1e-3f;
	stats.shadeTime = timing[3] * 1e-3f, stats.shadowTraceTime = timing[4] * 1e-3f;
	stats.primaryRayCount = (uint32_t)(pixels * scrspp);
}

void RenderCore::WaitForRender()
{
	if (stream) hipCheck(hipStreamSynchronize(stream), "sync");
}

void RenderCore::ReadImage(void* hostOut)
{
	if (!hostOut) throw CoreError(LH2HIP_ERR_INVALID, "read_image: null");
	if (scrw == 0) throw CoreError(LH2HIP_ERR_STATE, "read_image: no target");
	hipCheck(hipMemcpyAsync(hostOut, wf.image(), sizeof(float4) * scrw * scrh, hipMemcpyDeviceToHost, stream), "D2H");
	hipCheck(hipStreamSynchronize(stream), "sync");
}

void RenderCore::ReadBuffer(int which, void* hostOut, size_t bytes)
{
	if (!hostOut) throw CoreError(LH2HIP_ERR_INVALID, "read_buffer: null");
	if (scrw == 0) throw CoreError(LH2HIP_ERR_STATE, "read_buffer: no target");
	const size_t px = (size_t)scrw * scrh;
	const void* src = nullptr;
	size_t n = 0;
	if (which == 0) src = wf.acc64(), n = px * 8 * sizeof(double);
	else if (which == 1) src = wf.gFeatures(), n = px * 16;
	else if (which == 2) src = wf.gWorldPos(), n = px * 16;
	else if (which == 3) src = wf.gDeltaDepth(), n = px * 16;
	else if (which == 4) src = wf.acc32(), n = px * 2 * sizeof(float4);
	else throw CoreError(LH2HIP_ERR_INVALID, "read_buffer: which 0..4");
	if (bytes != n) throw CoreError(LH2HIP_ERR_INVALID, "read_buffer: size mismatch (expected " + std::to_string(n) + " bytes)");
	hipCheck(hipMemcpyAsync(hostOut, src, n, hipMemcpyDeviceToHost, stream), "D2H");
	hipCheck(hipStreamSynchronize(stream), "sync");
}

void RenderCore::ReadAccumulatorF64(double* hostOut)
{
	if (!hostOut) throw CoreError(LH2HIP_ERR_INVALID, "read_accumulator: null");
	if (scrw == 0) throw CoreError(LH2HIP_ERR_STATE, "read_accumulator: no target");
	hipCheck(hipMemcpyAsync(hostOut, wf.acc64(), sizeof(double) * 4 * scrw * scrh, hipMemcpyDeviceToHost, stream), "D2H");
	hipCheck(hipStreamSynchronize(stream), "sync");
}

} // namespace lh
