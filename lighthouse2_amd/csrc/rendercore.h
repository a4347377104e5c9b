/* rendercore.h - RenderCore_HIP host core: a CoreAPI_Base implementation that
 * keeps the scene in HBM, builds the BVHs on the GPU and drives the wavefront
 * kernels. Mirrors lib/RenderCore_PrimeRef/rendercore.h (state per core, one
 * device, single-threaded API) with headless output. */
#pragma once
#include "../../include/core_api_base.h"
#include "bvh.h"
#include "trace.h"
#include "waves.h"
#include "shade.h"
#include <string>
#include <vector>
#include <stdexcept>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <atomic>

namespace lh {

/* Raised by every failing call; the C ABI turns it into a status code. */
struct CoreError : std::runtime_error
{
	int code;
	CoreError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

/* Owning device allocation (CoreBuffer<T> ON_DEVICE, lib/CUDA/shared_host_code/cudatools.h:196-327). */
struct DevBuf
{
	void* p = nullptr;
	size_t bytes = 0;
	DevBuf() = default;
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr, o.bytes = 0; }
	DevBuf& operator=(DevBuf&& o) noexcept { release(); p = o.p, bytes = o.bytes; o.p = nullptr, o.bytes = 0; return *this; }
	~DevBuf() { release(); }
	void release();
	void ensure(size_t n);             /* grow-only, contents undefined after growth */
	void upload(const void* src, size_t n, hipStream_t s);
	template <typename T> T* as() const { return (T*)p; }
};

/* Path buffers of the wavefront loop (RenderCore_PrimeRef/rendercore.cpp:148-208
 * sizes them maxPixels * spp; here one pass holds at most PASS_CAP paths and a
 * frame with more samples runs several passes). */
struct Wavefront
{
	static constexpr size_t PASS_CAP = (size_t)1 << 24;
	DevBuf rays[2], state[2], pathIdB[2], auxB[2], hitsB, shadowRaysB, shadowPotB, occludedB, pathAccB, countersB;
	DevBuf acc32B, acc64B, imageB;
	/* Pinned host words the kernels write directly, read by the host after a
	 * stream synchronisation (no copy on the stream): [0] live path count after
	 * the last advance, [4..6] probe hit (instance, primitive, distance bits),
	 * [HOST_LOG..] paths per wave. */
	static constexpr int HOST_LOG = 16, HOST_WORDS = HOST_LOG + 1100;
	uint32_t* hostW = nullptr;
	void ensureHost();
	DevBuf gFeaturesB, gWorldPosB, gDeltaDepthB;   /* SVGF G-buffer of the filter estimator */
	size_t pathCap = 0;
	int pixelCap = 0;
	void resizeTarget(int pixels, hipStream_t s);
	void clearAccumulators(hipStream_t s);
	void resizePaths(size_t n, hipStream_t s);
	void release();
	PathBuffers buffers();
	int* probe() { return (int*)(hostW + 4); }
	uint32_t* hostLive() { return hostW; }
	float4* acc32() { return acc32B.as<float4>(); }
	double* acc64() { return acc64B.as<double>(); }
	float4* image() { return imageB.as<float4>(); }
	uint32_t* countLog() { return hostW + HOST_LOG; }
	uint4* gFeatures() { return gFeaturesB.as<uint4>(); }
	float4* gWorldPos() { return gWorldPosB.as<float4>(); }
	float4* gDeltaDepth() { return gDeltaDepthB.as<float4>(); }
};

struct MeshData
{
	int triCount = 0;
	DevBuf verts, coreTris, nodesF32, nodesQ8, tris;
	uint32_t rootF32 = 0, rootQ8 = 0;
	int depthF32 = 0, depthQ8 = 0, stackF32 = 0, stackQ8 = 0, buildDepth = 0;
	size_t nodeCountF32 = 0, nodeCountQ8 = 0, leafCount = 0;
	float lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
	/* the build tree stays on the device (refit, inspection): breadth-first
	 * nodes, primitive ranges, the layout maps and the leaf slot order */
	DevBuf dBuildNodes, dBuildPrims, dF32Kids, dQ8Map, dOrder, dLast;
	int buildNodeCount = 0;
	std::vector<int> levelStart;         /* build ids of depth d: [levelStart[d], levelStart[d + 1]) */
	float buildMs = 0;
	int maxMaterial = -1;                /* largest CoreTri::material, -1 without CoreTri data */
};

struct InstanceData { int mesh; float T[16]; };

/* Per-frame constants of the wavefront kernels. */
struct FrameSetup
{
	CamParams cam;
	ShadeParams P;
	uint32_t R0gen = 0;
	int samplesTaken = 0, spp = 1, maxDepth = 1, pixels = 0;
	/* ESTIMATOR_OPTIX7: per-bounce shade seeds R0 = xor32^k(camSeed) + k * 91771 */
	int estimator = 0;
	uint32_t camSeed = 0;
};

enum { ESTIMATOR_REFERENCE = 0, ESTIMATOR_OPTIX7 = 1, ESTIMATOR_FILTER = 2 };

/* One job's samples [s0, s1) in a pass of the wavefront loop, resolved into
 * its accumulators (a pass may carry several jobs' paths, see runPass). */
struct PassPart
{
	const FrameSetup* f;
	int s0, s1;
	float4* acc32;
	double* acc64;
	bool fresh = false;   /* the part's first samples since a restart: the resolve writes instead of adding */
};

class Svgf;

class RenderCore : public lighthouse2::CoreAPI_Base
{
public:
	RenderCore();
	~RenderCore();
	/* CoreAPI_Base */
	lighthouse2::CoreStats GetCoreStats() const override { return stats; }
	void Init() override;
	void SetProbePos(const lighthouse2::int2 pos) override;
	void SetTarget(lighthouse2::GLTexture* target, const uint32_t spp) override;
	void Setting(const char* name, float value) override;
	void Render(const lighthouse2::ViewPyramid& view, const lighthouse2::Convergence converge, bool async) override;
	void WaitForRender() override;
	void Shutdown() override;
	void SetTextures(const lighthouse2::CoreTexDesc* tex, const int textureCount) override;
	void SetMaterials(lighthouse2::CoreMaterial* mat, const int materialCount) override;
	void SetLights(const lighthouse2::CoreLightTri* triLights, const int triLightCount,
		const lighthouse2::CorePointLight* pointLights, const int pointLightCount,
		const lighthouse2::CoreSpotLight* spotLights, const int spotLightCount,
		const lighthouse2::CoreDirectionalLight* directionalLights, const int directionalLightCount) override;
	void SetSkyData(const lighthouse2::float3* pixels, const uint32_t width, const uint32_t height, const lighthouse2::mat4& worldToLight) override;
	/* Blue-noise tables of the optix7 estimator, 8 bits per entry: 65536
	 * sequence values, 131072 scrambling and 131072 ranking entries (the
	 * conversion of rendercore_optix7/rendercore.cpp:246-255). Without this
	 * call a seeded generator fills them. */
	void SetBlueNoise(const uint8_t* sequence, const uint8_t* scrambling, const uint8_t* ranking);
	void SetGeometry(const int meshIdx, const lighthouse2::float4* vertexData, const int vertexCount, const int triangleCount, const lighthouse2::CoreTri* triangles) override;
	void SetInstance(const int instanceIdx, const int modelIdx, const lighthouse2::mat4& transform) override;
	void FinalizeInstances() override;

	/* C-ABI extensions */
	void SetTargetSize(int w, int h, int spp);
	void TraceClosest(const void* rays, void* hits, int count, hipStream_t s);
	void TraceAny(const void* rays, uint32_t* mask, int count, hipStream_t s);
	void TraceClosestCounted(const void* rays, void* hits, int count, hipStream_t s, uint64_t* out2);
	void Bounce(const void* raysIn, const void* hits, const uint32_t* pathIn, int count, uint32_t R0, float tmin,
		void* raysOut, uint32_t* pathOut, uint32_t* countOut, hipStream_t s);
	void ReadImage(void* hostOut);
	void ReadAccumulatorF64(double* hostOut);
	void ReadBuffer(int which, void* hostOut, size_t bytes);
	void BvhInfo(int mesh, int64_t* out8) const;
	int BvhDownload(int mesh, void* nodesOut, int maxNodes, uint32_t* primsOut, int maxPrims) const;
	void BvhLayoutDownload(int mesh, void* f32Out, int maxF32, void* q8Out, int maxQ8, void* trisOut, int maxTris) const;
	const TraceScene& Scene();
	/* Render building blocks, shared with the job server (server.cpp). */
	void checkRenderable();
	FrameSetup frameParams(const lighthouse2::ViewPyramid& view, int w, int h, int spp, int depth, uint32_t R0gen, int samplesTakenAtStart,
		float jitterX = 0, float jitterY = 0, int estimatorOverride = -1);
	float clampDirectSetting() const { return clampDirect; }
	float clampIndirectSetting() const { return clampIndirect; }
	/* A second stream on which the shadow rays of wave k are traced and
	 * connected while wave k + 1's extension rays are traced (own ray feed,
	 * overflow stack and counter snapshot; timing pairs per wave). */
	struct SideLane
	{
		hipStream_t st = nullptr;
		uint32_t* counters = nullptr;   /* 4 words: snapshot of the wave's counters */
		uint32_t* queue = nullptr;
		uint32_t* ovf = nullptr;
		hipEvent_t shaded = nullptr, done = nullptr;
		hipEvent_t* timing = nullptr;   /* 2 per wave, SIDE_TIMED waves */
	};
	static const int SIDE_TIMED = 16;
	/* Returns the number of shade waves the deepest pass ran. */
	int runSamples(const FrameSetup& f, int s0, int s1, size_t passCap, Wavefront& w, float4* acc32, double* acc64,
		uint32_t* queue, uint32_t* ovf, hipStream_t st, hipEvent_t* evs, float* timing, const SideLane* side = nullptr, bool fresh = false);
	/* One pass over the paths of 1..JOB_SEGS_MAX parts (at most 2^24 paths in
	 * all, one estimator); several parts need a staging segment table on the
	 * host and its device copy. Returns the number of waves run. */
	int runPass(const PassPart* parts, int np, Wavefront& w, uint32_t* queue, uint32_t* ovf, hipStream_t st, hipEvent_t* evs, float* timing,
		const SideLane* side, const JobSeg* segsDev, JobSeg* segsHost, int* sideTimed = nullptr);
	int estimatorSetting() const { return estimator; }
	bool sceneIsReady() const { return sceneReady; }
	int traceBlocks() const { return persistentBlocks; }
	bool shadowLaneOn() const { return shadowLane; }
	float epsilon() const { return geometryEpsilon; }

	hipStream_t stream = nullptr;
	int deviceCUs = 0;
	/* Job servers (server.cpp) read this core's scene from their threads
	 * (submission and passes, shared locks); the calls that change the scene
	 * or render take it exclusively. While a server has jobs pending the
	 * calls that change the scene fail with LH2HIP_ERR_STATE, and the core
	 * cannot shut down while a server exists. */
	std::shared_mutex sceneMutex;
	std::mutex blueNoiseMutex;
	std::atomic<int> serverCount{ 0 }, pendingJobs{ 0 };
	void requireNoJobs(const char* what) const;

private:
	void buildMesh(MeshData& m);
	/* Refit: the build tree keeps its topology, node boxes are recomputed
	 * bottom-up from the new vertices and the laid-out nodes rewritten, all
	 * on the device. */
	void refitMesh(MeshData& m);
	void ensureTraceBuffers();
	void updateTopLevel();

	lighthouse2::CoreStats stats = {};
	std::vector<MeshData> meshes;
	std::vector<InstanceData> instances;
	bool instancesDirty = true, sceneReady = false;
	/* top level */
	DevBuf tlasNodes, instRecs, shadeInsts;
	TraceScene scene;
	/* trace scratch (overflow2 / sideCounters: the shadow-ray lane) */
	DevBuf overflow, counters, overflow2, sideCounters;
	hipStream_t stream2 = nullptr;
	hipEvent_t sideEv[2 + 2 * SIDE_TIMED];
	double wallClockNs = 10.0;         /* period of the device wall clock */
	int persistentBlocks = 0;
	float minRaysPerLane = 0.0f;       /* 0: always launch persistentBlocks */
	int blocksFor(int count) const;
	/* settings (RenderCore_PrimeRef rendercore.cpp:484-502 + extensions) */
	float geometryEpsilon = 1.0e-3f, clampValue = 10.0f;
	int maxDepth = 64, nodeFormat = FMT_Q8, traceLoop = LOOP_MERGED, triBias = 16, refillLanes = 16, chunk = 32;
	uint32_t camSeed = 0x12345678u, seedSetting = 0x12345678u;
	int estimator = ESTIMATOR_REFERENCE;
	bool refitSetting = false;         /* SetGeometry on a mesh of unchanged size: refit instead of rebuild */
	bool shadowLane = true;            /* shadow rays of wave k on a second stream, beside wave k + 1's extension */
	int bsdfSetting = 1;               /* Disney, as rendercore_optix7 ships (kernels/bsdf.h) */
	/* filter estimator (RenderCore_Optix7Filter/rendercore.h:32-35,124) */
	bool filterOn = true;
	float clampDirect = 2.5f, clampIndirect = 15.0f;
	int blueSlot = 0;
	/* TAA (rendercore.h:35,126; rendercore.cpp:754-758): on by default; the
	 * view is jittered by a 4-frame Halton cycle while TAA and the filter are on */
	bool taaOn = true;
	int frameCycle = 0;
	float j0 = 0, j1 = 0, prevj0 = 0, prevj1 = 0;
	lighthouse2::ViewPyramid prevView = {};
	std::unique_ptr<Svgf> svgf;
	uint32_t shiftSeed = 0x11331445u;  /* blue-noise shift of the optix7 estimator (rendercore_optix7/rendercore.h:123) */
	DevBuf blueNoise;
	/* render state */
	int scrw = 0, scrh = 0, scrspp = 1;
	int samplesTaken = 0;
	bool firstConvergingFrame = false;   /* rendercore.h:134 */
	lighthouse2::int2 probePos = { 0, 0 };
	DevBuf materials, triLights, pointLights, spotLights, dirLights;
	int lightCounts[4] = { 0, 0, 0, 0 };
	/* textures (continuous texel arrays per storage kind) and skydome */
	struct TexInfo { uint32_t width, height, flags, firstPixel, pixelCount; int storage; };
	std::vector<TexInfo> textures;
	DevBuf texARGB32, texARGB128, texNRM32, skyPixels;
	int skyW = 0, skyH = 0;
	float worldToSky[12] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0 };
	int materialCount = 0;
	Wavefront wf;
	hipEvent_t ev[8];
};

} // namespace lh
