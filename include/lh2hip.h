/* lh2hip.h - C ABI of librendercore_hip.so (RenderCore_HIP).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status returns
 * (0 = ok, negative = error; lh2hip_last_error() holds the message of the most
 * recent failure on the calling thread). Device pointers are hipMalloc'd memory
 * of the current device; a `stream` argument is a hipStream_t (NULL = the null
 * stream). Every entry point below cites the reference interface it replaces.
 * The same library exports the C++ plugin symbol CreateCore (core_api_base.h).
 */
#ifndef LH2HIP_H
#define LH2HIP_H

#include "lh2hip_types.h"

#ifdef __cplusplus
extern "C" {
using namespace lh2hip;
#endif

#define LH2HIP_API __attribute__((visibility("default")))

enum
{
	LH2HIP_OK = 0,
	LH2HIP_ERR_INVALID = -1,   /* bad argument or size */
	LH2HIP_ERR_DEVICE = -2,    /* a HIP call failed */
	LH2HIP_ERR_STATE = -3,     /* call out of order (e.g. render before scene) */
	LH2HIP_ERR_OOM = -4,       /* device allocation failed or budget exceeded */
	LH2HIP_ERR_MISSING = -5    /* no HIP device present */
};

typedef struct lh2hip_core lh2hip_core;

LH2HIP_API const char* lh2hip_last_error(void);
LH2HIP_API int lh2hip_version(void);

/* ---- render core (lib/RenderSystem/core_api_base.h:81-119) ---------------- */

/* CoreAPI_Base::Init (lib/RenderCore_PrimeRef/rendercore.cpp:100-142). */
LH2HIP_API int lh2hip_create(lh2hip_core** out);
/* CoreAPI_Base::Shutdown (rendercore.cpp:639-). Frees the handle. */
LH2HIP_API int lh2hip_destroy(lh2hip_core* core);
/* CoreAPI_Base::SetTarget, headless (rendercore.cpp:148-208). */
LH2HIP_API int lh2hip_set_target(lh2hip_core* core, int width, int height, int spp);
/* CoreAPI_Base::Setting (rendercore.cpp:484-502; keys "epsilon", "clampValue",
 * plus "maxDepth", "nodeFormat" (0 fp32 BVH2, 1 q8 BVH4), "seed": camera RNG
 * seed replacing 0x12345678, exact for values up to 2^24), "traceBlocksPerCU":
 * persistent traversal blocks of 256 threads per compute unit, default 8;
 * "traceMinRaysPerLane": shrink the persistent grid of a small wave so each
 * lane gets at least this many rays, 0 = off; "traceLoop": 0 merged loop
 * (one node visit and one leaf triangle per lane and iteration, default),
 * 1 while-while loop with postponed leaves, 2 greedy loop (each iteration a
 * node step or a leaf step, whichever more lanes wait for); "traceTriBias":
 * greedy loop weight of waiting leaf lanes in 1/16, default 16;
 * "traceRefillLanes": idle lanes of a wave that trigger fetching new rays,
 * 1..64, default 16; "traceChunk": rays a wave reserves per atomic, 1..1024,
 * default 32; "estimator": 0 the shade kernel of RenderCore_PrimeRef
 * (kernels/pathtracer.h, default), 1 that of rendercore_optix7 with its
 * Lambert BSDF (kernels/pathtracer.h:54-238, sharedBSDFs/lambert.h): blue
 * noise, MIS, firefly clamp "clampValue", at most 3 path vertices). */
LH2HIP_API int lh2hip_setting(lh2hip_core* core, const char* name, float value);
/* CoreAPI_Base::SetGeometry (rendercore.cpp:214-222; core_mesh.cpp:36-64).
 * vertexData: 3 float4 per triangle (w ignored); builds the mesh BVH on the GPU. */
LH2HIP_API int lh2hip_set_geometry(lh2hip_core* core, int meshIdx, const lh2_float4* vertexData,
	int vertexCount, int triangleCount, const CoreTri* triangles);
/* CoreAPI_Base::SetInstance (rendercore.cpp:228-242); meshIdx -1 truncates. */
LH2HIP_API int lh2hip_set_instance(lh2hip_core* core, int instanceIdx, int meshIdx, const lh2_mat4* transform);
/* CoreAPI_Base::FinalizeInstances (rendercore.cpp:248-277): builds the TLAS. */
LH2HIP_API int lh2hip_finalize_instances(lh2hip_core* core);
/* CoreAPI_Base::SetMaterials (rendercore.cpp:360-402). */
LH2HIP_API int lh2hip_set_materials(lh2hip_core* core, const CoreMaterial* mats, int count);
/* CoreAPI_Base::SetLights (rendercore.cpp:408-460). */
LH2HIP_API int lh2hip_set_lights(lh2hip_core* core,
	const CoreLightTri* triLights, int triLightCount,
	const CorePointLight* pointLights, int pointLightCount,
	const CoreSpotLight* spotLights, int spotLightCount,
	const CoreDirectionalLight* dirLights, int dirLightCount);
/* CoreAPI_Base::SetTextures (rendercore.cpp:283-310). */
LH2HIP_API int lh2hip_set_textures(lh2hip_core* core, const CoreTexDesc* tex, int count);
/* CoreAPI_Base::SetSkyData (rendercore.cpp:466-478). */
LH2HIP_API int lh2hip_set_sky(lh2hip_core* core, const lh2_float3* pixels, int width, int height,
	const lh2_mat4* worldToLight);
/* Blue-noise tables of the optix7 estimator (rendercore_optix7/rendercore.cpp:246-255
 * converts its 8-bit sob256_64 / scr256_64 / rnk256_64 tables the same way):
 * 65536 sequence, 131072 scrambling and 131072 ranking entries. Optional; a
 * seeded generator fills the tables otherwise. */
LH2HIP_API int lh2hip_set_blue_noise(lh2hip_core* core, const uint8_t* sequence, const uint8_t* scrambling, const uint8_t* ranking);
/* CoreAPI_Base::SetProbePos (rendercore.cpp:91-94). */
LH2HIP_API int lh2hip_set_probe_pos(lh2hip_core* core, int x, int y);
/* CoreAPI_Base::Render (rendercore.cpp:535-632): one frame of spp samples per
 * pixel. converge = LH2_CONVERGE / LH2_RESTART. async = 1 returns after enqueue. */
LH2HIP_API int lh2hip_render(lh2hip_core* core, const ViewPyramid* view, int converge, int async);
/* CoreAPI_Base::WaitForRender (rendercore_optix7/rendercore.cpp:945-957). */
LH2HIP_API int lh2hip_wait(lh2hip_core* core);
/* Reads the finalized image (accumulator / samplesTaken, float4 per pixel) to
 * host memory; replaces the GL surface write of finalizeRender
 * (lib/CUDA/shared_kernel_code/finalize_shared.h:29-45). */
LH2HIP_API int lh2hip_read_image(lh2hip_core* core, lh2_float4* hostOut);
/* Reads the FP64 accumulator (sum of all samples, 4 doubles per pixel). */
LH2HIP_API int lh2hip_read_accumulator_f64(lh2hip_core* core, double* hostOut);
/* Internal frame buffers (no reference counterpart; the reference keeps them
 * on the device): 0 fp64 accumulators, both halves (2 * w * h * 4 doubles:
 * direct, then indirect radiance of the filter estimator), 1 SVGF features
 * (w * h uint4, RenderCore_Optix7Filter/kernels/pathtracer.h:44-48), 2 world
 * positions (w * h float4), 3 depth derivatives (w * h float4), 4 fp32
 * accumulators (2 * w * h float4). bytes must equal the buffer size. */
LH2HIP_API int lh2hip_read_buffer(lh2hip_core* core, int which, void* hostOut, size_t bytes);
/* CoreAPI_Base::GetCoreStats (core_api_base.h:87). */
LH2HIP_API int lh2hip_get_stats(lh2hip_core* core, CoreStats* out);

/* Headless readback for a core obtained through the plugin entry point
 * CreateCore (lib/RenderSystem/core_api_base.cpp:119-130 loads it; the
 * reference's cores write the image into the GL texture of SetTarget,
 * lib/CUDA/shared_host_code/interoptexture.cpp:32-73): `core` is the
 * CoreAPI_Base* CreateCore returned. Same output as lh2hip_read_image and
 * lh2hip_read_accumulator_f64. */
LH2HIP_API int lh2hip_plugin_read_image(void* core, lh2_float4* hostOut);
LH2HIP_API int lh2hip_plugin_read_accumulator_f64(void* core, double* hostOut);

/* ---- traversal queries (OptiX Prime rtpQueryExecute replacement) ---------- */

/* Closest hit, RTP_QUERY_TYPE_CLOSEST (lib/RenderCore_PrimeRef/rendercore.cpp:559-573;
 * rendercore_optix7/optix/.optix.cu:125-140,174-184). rays/hits are device
 * pointers; hit = tmin < t < tmax nearest, ties to the smaller (inst, prim). */
LH2HIP_API int lh2hip_trace_closest(lh2hip_core* core, const Ray4* rays, HitRecord* hits, int count, void* stream);
/* Any hit, RTP_QUERY_TYPE_ANY + RTP_BUFFER_FORMAT_HIT_BITMASK
 * (rendercore.cpp:598-607): bit i of occluded[i/32] set when ray i is blocked. */
LH2HIP_API int lh2hip_trace_any(lh2hip_core* core, const Ray4* rays, uint32_t* occluded, int count, void* stream);

/* Closest hit as lh2hip_trace_closest, with the traversal's work counted
 * (host memory, accumulated): out6[0] interior nodes visited, [1] triangles
 * tested, [2]/[3] node steps of all waves / lanes visiting a node in them,
 * [4]/[5] the same for leaf steps, [6]/[7] traversal-loop iterations / lanes
 * holding a ray in them while rays remain to be fetched, [8]/[9] the same
 * once all rays are handed out, [10] ns from the first wave's start until the
 * last ray was handed out, [11] ns from the first start to the last wave's
 * end. A separate kernel variant; used to derive algorithmic bytes per ray,
 * SIMD utilisation and the share of the launch spent after the feed ran dry. */
LH2HIP_API int lh2hip_trace_closest_counted(lh2hip_core* core, const Ray4* rays, HitRecord* hits, int count, void* stream,
	uint64_t* out12);

/* ---- microbenchmark wave helpers (incoherent traversal workload) ----------
 * Primary rays: generateEyeRaysKernel (lib/RenderCore_PrimeRef/kernels/camera.h:39-68)
 * with pixel jitter from WangHash(index + R0). */
LH2HIP_API int lh2hip_generate_primary(const ViewPyramid* view, int width, int height, uint32_t R0,
	float tmin, Ray4* raysOut, uint32_t* pathIdOut, void* stream);
/* Next diffuse wave: for each hit ray, a cosine-weighted bounce about the
 * geometric normal flipped toward the incoming ray (seed WangHash(pathId*17+R0));
 * misses die; survivors compacted through an atomic counter (*countOut, device). */
LH2HIP_API int lh2hip_bounce_diffuse(lh2hip_core* core, const Ray4* raysIn, const HitRecord* hits,
	const uint32_t* pathIdIn, int count, uint32_t R0, float tmin,
	Ray4* raysOut, uint32_t* pathIdOut, uint32_t* countOut, void* stream);
/* Copies the core's leaf-ordered BVH statistics (node count, leaf count, depth,
 * bytes) for the current scene: out[0..7]. */
LH2HIP_API int lh2hip_bvh_info(lh2hip_core* core, int meshIdx, int64_t* out8);
/* Downloads the binary build tree of a mesh in canonical form (for
 * structure comparison): per node {lo.xyz, hi.xyz, leftOrFirst, count} as
 * 8 x 4 bytes, and the leaf primitive list. Returns node count in *nodeCount. */
LH2HIP_API int lh2hip_bvh_download(lh2hip_core* core, int meshIdx, void* nodesOut, int maxNodes,
	uint32_t* primsOut, int maxPrims, int* nodeCount);
/* Downloads the traversal layout the device built for a mesh (inspection of
 * the build that replaces OptiX's, lib/rendercore_optix7/core_mesh.cpp:105-123):
 * the f32 binary nodes (64 B each, bvh_info out[3] of them), the 4-wide q8
 * nodes (64 B each, out[4]) and the leaf triangles (3 x float4 per triangle,
 * v0.w = primitive id | bit 31 on a leaf's last triangle). A null pointer skips
 * that array; each array must hold the full count. */
LH2HIP_API int lh2hip_bvh_layout_download(lh2hip_core* core, int meshIdx, void* f32Out, int maxF32,
	void* q8Out, int maxQ8, void* trisOut, int maxTris);

/* ---- SVGF ----------------------------------------------------------------
 * prepareFilter + applyFilter x 3 + finalizeNoTAA of RenderCore_Optix7Filter
 * (lib/CUDA/shared_kernel_code/finalize_shared.h:169-484,587-600; driver
 * lib/RenderCore_Optix7Filter/rendercore.cpp:897-935). The handle keeps the
 * history (previous world positions, luminance moments, phase-1 output,
 * history counts). Frame inputs are device arrays of w*h pixels:
 * accumulator float4[2wh] (direct sums, then indirect sums), features uint4
 * {HDRtoRGB32 albedo, PackNormal2 normal, depth bits, specular << 4 | matid << 6},
 * worldPos float4 {xyz, packed normal | specular bits}, deltaDepth float4
 * (z, w used); out float4[wh]: sqrt of the gamma-2 filtered image (border
 * pixels left untouched, as the reference). */
typedef struct lh2hip_svgf lh2hip_svgf;
LH2HIP_API int lh2hip_svgf_create(int width, int height, lh2hip_svgf** out);
LH2HIP_API int lh2hip_svgf_destroy(lh2hip_svgf* f);
LH2HIP_API int lh2hip_svgf_frame(lh2hip_svgf* f, const lh2_float4* accumulator, const uint32_t* features, const lh2_float4* worldPos,
	const lh2_float4* deltaDepth, const ViewPyramid* prevView, float j0, float j1, float prevj0, float prevj1, int samplesTaken,
	float clampDirect, float clampIndirect, int camIsStationary, lh2_float4* out, void* stream);
/* TAA after the last pass (RenderCore_Optix7Filter "TAA" setting,
 * rendercore.cpp:677,922-926): 1 = TAApass + unsharpenTAA give `out`
 * (squared, border untouched), 0 (the default of this object) = finalizeNoTAA. */
LH2HIP_API int lh2hip_svgf_set_taa(lh2hip_svgf* f, int enabled);
/* Copies internal state to host memory: which 0 features (uint4), 1 luminance
 * moments (float4), 2 motion (float2), 3 phase-1 filtered history (float4),
 * 4 last-pass image (float4), 5 world positions of the last frame (float4),
 * 6 TAA history: the last TAA output (float4). */
LH2HIP_API int lh2hip_svgf_read(lh2hip_svgf* f, int which, void* hostOut);

/* ---- learned denoiser kernels --------------------------------------------
 * No counterpart in the reference. A 5-level 3x3 convolutional U-Net is
 * composed from these calls (lighthouse2_amd/denoiser.py); tensors are NHWC
 * device arrays, activations and weights bf16 (uint16 bits), accumulation
 * fp32 on the matrix cores. Channel counts are multiples of 16. */
enum { LH2HIP_DN_RELU = 1, LH2HIP_DN_OUT_F32 = 2 };
/* y = act(conv3x3(x, w) + bias): x [H][W][Cin] bf16, w [9][Cout][Cin] bf16
 * (tap = 3 dy + dx), bias [Cout] fp32 or NULL, y bf16 (fp32 with
 * LH2HIP_DN_OUT_F32); Cout <= 64. With the flipped weights of lh2hip_dn_adam
 * this is also the input gradient of a convolution. */
LH2HIP_API int lh2hip_dn_conv3x3(const void* x, int Cin, const void* w, const float* bias, void* y, int Cout, int H, int W, int flags, void* stream);
/* y = relu(conv3x3(x, w) + bias) as above (Cin 16 or 32, Cout 32, bf16) and,
 * written from the same tile, ypool = lh2hip_dn_pool2(y) (H, W even). */
LH2HIP_API int lh2hip_dn_conv3x3_pool(const void* x, int Cin, const void* w, const float* bias, void* y, int Cout, int H, int W, void* ypool,
	void* stream);
/* The convolution above (Cout 32) of the 64-channel concatenation
 * [nearest 2x upsample of coarse [H/2][W/2][32], skip [H][W][32]] without
 * materialising it (H, W even): lh2hip_dn_upcat then lh2hip_dn_conv3x3. */
LH2HIP_API int lh2hip_dn_conv3x3_cat(const void* coarse, const void* skip, const void* w, const float* bias, void* y, int Cout, int H, int W,
	int flags, void* stream);
LH2HIP_API uint64_t lh2hip_dn_dw_workspace_bytes(int Cin, int Cout);
/* dw [9][Cout][Cin] and db [Cout] (fp32, overwritten) of a 3x3 convolution
 * from its input x and output gradient dy; ws: workspace bytes above. */
LH2HIP_API int lh2hip_dn_conv3x3_dw(const void* x, int Cin, const void* dy, int Cout, int H, int W, float* dw, float* db, void* ws, void* stream);
/* No counterpart in the reference: lh2hip_dn_conv3x3_dw for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops). The workspace is a file-scope device array of
 * lh2hip_dn_dw_scratch_bytes() (37.8 MB, each device its own copy, held
 * until the process ends), enough for every supported shape; shapes beyond
 * it are rejected. No hipMalloc / hipFree. Calls that overlap on one device
 * share the array and race; results equal lh2hip_dn_conv3x3_dw's. */
LH2HIP_API uint64_t lh2hip_dn_dw_scratch_bytes(void);
LH2HIP_API int lh2hip_dn_conv3x3_dw_noalloc(const void* x, int Cin, const void* dy, int Cout, int H, int W, float* dw, float* db, void* stream);
LH2HIP_API int lh2hip_dn_relu_bwd(const void* y, const void* dy, void* dyPre, uint64_t n, void* stream);
LH2HIP_API int lh2hip_dn_pool2(const void* x, int C, int H, int W, void* y, void* stream);
LH2HIP_API int lh2hip_dn_pool2_bwd(const void* dy, int C, int H, int W, void* dx, int accumulate, void* stream);
LH2HIP_API int lh2hip_dn_upcat(const void* coarse, int C0, const void* skip, int C1, int H, int W, void* out, void* stream);
LH2HIP_API int lh2hip_dn_upcat_bwd(const void* dcat, int C0, int C1, int H, int W, void* dcoarse, void* dskip, void* stream);
LH2HIP_API int lh2hip_dn_pack_input(const lh2_float4* radiance, const lh2_float4* albedo, const lh2_float4* normal, const float* depth,
	int H, int W, void* out, void* stream);
/* Network input of a filter-estimator frame: acc32 float4 [2 w h] (direct,
 * then indirect sums), features uint4 [w h] (the SVGF G-buffer); out bf16
 * [Hp][Wp][16] = log(1 + radiance invSpp / max(albedo, 0.02)), albedo,
 * normal, 0.1 log(1 + min(depth, 1e4)), zero padding beyond w x h. */
LH2HIP_API int lh2hip_dn_frame_input(const lh2_float4* acc32, const uint32_t* features, float invSpp, int w, int h, int Hp, int Wp, void* out,
	void* stream);
/* Denoised display image from the network's fp32 output [Hp][Wp][16]:
 * (exp(o) - 1) max(albedo, 0.02) per channel, raised to 1/4 (the filter
 * core's display convention: the sqrt of SVGF's sqrt last pass). */
LH2HIP_API int lh2hip_dn_frame_output(const float* out, int Hp, int Wp, const uint32_t* features, int w, int h, lh2_float4* image, void* stream);
/* The output convolution fused with the frame conversion, for inference:
 * conv3x3(x) + bias (32 -> 16 channels) written as the display image of a
 * w x h frame inside the Hp x Wp network image, as lh2hip_dn_conv3x3 (fp32
 * out) then lh2hip_dn_frame_output compute it (bit for bit). No
 * counterpart in the reference. */
LH2HIP_API int lh2hip_dn_conv3x3_display(const void* x, const void* wts, const float* bias, const uint32_t* features, int w, int h,
	lh2_float4* image, int Hp, int Wp, void* stream);
LH2HIP_API int lh2hip_dn_relmse(const float* out, int Cout, const lh2_float4* target, int H, int W, void* dout, float* loss, void* stream);
/* Adam step (g NULL: no step, only the bf16 copies); wfwd / wbwd (NULL to
 * skip): bf16 weights in the forward layout and flipped for the input gradient. */
LH2HIP_API int lh2hip_dn_adam(float* p, float* m, float* v, const float* g, uint64_t n, float lr, float b1, float b2, float eps, int t,
	void* wfwd, void* wbwd, int Cout, int Cin, void* stream);

/* ---- render job server ---------------------------------------------------
 * No counterpart in the reference (its nearest is the asynchronous Render +
 * RenderThread of lib/rendercore_optix7/rendercore.cpp:803-847). A queue of
 * render jobs over the core's current scene: mixed resolutions, spp and
 * depths, admitted while their buffers (64 B per pixel) fit budgetBytes of
 * device memory (a job that does not fit lets smaller later ones pass, at
 * most 8 times), run by a worker thread in passes of at most passPaths paths
 * that carry up to 16 jobs' paths together, least remaining work first. A
 * finished job's image, then its fp64 accumulator, are copied to host
 * memory by a second server thread while the next passes run (the job is
 * done once the image is there); its device buffers are then kept in a pool
 * for later jobs (at most budgetBytes more, freed with the server). A
 * job's image equals RenderCore::Render(view, Restart) with the same size,
 * spp, depth and seed, bit for bit. Submit from any thread. While jobs are
 * pending, the calls that change the scene (SetGeometry, SetInstance,
 * FinalizeInstances, SetMaterials, SetLights, SetTextures, SetSkyData,
 * SetBlueNoise, the nodeFormat and refit settings) return LH2HIP_ERR_STATE,
 * and lh2hip_destroy of the core fails while a server of it exists. w * h *
 * spp is at most 2^32 (so is SetTarget's). */
typedef struct lh2hip_server lh2hip_server;
LH2HIP_API int lh2hip_server_create(lh2hip_core* core, uint64_t budgetBytes, uint64_t passPaths, lh2hip_server** out);
LH2HIP_API int lh2hip_server_destroy(lh2hip_server* server);
LH2HIP_API int lh2hip_server_submit(lh2hip_server* server, const ViewPyramid* view, int width, int height, int spp,
	int maxDepth, uint32_t seed, uint64_t* jobId);
/* As lh2hip_server_submit, with denoise 1: the job is rendered by the filter
 * estimator with its own G-buffer and filtered by SVGF, its image equal to
 * RenderCore::Render(view, Restart) with the filter estimator, TAA off and
 * camera seed `seed` on a fresh core (RenderCore_Optix7Filter's
 * FinalizeRender for a first frame); budget 160 B per pixel. denoise 2: the
 * same frame denoised by the learned U-Net (lh2hip_server_set_denoiser).
 * denoise 0 is lh2hip_server_submit. No counterpart in the reference (its
 * cores have no job server). */
LH2HIP_API int lh2hip_server_submit_denoised(lh2hip_server* server, const ViewPyramid* view, int width, int height, int spp,
	int maxDepth, uint32_t seed, int denoise, uint64_t* jobId);
/* Weights of the learned denoiser for jobs submitted with denoise 2
 * (lh2hip_dn_weights_bytes() bytes: per layer of the U-Net in forward order,
 * bf16 weights [9][Cout][Cin] then fp32 bias [Cout]). With denoise 2 a job is
 * rendered as with denoise 1 and its frame goes through lh2hip_dn_frame_input,
 * the U-Net forward pass and lh2hip_dn_frame_output instead of SVGF. No
 * counterpart in the reference. */
LH2HIP_API int lh2hip_server_set_denoiser(lh2hip_server* server, const void* weights, uint64_t bytes);
LH2HIP_API uint64_t lh2hip_dn_weights_bytes(void);
/* paused 1: the worker finishes the pass it runs and starts no other until
 * paused 0 (a burst of jobs submitted meanwhile shares passes from the start). */
LH2HIP_API int lh2hip_server_pause(lh2hip_server* server, int paused);
/* Batching counters (out8): passes run, job parts over all passes, passes
 * shared by more than one job, the most jobs in one pass, paths (eye rays),
 * extension rays traced over all waves, and the device time of the
 * extension-traversal and shade launches in microseconds. */
LH2HIP_API int lh2hip_server_stats(lh2hip_server* server, uint64_t* out8);
/* *done = 1 once the job finished (or failed: lh2hip_server_read reports it). */
LH2HIP_API int lh2hip_server_done(lh2hip_server* server, uint64_t jobId, int* done);
/* Blocks until the job is done; timeoutSeconds < 0 waits without limit.
 * Returns LH2HIP_ERR_STATE on timeout. */
LH2HIP_API int lh2hip_server_wait(lh2hip_server* server, uint64_t jobId, double timeoutSeconds);
/* Copies the finished job's image (w*h float4, host) and optionally its fp64
 * accumulator (w*h*4 doubles, may be NULL; waits until it is on the host),
 * then releases the job. */
LH2HIP_API int lh2hip_server_read(lh2hip_server* server, uint64_t jobId, lh2_float4* image, double* acc64);
/* Seconds since the server started at which the job was queued, admitted
 * and finished (out3). */
LH2HIP_API int lh2hip_server_times(lh2hip_server* server, uint64_t jobId, double* out3);

#ifdef __cplusplus
}
#endif

#endif /* LH2HIP_H */
