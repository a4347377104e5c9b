/* capi.cpp - C ABI of librendercore_hip.so and the CreateCore plugin export.
 * Every entry point catches the core's exceptions and returns a status code;
 * the message is kept per thread for lh2hip_last_error(). */
#include "../../include/lh2hip.h"
#include "rendercore.h"
#include "server.h"
#include "denoise.h"
#include "svgf.h"
#include <cstring>
#include <new>

using lh::RenderCore;
using lh::CoreError;

struct lh2hip_core { RenderCore core; };
struct lh2hip_server { lh::RenderServer* srv; };
struct lh2hip_svgf { lh::Svgf* f; };

static thread_local std::string g_lastError;

static int fail(int code, const char* msg)
{
	g_lastError = msg;
	return code;
}

template <typename F>
static int guarded(lh2hip_core* c, F f)
{
	if (!c) return fail(LH2HIP_ERR_INVALID, "null core handle");
	try
	{
		f(c->core);
		return LH2HIP_OK;
	}
	catch (const CoreError& e) { return fail(e.code, e.what()); }
	catch (const std::bad_alloc&) { return fail(LH2HIP_ERR_OOM, "host allocation failed"); }
	catch (const std::exception& e) { return fail(LH2HIP_ERR_DEVICE, e.what()); }
}

template <typename F>
static int guarded_server(lh2hip_server* s, F f)
{
	if (!s || !s->srv) return fail(LH2HIP_ERR_INVALID, "null server handle");
	try
	{
		f(*s->srv);
		return LH2HIP_OK;
	}
	catch (const CoreError& e) { return fail(e.code, e.what()); }
	catch (const std::bad_alloc&) { return fail(LH2HIP_ERR_OOM, "host allocation failed"); }
	catch (const std::exception& e) { return fail(LH2HIP_ERR_DEVICE, e.what()); }
}

extern "C" {

LH2HIP_API const char* lh2hip_last_error(void) { return g_lastError.c_str(); }
LH2HIP_API int lh2hip_version(void) { return 1; }

LH2HIP_API int lh2hip_create(lh2hip_core** out)
{
	if (!out) return fail(LH2HIP_ERR_INVALID, "null output pointer");
	*out = nullptr;
	lh2hip_core* c = nullptr;
	try
	{
		c = new lh2hip_core();
		c->core.Init();
	}
	catch (const CoreError& e) { delete c; return fail(e.code, e.what()); }
	catch (const std::exception& e) { delete c; return fail(LH2HIP_ERR_DEVICE, e.what()); }
	*out = c;
	return LH2HIP_OK;
}

LH2HIP_API int lh2hip_destroy(lh2hip_core* core)
{
	if (!core) return LH2HIP_OK;
	if (core->core.serverCount.load() > 0) return fail(LH2HIP_ERR_STATE, "destroy: destroy the job servers of this core first");
	int r = guarded(core, [](RenderCore& c) { c.Shutdown(); });
	delete core;
	return r;
}

LH2HIP_API int lh2hip_set_target(lh2hip_core* core, int w, int h, int spp)
{
	return guarded(core, [&](RenderCore& c) { c.SetTargetSize(w, h, spp); });
}

LH2HIP_API int lh2hip_setting(lh2hip_core* core, const char* name, float value)
{
	return guarded(core, [&](RenderCore& c) { c.Setting(name, value); });
}

LH2HIP_API int lh2hip_set_geometry(lh2hip_core* core, int meshIdx, const lh2_float4* v, int vc, int tc, const CoreTri* tris)
{
	return guarded(core, [&](RenderCore& c) { c.SetGeometry(meshIdx, v, vc, tc, tris); });
}

LH2HIP_API int lh2hip_set_instance(lh2hip_core* core, int instanceIdx, int meshIdx, const lh2_mat4* transform)
{
	return guarded(core, [&](RenderCore& c) {
		lh2_mat4 I = { { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 } };
		c.SetInstance(instanceIdx, meshIdx, transform ? *transform : I);
	});
}

LH2HIP_API int lh2hip_finalize_instances(lh2hip_core* core)
{
	return guarded(core, [&](RenderCore& c) { c.FinalizeInstances(); });
}

LH2HIP_API int lh2hip_set_materials(lh2hip_core* core, const CoreMaterial* mats, int count)
{
	return guarded(core, [&](RenderCore& c) { c.SetMaterials(const_cast<CoreMaterial*>(mats), count); });
}

LH2HIP_API int lh2hip_set_lights(lh2hip_core* core, const CoreLightTri* t, int nt, const CorePointLight* p, int np,
	const CoreSpotLight* s, int ns, const CoreDirectionalLight* d, int nd)
{
	return guarded(core, [&](RenderCore& c) { c.SetLights(t, nt, p, np, s, ns, d, nd); });
}

LH2HIP_API int lh2hip_set_textures(lh2hip_core* core, const CoreTexDesc* tex, int count)
{
	return guarded(core, [&](RenderCore& c) { c.SetTextures(tex, count); });
}

LH2HIP_API int lh2hip_set_sky(lh2hip_core* core, const lh2_float3* pixels, int w, int h, const lh2_mat4* worldToLight)
{
	return guarded(core, [&](RenderCore& c) {
		lh2_mat4 I = { { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 } };
		if (w < 0 || h < 0) throw CoreError(LH2HIP_ERR_INVALID, "set_sky: negative size");
		c.SetSkyData(pixels, (uint32_t)w, (uint32_t)h, worldToLight ? *worldToLight : I);
	});
}

LH2HIP_API int lh2hip_set_blue_noise(lh2hip_core* core, const uint8_t* sequence, const uint8_t* scrambling, const uint8_t* ranking)
{
	return guarded(core, [&](RenderCore& c) { c.SetBlueNoise(sequence, scrambling, ranking); });
}

LH2HIP_API int lh2hip_set_probe_pos(lh2hip_core* core, int x, int y)
{
	return guarded(core, [&](RenderCore& c) { lh2_int2 p; p.x = x, p.y = y; c.SetProbePos(p); });
}

LH2HIP_API int lh2hip_render(lh2hip_core* core, const ViewPyramid* view, int converge, int async)
{
	return guarded(core, [&](RenderCore& c) {
		if (!view) throw CoreError(LH2HIP_ERR_INVALID, "render: null view");
		c.Render(*view, converge == LH2_RESTART ? lighthouse2::Restart : lighthouse2::Converge, async != 0);
	});
}

LH2HIP_API int lh2hip_wait(lh2hip_core* core)
{
	return guarded(core, [&](RenderCore& c) { c.WaitForRender(); });
}

LH2HIP_API int lh2hip_read_image(lh2hip_core* core, lh2_float4* out)
{
	return guarded(core, [&](RenderCore& c) { c.ReadImage(out); });
}

LH2HIP_API int lh2hip_read_accumulator_f64(lh2hip_core* core, double* out)
{
	return guarded(core, [&](RenderCore& c) { c.ReadAccumulatorF64(out); });
}

LH2HIP_API int lh2hip_read_buffer(lh2hip_core* core, int which, void* out, size_t bytes)
{
	return guarded(core, [&](RenderCore& c) { c.ReadBuffer(which, out, bytes); });
}

LH2HIP_API int lh2hip_get_stats(lh2hip_core* core, CoreStats* out)
{
	return guarded(core, [&](RenderCore& c) {
		if (!out) throw CoreError(LH2HIP_ERR_INVALID, "get_stats: null");
		*out = c.GetCoreStats();
	});
}

LH2HIP_API int lh2hip_trace_closest(lh2hip_core* core, const Ray4* rays, HitRecord* hits, int count, void* stream)
{
	return guarded(core, [&](RenderCore& c) { c.TraceClosest(rays, hits, count, (hipStream_t)stream); });
}

LH2HIP_API int lh2hip_trace_closest_counted(lh2hip_core* core, const Ray4* rays, HitRecord* hits, int count, void* stream, uint64_t* out2)
{
	return guarded(core, [&](RenderCore& c) {
		if (!out2) throw CoreError(LH2HIP_ERR_INVALID, "trace_closest_counted: null output");
		c.TraceClosestCounted(rays, hits, count, (hipStream_t)stream, out2);
	});
}

LH2HIP_API int lh2hip_trace_any(lh2hip_core* core, const Ray4* rays, uint32_t* occluded, int count, void* stream)
{
	return guarded(core, [&](RenderCore& c) { c.TraceAny(rays, occluded, count, (hipStream_t)stream); });
}

LH2HIP_API int lh2hip_generate_primary(const ViewPyramid* view, int w, int h, uint32_t R0, float tmin, Ray4* raysOut,
	uint32_t* pathIdOut, void* stream)
{
	if (!view || !raysOut || w <= 0 || h <= 0) return fail(LH2HIP_ERR_INVALID, "generate_primary: bad arguments");
	lh::CamParams c;
	memcpy(c.pos, &view->pos, 12), memcpy(c.p1, &view->p1, 12), memcpy(c.p2, &view->p2, 12), memcpy(c.p3, &view->p3, 12);
	c.aperture = view->aperture, c.spreadAngle = view->spreadAngle, c.w = w, c.h = h;
	const hipError_t e = lh::launch_primary(c, R0, tmin, (float4*)raysOut, pathIdOut, (hipStream_t)stream);
	if (e != hipSuccess) return fail(LH2HIP_ERR_DEVICE, hipGetErrorString(e));
	return LH2HIP_OK;
}

LH2HIP_API int lh2hip_bounce_diffuse(lh2hip_core* core, const Ray4* raysIn, const HitRecord* hits, const uint32_t* pathIdIn,
	int count, uint32_t R0, float tmin, Ray4* raysOut, uint32_t* pathIdOut, uint32_t* countOut, void* stream)
{
	return guarded(core, [&](RenderCore& c) {
		if (count < 0 || (count > 0 && (!raysIn || !hits || !pathIdIn || !raysOut || !pathIdOut)) || !countOut)
			throw CoreError(LH2HIP_ERR_INVALID, "bounce_diffuse: bad arguments");
		c.Bounce(raysIn, hits, pathIdIn, count, R0, tmin, raysOut, pathIdOut, countOut, (hipStream_t)stream);
	});
}

LH2HIP_API int lh2hip_bvh_info(lh2hip_core* core, int meshIdx, int64_t* out8)
{
	return guarded(core, [&](RenderCore& c) {
		if (!out8) throw CoreError(LH2HIP_ERR_INVALID, "bvh_info: null");
		c.BvhInfo(meshIdx, out8);
	});
}

LH2HIP_API int lh2hip_bvh_download(lh2hip_core* core, int meshIdx, void* nodesOut, int maxNodes, uint32_t* primsOut, int maxPrims, int* nodeCount)
{
	return guarded(core, [&](RenderCore& c) {
		const int n = c.BvhDownload(meshIdx, nodesOut, maxNodes, primsOut, maxPrims);
		if (nodeCount) *nodeCount = n;
	});
}

LH2HIP_API int lh2hip_bvh_layout_download(lh2hip_core* core, int meshIdx, void* f32Out, int maxF32, void* q8Out, int maxQ8, void* trisOut, int maxTris)
{
	return guarded(core, [&](RenderCore& c) { c.BvhLayoutDownload(meshIdx, f32Out, maxF32, q8Out, maxQ8, trisOut, maxTris); });
}


LH2HIP_API int lh2hip_server_create(lh2hip_core* core, uint64_t budgetBytes, uint64_t passPaths, lh2hip_server** out)
{
	if (!out) return fail(LH2HIP_ERR_INVALID, "null output");
	*out = nullptr;
	return guarded(core, [&](RenderCore& c) {
		lh2hip_server* s = new lh2hip_server;
		try { s->srv = new lh::RenderServer(c, budgetBytes, (size_t)passPaths); }
		catch (...) { delete s; throw; }
		*out = s;
	});
}

LH2HIP_API int lh2hip_server_destroy(lh2hip_server* server)
{
	if (!server) return LH2HIP_OK;
	delete server->srv;
	delete server;
	return LH2HIP_OK;
}

LH2HIP_API int lh2hip_server_submit(lh2hip_server* server, const ViewPyramid* view, int width, int height, int spp,
	int maxDepth, uint32_t seed, uint64_t* jobId)
{
	return guarded_server(server, [&](lh::RenderServer& s) {
		if (!view || !jobId) throw CoreError(LH2HIP_ERR_INVALID, "server submit: null argument");
		lighthouse2::ViewPyramid v;
		memcpy(&v, view, sizeof(v));
		*jobId = s.Submit(v, width, height, spp, maxDepth, seed);
	});
}

LH2HIP_API int lh2hip_server_submit_denoised(lh2hip_server* server, const ViewPyramid* view, int width, int height, int spp,
	int maxDepth, uint32_t seed, int denoise, uint64_t* jobId)
{
	return guarded_server(server, [&](lh::RenderServer& s) {
		if (!view || !jobId) throw CoreError(LH2HIP_ERR_INVALID, "server submit: null argument");
		lighthouse2::ViewPyramid v;
		memcpy(&v, view, sizeof(v));
		*jobId = s.Submit(v, width, height, spp, maxDepth, seed, denoise);
	});
}

LH2HIP_API int lh2hip_server_done(lh2hip_server* server, uint64_t jobId, int* done)
{
	return guarded_server(server, [&](lh::RenderServer& s) {
		if (!done) throw CoreError(LH2HIP_ERR_INVALID, "null output");
		*done = s.Done(jobId) ? 1 : 0;
	});
}

LH2HIP_API int lh2hip_server_wait(lh2hip_server* server, uint64_t jobId, double timeoutSeconds)
{
	return guarded_server(server, [&](lh::RenderServer& s) {
		if (!s.Wait(jobId, timeoutSeconds)) throw CoreError(LH2HIP_ERR_STATE, "server wait: timed out");
	});
}

LH2HIP_API int lh2hip_server_read(lh2hip_server* server, uint64_t jobId, lh2_float4* image, double* acc64)
{
	return guarded_server(server, [&](lh::RenderServer& s) { s.Read(jobId, image, acc64); });
}

LH2HIP_API int lh2hip_server_set_denoiser(lh2hip_server* server, const void* weights, uint64_t bytes)
{
	return guarded_server(server, [&](lh::RenderServer& s) { s.SetDenoiser(weights, (size_t)bytes); });
}

LH2HIP_API int lh2hip_server_pause(lh2hip_server* server, int paused)
{
	return guarded_server(server, [&](lh::RenderServer& s) { s.Pause(paused != 0); });
}

LH2HIP_API uint64_t lh2hip_dn_weights_bytes(void) { return (uint64_t)lh::UNetForward::BlobBytes(); }

LH2HIP_API int lh2hip_server_stats(lh2hip_server* server, uint64_t* out8)
{
	return guarded_server(server, [&](lh::RenderServer& s) {
		if (!out8) throw CoreError(LH2HIP_ERR_INVALID, "null output");
		const lh::ServerStats t = s.Stats();
		out8[0] = t.passes, out8[1] = t.parts, out8[2] = t.sharedPasses, out8[3] = t.maxParts, out8[4] = t.paths, out8[5] = t.rays;
		out8[6] = (uint64_t)(t.traceMs * 1e3), out8[7] = (uint64_t)(t.shadeMs * 1e3);
	});
}

LH2HIP_API int lh2hip_server_times(lh2hip_server* server, uint64_t jobId, double* out3)
{
	return guarded_server(server, [&](lh::RenderServer& s) {
		if (!out3) throw CoreError(LH2HIP_ERR_INVALID, "null output");
		const lh::JobTimes t = s.Times(jobId);
		out3[0] = t.queued, out3[1] = t.admitted, out3[2] = t.done;
	});
}


static int dn_status(hipError_t e, const char* what)
{
	if (e == hipSuccess) return LH2HIP_OK;
	if (e == hipErrorInvalidValue) return fail(LH2HIP_ERR_INVALID, what);
	g_lastError = std::string(what) + ": " + hipGetErrorString(e);
	return LH2HIP_ERR_DEVICE;
}

LH2HIP_API int lh2hip_dn_conv3x3(const void* x, int Cin, const void* w, const float* bias, void* y, int Cout, int H, int W, int flags, void* stream)
{
	if (!x || !w || !y) return fail(LH2HIP_ERR_INVALID, "dn_conv3x3: null pointer");
	return dn_status(lh::dn_conv3x3((const lh::bf16_t*)x, Cin, (const lh::bf16_t*)w, bias, y, Cout, H, W, flags, (hipStream_t)stream),
		"dn_conv3x3 (Cin, Cout multiples of 16, Cout <= 64)");
}

LH2HIP_API int lh2hip_dn_conv3x3_pool(const void* x, int Cin, const void* w, const float* bias, void* y, int Cout, int H, int W, void* ypool,
	void* stream)
{
	if (!x || !w || !y || !ypool) return fail(LH2HIP_ERR_INVALID, "dn_conv3x3_pool: null pointer");
	return dn_status(lh::dn_conv3x3_pool((const lh::bf16_t*)x, Cin, (const lh::bf16_t*)w, bias, (lh::bf16_t*)y, Cout, H, W, (lh::bf16_t*)ypool,
		(hipStream_t)stream), "dn_conv3x3_pool (Cin 16 or 32, Cout 32, even H, W)");
}

LH2HIP_API int lh2hip_dn_conv3x3_cat(const void* coarse, const void* skip, const void* w, const float* bias, void* y, int Cout, int H, int W,
	int flags, void* stream)
{
	if (!coarse || !skip || !w || !y) return fail(LH2HIP_ERR_INVALID, "dn_conv3x3_cat: null pointer");
	return dn_status(lh::dn_conv3x3_cat((const lh::bf16_t*)coarse, (const lh::bf16_t*)skip, (const lh::bf16_t*)w, bias, y, Cout, H, W, flags,
		(hipStream_t)stream), "dn_conv3x3_cat (32 + 32 channels in, Cout 32, even H, W)");
}

LH2HIP_API uint64_t lh2hip_dn_dw_workspace_bytes(int Cin, int Cout) { return lh::dn_dw_workspace_bytes(Cin, Cout); }

LH2HIP_API int lh2hip_dn_conv3x3_dw(const void* x, int Cin, const void* dy, int Cout, int H, int W, float* dw, float* db, void* ws, void* stream)
{
	if (!x || !dy || !dw || !ws) return fail(LH2HIP_ERR_INVALID, "dn_conv3x3_dw: null pointer");
	return dn_status(lh::dn_conv3x3_dw((const lh::bf16_t*)x, Cin, (const lh::bf16_t*)dy, Cout, H, W, dw, db, ws, (hipStream_t)stream),
		"dn_conv3x3_dw (Cin x Cout in 16x32, 32x32, 64x32, 32x16, 16x16)");
}

LH2HIP_API int lh2hip_dn_conv3x3_display(const void* x, const void* wts, const float* bias, const uint32_t* features, int w, int h,
	lh2_float4* image, int Hp, int Wp, void* stream)
{
	if (!x || !wts || !features || !image) return fail(LH2HIP_ERR_INVALID, "dn_conv3x3_display: null pointer");
	return dn_status(lh::dn_conv3x3_display((const lh::bf16_t*)x, (const lh::bf16_t*)wts, bias, (const uint4*)features, w, h, (float4*)image,
		Hp, Wp, (hipStream_t)stream), "dn_conv3x3_display (frame inside the Hp x Wp image)");
}

LH2HIP_API uint64_t lh2hip_dn_dw_scratch_bytes(void) { return lh::dn_dw_scratch_bytes(); }

LH2HIP_API int lh2hip_dn_conv3x3_dw_noalloc(const void* x, int Cin, const void* dy, int Cout, int H, int W, float* dw, float* db, void* stream)
{
	if (!x || !dy || !dw) return fail(LH2HIP_ERR_INVALID, "dn_conv3x3_dw_noalloc: null pointer");
	if (lh::dn_dw_workspace_bytes(Cin, Cout) > lh::dn_dw_scratch_bytes())
		return fail(LH2HIP_ERR_INVALID, "dn_conv3x3_dw_noalloc: the shape's workspace exceeds the built-in scratch array");
	return dn_status(lh::dn_conv3x3_dw_noalloc((const lh::bf16_t*)x, Cin, (const lh::bf16_t*)dy, Cout, H, W, dw, db, (hipStream_t)stream),
		"dn_conv3x3_dw_noalloc (Cin x Cout in 16x32, 32x32, 64x32, 32x16, 16x16; W <= 32768)");
}

LH2HIP_API int lh2hip_dn_relu_bwd(const void* y, const void* dy, void* dyPre, uint64_t n, void* stream)
{
	return dn_status(lh::dn_relu_bwd((const lh::bf16_t*)y, (const lh::bf16_t*)dy, (lh::bf16_t*)dyPre, (size_t)n, (hipStream_t)stream), "dn_relu_bwd");
}

LH2HIP_API int lh2hip_dn_pool2(const void* x, int C, int H, int W, void* y, void* stream)
{
	return dn_status(lh::dn_pool2((const lh::bf16_t*)x, C, H, W, (lh::bf16_t*)y, (hipStream_t)stream), "dn_pool2 (even H, W)");
}

LH2HIP_API int lh2hip_dn_pool2_bwd(const void* dy, int C, int H, int W, void* dx, int accumulate, void* stream)
{
	return dn_status(lh::dn_pool2_bwd((const lh::bf16_t*)dy, C, H, W, (lh::bf16_t*)dx, accumulate, (hipStream_t)stream), "dn_pool2_bwd (even H, W)");
}

LH2HIP_API int lh2hip_dn_upcat(const void* coarse, int C0, const void* skip, int C1, int H, int W, void* out, void* stream)
{
	return dn_status(lh::dn_upcat((const lh::bf16_t*)coarse, C0, (const lh::bf16_t*)skip, C1, H, W, (lh::bf16_t*)out, (hipStream_t)stream), "dn_upcat (even H, W)");
}

LH2HIP_API int lh2hip_dn_upcat_bwd(const void* dcat, int C0, int C1, int H, int W, void* dcoarse, void* dskip, void* stream)
{
	return dn_status(lh::dn_upcat_bwd((const lh::bf16_t*)dcat, C0, C1, H, W, (lh::bf16_t*)dcoarse, (lh::bf16_t*)dskip, (hipStream_t)stream),
		"dn_upcat_bwd (even H, W)");
}

LH2HIP_API int lh2hip_dn_pack_input(const lh2_float4* radiance, const lh2_float4* albedo, const lh2_float4* normal, const float* depth,
	int H, int W, void* out, void* stream)
{
	return dn_status(lh::dn_pack_input((const float4*)radiance, (const float4*)albedo, (const float4*)normal, depth, H, W, (lh::bf16_t*)out,
		(hipStream_t)stream), "dn_pack_input");
}

LH2HIP_API int lh2hip_dn_frame_input(const lh2_float4* acc32, const uint32_t* features, float invSpp, int w, int h, int Hp, int Wp, void* out,
	void* stream)
{
	if (!acc32 || !features || !out) return fail(LH2HIP_ERR_INVALID, "dn_frame_input: null pointer");
	return dn_status(lh::dn_frame_input((const float4*)acc32, (const uint4*)features, invSpp, w, h, Hp, Wp, (lh::bf16_t*)out, (hipStream_t)stream),
		"dn_frame_input (Hp >= h, Wp >= w)");
}

LH2HIP_API int lh2hip_dn_frame_output(const float* out, int Hp, int Wp, const uint32_t* features, int w, int h, lh2_float4* image, void* stream)
{
	if (!out || !features || !image) return fail(LH2HIP_ERR_INVALID, "dn_frame_output: null pointer");
	return dn_status(lh::dn_frame_output(out, Hp, Wp, (const uint4*)features, w, h, (float4*)image, (hipStream_t)stream),
		"dn_frame_output (Hp >= h, Wp >= w)");
}

LH2HIP_API int lh2hip_dn_relmse(const float* out, int Cout, const lh2_float4* target, int H, int W, void* dout, float* loss, void* stream)
{
	return dn_status(lh::dn_relmse(out, Cout, (const float4*)target, H, W, (lh::bf16_t*)dout, loss, (hipStream_t)stream), "dn_relmse");
}

LH2HIP_API int lh2hip_dn_adam(float* p, float* m, float* v, const float* g, uint64_t n, float lr, float b1, float b2, float eps, int t,
	void* wfwd, void* wbwd, int Cout, int Cin, void* stream)
{
	return dn_status(lh::dn_adam(p, m, v, g, (size_t)n, lr, b1, b2, eps, t, (lh::bf16_t*)wfwd, (lh::bf16_t*)wbwd, Cout, Cin, (hipStream_t)stream),
		"dn_adam (a flipped copy needs n = 9 Cout Cin)");
}


LH2HIP_API int lh2hip_svgf_create(int width, int height, lh2hip_svgf** out)
{
	if (!out) return fail(LH2HIP_ERR_INVALID, "null output");
	*out = nullptr;
	try
	{
		lh2hip_svgf* h = new lh2hip_svgf;
		try { h->f = new lh::Svgf(width, height); }
		catch (...) { delete h; throw; }
		*out = h;
		return LH2HIP_OK;
	}
	catch (const CoreError& e) { return fail(e.code, e.what()); }
	catch (const std::exception& e) { return fail(LH2HIP_ERR_DEVICE, e.what()); }
}

LH2HIP_API int lh2hip_svgf_destroy(lh2hip_svgf* f)
{
	if (f) delete f->f, delete f;
	return LH2HIP_OK;
}

LH2HIP_API int lh2hip_svgf_frame(lh2hip_svgf* f, const lh2_float4* accumulator, const uint32_t* features, const lh2_float4* worldPos,
	const lh2_float4* deltaDepth, const ViewPyramid* prevView, float j0, float j1, float prevj0, float prevj1, int samplesTaken,
	float clampDirect, float clampIndirect, int camIsStationary, lh2_float4* out, void* stream)
{
	if (!f || !f->f) return fail(LH2HIP_ERR_INVALID, "null svgf handle");
	if (!accumulator || !features || !worldPos || !deltaDepth || !prevView || !out) return fail(LH2HIP_ERR_INVALID, "svgf frame: null pointer");
	try
	{
		lighthouse2::ViewPyramid v;
		memcpy(&v, prevView, sizeof(v));
		f->f->Frame((const float4*)accumulator, (const uint4*)features, (const float4*)worldPos, (const float4*)deltaDepth, v, j0, j1, prevj0, prevj1,
			samplesTaken, clampDirect, clampIndirect, camIsStationary, (float4*)out, (hipStream_t)stream);
		return LH2HIP_OK;
	}
	catch (const CoreError& e) { return fail(e.code, e.what()); }
	catch (const std::exception& e) { return fail(LH2HIP_ERR_DEVICE, e.what()); }
}

LH2HIP_API int lh2hip_svgf_set_taa(lh2hip_svgf* f, int enabled)
{
	if (!f || !f->f) return fail(LH2HIP_ERR_INVALID, "null svgf handle");
	if (enabled != 0 && enabled != 1) return fail(LH2HIP_ERR_INVALID, "svgf set_taa: 0 or 1");
	f->f->taa = enabled;
	return LH2HIP_OK;
}

LH2HIP_API int lh2hip_svgf_read(lh2hip_svgf* f, int which, void* hostOut)
{
	if (!f || !f->f || !hostOut) return fail(LH2HIP_ERR_INVALID, "svgf read: null argument");
	lh::Svgf& s = *f->f;
	lh::DevBuf* b = which == 0 ? &s.features : which == 1 ? &s.prevMoments : which == 2 ? &s.motion : which == 3 ? &s.filteredIn
		: which == 4 ? &s.lastPass : which == 5 ? &s.prevWorldPos : which == 6 ? &s.prevPixels : nullptr;
	if (!b) return fail(LH2HIP_ERR_INVALID, "svgf read: which 0..6");
	if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(hostOut, b->p, b->bytes, hipMemcpyDeviceToHost) != hipSuccess)
		return fail(LH2HIP_ERR_DEVICE, "svgf read: copy failed");
	return LH2HIP_OK;
}

} /* extern "C" */

/* Headless readback for a core made by CreateCore (the plugin path has no C
 * handle): the finalized image and the fp64 accumulator, as
 * lh2hip_read_image / lh2hip_read_accumulator_f64 read them. */
template <typename F>
static int guarded_plugin(void* core, F f)
{
	lh::RenderCore* c = core ? dynamic_cast<lh::RenderCore*>((lighthouse2::CoreAPI_Base*)core) : nullptr;
	if (!c) return fail(LH2HIP_ERR_INVALID, "not a core made by CreateCore");
	try
	{
		f(*c);
		return LH2HIP_OK;
	}
	catch (const CoreError& e) { return fail(e.code, e.what()); }
	catch (const std::exception& e) { return fail(LH2HIP_ERR_DEVICE, e.what()); }
}

extern "C" {
LH2HIP_API int lh2hip_plugin_read_image(void* core, lh2_float4* hostOut)
{
	return guarded_plugin(core, [&](RenderCore& c) { c.ReadImage(hostOut); });
}
LH2HIP_API int lh2hip_plugin_read_accumulator_f64(void* core, double* hostOut)
{
	return guarded_plugin(core, [&](RenderCore& c) { c.ReadAccumulatorF64(hostOut); });
}
}

/* Plugin entry point: RenderSystem dlopens the core and calls Init itself
 * (lib/RenderSystem/core_api_base.cpp:119-130). */
extern "C" __attribute__((visibility("default"))) lighthouse2::CoreAPI_Base* CreateCore()
{
	return new RenderCore();
}
