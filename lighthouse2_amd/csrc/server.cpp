/* server.cpp - render job server (see server.h).
 *
 * One worker thread owns a HIP stream (plus a second one for the shadow
 * lane, as in RenderCore::Render), its own path buffers and traversal
 * scratch. Jobs wait in arrival order until their per-job buffers fit the
 * HBM budget; among admitted jobs the one with the least remaining work
 * (samples x pixels) runs next, one quantum of at most passPaths paths at a
 * time, so small jobs overtake large ones (shortest remaining processing time
 * keeps p50 and p99 latency low under mixed load). A job renders exactly what
 * RenderCore::Render(view, Restart) with its size, spp, depth and seed
 * renders: the quanta split its samples, and a path's random numbers depend on
 * its index in the frame only. The scene must not change while jobs run. */
#include "server.h"
#include "svgf.h"
#include "../../include/lh2hip.h"
#include <algorithm>
#include <cstring>

namespace lh {

static inline void check(hipError_t e, const char* what)
{
	if (e == hipErrorOutOfMemory) throw CoreError(LH2HIP_ERR_OOM, std::string(what) + ": out of device memory");
	if (e != hipSuccess) throw CoreError(LH2HIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

static uint32_t xor32_seed(uint32_t s) { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; } /* RandomUInt, system.cpp:51 */

RenderServer::RenderServer(RenderCore& c, uint64_t budgetBytes, size_t pp) : core(c), budget(budgetBytes), passPaths(std::max<size_t>(pp, 1))
{
	if (!core.sceneIsReady()) throw CoreError(LH2HIP_ERR_STATE, "server: load and finalize a scene first");
	core.checkRenderable();
	check(hipGetDevice(&device), "hipGetDevice");
	check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
	queue.ensure(TRACE_QUEUE_WORDS * sizeof(uint32_t));
	ovf.ensure(trace_overflow_bytes(core.traceBlocks()));
	check(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking), "hipStreamCreate");
	for (int i = 0; i < 2; i++) check(hipEventCreateWithFlags(&laneEv[i], hipEventDisableTiming), "hipEventCreate");
	laneScratch.ensure((4 + TRACE_QUEUE_WORDS) * sizeof(uint32_t));
	laneOvf.ensure(trace_overflow_bytes(core.traceBlocks()));
	t0 = std::chrono::steady_clock::now();
	worker = std::thread([this] { loop(); });
}

RenderServer::~RenderServer()
{
	{
		std::lock_guard<std::mutex> g(lock);
		stopping = true;
	}
	wake.notify_all();
	if (worker.joinable()) worker.join();
	jobs.clear(), waiting.clear();
	wf.release(), queue.release(), ovf.release(), laneScratch.release(), laneOvf.release();
	for (int i = 0; i < 2; i++) if (laneEv[i]) (void)hipEventDestroy(laneEv[i]);
	if (stream2) (void)hipStreamDestroy(stream2);
	if (stream) (void)hipStreamDestroy(stream);
}

double RenderServer::now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

uint64_t RenderServer::Submit(const lighthouse2::ViewPyramid& view, int w, int h, int spp, int maxDepth, uint32_t seed, int denoise)
{
	if (w <= 0 || h <= 0 || spp <= 0 || maxDepth < 1 || maxDepth > 1024 || (int64_t)w * h > (1 << 24))
		throw CoreError(LH2HIP_ERR_INVALID, "server submit: bad job size (w*h <= 2^24, spp >= 1, depth 1..1024)");
	if (denoise != 0 && denoise != 1) throw CoreError(LH2HIP_ERR_INVALID, "server submit: denoise 0 (none) or 1 (SVGF)");
	if (denoise && (w < 3 || h < 3)) throw CoreError(LH2HIP_ERR_INVALID, "server submit: SVGF needs at least 3 x 3 pixels");
	auto j = std::make_shared<Job>();
	j->view = view, j->w = w, j->h = h, j->spp = spp, j->depth = maxDepth, j->seed = seed, j->denoise = denoise;
	if (j->bytes() > budget) throw CoreError(LH2HIP_ERR_INVALID, "server submit: job needs more device memory than the budget");
	/* the frame constants only read scene state, which must stay fixed */
	j->frame = core.frameParams(view, w, h, spp, maxDepth, xor32_seed(seed), 0, 0.0f, 0.0f, denoise ? ESTIMATOR_FILTER : -1);
	/* the core's G-buffer belongs to Render's target; an SVGF job gets its own
	 * when admitted (rendercore.cpp:780-797: aperture 0 while filtering) */
	j->frame.P.gFeatures = nullptr, j->frame.P.gWorldPos = nullptr, j->frame.P.gDeltaDepth = nullptr;
	if (denoise) j->frame.cam.aperture = 0, j->frame.camSeed = seed;
	std::lock_guard<std::mutex> g(lock);
	j->id = nextId++;
	j->t.queued = now();
	waiting.push_back(j);
	jobs[j->id] = j;
	wake.notify_all();
	return j->id;
}

void RenderServer::admit()
{	/* called with the lock held */
	while (!waiting.empty() && inUse + waiting.front()->bytes() <= budget)
	{
		auto j = waiting.front();
		waiting.pop_front();
		inUse += j->bytes();
		j->admitted = true;
		j->t.admitted = now();
	}
}

void RenderServer::loop()
{
	check(hipSetDevice(device), "hipSetDevice");
	while (true)
	{
		std::shared_ptr<Job> j;
		{
			std::unique_lock<std::mutex> g(lock);
			wake.wait(g, [&] {
				if (stopping) return true;
				admit();
				for (auto& kv : jobs) if (kv.second->admitted && !kv.second->done) return true;
				return false;
			});
			if (stopping) return;
			for (auto& kv : jobs)
			{
				Job* c = kv.second.get();
				if (!c->admitted || c->done) continue;
				if (!j || c->remaining() < j->remaining()) j = kv.second;
			}
		}
		try
		{
			const int pixels = j->w * j->h;
			const int halves = j->denoise ? 2 : 1;
			if (j->next == 0)
			{	/* first quantum: per-job buffers, cleared */
				j->acc32.ensure((size_t)halves * pixels * sizeof(float4)), j->acc64.ensure((size_t)halves * pixels * 4 * sizeof(double));
				j->image.ensure((size_t)pixels * sizeof(float4));
				check(hipMemsetAsync(j->acc32.p, 0, j->acc32.bytes, stream), "memset");
				check(hipMemsetAsync(j->acc64.p, 0, j->acc64.bytes, stream), "memset");
				if (j->denoise)
				{	/* SVGF writes the interior pixels only: the border stays 0, as in a fresh render target */
					check(hipMemsetAsync(j->image.p, 0, j->image.bytes, stream), "memset");
					j->gFeatures.ensure((size_t)pixels * 16), j->gWorldPos.ensure((size_t)pixels * 16), j->gDeltaDepth.ensure((size_t)pixels * 16);
					check(hipMemsetAsync(j->gFeatures.p, 0, j->gFeatures.bytes, stream), "memset");
					j->frame.P.gFeatures = j->gFeatures.as<uint4>(), j->frame.P.gWorldPos = j->gWorldPos.as<float4>();
					j->frame.P.gDeltaDepth = j->gDeltaDepth.as<float4>();
				}
			}
			const size_t need = trace_overflow_bytes(core.traceBlocks());
			if (ovf.bytes < need || laneOvf.bytes < need)
			{	/* the core's persistent block count grew (traceBlocksPerCU): grow the overflow stacks */
				check(hipStreamSynchronize(stream), "sync");
				check(hipStreamSynchronize(stream2), "sync");
				ovf.ensure(need), laneOvf.ensure(need);
			}
			const int quantum = std::max(1, (int)std::min<size_t>(passPaths / (size_t)pixels, (size_t)(j->spp - j->next)));
			const int s1 = std::min(j->spp, j->next + quantum);
			RenderCore::SideLane lane;
			lane.st = stream2, lane.counters = laneScratch.as<uint32_t>(), lane.queue = laneScratch.as<uint32_t>() + 4;
			lane.ovf = laneOvf.as<uint32_t>(), lane.shaded = laneEv[0], lane.done = laneEv[1];
			core.runSamples(j->frame, j->next, s1, passPaths, wf, j->acc32.as<float4>(), j->acc64.as<double>(),
				queue.as<uint32_t>(), ovf.as<uint32_t>(), stream, nullptr, nullptr, core.shadowLaneOn() ? &lane : nullptr);
			j->next = s1;
			if (j->next == j->spp)
			{
				if (j->denoise)
				{	/* FinalizeRender's filter path for a first frame (RenderCore_Optix7Filter/rendercore.cpp:897-935) */
					if (!filters.count({ j->w, j->h }) && filters.size() >= 8)
					{	/* at most 8 job sizes keep a filter */
						check(hipStreamSynchronize(stream), "sync");
						filters.clear();
					}
					auto& f = filters[{ j->w, j->h }];
					if (!f) f.reset(new Svgf(j->w, j->h));
					else f->Reset(stream);
					f->taa = 0;
					const lighthouse2::ViewPyramid none = {};
					f->Frame(j->acc32.as<float4>(), j->gFeatures.as<uint4>(), j->gWorldPos.as<float4>(), j->gDeltaDepth.as<float4>(), none,
						0.0f, 0.0f, 0.0f, 0.0f, j->spp, core.clampDirectSetting(), core.clampIndirectSetting(), 0, j->image.as<float4>(), stream);
					j->gFeatures.release(), j->gWorldPos.release(), j->gDeltaDepth.release();
				}
				else check(launch_finalize(j->acc32.as<float4>(), pixels, 1.0f / (float)j->spp, j->image.as<float4>(), stream), "finalize");
				j->hostImage.resize((size_t)pixels * 4), j->hostAcc.resize((size_t)pixels * 4);
				check(hipMemcpyAsync(j->hostImage.data(), j->image.p, j->image.bytes, hipMemcpyDeviceToHost, stream), "D2H");
				check(hipMemcpyAsync(j->hostAcc.data(), j->acc64.p, (size_t)pixels * 4 * sizeof(double), hipMemcpyDeviceToHost, stream), "D2H");
				check(hipStreamSynchronize(stream), "sync");
				j->acc32.release(), j->acc64.release(), j->image.release();
				std::lock_guard<std::mutex> g(lock);
				j->done = true;
				j->t.done = now();
				inUse -= j->bytes();
			}
		}
		catch (const std::exception& e)
		{
			std::lock_guard<std::mutex> g(lock);
			j->error = e.what(), j->done = true, j->t.done = now();
			j->acc32.release(), j->acc64.release(), j->image.release();
			inUse -= j->bytes();
		}
		if (j->done) finished.notify_all();
	}
}

bool RenderServer::Done(uint64_t id)
{
	std::lock_guard<std::mutex> g(lock);
	auto it = jobs.find(id);
	if (it == jobs.end()) throw CoreError(LH2HIP_ERR_INVALID, "server: unknown job");
	return it->second->done;
}

bool RenderServer::Wait(uint64_t id, double timeoutSeconds)
{
	std::unique_lock<std::mutex> g(lock);
	auto it = jobs.find(id);
	if (it == jobs.end()) throw CoreError(LH2HIP_ERR_INVALID, "server: unknown job");
	auto j = it->second;
	if (timeoutSeconds < 0) { finished.wait(g, [&] { return j->done; }); return true; }
	return finished.wait_for(g, std::chrono::duration<double>(timeoutSeconds), [&] { return j->done; });
}

void RenderServer::Read(uint64_t id, void* image, double* acc64)
{
	std::shared_ptr<Job> j;
	{
		std::lock_guard<std::mutex> g(lock);
		auto it = jobs.find(id);
		if (it == jobs.end()) throw CoreError(LH2HIP_ERR_INVALID, "server: unknown job");
		if (!it->second->done) throw CoreError(LH2HIP_ERR_STATE, "server: job not done");
		j = it->second;
		jobs.erase(it);
	}
	if (!j->error.empty()) throw CoreError(LH2HIP_ERR_DEVICE, "server job failed: " + j->error);
	if (image) memcpy(image, j->hostImage.data(), j->hostImage.size() * sizeof(float));
	if (acc64) memcpy(acc64, j->hostAcc.data(), j->hostAcc.size() * sizeof(double));
}

JobTimes RenderServer::Times(uint64_t id)
{
	std::lock_guard<std::mutex> g(lock);
	auto it = jobs.find(id);
	if (it == jobs.end()) throw CoreError(LH2HIP_ERR_INVALID, "server: unknown job");
	return it->second->t;
}

uint64_t RenderServer::BytesInUse()
{
	std::lock_guard<std::mutex> g(lock);
	return inUse;
}

} // namespace lh
