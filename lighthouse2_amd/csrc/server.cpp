/* server.cpp - render job server (see server.h).
 *
 * One worker thread owns a HIP stream (plus a second one for the shadow
 * lane, as in RenderCore::Render), its own path buffers and traversal
 * scratch. Admission: waiting jobs are admitted in arrival order while their
 * per-job buffers fit the HBM budget; a job that does not fit lets later,
 * smaller jobs pass, until it has been passed MAX_BYPASS times, after which
 * it holds the line until memory frees up (no starvation of large jobs).
 * Batching: every pass of the wavefront loop carries the paths of several
 * admitted jobs at once (up to JOB_SEGS_MAX jobs of one estimator and
 * passPaths paths): the job with the least remaining work (samples x pixels)
 * takes all it needs, the next ones fill the rest, whole samples each; the
 * jobs' paths share every generate / extend / shade / connect launch and
 * are resolved into each job's own accumulators. A finished job's results
 * reach the host through a drain thread (its own stream and pinned staging),
 * so the worker never waits on a read-back; freed job buffers are pooled by
 * size, so admission rarely calls hipMalloc and never hipFree. A job renders exactly what
 * RenderCore::Render(view, Restart) with its size, spp, depth and seed
 * renders: a path's random numbers depend on its index in its job's frame
 * only, and each path slot has one writer.
 *
 * The scene is shared with the core: calls that change it fail while jobs
 * are pending (RenderCore::requireNoJobs); submission and the worker's
 * passes hold the core's scene lock shared, the calls that change the scene
 * (or render on the core) hold it exclusively. */
#include "server.h"
#include "svgf.h"
#include "../../include/lh2hip.h"
#include <algorithm>
#include <cstring>

namespace lh {

static const int MAX_BYPASS = 8;

static inline void check(hipError_t e, const char* what)
{
	if (e == hipErrorOutOfMemory) throw CoreError(LH2HIP_ERR_OOM, std::string(what) + ": out of device memory");
	if (e != hipSuccess) throw CoreError(LH2HIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

static uint32_t xor32_seed(uint32_t s) { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; } /* RandomUInt, system.cpp:51 */

RenderServer::RenderServer(RenderCore& c, uint64_t budgetBytes, size_t pp) : core(c), budget(budgetBytes), passPaths(std::max<size_t>(pp, 1))
{
	std::shared_lock<std::shared_mutex> g(core.sceneMutex);
	if (!core.sceneIsReady()) throw CoreError(LH2HIP_ERR_STATE, "server: load and finalize a scene first");
	core.checkRenderable();
	check(hipGetDevice(&device), "hipGetDevice");
	check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
	queue.ensure(TRACE_QUEUE_WORDS * sizeof(uint32_t));
	ovf.ensure(trace_overflow_bytes(core.traceBlocks()));
	check(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking), "hipStreamCreate");
	for (int i = 0; i < 2; i++) check(hipEventCreateWithFlags(&laneEv[i], hipEventDisableTiming), "hipEventCreate");
	for (int i = 0; i < 4; i++) check(hipEventCreate(&passEv[i]), "hipEventCreate");
	laneScratch.ensure((4 + TRACE_QUEUE_WORDS) * sizeof(uint32_t));
	laneOvf.ensure(trace_overflow_bytes(core.traceBlocks()));
	segs.ensure(sizeof(JobSeg) * JOB_SEGS_MAX);
	check(hipHostMalloc((void**)&segsHost, sizeof(JobSeg) * JOB_SEGS_MAX, hipHostMallocDefault), "hipHostMalloc");
	check(hipStreamCreateWithFlags(&cstream, hipStreamNonBlocking), "hipStreamCreate");
	for (int i = 0; i < 2; i++)
	{
		check(hipHostMalloc(&stage[i], STAGE_BYTES, hipHostMallocDefault), "hipHostMalloc");
		check(hipEventCreateWithFlags(&stageEv[i], hipEventBlockingSync | hipEventDisableTiming), "hipEventCreate");
	}
	t0 = std::chrono::steady_clock::now();
	core.serverCount++;
	drainer = std::thread([this] { drainLoop(); });
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
	{	/* the drain thread empties its queue, then ends */
		std::lock_guard<std::mutex> g(drainLock);
		drainStop = true;
	}
	drainWake.notify_all();
	if (drainer.joinable()) drainer.join();
	for (auto& kv : jobs) if (!kv.second->done) core.pendingJobs--;
	jobs.clear(), waiting.clear();
	filters.clear();
	pool.clear();
	for (int i = 0; i < 2; i++)
	{
		if (stage[i]) (void)hipHostFree(stage[i]);
		if (stageEv[i]) (void)hipEventDestroy(stageEv[i]);
	}
	if (cstream) (void)hipStreamDestroy(cstream);
	wf.release(), queue.release(), ovf.release(), laneScratch.release(), laneOvf.release(), segs.release();
	if (segsHost) (void)hipHostFree(segsHost);
	for (int i = 0; i < 2; i++) if (laneEv[i]) (void)hipEventDestroy(laneEv[i]);
	for (int i = 0; i < 4; i++) if (passEv[i]) (void)hipEventDestroy(passEv[i]);
	if (stream2) (void)hipStreamDestroy(stream2);
	if (stream) (void)hipStreamDestroy(stream);
	core.serverCount--;
}

double RenderServer::now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

uint64_t RenderServer::Submit(const lighthouse2::ViewPyramid& view, int w, int h, int spp, int maxDepth, uint32_t seed, int denoise)
{
	if (w <= 0 || h <= 0 || spp <= 0 || maxDepth < 1 || maxDepth > 1024 || (int64_t)w * h > (1 << 24))
		throw CoreError(LH2HIP_ERR_INVALID, "server submit: bad job size (w*h <= 2^24, spp >= 1, depth 1..1024)");
	if ((uint64_t)w * (uint64_t)h * (uint64_t)spp > ((uint64_t)1 << 32))
		throw CoreError(LH2HIP_ERR_INVALID, "server submit: w * h * spp must not exceed 2^32 (path indices are 32-bit)");
	if (denoise < DENOISE_NONE || denoise > DENOISE_LEARNED) throw CoreError(LH2HIP_ERR_INVALID, "server submit: denoise 0 (none), 1 (SVGF) or 2 (learned)");
	if (denoise && (w < 3 || h < 3)) throw CoreError(LH2HIP_ERR_INVALID, "server submit: denoising needs at least 3 x 3 pixels");
	if (denoise == DENOISE_LEARNED)
	{
		std::lock_guard<std::mutex> g(unetLock);
		if (!unet.Ready()) throw CoreError(LH2HIP_ERR_STATE, "server submit: learned denoising needs the denoiser's weights (SetDenoiser)");
	}
	auto j = std::make_shared<Job>();
	j->view = view, j->w = w, j->h = h, j->spp = spp, j->depth = maxDepth, j->seed = seed, j->denoise = denoise;
	if (j->bytes() > budget) throw CoreError(LH2HIP_ERR_INVALID, "server submit: job needs more device memory than the budget");
	{
		/* the frame constants point at scene buffers: taken under the scene
		 * lock, and the job counts as pending before the lock is released */
		std::shared_lock<std::shared_mutex> g(core.sceneMutex);
		j->frame = core.frameParams(view, w, h, spp, maxDepth, xor32_seed(seed), 0, 0.0f, 0.0f, denoise ? ESTIMATOR_FILTER : -1);
		core.pendingJobs++;
	}
	/* the core's G-buffer belongs to Render's target; a denoised job gets its
	 * own when it starts (rendercore.cpp:780-797: aperture 0 while filtering) */
	j->frame.P.gFeatures = nullptr, j->frame.P.gWorldPos = nullptr, j->frame.P.gDeltaDepth = nullptr;
	j->frame.P.probe = nullptr;
	j->frame.camSeed = seed;   /* Render's camera seed after Setting("seed", seed) and a restart */
	if (denoise) j->frame.cam.aperture = 0;
	std::lock_guard<std::mutex> g(lock);
	j->id = nextId++;
	j->t.queued = now();
	waiting.push_back(j);
	jobs[j->id] = j;
	wake.notify_all();
	return j->id;
}

void RenderServer::Pause(bool p)
{
	{
		std::lock_guard<std::mutex> g(lock);
		paused = p;
	}
	wake.notify_all();
}

void RenderServer::SetDenoiser(const void* blob, size_t bytes)
{
	std::lock_guard<std::mutex> g(unetLock);
	unet.SetWeights(blob, bytes, stream);
}

void RenderServer::admit()
{	/* called with the lock held */
	for (auto it = waiting.begin(); it != waiting.end();)
	{
		Job& j = **it;
		if (inUse + j.bytes() <= budget)
		{
			inUse += j.bytes();
			j.admitted = true;
			j.t.admitted = now();
			for (auto e = waiting.begin(); e != it; ++e) (*e)->bypassed++;
			it = waiting.erase(it);
			continue;
		}
		if (j.bypassed >= MAX_BYPASS) break;   /* it has waited long enough: no more overtaking */
		++it;
	}
}

void RenderServer::startJob(Job& j)
{	/* per-job buffers, cleared (worker thread, on the server's stream) */
	const int pixels = j.w * j.h;
	const int halves = j.denoise ? 2 : 1;
	takeBuf(j.acc32, (size_t)halves * pixels * sizeof(float4)), takeBuf(j.acc64, (size_t)halves * pixels * 4 * sizeof(double));
	takeBuf(j.image, (size_t)pixels * sizeof(float4));
	/* no clear of the accumulators: the job's first pass writes them (PassPart::fresh) */
	if (j.denoise)
	{	/* SVGF writes the interior pixels only: the border stays 0, as in a fresh render target */
		check(hipMemsetAsync(j.image.p, 0, j.image.bytes, stream), "memset");
		takeBuf(j.gFeatures, (size_t)pixels * 16), takeBuf(j.gWorldPos, (size_t)pixels * 16), takeBuf(j.gDeltaDepth, (size_t)pixels * 16);
		/* pixels whose path reaches no diffuse surface keep these zeros, as in a fresh render target */
		check(hipMemsetAsync(j.gFeatures.p, 0, j.gFeatures.bytes, stream), "memset");
		check(hipMemsetAsync(j.gWorldPos.p, 0, j.gWorldPos.bytes, stream), "memset");
		check(hipMemsetAsync(j.gDeltaDepth.p, 0, j.gDeltaDepth.bytes, stream), "memset");
		j.frame.P.gFeatures = j.gFeatures.as<uint4>(), j.frame.P.gWorldPos = j.gWorldPos.as<float4>();
		j.frame.P.gDeltaDepth = j.gDeltaDepth.as<float4>();
	}
	j.started = true;
}

void RenderServer::finishJob(Job& j)
{	/* all samples resolved: denoise or finalize; the drain thread copies the results once `ready` is reached */
	const int pixels = j.w * j.h;
	if (j.denoise == DENOISE_SVGF)
	{	/* FinalizeRender's filter path for a first frame (RenderCore_Optix7Filter/rendercore.cpp:897-935) */
		if (!filters.count({ j.w, j.h }) && filters.size() >= 8)
		{	/* at most 8 job sizes keep a filter */
			check(hipStreamSynchronize(stream), "sync");
			filters.clear();
		}
		auto& f = filters[{ j.w, j.h }];
		if (!f) f.reset(new Svgf(j.w, j.h, stream));
		else f->Reset(stream);
		f->taa = 0;
		const lighthouse2::ViewPyramid none = {};
		f->Frame(j.acc32.as<float4>(), j.gFeatures.as<uint4>(), j.gWorldPos.as<float4>(), j.gDeltaDepth.as<float4>(), none,
			0.0f, 0.0f, 0.0f, 0.0f, j.spp, core.clampDirectSetting(), core.clampIndirectSetting(), 0, j.image.as<float4>(), stream);
	}
	else if (j.denoise == DENOISE_LEARNED)
	{	/* the U-Net on the job's frame, padded to multiples of 16; the frame is read by the
		 * first convolution and the display image written by the last one */
		std::lock_guard<std::mutex> g(unetLock);
		unet.DenoiseFrame(j.acc32.as<float4>(), j.gFeatures.as<uint4>(), 1.0f / (float)j.spp, j.w, j.h, j.image.as<float4>(), stream);
	}
	else check(launch_finalize(j.acc32.as<float4>(), pixels, 1.0f / (float)j.spp, j.image.as<float4>(), stream), "finalize");
	check(hipEventCreateWithFlags(&j.ready, hipEventDisableTiming), "hipEventCreate");
	check(hipEventRecord(j.ready, stream), "hipEventRecord");
}

void RenderServer::releaseJob(Job& j)
{
	giveBuf(j.acc32), giveBuf(j.acc64), giveBuf(j.image);
	giveBuf(j.gFeatures), giveBuf(j.gWorldPos), giveBuf(j.gDeltaDepth);
}

void RenderServer::takeBuf(DevBuf& b, size_t bytes)
{
	{
		std::lock_guard<std::mutex> g(poolLock);
		auto it = pool.find(bytes);
		if (it != pool.end())
		{
			b = std::move(it->second);
			pool.erase(it);
			pooled -= bytes;
			return;
		}
	}
	if (hipMalloc(&b.p, bytes) == hipSuccess) { b.bytes = bytes; return; }
	(void)hipGetLastError();
	b.p = nullptr;
	{	/* out of memory: give the pooled buffers back and try once more */
		std::lock_guard<std::mutex> g(poolLock);
		pool.clear(), pooled = 0;
	}
	b.ensure(bytes);
}

void RenderServer::giveBuf(DevBuf& b)
{
	if (!b.p) return;
	std::lock_guard<std::mutex> g(poolLock);
	if (pooled + b.bytes > budget) { b.release(); return; }
	pooled += b.bytes;
	const size_t n = b.bytes;
	pool.emplace(n, std::move(b));
}

void RenderServer::drainCopy(void* dst, const void* src, size_t bytes)
{	/* the copy engine fills one staging buffer while this thread empties the other */
	size_t next = 0, off[2] = { 0, 0 }, len[2] = { 0, 0 };
	for (int k = 0; next < bytes || len[0] || len[1]; k ^= 1)
	{
		if (len[k])
		{
			check(hipEventSynchronize(stageEv[k]), "hipEventSynchronize");
			memcpy((char*)dst + off[k], stage[k], len[k]);
			len[k] = 0;
		}
		if (next < bytes)
		{
			const size_t n = std::min(STAGE_BYTES, bytes - next);
			check(hipMemcpyAsync(stage[k], (const char*)src + next, n, hipMemcpyDeviceToHost, cstream), "D2H");
			check(hipEventRecord(stageEv[k], cstream), "hipEventRecord");
			off[k] = next, len[k] = n, next += n;
		}
	}
}

void RenderServer::drainLoop()
{
	(void)hipSetDevice(device);
	while (true)
	{
		std::shared_ptr<Job> j;
		{
			std::unique_lock<std::mutex> g(drainLock);
			drainWake.wait(g, [&] { return drainStop || !draining.empty(); });
			if (draining.empty()) return;
			j = draining.front();
			draining.pop_front();
		}
		try
		{
			check(hipStreamWaitEvent(cstream, j->ready, 0), "hipStreamWaitEvent");
			const size_t pixels = (size_t)j->w * j->h;
			j->hostImage.resize(pixels * 4);
			drainCopy(j->hostImage.data(), j->image.p, pixels * sizeof(float4));
		}
		catch (const std::exception& e) { j->error = e.what(); }
		{	/* the image is on the host: the job is done and reads the scene no more */
			std::lock_guard<std::mutex> g(lock);
			j->done = true;
			j->t.done = now();
			core.pendingJobs--;
		}
		finished.notify_all();
		if (j->error.empty())
			try
			{	/* denoised jobs: the direct half, as RenderCore::ReadAccumulatorF64 */
				const size_t pixels = (size_t)j->w * j->h;
				j->hostAcc.resize(pixels * 4);
				drainCopy(j->hostAcc.data(), j->acc64.p, pixels * 4 * sizeof(double));
			}
			catch (const std::exception& e) { j->accError = e.what(); }
		(void)hipEventDestroy(j->ready);
		j->ready = nullptr;
		releaseJob(*j);
		{
			std::lock_guard<std::mutex> g(lock);
			j->accReady = true;
			inUse -= j->bytes();
		}
		finished.notify_all();
		wake.notify_all();   /* the freed budget may admit waiting jobs */
	}
}

void RenderServer::loop()
{
	check(hipSetDevice(device), "hipSetDevice");
	std::vector<std::shared_ptr<Job>> batch;
	std::vector<PassPart> parts;
	while (true)
	{
		batch.clear(), parts.clear();
		{
			std::unique_lock<std::mutex> g(lock);
			wake.wait(g, [&] {
				if (stopping) return true;
				if (paused) return false;
				admit();
				for (auto& kv : jobs) if (kv.second->admitted && !kv.second->rendered) return true;
				return false;
			});
			if (stopping) return;
			/* least remaining work first; the pass takes jobs of the first one's estimator */
			std::vector<std::shared_ptr<Job>> ready;
			for (auto& kv : jobs) if (kv.second->admitted && !kv.second->rendered) ready.push_back(kv.second);
			std::stable_sort(ready.begin(), ready.end(), [](const std::shared_ptr<Job>& a, const std::shared_ptr<Job>& b) {
				return a->remaining() < b->remaining();
			});
			const int est = ready[0]->estimator();
			size_t room = std::min(passPaths, Wavefront::PASS_CAP);
			for (auto& c : ready)
			{
				if ((int)batch.size() == JOB_SEGS_MAX) break;
				if (c->estimator() != est) continue;
				const size_t px = (size_t)c->w * c->h;
				int k = (int)std::min<size_t>((size_t)(c->spp - c->next), room / px);
				if (k == 0)
				{
					if (!batch.empty()) continue;
					k = 1;   /* one sample larger than a pass: it runs alone */
				}
				batch.push_back(c);
				room -= std::min(room, (size_t)k * px);
				parts.push_back(PassPart{ &c->frame, c->next, c->next + k, nullptr, nullptr, c->next == 0 });
			}
		}
		std::vector<std::shared_ptr<Job>> completed;   /* finish enqueued, or failed (error set) */
		try
		{
			std::shared_lock<std::shared_mutex> sg(core.sceneMutex);
			const size_t need = trace_overflow_bytes(core.traceBlocks());
			if (ovf.bytes < need || laneOvf.bytes < need)
			{	/* the core's persistent block count grew (traceBlocksPerCU): grow the overflow stacks */
				check(hipStreamSynchronize(stream), "sync");
				check(hipStreamSynchronize(stream2), "sync");
				ovf.ensure(need), laneOvf.ensure(need);
			}
			for (size_t i = 0; i < batch.size(); i++)
			{
				if (!batch[i]->started) startJob(*batch[i]);
				parts[i].f = &batch[i]->frame;
				parts[i].acc32 = batch[i]->acc32.as<float4>(), parts[i].acc64 = batch[i]->acc64.as<double>();
			}
			RenderCore::SideLane lane;
			lane.st = stream2, lane.counters = laneScratch.as<uint32_t>(), lane.queue = laneScratch.as<uint32_t>() + 4;
			lane.ovf = laneOvf.as<uint32_t>(), lane.shaded = laneEv[0], lane.done = laneEv[1];
			float timing[5] = { 0, 0, 0, 0, 0 };
			const int waves = core.runPass(parts.data(), (int)parts.size(), wf, queue.as<uint32_t>(), ovf.as<uint32_t>(), stream, passEv, timing,
				core.shadowLaneOn() ? &lane : nullptr, segs.as<JobSeg>(), segsHost);
			/* wave sizes: the advance after wave k logged the size of wave k + 1 */
			std::vector<uint32_t> logged((size_t)waves + 1, 0u);
			check(hipStreamSynchronize(stream), "sync");
			memcpy(logged.data(), (const void*)wf.countLog(), sizeof(uint32_t) * logged.size());   /* pinned, written by the advances */
			uint64_t paths = 0, rays = 0;
			for (int k = 1; k < waves; k++) rays += logged[k];
			for (size_t i = 0; i < batch.size(); i++)
			{
				Job& j = *batch[i];
				paths += (uint64_t)(parts[i].s1 - parts[i].s0) * (uint64_t)j.w * j.h;
				j.next = parts[i].s1;
				if (j.next == j.spp)
				{
					try { finishJob(j); }
					catch (const std::exception& e) { j.error = e.what(); }
					completed.push_back(batch[i]);
				}
			}
			std::lock_guard<std::mutex> g(lock);
			stats.passes++, stats.parts += batch.size(), stats.paths += paths, stats.rays += paths + rays;
			stats.traceMs += timing[0] + timing[1] + timing[2], stats.shadeMs += timing[3];
			if (batch.size() > 1) stats.sharedPasses++;
			stats.maxParts = std::max<uint64_t>(stats.maxParts, batch.size());
		}
		catch (const std::exception& e)
		{	/* the pass failed: every job in it fails */
			completed.clear();
			for (auto& c : batch) if (!c->rendered) c->error = e.what(), completed.push_back(c);
		}
		if (!completed.empty())
		{
			std::vector<std::shared_ptr<Job>> failed;
			{
				std::lock_guard<std::mutex> g(lock);
				for (auto& j : completed)
				{
					j->rendered = true;
					if (!j->error.empty()) failed.push_back(j);
				}
			}
			if (!failed.empty())
			{	/* kernels of a failed pass may still use its jobs' buffers: let the streams drain before they are pooled */
				(void)hipStreamSynchronize(stream), (void)hipStreamSynchronize(stream2);
				(void)hipGetLastError();
				for (auto& j : failed)
				{
					if (j->ready) (void)hipEventDestroy(j->ready), j->ready = nullptr;
					releaseJob(*j);
				}
				{
					std::lock_guard<std::mutex> g(lock);
					for (auto& j : failed)
					{
						j->done = j->accReady = true;
						j->t.done = now();
						inUse -= j->bytes();
						core.pendingJobs--;
					}
				}
				finished.notify_all();
				wake.notify_all();
			}
			{
				std::lock_guard<std::mutex> g(drainLock);
				for (auto& j : completed) if (j->error.empty()) draining.push_back(j);
			}
			drainWake.notify_all();
		}
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
	if (acc64)
	{	/* the accumulator reaches the host shortly after the image */
		std::unique_lock<std::mutex> g(lock);
		finished.wait(g, [&] { return j->accReady; });
		g.unlock();
		if (!j->accError.empty()) throw CoreError(LH2HIP_ERR_DEVICE, "server job failed: " + j->accError);
		memcpy(acc64, j->hostAcc.data(), j->hostAcc.size() * sizeof(double));
	}
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

ServerStats RenderServer::Stats()
{
	std::lock_guard<std::mutex> g(lock);
	return stats;
}

} // namespace lh
