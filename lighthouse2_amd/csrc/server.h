/* server.h - render job server of RenderCore_HIP: a device-resident job queue
 * with continuous batching of mixed-resolution / mixed-spp jobs, admitted by
 * an HBM budget. No counterpart in the reference; the nearest is the
 * asynchronous Render + RenderThread of lib/rendercore_optix7/rendercore.cpp:803-847. */
#pragma once
#include "rendercore.h"
#include "unet.h"
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace lh {

struct JobTimes { double queued, admitted, done; };   /* seconds since the server started */

/* Counters of the batching: passes run, job parts over all passes, passes
 * that carried more than one job, the most jobs one pass carried, paths
 * (eye rays) over all passes, extension rays traced (every wave), and the
 * device time of the extension traversal and shade launches (ms). */
struct ServerStats { uint64_t passes, parts, sharedPasses, maxParts, paths, rays; double traceMs, shadeMs; };

enum { DENOISE_NONE = 0, DENOISE_SVGF = 1, DENOISE_LEARNED = 2 };

class RenderServer
{
public:
	/* budgetBytes: device memory the per-job buffers of admitted jobs may hold
	 * (accumulators + image: 64 B per pixel; denoised jobs 160 B per pixel);
	 * passPaths: paths per pass (the admitted jobs with the least remaining
	 * work share a pass up to this many paths; a job whose single sample is
	 * larger runs a pass of its own). */
	RenderServer(RenderCore& core, uint64_t budgetBytes, size_t passPaths);
	~RenderServer();
	/* denoise 0: the core's estimator, image = accumulator / spp; 1: the
	 * filter estimator (G-buffer per job) and SVGF on the job's samples, as
	 * RenderCore::Render draws a first frame with the filter on and TAA off;
	 * 2: the same frame denoised by the learned U-Net (SetDenoiser). */
	uint64_t Submit(const lighthouse2::ViewPyramid& view, int w, int h, int spp, int maxDepth, uint32_t seed, int denoise = 0);
	/* Paused: the worker finishes its pass and starts no other until resumed
	 * (jobs submitted meanwhile then share passes from the start). */
	void Pause(bool paused);
	/* Weights of the learned denoiser (UNetForward::BlobBytes() bytes). */
	void SetDenoiser(const void* blob, size_t bytes);
	bool Done(uint64_t id);
	/* Blocks until the job is done (timeout < 0: no limit); false on timeout. */
	bool Wait(uint64_t id, double timeoutSeconds);
	/* Copies the final image (w*h float4) and, optionally, the fp64 accumulator
	 * (w*h*4 doubles; denoised jobs: the direct half), then forgets the job. */
	void Read(uint64_t id, void* image, double* acc64);
	JobTimes Times(uint64_t id);
	uint64_t BytesInUse();
	ServerStats Stats();

private:
	struct Job
	{
		uint64_t id;
		lighthouse2::ViewPyramid view;
		int w, h, spp, depth, denoise = 0;
		uint32_t seed;
		int next = 0;                   /* next sample to render */
		int bypassed = 0;               /* later jobs admitted ahead of this one */
		bool admitted = false, started = false, done = false;
		bool rendered = false;          /* every sample resolved and the finish enqueued (or failed) */
		hipEvent_t ready = nullptr;     /* recorded after the finish: the results may be copied */
		std::string error;
		DevBuf acc32, acc64, image, gFeatures, gWorldPos, gDeltaDepth;
		std::vector<float> hostImage;
		std::vector<double> hostAcc;
		bool accReady = false;          /* hostAcc filled (after done), the job's buffers and budget released */
		std::string accError;
		JobTimes t = { 0, 0, 0 };
		FrameSetup frame;
		/* accumulators + image: 64 B/px; denoised: direct and indirect halves
		 * (another 48 B/px) and the G-buffer (48 B/px) */
		uint64_t bytes() const { return (uint64_t)w * h * (denoise ? 160 : 64); }
		uint64_t remaining() const { return (uint64_t)(spp - next) * (uint64_t)w * h; }
		int estimator() const { return frame.estimator; }
	};
	void loop();
	void admit();
	void startJob(Job& j);
	void finishJob(Job& j);
	void releaseJob(Job& j);
	/* The drain thread: copies a finished job's image to the host on its own
	 * stream through two pinned staging buffers and marks the job done, then
	 * copies its fp64 accumulator, returns its buffers to the pool and
	 * releases its budget, while the worker goes on with the next pass. */
	void drainLoop();
	void drainCopy(void* dst, const void* src, size_t bytes);
	/* Per-job device buffers are kept for reuse, keyed by size, up to the
	 * budget: hipFree synchronises the whole device. */
	void takeBuf(DevBuf& b, size_t bytes);
	void giveBuf(DevBuf& b);
	double now() const;
	RenderCore& core;
	uint64_t budget;
	size_t passPaths;
	std::mutex lock;
	std::condition_variable wake, finished;
	std::deque<std::shared_ptr<Job>> waiting;            /* not yet admitted, arrival order */
	std::map<uint64_t, std::shared_ptr<Job>> jobs;       /* every job not yet read */
	uint64_t nextId = 1, inUse = 0;
	bool stopping = false, paused = false;
	ServerStats stats = { 0, 0, 0, 0, 0, 0, 0.0, 0.0 };
	hipEvent_t passEv[4] = { nullptr, nullptr, nullptr, nullptr };   /* per-wave timing of the pass launches */
	hipStream_t stream = nullptr, stream2 = nullptr;   /* stream2: the shadow lane */
	hipEvent_t laneEv[2] = { nullptr, nullptr };
	int device = 0;
	Wavefront wf;
	std::map<std::pair<int, int>, std::unique_ptr<Svgf>> filters;   /* per job size, reused */
	DevBuf queue, ovf, laneScratch, laneOvf, segs;
	JobSeg* segsHost = nullptr;                          /* pinned staging of the pass's segment table */
	UNetForward unet;                                   /* learned denoiser (worker thread only after SetDenoiser) */
	std::mutex unetLock;
	std::chrono::steady_clock::time_point t0;
	std::thread worker;
	std::mutex drainLock, poolLock;
	std::condition_variable drainWake;
	std::deque<std::shared_ptr<Job>> draining;          /* rendered, results not yet on the host */
	bool drainStop = false;
	std::multimap<size_t, DevBuf> pool;                  /* freed job buffers by size */
	uint64_t pooled = 0;
	static constexpr size_t STAGE_BYTES = (size_t)32 << 20;
	hipStream_t cstream = nullptr;                       /* the drain thread's copies */
	void* stage[2] = { nullptr, nullptr };               /* pinned staging of the read-back */
	hipEvent_t stageEv[2] = { nullptr, nullptr };
	std::thread drainer;
};

} // namespace lh
