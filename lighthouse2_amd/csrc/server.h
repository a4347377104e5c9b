/* server.h - render job server of RenderCore_HIP: a device-resident job queue
 * with continuous batching of mixed-resolution / mixed-spp jobs, admitted by
 * an HBM budget. No counterpart in the reference; the nearest is the
 * asynchronous Render + RenderThread of lib/rendercore_optix7/rendercore.cpp:803-847. */
#pragma once
#include "rendercore.h"
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace lh {

struct JobTimes { double queued, admitted, done; };   /* seconds since the server started */

class RenderServer
{
public:
	/* budgetBytes: device memory the per-job buffers of admitted jobs may hold
	 * (accumulators + image, 64 B per pixel); passPaths: paths per scheduling
	 * quantum (a job runs at most this many paths before another job may run). */
	RenderServer(RenderCore& core, uint64_t budgetBytes, size_t passPaths);
	~RenderServer();
	/* denoise 0: the core's estimator, image = accumulator / spp; 1: the
	 * filter estimator (G-buffer per job) and SVGF on the job's samples, as
	 * RenderCore::Render draws a first frame with the filter on and TAA off. */
	uint64_t Submit(const lighthouse2::ViewPyramid& view, int w, int h, int spp, int maxDepth, uint32_t seed, int denoise = 0);
	bool Done(uint64_t id);
	/* Blocks until the job is done (timeout < 0: no limit); false on timeout. */
	bool Wait(uint64_t id, double timeoutSeconds);
	/* Copies the final image (w*h float4) and, optionally, the fp64 accumulator
	 * (w*h*4 doubles), then forgets the job. */
	void Read(uint64_t id, void* image, double* acc64);
	JobTimes Times(uint64_t id);
	uint64_t BytesInUse();

private:
	struct Job
	{
		uint64_t id;
		lighthouse2::ViewPyramid view;
		int w, h, spp, depth, denoise = 0;
		uint32_t seed;
		int next = 0;                   /* next sample to render */
		bool admitted = false, done = false;
		std::string error;
		DevBuf acc32, acc64, image, gFeatures, gWorldPos, gDeltaDepth;
		std::vector<float> hostImage;
		std::vector<double> hostAcc;
		JobTimes t = { 0, 0, 0 };
		FrameSetup frame;
		/* accumulators + image: 64 B/px; with SVGF, direct and indirect halves
		 * (another 48 B/px) and the G-buffer (48 B/px) */
		uint64_t bytes() const { return (uint64_t)w * h * (denoise ? 160 : 64); }
		uint64_t remaining() const { return (uint64_t)(spp - next) * (uint64_t)w * h; }
	};
	void loop();
	void admit();
	double now() const;
	RenderCore& core;
	uint64_t budget;
	size_t passPaths;
	std::mutex lock;
	std::condition_variable wake, finished;
	std::deque<std::shared_ptr<Job>> waiting;            /* not yet admitted, arrival order */
	std::map<uint64_t, std::shared_ptr<Job>> jobs;       /* every job not yet read */
	uint64_t nextId = 1, inUse = 0;
	bool stopping = false;
	hipStream_t stream = nullptr, stream2 = nullptr;   /* stream2: the shadow lane */
	hipEvent_t laneEv[2] = { nullptr, nullptr };
	int device = 0;
	Wavefront wf;
	std::map<std::pair<int, int>, std::unique_ptr<Svgf>> filters;   /* per job size, reused */
	DevBuf queue, ovf, laneScratch, laneOvf;
	std::chrono::steady_clock::time_point t0;
	std::thread worker;
};

} // namespace lh
