/* trace.h - traversal kernels of RenderCore_HIP (host-side launch API). */
#pragma once
#include "bvh.h"

namespace lh {

enum NodeFormat { FMT_F32 = 0, FMT_Q8 = 1 };
enum TraceMode { MODE_CLOSEST = 0, MODE_ANY_MASK = 1, MODE_ANY_FLAG = 2 };
enum TraceLoop { LOOP_MERGED = 0, LOOP_WHILE_WHILE = 1, LOOP_GREEDY = 2 };

struct TraceScene
{
	int format = FMT_F32;
	int twoLevel = 0;
	/* single level: one mesh, instance 0 with identity transform */
	const void* nodes = nullptr;
	const float4* tris = nullptr;
	uint32_t root = 0;
	/* two level */
	const NodeF32* tlas = nullptr;
	uint32_t tlasRoot = 0;
	const InstanceRec* inst = nullptr;
	int maxStack = 0;
	int loop = LOOP_MERGED;
	int refillLanes = 16; /* a wave fetches new rays once this many lanes are idle */
	int chunk = 32;       /* rays a wave reserves from a queue counter at once */
	int triBias = 16;   /* greedy loop: a leaf step when (triBias / 16) x pending-leaf lanes > node lanes */
};

/* Work counters of the counting kernel variant: node visits and triangle
 * tests per ray, and per wave the loop iterations and the lanes active in
 * them (SIMD utilisation = lanes / (64 * iterations)). */
struct TraceStats
{
	unsigned long long nodes, tris, nodeIters, nodeLanes, leafIters, leafLanes, loopIters, loopLanes, tailIters, tailLanes;
	unsigned long long tEnd, tStart, tDrained;   /* wall clock (100 MHz): last wave end, first wave start, first wave to find the feed empty */
};

/* rays: Ray4 {O.xyz, tmin; D.xyz, tmax}. out: MODE_CLOSEST -> float4 hit records;
 * MODE_ANY_MASK -> uint32 bitmask (must be zeroed); MODE_ANY_FLAG -> uint32 per ray.
 * overflow: device overflow stack (see trace_overflow_bytes), queue:
 * TRACE_QUEUE_WORDS uint32 ray-feed counters (cleared by the call). */
hipError_t launch_trace(const TraceScene& s, int mode, const float4* rays, void* out, int count,
	uint32_t* queue, uint32_t* overflow, int persistentBlocks, hipStream_t stream, TraceStats* statsDev = nullptr);
/* As launch_trace, with the ray count read by the kernel from countDev
 * (device memory) and maxCount an upper bound: no host round trip. */
hipError_t launch_trace_dev(const TraceScene& s, int mode, const float4* rays, void* out, int maxCount, const uint32_t* countDev,
	uint32_t* queue, uint32_t* overflow, int persistentBlocks, hipStream_t stream, TraceStats* statsDev = nullptr);
size_t trace_overflow_bytes(int persistentBlocks);

static const int TRACE_BLOCK = 256;
static const int TRACE_LDS_STACK = 16;
static const int TRACE_OVF_STACK = 112;
static const int TRACE_MAX_STACK = TRACE_LDS_STACK + TRACE_OVF_STACK;
static const uint32_t TRACE_MAX_STEPS = 1u << 26;   /* node visits per ray (safety bound) */
/* Ray feed: the rays of a launch are split into TRACE_QUEUES contiguous
 * ranges, each handed out by its own counter (TRACE_QSTRIDE words apart, so
 * the counters sit on separate cache lines). A launch needs a zeroed area of
 * TRACE_QUEUE_WORDS words; launch_trace clears it on the stream. */
static const int TRACE_QUEUES = 8;
static const int TRACE_QSTRIDE = 64;
static const int TRACE_QUEUE_WORDS = TRACE_QUEUES * TRACE_QSTRIDE;

} // namespace lh
