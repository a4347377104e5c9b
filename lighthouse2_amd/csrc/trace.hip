/* trace.hip - closest-hit and any-hit traversal for gfx950.
 *
 * Replaces rtpQueryExecute (lib/RenderCore_PrimeRef/rendercore.cpp:559-573,
 * 598-607) and optixTrace + __closesthit__radiance / shadow any-hit
 * (lib/rendercore_optix7/optix/.optix.cu:125-154,174-184).
 *
 * Persistent waves: a wave hands rays to its idle lanes from a chunk it
 * reserved earlier with one atomicAdd on one of TRACE_QUEUES counters (ballot
 * -> popcount -> mbcnt gives each idle lane its ray), once `refillLanes` lanes
 * are idle. Three loop shapes share the node and triangle steps: merged (a
 * node visit and a leaf triangle per lane and iteration), while-while with
 * postponed leaves, and greedy (per iteration the step more lanes wait for).
 * Each lane walks the tree with a short stack of TRACE_LDS_STACK entries in
 * LDS (lane-interleaved, conflict free) that spills to a per-thread global
 * area beyond that. Box tests are conservative (fp32 slabs widened by 2^-16
 * relative; quantised slabs widened by a bound of their rounding), so
 * rounding never prunes a box the exact interval reaches; only the watertight
 * triangle test (lh_math.h) decides a hit, hence results equal a brute-force
 * scan bit for bit.
 */
#include "trace.h"
#include <algorithm>
#include <map>
#include <mutex>
#include "lh_math.h"

namespace lh {

struct Lane
{
	v3 wO, wD;          /* world ray */
	v3 O, D, inv;       /* current (object-space in a BLAS) ray */
	RayPre rp;
	float tmin, tbest;
	float bU, bV, bdet;
	int binst, bprim;
	int inst;           /* current instance, -1 at the top level */
	const void* nodes;
	const float4* tris;
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ bool slab(float lx, float hx, float ly, float hy, float lz, float hz, const v3& O, const v3& inv,
	float tmin, float tmax, float& tnear)
{
	const float tx1 = (lx - O.x) * inv.x, tx2 = (hx - O.x) * inv.x;
	const float ty1 = (ly - O.y) * inv.y, ty2 = (hy - O.y) * inv.y;
	const float tz1 = (lz - O.z) * inv.z, tz2 = (hz - O.z) * inv.z;
	const float tn = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fmaxf(fminf(tz1, tz2), tmin));
	float tf = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fminf(fmaxf(tz1, tz2), tmax));
	tf = tf + fabsf(tf) * 1.52587890625e-5f;
	tnear = tn;
	return tn <= tf;
}

/* An LDS pointer typed as such, so that loads through it stay LDS loads. */
typedef __attribute__((address_space(3))) uint32_t lds_u32;

template <int BLOCK>
struct Stack
{
	lds_u32* lds;       /* TRACE_LDS_STACK x BLOCK */
	uint32_t* ovf;      /* this thread's overflow area */
	int tid, sp;
	/* The host checks that the tree depth fits TRACE_MAX_STACK; a push beyond
	 * it is dropped (never written out of bounds). */
	__device__ __forceinline__ void push(uint32_t v)
	{
		if (sp < TRACE_LDS_STACK) lds[sp * BLOCK + tid] = v, sp++;
		else if (sp < TRACE_MAX_STACK) ovf[sp - TRACE_LDS_STACK] = v, sp++;
	}
	/* Pushes c (bottom) if hc, then b if hb, then a (top) if ha. With room
	 * for three entries in LDS, every lane writes all three slots and only
	 * moves its stack pointer for the pushed ones: no divergent branch. */
	__device__ __forceinline__ void push3(uint32_t c, bool hc, uint32_t b, bool hb, uint32_t a, bool ha)
	{
		if (sp + 3 <= TRACE_LDS_STACK)
		{
			lds[sp * BLOCK + tid] = c;
			sp += hc ? 1 : 0;
			lds[sp * BLOCK + tid] = b;
			sp += hb ? 1 : 0;
			lds[sp * BLOCK + tid] = a;
			sp += ha ? 1 : 0;
		}
		else
		{
			if (hc) push(c);
			if (hb) push(b);
			if (ha) push(a);
		}
	}
	/* LDS read for every lane (the row clamped into the LDS part), the
	 * overflow area read only by lanes below it. */
	__device__ __forceinline__ uint32_t pop()
	{
		sp--;
		uint32_t v = lds[min(sp, TRACE_LDS_STACK - 1) * BLOCK + tid];
		if (sp >= TRACE_LDS_STACK) v = ovf[sp - TRACE_LDS_STACK];
		return v;
	}
};

/* Sort up to 4 (tnear, ref) pairs ascending (misses carry +inf). */
__device__ __forceinline__ void cswap(float& a, float& b, uint32_t& ra, uint32_t& rb)
{
	const bool s = b < a;
	const float ta = s ? b : a, tb = s ? a : b;
	const uint32_t qa = s ? rb : ra, qb = s ? ra : rb;
	a = ta, b = tb, ra = qa, rb = qb;
}

static const uint32_t REF_DONE = 0xffffffffu;

/* Direction component kept away from zero for the box tests only (the
 * triangle test uses the exact direction): |D| >= 2^-66 keeps 1/D finite. */
__device__ __forceinline__ float safe_rcp(float d) { return rcp(fabsf(d) < 1.3552527e-20f ? copysignf(1.3552527e-20f, d) : d); }

__device__ __forceinline__ float ubyte(uint32_t w, int c) { return (float)((w >> (8 * c)) & 255u); }

__device__ __forceinline__ bool is_blas_leaf(uint32_t r) { return (r & (REF_LEAF | REF_TLAS)) == REF_LEAF; }
__device__ __forceinline__ bool is_inner(uint32_t r) { return !(r & (REF_LEAF | REF_TLAS)); }
__device__ __forceinline__ bool is_top(uint32_t r) { return (r & REF_TLAS) && r != REF_DONE; }

/* One 4-wide quantised node. Child planes are origin + q * 2^e per axis; per
 * node and axis b = (origin - O) / D and d = 2^e / D, so a plane lies at
 * t = q * d + b. Near and far planes are picked by the sign of D, and
 * e = 2^-20 (|b| + 255 |d|), an upper bound of the rounding of this form, is
 * subtracted from near and added to far planes: every box that the exact
 * interval [tmin, tbest] reaches is kept. Sets ref to the nearest hit child
 * (the others pushed far to near) or pops. */
template <int BLOCK>
__device__ __forceinline__ void visit_q8(const NodeQ8& n, const Lane& L, Stack<BLOCK>& st, uint32_t& ref)
{
	const float dx = bitsf(n.qb.z) * L.inv.x;
	const float dy = bitsf(n.qb.w) * L.inv.y;
	const float dz = n.origin.w * L.inv.z;
	const float bx = (n.origin.x - L.O.x) * L.inv.x;
	const float by = (n.origin.y - L.O.y) * L.inv.y;
	const float bz = (n.origin.z - L.O.z) * L.inv.z;
	const float ex = fmaf(fabsf(dx), 2.43186950683593750e-4f, fabsf(bx) * 9.5367431640625e-7f);   /* 255 * 2^-20, 2^-20 */
	const float ey = fmaf(fabsf(dy), 2.43186950683593750e-4f, fabsf(by) * 9.5367431640625e-7f);
	const float ez = fmaf(fabsf(dz), 2.43186950683593750e-4f, fabsf(bz) * 9.5367431640625e-7f);
	const float bnx = bx - ex, bfx = bx + ex, bny = by - ey, bfy = by + ey, bnz = bz - ez, bfz = bz + ez;
	const bool px = L.inv.x >= 0.0f, py = L.inv.y >= 0.0f, pz = L.inv.z >= 0.0f;
	const uint32_t qnx = px ? n.qa.x : n.qa.w, qfx = px ? n.qa.w : n.qa.x;
	const uint32_t qny = py ? n.qa.y : n.qb.x, qfy = py ? n.qb.x : n.qa.y;
	const uint32_t qnz = pz ? n.qa.z : n.qb.y, qfz = pz ? n.qb.y : n.qa.z;
	float tk[4];
	uint32_t rk[4] = { n.refs.x, n.refs.y, n.refs.z, n.refs.w };
#pragma unroll
	for (int c = 0; c < 4; c++)
	{
		const float tn = fmaxf(fmaxf(fmaf(ubyte(qnx, c), dx, bnx), fmaf(ubyte(qny, c), dy, bny)), fmaxf(fmaf(ubyte(qnz, c), dz, bnz), L.tmin));
		const float tf = fminf(fminf(fmaf(ubyte(qfx, c), dx, bfx), fmaf(ubyte(qfy, c), dy, bfy)), fminf(fmaf(ubyte(qfz, c), dz, bfz), L.tbest));
		tk[c] = (rk[c] != Q8_EMPTY && tn <= tf) ? tn : __builtin_huge_valf();
	}
	cswap(tk[0], tk[1], rk[0], rk[1]);
	cswap(tk[2], tk[3], rk[2], rk[3]);
	cswap(tk[0], tk[2], rk[0], rk[2]);
	cswap(tk[1], tk[3], rk[1], rk[3]);
	cswap(tk[1], tk[2], rk[1], rk[2]);
	if (tk[0] < __builtin_huge_valf())
	{
		st.push3(rk[3], tk[3] < __builtin_huge_valf(), rk[2], tk[2] < __builtin_huge_valf(), rk[1], tk[1] < __builtin_huge_valf());
		ref = rk[0];
	}
	else ref = st.sp ? st.pop() : REF_DONE;
}

/* One binary fp32 node (both child boxes stored in the parent). */
template <int BLOCK>
__device__ __forceinline__ void visit_f32(const NodeF32& n, const Lane& L, Stack<BLOCK>& st, uint32_t& ref, uint32_t flag)
{
	float tl, tr;
	const bool hl = slab(n.lxy.x, n.lxy.y, n.lxy.z, n.lxy.w, n.lrz.x, n.lrz.y, L.O, L.inv, L.tmin, L.tbest, tl);
	const bool hr = slab(n.rxy.x, n.rxy.y, n.rxy.z, n.rxy.w, n.lrz.z, n.lrz.w, L.O, L.inv, L.tmin, L.tbest, tr);
	const uint32_t rl = n.refs.x | flag, rr = n.refs.y | flag;
	if (hl && hr) { if (tr < tl) { st.push(rl); ref = rr; } else { st.push(rr); ref = rl; } }
	else if (hl) ref = rl;
	else if (hr) ref = rr;
	else ref = st.sp ? st.pop() : REF_DONE;
}

/* A top-level entry: restore the world ray if a BLAS was being walked, then
 * either enter an instance (object-space ray, BLAS root) or walk a TLAS node. */
template <int BLOCK>
__device__ __forceinline__ void visit_top(const TraceScene& sc, Lane& L, Stack<BLOCK>& st, uint32_t& ref)
{
	if (L.inst >= 0)
	{
		L.O = L.wO, L.D = L.wD;
		L.inv = mk3(safe_rcp(L.D.x), safe_rcp(L.D.y), safe_rcp(L.D.z));
		L.inst = -1;
	}
	const uint32_t r = ref & ~REF_TLAS;
	if (r & REF_LEAF)
	{
		const int id = (int)(r & REF_MASK);
		const InstanceRec& I = sc.inst[id];
		if (!(I.flags & 1u))
		{
			const float4 a = I.inv0, b = I.inv1, c = I.inv2;
			const v3 wO = L.wO, wD = L.wD;
			L.O = mk3(a.x * wO.x + a.y * wO.y + a.z * wO.z + a.w, b.x * wO.x + b.y * wO.y + b.z * wO.z + b.w, c.x * wO.x + c.y * wO.y + c.z * wO.z + c.w);
			L.D = mk3(a.x * wD.x + a.y * wD.y + a.z * wD.z, b.x * wD.x + b.y * wD.y + b.z * wD.z, c.x * wD.x + c.y * wD.y + c.z * wD.z);
			L.inv = mk3(safe_rcp(L.D.x), safe_rcp(L.D.y), safe_rcp(L.D.z));
		}
		L.rp = ray_setup(L.D);
		L.inst = id;
		L.nodes = I.nodes, L.tris = I.tris;
		ref = I.root;
	}
	else visit_f32(sc.tlas[r], L, st, ref, REF_TLAS);
}

/* Watertight test of one leaf triangle; true when an any-hit query is done. */
template <int MODE>
__device__ __forceinline__ bool test_tri(Lane& L, const float4& a, const float4& b, const float4& c, int instId)
{
	const int prim = (int)(fbits(a.w) & 0x7fffffffu);
	float t, U, V, det;
	if (tri_intersect(L.rp, L.O, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), t, U, V, det) && t > L.tmin)
	{
		if (MODE != MODE_CLOSEST)
		{
			if (t < L.tbest) { L.binst = instId, L.bprim = prim; return true; }
		}
		else if (better_hit(t, instId, prim, L.tbest, L.binst, L.bprim))
			L.tbest = t, L.binst = instId, L.bprim = prim, L.bU = U, L.bV = V, L.bdet = det;
	}
	return false;
}

template <int MODE>
__device__ __forceinline__ void write_result(const Lane& L, void* out, int rayIdx)
{
	if (MODE == MODE_CLOSEST)
	{
		float4 h;
		if (L.bprim < 0) h = make_float4(0.0f, 0.0f, bitsf(0xffffffffu), L.tbest);
		else h = make_float4(bitsf(pack_uv(L.bU, L.bV, L.bdet)), bitsf((uint32_t)L.binst), bitsf((uint32_t)L.bprim), L.tbest);
		((float4*)out)[rayIdx] = h;
	}
	else if (MODE == MODE_ANY_MASK)
	{
		if (L.bprim >= 0) atomicOr(&((uint32_t*)out)[rayIdx >> 5], 1u << (rayIdx & 31));
	}
	else ((uint32_t*)out)[rayIdx] = L.bprim >= 0 ? 1u : 0u;
}

/* Wave-uniform ray feed. The rays are split into TRACE_QUEUES contiguous
 * ranges with a counter each; a wave starts on the queue of its block's XCD
 * (blocks are dealt to the XCDs round robin) and reserves `chunk` rays per
 * atomic, which later refills hand out without touching memory. A drained
 * queue sends the wave to the next one; after TRACE_QUEUES empty queues in a
 * row the feed is empty. */
struct Feed
{
	uint32_t cb, ce;    /* reserved, not yet handed out: [cb, ce) */
	int q, tried;
	bool empty;
};

__device__ __forceinline__ void feed_init(Feed& f)
{
	f.cb = f.ce = 0, f.q = (int)(blockIdx.x % TRACE_QUEUES), f.tried = 0, f.empty = false;
}

/* Hands rays to the idle lanes (ballot -> popcount -> mbcnt rank). Returns
 * true for a lane that received one; sets `drained` once every ray of the
 * launch has been handed out. */
__device__ __forceinline__ bool refill(uint64_t idle, Feed& f, uint32_t* queue, int count, int chunk, bool& drained, uint32_t& idx)
{
	const uint32_t nIdle = (uint32_t)__popcll(idle);
	const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
	const uint32_t rem = f.ce - f.cb;
	uint32_t nb = 0, ne = 0;
	if (rem < nIdle && !f.empty)
	{
		const uint32_t lo = (uint32_t)(((uint64_t)count * (uint32_t)f.q) / TRACE_QUEUES);
		const uint32_t hi = (uint32_t)(((uint64_t)count * (uint32_t)(f.q + 1)) / TRACE_QUEUES);
		const int first = __ffsll((unsigned long long)idle) - 1;
		uint32_t base = 0;
		if (__lane_id() == first) base = atomicAdd(queue + f.q * TRACE_QSTRIDE, (uint32_t)chunk);
		base = __builtin_amdgcn_readlane(base, first);
		if (base < hi - lo)
		{
			nb = lo + base, ne = min(nb + (uint32_t)chunk, hi);
			f.tried = 0;
			if (ne == hi) f.q = (f.q + 1) % TRACE_QUEUES;
		}
		else
		{
			f.q = (f.q + 1) % TRACE_QUEUES;
			if (++f.tried >= TRACE_QUEUES) f.empty = true;
		}
	}
	bool got = true;
	if (below < rem) idx = f.cb + below;
	else if (below - rem < ne - nb) idx = nb + (below - rem);
	else got = false;
	const uint32_t used = min(nIdle, rem + (ne - nb));
	if (used <= rem) f.cb += used;
	else f.cb = nb + (used - rem), f.ce = ne;
	drained = f.empty && f.cb == f.ce;
	return got;
}

template <bool TWO>
__device__ __forceinline__ void start_ray(const TraceScene& sc, const float4* rays, uint32_t idx, Lane& L, uint32_t& ref)
{
	const float4 o4 = rays[2 * idx], d4 = rays[2 * idx + 1];
	L.wO = mk3(o4.x, o4.y, o4.z), L.wD = mk3(d4.x, d4.y, d4.z);
	L.tmin = o4.w, L.tbest = d4.w;
	L.binst = -1, L.bprim = -1, L.bU = L.bV = 0, L.bdet = 1;
	L.O = L.wO, L.D = L.wD;
	L.inv = mk3(safe_rcp(L.D.x), safe_rcp(L.D.y), safe_rcp(L.D.z));
	if (TWO) { L.inst = -1; ref = sc.tlasRoot | REF_TLAS; }
	else { L.rp = ray_setup(L.D); ref = sc.root; }
}

#define COUNT_WAVE(pred, iters, lanes) do { if (COUNT) { const uint64_t m_ = __ballot(pred); \
	if (m_ && __lane_id() == __ffsll((unsigned long long)m_) - 1) iters++, lanes += __popcll(m_); } } while (0)

/* Merged loop (default): every iteration each lane tests one pending leaf
 * triangle and visits one node, both predicated, with both loads issued at
 * the top of the iteration. Lanes never wait for one another between a node
 * phase and a leaf phase; finished lanes are refilled while rays remain. */
template <int FMT, int MODE, bool TWO, bool COUNT>
__global__ void __launch_bounds__(TRACE_BLOCK, TWO ? 5 : 6) k_trace_merged(const TraceScene sc, const float4* __restrict__ rays, void* __restrict__ out,
	int maxCount, const uint32_t* __restrict__ countDev, uint32_t* __restrict__ queue, uint32_t* __restrict__ overflow, TraceStats* __restrict__ stats)
{
	__shared__ uint32_t sstack[TRACE_LDS_STACK * TRACE_BLOCK];
	const int count = countDev ? (int)*countDev : maxCount;
	Stack<TRACE_BLOCK> st;
	st.lds = (lds_u32*)sstack, st.ovf = overflow + ((size_t)blockIdx.x * TRACE_BLOCK + threadIdx.x) * TRACE_OVF_STACK, st.tid = threadIdx.x, st.sp = 0;
	Lane L;
	L.inst = -1;
	L.nodes = sc.nodes, L.tris = sc.tris;
	int rayIdx = -1;
	bool active = false, exhausted = false;
	uint32_t ref = REF_DONE, pend = REF_DONE, steps = 0;
	unsigned long long cn = 0, ct = 0, wn = 0, wl = 0, fn = 0, fl = 0, li = 0, ll = 0, ti = 0, tl = 0;
	Feed feed;
	feed_init(feed);
	if (COUNT && threadIdx.x == 0) atomicMin(&stats->tStart, (unsigned long long)wall_clock64());
	while (true)
	{
		const uint64_t idle = __ballot(!active);
		if (!exhausted && __popcll(idle) >= sc.refillLanes)
		{
			uint32_t idx;
			const bool got = refill(idle, feed, queue, count, sc.chunk, exhausted, idx);
			if (COUNT && exhausted && __lane_id() == 0) atomicMin(&stats->tDrained, (unsigned long long)wall_clock64());
			if (!active && got)
			{
				rayIdx = (int)idx;
				start_ray<TWO>(sc, rays, idx, L, ref);
				st.sp = 0, steps = 0, pend = REF_DONE;
				if (is_blas_leaf(ref)) pend = ref & REF_MASK, ref = REF_DONE;
				active = true;
			}
		}
		if (!__any(active))
		{
			if (exhausted) break;
			continue;   /* this queue was empty: the next refill tries another */
		}
		if (!active) continue;
		bool finished = false;
		while (true)
		{
			if (!exhausted && __popcll(__ballot(true)) <= 64 - sc.refillLanes) break;
			if (exhausted) COUNT_WAVE(true, ti, tl); else COUNT_WAVE(true, li, ll);
			const bool doTri = pend != REF_DONE;
			const bool doNode = is_inner(ref) || (TWO && is_top(ref) && !doTri);
			COUNT_WAVE(doNode, wn, wl);
			COUNT_WAVE(doTri, fn, fl);
			/* issue the triangle and node loads together */
			float4 ta, tb, tc;
			if (doTri) ta = L.tris[3 * pend], tb = L.tris[3 * pend + 1], tc = L.tris[3 * pend + 2];
			const bool bl = doNode && !(TWO && is_top(ref));
			NodeQ8 nq;
			NodeF32 nf;
			if (bl)
			{
				if (FMT == FMT_Q8) nq = ((const NodeQ8*)L.nodes)[ref];
				else nf = ((const NodeF32*)L.nodes)[ref];
			}
			if (doTri)
			{
				if (COUNT) ct++;
				const bool last = (fbits(ta.w) & TRI_LAST) != 0;
				if (test_tri<MODE>(L, ta, tb, tc, TWO ? L.inst : 0)) { finished = true; break; }
				pend = last ? REF_DONE : pend + 1;
			}
			if (doNode)
			{
				if (++steps > TRACE_MAX_STEPS) { ref = REF_DONE, st.sp = 0; }
				else if (!bl) visit_top(sc, L, st, ref);
				else
				{
					if (COUNT) cn++;
					if (FMT == FMT_Q8) visit_q8(nq, L, st, ref);
					else visit_f32(nf, L, st, ref, 0u);
				}
			}
			if (pend == REF_DONE && is_blas_leaf(ref)) pend = ref & REF_MASK, ref = st.sp ? st.pop() : REF_DONE;
			if (pend == REF_DONE && ref == REF_DONE) { finished = true; break; }
		}
		if (!finished) continue;
		write_result<MODE>(L, out, rayIdx);
		active = false;
	}
	if (COUNT)
	{
		atomicAdd(&stats->nodes, cn), atomicAdd(&stats->tris, ct);
		if (wn) atomicAdd(&stats->nodeIters, wn), atomicAdd(&stats->nodeLanes, wl);
		if (fn) atomicAdd(&stats->leafIters, fn), atomicAdd(&stats->leafLanes, fl);
		if (li) atomicAdd(&stats->loopIters, li), atomicAdd(&stats->loopLanes, ll);
		if (ti) atomicAdd(&stats->tailIters, ti), atomicAdd(&stats->tailLanes, tl);
		if (__lane_id() == 0) atomicMax(&stats->tEnd, (unsigned long long)wall_clock64());
	}
}

/* Greedy loop: each iteration the wave does one step of the kind more of
 * its lanes are waiting for, a node visit (lanes holding an interior node)
 * or a leaf triangle test (lanes holding a pending leaf), weighted by
 * sc.triBias / 16. A lane keeps walking nodes while a leaf waits; it stalls
 * only when it reaches a second leaf. */
template <int FMT, int MODE, bool TWO, bool COUNT>
__global__ void __launch_bounds__(TRACE_BLOCK, TWO ? 6 : 8) k_trace_greedy(const TraceScene sc, const float4* __restrict__ rays, void* __restrict__ out,
	int maxCount, const uint32_t* __restrict__ countDev, uint32_t* __restrict__ queue, uint32_t* __restrict__ overflow, TraceStats* __restrict__ stats)
{
	__shared__ uint32_t sstack[TRACE_LDS_STACK * TRACE_BLOCK];
	const int count = countDev ? (int)*countDev : maxCount;
	Stack<TRACE_BLOCK> st;
	st.lds = (lds_u32*)sstack, st.ovf = overflow + ((size_t)blockIdx.x * TRACE_BLOCK + threadIdx.x) * TRACE_OVF_STACK, st.tid = threadIdx.x, st.sp = 0;
	Lane L;
	L.inst = -1;
	L.nodes = sc.nodes, L.tris = sc.tris;
	int rayIdx = -1;
	bool active = false, exhausted = false;
	uint32_t ref = REF_DONE, pend = REF_DONE, steps = 0;
	unsigned long long cn = 0, ct = 0, wn = 0, wl = 0, fn = 0, fl = 0, li = 0, ll = 0, ti = 0, tl = 0;
	Feed feed;
	feed_init(feed);
	if (COUNT && threadIdx.x == 0) atomicMin(&stats->tStart, (unsigned long long)wall_clock64());
	while (true)
	{
		const uint64_t idle = __ballot(!active);
		if (!exhausted && __popcll(idle) >= sc.refillLanes)
		{
			uint32_t idx;
			const bool got = refill(idle, feed, queue, count, sc.chunk, exhausted, idx);
			if (COUNT && exhausted && __lane_id() == 0) atomicMin(&stats->tDrained, (unsigned long long)wall_clock64());
			if (!active && got)
			{
				rayIdx = (int)idx;
				start_ray<TWO>(sc, rays, idx, L, ref);
				st.sp = 0, steps = 0, pend = REF_DONE;
				if (is_blas_leaf(ref)) pend = ref & REF_MASK, ref = REF_DONE;
				active = true;
			}
		}
		if (!__any(active))
		{
			if (exhausted) break;
			continue;   /* this queue was empty: the next refill tries another */
		}
		if (!active) continue;
		bool finished = false;
		while (true)
		{
			if (!exhausted && __popcll(__ballot(true)) <= 64 - sc.refillLanes) break;
			if (exhausted) COUNT_WAVE(true, ti, tl); else COUNT_WAVE(true, li, ll);
			const bool wantTri = pend != REF_DONE;
			const bool wantNode = is_inner(ref) || (TWO && is_top(ref) && !wantTri);
			const int nT = __popcll(__ballot(wantTri)), nN = __popcll(__ballot(wantNode));
			if (nT * sc.triBias > nN * 16)
			{
				COUNT_WAVE(wantTri, fn, fl);
				if (wantTri)
				{
					const float4 a = L.tris[3 * pend], b = L.tris[3 * pend + 1], c = L.tris[3 * pend + 2];
					if (COUNT) ct++;
					if (test_tri<MODE>(L, a, b, c, TWO ? L.inst : 0)) { finished = true; break; }
					pend = (fbits(a.w) & TRI_LAST) ? REF_DONE : pend + 1;
				}
			}
			else
			{
				COUNT_WAVE(wantNode, wn, wl);
				if (wantNode)
				{
					if (++steps > TRACE_MAX_STEPS) { ref = REF_DONE, st.sp = 0; }
					else if (TWO && is_top(ref)) visit_top(sc, L, st, ref);
					else
					{
						if (COUNT) cn++;
						if (FMT == FMT_Q8) visit_q8(((const NodeQ8*)L.nodes)[ref], L, st, ref);
						else visit_f32(((const NodeF32*)L.nodes)[ref], L, st, ref, 0u);
					}
				}
			}
			if (pend == REF_DONE && is_blas_leaf(ref)) pend = ref & REF_MASK, ref = st.sp ? st.pop() : REF_DONE;
			if (pend == REF_DONE && ref == REF_DONE) { finished = true; break; }
		}
		if (!finished) continue;
		write_result<MODE>(L, out, rayIdx);
		active = false;
	}
	if (COUNT)
	{
		atomicAdd(&stats->nodes, cn), atomicAdd(&stats->tris, ct);
		if (wn) atomicAdd(&stats->nodeIters, wn), atomicAdd(&stats->nodeLanes, wl);
		if (fn) atomicAdd(&stats->leafIters, fn), atomicAdd(&stats->leafLanes, fl);
		if (li) atomicAdd(&stats->loopIters, li), atomicAdd(&stats->loopLanes, ll);
		if (ti) atomicAdd(&stats->tailIters, ti), atomicAdd(&stats->tailLanes, tl);
		if (__lane_id() == 0) atomicMax(&stats->tEnd, (unsigned long long)wall_clock64());
	}
}

/* While-while loop with postponed leaves (Aila & Laine 2009): lanes walk
 * interior nodes holding the first leaf they reach until every lane of the
 * wave holds one, then all test leaves together. */
template <int FMT, int MODE, bool TWO, bool COUNT>
__global__ void __launch_bounds__(TRACE_BLOCK, TWO ? 6 : 8) k_trace_ww(const TraceScene sc, const float4* __restrict__ rays, void* __restrict__ out,
	int maxCount, const uint32_t* __restrict__ countDev, uint32_t* __restrict__ queue, uint32_t* __restrict__ overflow, TraceStats* __restrict__ stats)
{
	__shared__ uint32_t sstack[TRACE_LDS_STACK * TRACE_BLOCK];
	const int count = countDev ? (int)*countDev : maxCount;
	Stack<TRACE_BLOCK> st;
	st.lds = (lds_u32*)sstack, st.ovf = overflow + ((size_t)blockIdx.x * TRACE_BLOCK + threadIdx.x) * TRACE_OVF_STACK, st.tid = threadIdx.x, st.sp = 0;
	Lane L;
	L.inst = -1;
	L.nodes = sc.nodes, L.tris = sc.tris;
	int rayIdx = -1;
	bool active = false, exhausted = false;
	uint32_t ref = REF_DONE, leaf = REF_DONE, steps = 0;
	unsigned long long cn = 0, ct = 0, wn = 0, wl = 0, fn = 0, fl = 0, li = 0, ll = 0, ti = 0, tl = 0;
	Feed feed;
	feed_init(feed);
	if (COUNT && threadIdx.x == 0) atomicMin(&stats->tStart, (unsigned long long)wall_clock64());
	while (true)
	{
		const uint64_t idle = __ballot(!active);
		if (!exhausted && __popcll(idle) >= sc.refillLanes)
		{
			uint32_t idx;
			const bool got = refill(idle, feed, queue, count, sc.chunk, exhausted, idx);
			if (COUNT && exhausted && __lane_id() == 0) atomicMin(&stats->tDrained, (unsigned long long)wall_clock64());
			if (!active && got)
			{
				rayIdx = (int)idx;
				start_ray<TWO>(sc, rays, idx, L, ref);
				st.sp = 0, steps = 0, leaf = REF_DONE;
				active = true;
			}
		}
		if (!__any(active))
		{
			if (exhausted) break;
			continue;   /* this queue was empty: the next refill tries another */
		}
		if (!active) continue;
		bool done = false, finished = false;
		while (true)
		{
			if (!exhausted && __popcll(__ballot(true)) <= 64 - sc.refillLanes) break;
			if (exhausted) COUNT_WAVE(true, ti, tl); else COUNT_WAVE(true, li, ll);
			if (leaf == REF_DONE && is_blas_leaf(ref)) leaf = ref, ref = st.sp ? st.pop() : REF_DONE;
			while (true)
			{
				const bool top = TWO && is_top(ref);
				if (!(is_inner(ref) || (top && leaf == REF_DONE))) break;
				COUNT_WAVE(true, wn, wl);
				if (++steps > TRACE_MAX_STEPS) { ref = REF_DONE, leaf = REF_DONE, st.sp = 0; break; }
				if (top) visit_top(sc, L, st, ref);
				else
				{
					if (COUNT) cn++;
					if (FMT == FMT_Q8) visit_q8(((const NodeQ8*)L.nodes)[ref], L, st, ref);
					else visit_f32(((const NodeF32*)L.nodes)[ref], L, st, ref, 0u);
				}
				if (leaf == REF_DONE && is_blas_leaf(ref)) leaf = ref, ref = st.sp ? st.pop() : REF_DONE;
				if (!__any(leaf == REF_DONE)) break;
			}
			while (leaf != REF_DONE)
			{
				COUNT_WAVE(true, fn, fl);
				uint32_t i = leaf & REF_MASK;
				while (true)
				{
					const float4 a = L.tris[3 * i], b = L.tris[3 * i + 1], c = L.tris[3 * i + 2];
					if (COUNT) ct++;
					if (test_tri<MODE>(L, a, b, c, TWO ? L.inst : 0)) { done = true; break; }
					if (fbits(a.w) & TRI_LAST) break;
					i++;
				}
				if (done) break;
				leaf = REF_DONE;
				if (is_blas_leaf(ref)) leaf = ref, ref = st.sp ? st.pop() : REF_DONE;
			}
			if (done || ref == REF_DONE) { finished = true; break; }
		}
		if (!finished) continue;
		write_result<MODE>(L, out, rayIdx);
		active = false;
		leaf = REF_DONE;
	}
	if (COUNT)
	{
		atomicAdd(&stats->nodes, cn), atomicAdd(&stats->tris, ct);
		if (wn) atomicAdd(&stats->nodeIters, wn), atomicAdd(&stats->nodeLanes, wl);
		if (fn) atomicAdd(&stats->leafIters, fn), atomicAdd(&stats->leafLanes, fl);
		if (li) atomicAdd(&stats->loopIters, li), atomicAdd(&stats->loopLanes, ll);
		if (ti) atomicAdd(&stats->tailIters, ti), atomicAdd(&stats->tailLanes, tl);
		if (__lane_id() == 0) atomicMax(&stats->tEnd, (unsigned long long)wall_clock64());
	}
}

size_t trace_overflow_bytes(int persistentBlocks)
{
	return (size_t)persistentBlocks * TRACE_BLOCK * TRACE_OVF_STACK * sizeof(uint32_t);
}

__global__ void k_reset_queue(uint32_t* q)
{
	for (int i = threadIdx.x; i < TRACE_QUEUE_WORDS; i += blockDim.x) q[i] = 0u;
}

/* Blocks of a kernel that stay resident on the whole device: a persistent
 * grid larger than this only queues blocks that find the ray counter spent. */
static int resident_blocks(const void* fn)
{
	static std::mutex lock;
	static std::map<std::pair<const void*, int>, int> cache;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) dev = 0;
	std::lock_guard<std::mutex> g(lock);
	auto it = cache.find({ fn, dev });
	if (it != cache.end()) return it->second;
	int perCU = 0, cus = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, TRACE_BLOCK, 0) != hipSuccess || perCU <= 0) perCU = 1 << 20;
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
	const int n = perCU >= (1 << 20) ? (1 << 30) : perCU * cus;
	cache[{ fn, dev }] = n;
	return n;
}

template <typename K>
static void go(K kernel, const TraceScene& s, const float4* rays, void* out, int count, const uint32_t* countDev, uint32_t* queue, uint32_t* overflow,
	int blocks, hipStream_t stream, TraceStats* stats)
{
	blocks = std::max(1, std::min(blocks, resident_blocks((const void*)kernel)));
	kernel<<<blocks, TRACE_BLOCK, 0, stream>>>(s, rays, out, count, countDev, queue, overflow, stats);
}

template <int FMT, int MODE, bool TWO>
static void launch3(const TraceScene& s, const float4* rays, void* out, int count, const uint32_t* countDev, uint32_t* queue, uint32_t* overflow,
	int blocks, hipStream_t stream, TraceStats* stats)
{
	if (s.loop == LOOP_GREEDY)
	{
		if (stats) go(k_trace_greedy<FMT, MODE, TWO, true>, s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
		else go(k_trace_greedy<FMT, MODE, TWO, false>, s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
	}
	else if (s.loop == LOOP_WHILE_WHILE)
	{
		if (stats) go(k_trace_ww<FMT, MODE, TWO, true>, s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
		else go(k_trace_ww<FMT, MODE, TWO, false>, s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
	}
	else
	{
		if (stats) go(k_trace_merged<FMT, MODE, TWO, true>, s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
		else go(k_trace_merged<FMT, MODE, TWO, false>, s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
	}
}

template <int FMT, int MODE>
static void launch2(const TraceScene& s, const float4* rays, void* out, int count, const uint32_t* countDev, uint32_t* queue, uint32_t* overflow,
	int blocks, hipStream_t stream, TraceStats* stats)
{
	if (s.twoLevel) launch3<FMT, MODE, true>(s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
	else launch3<FMT, MODE, false>(s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
}

template <int FMT>
static void launch1(const TraceScene& s, int mode, const float4* rays, void* out, int count, const uint32_t* countDev, uint32_t* queue, uint32_t* overflow,
	int blocks, hipStream_t stream, TraceStats* stats)
{
	if (mode == MODE_CLOSEST) launch2<FMT, MODE_CLOSEST>(s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
	else if (mode == MODE_ANY_MASK) launch2<FMT, MODE_ANY_MASK>(s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
	else launch2<FMT, MODE_ANY_FLAG>(s, rays, out, count, countDev, queue, overflow, blocks, stream, stats);
}

hipError_t launch_trace_dev(const TraceScene& s, int mode, const float4* rays, void* out, int maxCount, const uint32_t* countDev,
	uint32_t* queue, uint32_t* overflow, int blocks, hipStream_t stream, TraceStats* stats)
{
	if (maxCount <= 0) return hipSuccess;
	if (s.maxStack > TRACE_MAX_STACK) return hipErrorInvalidValue;
	k_reset_queue<<<1, 256, 0, stream>>>(queue);
	if (s.format == FMT_Q8) launch1<FMT_Q8>(s, mode, rays, out, maxCount, countDev, queue, overflow, blocks, stream, stats);
	else launch1<FMT_F32>(s, mode, rays, out, maxCount, countDev, queue, overflow, blocks, stream, stats);
	return hipGetLastError();
}

hipError_t launch_trace(const TraceScene& s, int mode, const float4* rays, void* out, int count,
	uint32_t* queue, uint32_t* overflow, int blocks, hipStream_t stream, TraceStats* stats)
{
	return launch_trace_dev(s, mode, rays, out, count, nullptr, queue, overflow, blocks, stream, stats);
}

} // namespace lh
