/* trace.hip - closest-hit and any-hit traversal for gfx950.
 *
 * Replaces rtpQueryExecute (lib/RenderCore_PrimeRef/rendercore.cpp:559-573,
 * 598-607) and optixTrace + __closesthit__radiance / shadow any-hit
 * (lib/rendercore_optix7/optix/.optix.cu:125-154,174-184).
 *
 * Persistent waves: a wave refills its idle lanes with one atomicAdd on a
 * global ray counter (ballot -> popcount -> mbcnt gives each idle lane its
 * ray), whenever a quarter of the wave is idle or every lane has finished.
 * Each lane walks the tree with a short stack of TRACE_LDS_STACK entries in
 * LDS (lane-interleaved, conflict free) that spills to a per-thread global
 * area beyond that. Box tests compute t = (plane - O) * rcp(D) and widen the
 * far distance by 2^-16 relative, so rounding never prunes a box the exact
 * interval reaches; only the watertight triangle test (lh_math.h) decides a
 * hit, hence results equal a brute-force scan bit for bit.
 */
#include "trace.h"
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

template <int BLOCK>
struct Stack
{
	uint32_t* lds;      /* TRACE_LDS_STACK x BLOCK */
	uint32_t* ovf;      /* this thread's overflow area */
	int tid, sp;
	/* The host checks that the tree depth fits TRACE_MAX_STACK; a push beyond
	 * it is dropped (never written out of bounds). */
	__device__ __forceinline__ void push(uint32_t v)
	{
		if (sp < TRACE_LDS_STACK) lds[sp * BLOCK + tid] = v, sp++;
		else if (sp < TRACE_MAX_STACK) ovf[sp - TRACE_LDS_STACK] = v, sp++;
	}
	__device__ __forceinline__ uint32_t pop()
	{
		sp--;
		return sp < TRACE_LDS_STACK ? lds[sp * BLOCK + tid] : ovf[sp - TRACE_LDS_STACK];
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

__device__ __forceinline__ float ubyte(uint32_t w, int c)
{
	return (float)((w >> (8 * c)) & 255u);
}

template <int FMT, int MODE, bool TWO, bool COUNT>
__global__ void __launch_bounds__(TRACE_BLOCK, TWO ? 6 : 8) k_trace(const TraceScene sc, const float4* __restrict__ rays, void* __restrict__ out,
	int maxCount, const uint32_t* __restrict__ countDev, uint32_t* __restrict__ counter, uint32_t* __restrict__ overflow, TraceStats* __restrict__ stats)
{
	__shared__ uint32_t sstack[TRACE_LDS_STACK * TRACE_BLOCK];
	const int count = countDev ? (int)*countDev : maxCount;
	const int tid = threadIdx.x;
	const int lane = __lane_id();
	Stack<TRACE_BLOCK> st;
	st.lds = sstack, st.ovf = overflow + ((size_t)blockIdx.x * TRACE_BLOCK + tid) * TRACE_OVF_STACK, st.tid = tid, st.sp = 0;
	Lane L;
	L.inst = -1;
	L.nodes = sc.nodes, L.tris = sc.tris;
	int rayIdx = -1;
	bool active = false, exhausted = false;
	uint32_t ref = REF_DONE, leaf = REF_DONE, steps = 0;
	unsigned long long cn = 0, ct = 0, wn = 0, wl = 0, fn = 0, fl = 0;
	while (true)
	{
		/* refill: one atomic per wave hands out rays to the idle lanes */
		const uint64_t idle = __ballot(!active);
		const int nIdle = __popcll(idle);
		if (!exhausted && nIdle >= 16)
		{
			const int first = __ffsll((unsigned long long)idle) - 1;
			uint32_t base = 0;
			if (lane == first) base = atomicAdd(counter, (uint32_t)nIdle);
			base = __builtin_amdgcn_readlane(base, first);
			exhausted = base + (uint32_t)nIdle >= (uint32_t)count;
			if (!active)
			{
				const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
				const uint32_t idx = base + below;
				if (idx < (uint32_t)count)
				{
					rayIdx = (int)idx;
					const float4 o4 = rays[2 * idx], d4 = rays[2 * idx + 1];
					L.wO = mk3(o4.x, o4.y, o4.z), L.wD = mk3(d4.x, d4.y, d4.z);
					L.tmin = o4.w, L.tbest = d4.w;
					L.binst = -1, L.bprim = -1, L.bU = L.bV = 0, L.bdet = 1;
					L.O = L.wO, L.D = L.wD;
					L.inv = mk3(safe_rcp(L.D.x), safe_rcp(L.D.y), safe_rcp(L.D.z));
					st.sp = 0;
					steps = 0;
					leaf = REF_DONE;
					if (TWO) { L.inst = -1; ref = sc.tlasRoot | REF_TLAS; }
					else { L.rp = ray_setup(L.D); ref = sc.root; }
					active = true;
				}
			}
		}
		if (!__any(active)) break;
		if (!active) continue;
		bool done = false, finished = false;
		while (true)
		{
			/* leave to refill when a quarter of the wave has gone idle */
			if (!exhausted && __popcll(__ballot(true)) < 48) break;
			/* a leaf as the current ref (a root that is a leaf, or a ref
			 * popped as the node loop ended) is handed to the leaf phase */
			if (leaf == REF_DONE && (ref & (REF_LEAF | REF_TLAS)) == REF_LEAF)
			{
				leaf = ref;
				ref = st.sp ? st.pop() : REF_DONE;
			}
			/* node phase: walk interior nodes; the first leaf reached is
			 * postponed and the lane keeps walking until every lane of the
			 * wave holds a leaf (Aila & Laine 2009, "while-while" with
			 * postponed leaves), so leaf tests run with the wave converged.
			 * A lane never crosses into the top level holding a leaf. */
			while (true)
			{
				const bool top = TWO && (ref & REF_TLAS) && ref != REF_DONE;
				const bool inner = !(ref & (REF_LEAF | REF_TLAS));
				if (!(inner || (top && leaf == REF_DONE))) break;
				if (COUNT && __lane_id() == __ffsll((unsigned long long)__ballot(true)) - 1)
					wn++, wl += __popcll(__ballot(true));
				if (++steps > TRACE_MAX_STEPS)
				{	/* safety bound: a ray never visits more nodes than the trees hold */
					ref = REF_DONE, leaf = REF_DONE, st.sp = 0;
					break;
				}
				if (top)
				{
					if (L.inst >= 0)
					{	/* back at the top level: restore the world ray */
						L.O = L.wO, L.D = L.wD;
						L.inv = mk3(safe_rcp(L.D.x), safe_rcp(L.D.y), safe_rcp(L.D.z));
						L.inst = -1;
					}
					const uint32_t r = ref & ~REF_TLAS;
					if (r & REF_LEAF)
					{	/* enter an instance */
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
					else
					{
						const NodeF32 n = sc.tlas[r];
						float tl, tr;
						const bool hl = slab(n.lxy.x, n.lxy.y, n.lxy.z, n.lxy.w, n.lrz.x, n.lrz.y, L.O, L.inv, L.tmin, L.tbest, tl);
						const bool hr = slab(n.rxy.x, n.rxy.y, n.rxy.z, n.rxy.w, n.lrz.z, n.lrz.w, L.O, L.inv, L.tmin, L.tbest, tr);
						const uint32_t rl = n.refs.x | REF_TLAS, rr = n.refs.y | REF_TLAS;
						if (hl && hr) { if (tr < tl) { st.push(rl); ref = rr; } else { st.push(rr); ref = rl; } }
						else if (hl) ref = rl;
						else if (hr) ref = rr;
						else ref = st.sp ? st.pop() : REF_DONE;
					}
				}
				else if (FMT == FMT_F32)
				{
					if (COUNT) cn++;
					const NodeF32 n = ((const NodeF32*)L.nodes)[ref];
					float tl, tr;
					const bool hl = slab(n.lxy.x, n.lxy.y, n.lxy.z, n.lxy.w, n.lrz.x, n.lrz.y, L.O, L.inv, L.tmin, L.tbest, tl);
					const bool hr = slab(n.rxy.x, n.rxy.y, n.rxy.z, n.rxy.w, n.lrz.z, n.lrz.w, L.O, L.inv, L.tmin, L.tbest, tr);
					if (hl && hr) { if (tr < tl) { st.push(n.refs.x); ref = n.refs.y; } else { st.push(n.refs.y); ref = n.refs.x; } }
					else if (hl) ref = n.refs.x;
					else if (hr) ref = n.refs.y;
					else ref = st.sp ? st.pop() : REF_DONE;
				}
				else
				{
					/* 4-wide node, child planes = origin + q * 2^e per axis. Per node
					 * and axis: b = (origin - O) / D and d = 2^e / D; a plane is at
					 * t = q * d + b. Near and far planes are picked by the sign of
					 * D; e = 2^-20 (|b| + 255 |d|) bounds the rounding of the
					 * reformulation and is subtracted from near and added to far
					 * planes, so every box the exact interval reaches is kept. */
					if (COUNT) cn++;
					const NodeQ8 n = ((const NodeQ8*)L.nodes)[ref];
					/* keep the child refs in the same memory round trip as the
					 * boxes (the compiler would otherwise fetch them after the
					 * box test, adding a dependent load per node visit) */
					asm volatile("" : : "v"(n.refs.x), "v"(n.refs.y), "v"(n.refs.z), "v"(n.refs.w));
					const uint32_t e = fbits(n.origin.w);
					const float dx = bitsf((e & 255u) << 23) * L.inv.x;
					const float dy = bitsf(((e >> 8) & 255u) << 23) * L.inv.y;
					const float dz = bitsf(((e >> 16) & 255u) << 23) * L.inv.z;
					const float bx = (n.origin.x - L.O.x) * L.inv.x;
					const float by = (n.origin.y - L.O.y) * L.inv.y;
					const float bz = (n.origin.z - L.O.z) * L.inv.z;
					const float ex = (fabsf(bx) + 255.0f * fabsf(dx)) * 9.5367431640625e-7f;
					const float ey = (fabsf(by) + 255.0f * fabsf(dy)) * 9.5367431640625e-7f;
					const float ez = (fabsf(bz) + 255.0f * fabsf(dz)) * 9.5367431640625e-7f;
					const float bnx = bx - ex, bfx = bx + ex, bny = by - ey, bfy = by + ey, bnz = bz - ez, bfz = bz + ez;
					const bool px = L.inv.x >= 0.0f, py = L.inv.y >= 0.0f, pz = L.inv.z >= 0.0f;
					const uint32_t qnx = px ? n.qa.x : n.qa.w, qfx = px ? n.qa.w : n.qa.x;
					const uint32_t qny = py ? n.qa.y : n.qb.x, qfy = py ? n.qb.x : n.qa.y;
					const uint32_t qnz = pz ? n.qa.z : n.qb.y, qfz = pz ? n.qb.y : n.qa.z;
					const int nc = (int)(e >> 24);
					float tk[4];
					uint32_t rk[4] = { n.refs.x, n.refs.y, n.refs.z, n.refs.w };
#pragma unroll
					for (int c = 0; c < 4; c++)
					{
						const float tn = fmaxf(fmaxf(fmaf(ubyte(qnx, c), dx, bnx), fmaf(ubyte(qny, c), dy, bny)), fmaxf(fmaf(ubyte(qnz, c), dz, bnz), L.tmin));
						const float tf = fminf(fminf(fmaf(ubyte(qfx, c), dx, bfx), fmaf(ubyte(qfy, c), dy, bfy)), fminf(fmaf(ubyte(qfz, c), dz, bfz), L.tbest));
						tk[c] = (c < nc && tn <= tf) ? tn : __builtin_huge_valf();
					}
					cswap(tk[0], tk[1], rk[0], rk[1]);
					cswap(tk[2], tk[3], rk[2], rk[3]);
					cswap(tk[0], tk[2], rk[0], rk[2]);
					cswap(tk[1], tk[3], rk[1], rk[3]);
					cswap(tk[1], tk[2], rk[1], rk[2]);
					if (tk[0] < __builtin_huge_valf())
					{
						if (tk[3] < __builtin_huge_valf()) st.push(rk[3]);
						if (tk[2] < __builtin_huge_valf()) st.push(rk[2]);
						if (tk[1] < __builtin_huge_valf()) st.push(rk[1]);
						ref = rk[0];
					}
					else ref = st.sp ? st.pop() : REF_DONE;
				}
				/* postpone the first leaf reached */
				if (leaf == REF_DONE && (ref & (REF_LEAF | REF_TLAS)) == REF_LEAF)
				{
					leaf = ref;
					ref = st.sp ? st.pop() : REF_DONE;
				}
				if (!__any(leaf == REF_DONE)) break;
			}
			/* leaf phase: the postponed leaf, then any leaves that follow */
			while (leaf != REF_DONE)
			{
				if (COUNT && __lane_id() == __ffsll((unsigned long long)__ballot(true)) - 1)
					fn++, fl += __popcll(__ballot(true));
				uint32_t i = leaf & REF_MASK;
				const int instId = TWO ? L.inst : 0;
				float4 a = L.tris[3 * i], b = L.tris[3 * i + 1], c = L.tris[3 * i + 2];
				while (true)
				{
					const uint32_t w = fbits(a.w);
					const bool last = (w & TRI_LAST) != 0;
					/* fetch the next triangle of the leaf while testing this one */
					float4 na = a, nb = b, nc = c;
					if (!last) na = L.tris[3 * i + 3], nb = L.tris[3 * i + 4], nc = L.tris[3 * i + 5];
					const int prim = (int)(w & 0x7fffffffu);
					if (COUNT) ct++;
					float t, U, V, det;
					if (tri_intersect(L.rp, L.O, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), t, U, V, det) && t > L.tmin)
					{
						if (MODE != MODE_CLOSEST)
						{
							if (t < L.tbest) { L.binst = instId, L.bprim = prim; done = true; break; }
						}
						else if (better_hit(t, instId, prim, L.tbest, L.binst, L.bprim))
							L.tbest = t, L.binst = instId, L.bprim = prim, L.bU = U, L.bV = V, L.bdet = det;
					}
					if (last) break;
					a = na, b = nb, c = nc;
					i++;
				}
				if (done) break;
				leaf = REF_DONE;
				if ((ref & (REF_LEAF | REF_TLAS)) == REF_LEAF)
				{
					leaf = ref;
					ref = st.sp ? st.pop() : REF_DONE;
				}
			}
			if (done || ref == REF_DONE) { finished = true; break; }
		}
		if (!finished) continue;
		/* write the result of ray rayIdx */
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
		else
		{
			((uint32_t*)out)[rayIdx] = L.bprim >= 0 ? 1u : 0u;
		}
		active = false;
		leaf = REF_DONE;
	}
	if (COUNT)
	{
		atomicAdd(&stats->nodes, cn);
		atomicAdd(&stats->tris, ct);
		if (wn) atomicAdd(&stats->nodeIters, wn), atomicAdd(&stats->nodeLanes, wl);
		if (fn) atomicAdd(&stats->leafIters, fn), atomicAdd(&stats->leafLanes, fl);
	}
}

size_t trace_overflow_bytes(int persistentBlocks)
{
	return (size_t)persistentBlocks * TRACE_BLOCK * TRACE_OVF_STACK * sizeof(uint32_t);
}

__global__ void k_reset_counter(uint32_t* c) { *c = 0u; }

template <int FMT, int MODE, bool TWO>
static void launch3(const TraceScene& s, const float4* rays, void* out, int count, const uint32_t* countDev, uint32_t* counter, uint32_t* overflow,
	int blocks, hipStream_t stream, TraceStats* stats)
{
	if (stats) k_trace<FMT, MODE, TWO, true><<<blocks, TRACE_BLOCK, 0, stream>>>(s, rays, out, count, countDev, counter, overflow, stats);
	else k_trace<FMT, MODE, TWO, false><<<blocks, TRACE_BLOCK, 0, stream>>>(s, rays, out, count, countDev, counter, overflow, stats);
}

template <int FMT, int MODE>
static void launch2(const TraceScene& s, const float4* rays, void* out, int count, const uint32_t* countDev, uint32_t* counter, uint32_t* overflow,
	int blocks, hipStream_t stream, TraceStats* stats)
{
	if (s.twoLevel) launch3<FMT, MODE, true>(s, rays, out, count, countDev, counter, overflow, blocks, stream, stats);
	else launch3<FMT, MODE, false>(s, rays, out, count, countDev, counter, overflow, blocks, stream, stats);
}

template <int FMT>
static void launch1(const TraceScene& s, int mode, const float4* rays, void* out, int count, const uint32_t* countDev, uint32_t* counter, uint32_t* overflow,
	int blocks, hipStream_t stream, TraceStats* stats)
{
	if (mode == MODE_CLOSEST) launch2<FMT, MODE_CLOSEST>(s, rays, out, count, countDev, counter, overflow, blocks, stream, stats);
	else if (mode == MODE_ANY_MASK) launch2<FMT, MODE_ANY_MASK>(s, rays, out, count, countDev, counter, overflow, blocks, stream, stats);
	else launch2<FMT, MODE_ANY_FLAG>(s, rays, out, count, countDev, counter, overflow, blocks, stream, stats);
}

hipError_t launch_trace_dev(const TraceScene& s, int mode, const float4* rays, void* out, int maxCount, const uint32_t* countDev,
	uint32_t* counter, uint32_t* overflow, int blocks, hipStream_t stream, TraceStats* stats)
{
	if (maxCount <= 0) return hipSuccess;
	if (s.maxStack > TRACE_MAX_STACK) return hipErrorInvalidValue;
	k_reset_counter<<<1, 1, 0, stream>>>(counter);
	if (s.format == FMT_Q8) launch1<FMT_Q8>(s, mode, rays, out, maxCount, countDev, counter, overflow, blocks, stream, stats);
	else launch1<FMT_F32>(s, mode, rays, out, maxCount, countDev, counter, overflow, blocks, stream, stats);
	return hipGetLastError();
}

hipError_t launch_trace(const TraceScene& s, int mode, const float4* rays, void* out, int count,
	uint32_t* counter, uint32_t* overflow, int blocks, hipStream_t stream, TraceStats* stats)
{
	return launch_trace_dev(s, mode, rays, out, count, nullptr, counter, overflow, blocks, stream, stats);
}

} // namespace lh
