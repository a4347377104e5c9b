/* bvh_build.hip - GPU binned-SAH BVH builder (16 bins) for gfx950.
 *
 * Replaces the BLAS/TLAS builds hidden in OptiX (rtpModelUpdate,
 * lib/RenderCore_PrimeRef/core_mesh.cpp:63-64; optixAccelBuild,
 * lib/rendercore_optix7/core_mesh.cpp:123). Breadth-first: each level of the
 * tree is one round of
 *   centroid bounds -> binning (3 axes x 16 bins) -> per-node split decision
 *   -> partition,
 * over all primitive positions, with min/max taken on order-preserving integer
 * keys (exact, so the atomics' order does not matter). A block whose 256
 * positions all belong to one node reduces in LDS first and issues one global
 * atomic per bin field; mixed blocks go straight to global atomics.
 *
 * Split rules (identical in every builder of this project): centroid
 * c = (lo + hi) * 0.5; axis valid when its centroid extent e > 0 and 16/e is
 * finite; bin = min(15, int((c - cmin) * (16 / e))); cost of the split after
 * bin i = nL * A(L) + nR * A(R) with A = dx*dy + dy*dz + dz*dx; first minimum
 * over axes 0..2 and i = 0..14; one primitive -> leaf; n <= maxLeaf -> leaf
 * unless A(node) + cost < n * A(node); no valid axis -> leaf up to
 * maxDegenerateLeaf primitives, else the range is halved.
 */
#include "bvh.h"
#include "lh_math.h"
#include <algorithm>

namespace lh {

static const int BINS = 16;
static const int NSPLIT = BINS - 1;
/* per-level-node scratch (ints) */
static const int S_CMIN = 0, S_CMAX = 3, S_CNT = 6, S_BLO = S_CNT + 3 * BINS, S_BHI = S_BLO + 3 * BINS * 3;
static const int S_SIZE = S_BHI + 3 * BINS * 3; /* 342 */
static const int S_STRIDE = 344;
static const int BLOCK = 256;

__device__ __forceinline__ float half_area(v3 lo, v3 hi)
{
	const v3 d = hi - lo;
	return d.x * d.y + d.y * d.z + d.z * d.x;
}

__device__ __forceinline__ v3 centroid(const float4* lo, const float4* hi, uint32_t p)
{
	const float4 a = lo[p], b = hi[p];
	return mk3((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f);
}

__global__ void k_tri_bounds(const float4* __restrict__ v, int n, float4* __restrict__ lo, float4* __restrict__ hi)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const v3 a = xyz(v[3 * i]), b = xyz(v[3 * i + 1]), c = xyz(v[3 * i + 2]);
	const v3 l = kmin3(a, kmin3(b, c)), h = kmax3(a, kmax3(b, c));
	lo[i] = make_float4(l.x, l.y, l.z, 0.0f);
	hi[i] = make_float4(h.x, h.y, h.z, 0.0f);
}

__global__ void k_scratch_init(int* scr, int count)
{
	const int64_t total = (int64_t)count * S_STRIDE;
	const int32_t kinf = fkey(__builtin_huge_valf()), kninf = fkey(-__builtin_huge_valf());
	for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
	{
		const int f = (int)(i % S_STRIDE);
		int v = 0;
		if (f < S_CMAX) v = kinf;
		else if (f < S_CNT) v = kninf;
		else if (f < S_BLO) v = 0;
		else if (f < S_BHI) v = kinf;
		else if (f < S_SIZE) v = kninf;
		scr[i] = v;
	}
}

/* Root box: reduction of all primitive boxes into 6 keys. */
__global__ void k_root_box(const float4* __restrict__ lo, const float4* __restrict__ hi, int n, int* keys)
{
	__shared__ int s[6];
	if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? fkey(__builtin_huge_valf()) : fkey(-__builtin_huge_valf());
	__syncthreads();
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
	{
		const float4 a = lo[i], b = hi[i];
		atomicMin(&s[0], fkey(a.x)), atomicMin(&s[1], fkey(a.y)), atomicMin(&s[2], fkey(a.z));
		atomicMax(&s[3], fkey(b.x)), atomicMax(&s[4], fkey(b.y)), atomicMax(&s[5], fkey(b.z));
	}
	__syncthreads();
	if (threadIdx.x < 3) atomicMin(&keys[threadIdx.x], s[threadIdx.x]);
	else if (threadIdx.x < 6) atomicMax(&keys[threadIdx.x], s[threadIdx.x]);
}

__global__ void k_root_node(BuildNode* nodes, const int* keys, int n)
{
	BuildNode r;
	for (int c = 0; c < 3; c++) r.lo[c] = keyf(keys[c]), r.hi[c] = keyf(keys[3 + c]);
	r.leftFirst = 0, r.count = n;
	nodes[0] = r;
}

/* Centroid bounds of every level node (chunk [levelBegin, levelBegin+levelCount)). */
__global__ void __launch_bounds__(BLOCK) k_centroid_bounds(const int* __restrict__ posNode, const uint32_t* __restrict__ prims,
	const float4* __restrict__ lo, const float4* __restrict__ hi, int n, int levelBegin, int levelCount, int* __restrict__ scr)
{
	__shared__ int sFirst;
	__shared__ int s[6];
	const int p = blockIdx.x * BLOCK + threadIdx.x;
	const int nd = p < n ? posNode[p] : -1;
	const int local = nd - levelBegin;
	const bool in = nd >= 0 && local >= 0 && local < levelCount;
	if (threadIdx.x == 0) sFirst = in ? nd : -2;
	if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? fkey(__builtin_huge_valf()) : fkey(-__builtin_huge_valf());
	__syncthreads();
	const bool uniform = __syncthreads_and(in && nd == sFirst);
	v3 c = mk3(0, 0, 0);
	if (in) c = centroid(lo, hi, prims[p]);
	if (uniform)
	{
		atomicMin(&s[0], fkey(c.x)), atomicMin(&s[1], fkey(c.y)), atomicMin(&s[2], fkey(c.z));
		atomicMax(&s[3], fkey(c.x)), atomicMax(&s[4], fkey(c.y)), atomicMax(&s[5], fkey(c.z));
		__syncthreads();
		int* d = scr + (int64_t)(sFirst - levelBegin) * S_STRIDE;
		if (threadIdx.x < 3) atomicMin(&d[S_CMIN + threadIdx.x], s[threadIdx.x]);
		else if (threadIdx.x < 6) atomicMax(&d[S_CMAX + threadIdx.x - 3], s[threadIdx.x]);
	}
	else if (in)
	{
		int* d = scr + (int64_t)local * S_STRIDE;
		atomicMin(&d[S_CMIN + 0], fkey(c.x)), atomicMin(&d[S_CMIN + 1], fkey(c.y)), atomicMin(&d[S_CMIN + 2], fkey(c.z));
		atomicMax(&d[S_CMAX + 0], fkey(c.x)), atomicMax(&d[S_CMAX + 1], fkey(c.y)), atomicMax(&d[S_CMAX + 2], fkey(c.z));
	}
}

__device__ __forceinline__ bool axis_scale(const int* d, int a, float& cmin, float& scale)
{
	cmin = keyf(d[S_CMIN + a]);
	const float ext = keyf(d[S_CMAX + a]) - cmin;
	if (!(ext > 0.0f)) return false;
	scale = 16.0f / ext;
	return scale < __builtin_huge_valf();
}

__device__ __forceinline__ int bin_of(float c, float cmin, float scale)
{
	int b = (int)((c - cmin) * scale);
	return b > BINS - 1 ? BINS - 1 : b;
}

__global__ void __launch_bounds__(BLOCK) k_bin(const int* __restrict__ posNode, const uint32_t* __restrict__ prims,
	const float4* __restrict__ lo, const float4* __restrict__ hi, int n, int levelBegin, int levelCount, int* __restrict__ scr)
{
	__shared__ int sFirst;
	__shared__ int s[S_SIZE];
	const int p = blockIdx.x * BLOCK + threadIdx.x;
	const int nd = p < n ? posNode[p] : -1;
	const int local = nd - levelBegin;
	const bool in = nd >= 0 && local >= 0 && local < levelCount;
	if (threadIdx.x == 0) sFirst = in ? nd : -2;
	__syncthreads();
	const bool uniform = __syncthreads_and(in && nd == sFirst);
	if (!uniform && !in) return;
	int* d = scr + (int64_t)(uniform ? sFirst - levelBegin : local) * S_STRIDE;
	if (uniform)
	{
		for (int i = threadIdx.x; i < S_SIZE; i += BLOCK) s[i] = (i < S_BLO) ? 0 : (i < S_BHI ? fkey(__builtin_huge_valf()) : fkey(-__builtin_huge_valf()));
		__syncthreads();
	}
	int* tgt = uniform ? s : d;
	const uint32_t prim = prims[p];
	const v3 c = centroid(lo, hi, prim);
	const float4 pl = lo[prim], ph = hi[prim];
	for (int a = 0; a < 3; a++)
	{
		float cmin, scale;
		if (!axis_scale(d, a, cmin, scale)) continue;
		const int b = bin_of(comp(c, a), cmin, scale);
		const int slot = a * BINS + b;
		atomicAdd(&tgt[S_CNT + slot], 1);
		atomicMin(&tgt[S_BLO + slot * 3 + 0], fkey(pl.x)), atomicMin(&tgt[S_BLO + slot * 3 + 1], fkey(pl.y)), atomicMin(&tgt[S_BLO + slot * 3 + 2], fkey(pl.z));
		atomicMax(&tgt[S_BHI + slot * 3 + 0], fkey(ph.x)), atomicMax(&tgt[S_BHI + slot * 3 + 1], fkey(ph.y)), atomicMax(&tgt[S_BHI + slot * 3 + 2], fkey(ph.z));
	}
	if (uniform)
	{
		__syncthreads();
		for (int i = threadIdx.x; i < S_SIZE - S_CNT; i += BLOCK)
		{
			const int f = S_CNT + i;
			const int v = s[f];
			if (f < S_BLO) { if (v) atomicAdd(&d[f], v); }
			else if (f < S_BHI) { if (v != fkey(__builtin_huge_valf())) atomicMin(&d[f], v); }
			else { if (v != fkey(-__builtin_huge_valf())) atomicMax(&d[f], v); }
		}
	}
}

__device__ __forceinline__ void bin_box(const int* d, int slot, v3& lo, v3& hi)
{
	lo = mk3(keyf(d[S_BLO + slot * 3]), keyf(d[S_BLO + slot * 3 + 1]), keyf(d[S_BLO + slot * 3 + 2]));
	hi = mk3(keyf(d[S_BHI + slot * 3]), keyf(d[S_BHI + slot * 3 + 1]), keyf(d[S_BHI + slot * 3 + 2]));
}

/* dec: x = axis (-1 leaf, 0..2 binned split, 3 halved), y = split bin, z = left count, w = left child. */
__global__ void k_decide(BuildNode* __restrict__ nodes, int levelBegin, int levelCount, const int* __restrict__ scr,
	int4* __restrict__ dec, int2* __restrict__ cursors, int* __restrict__ nodeCounter, int maxLeaf, int maxDegen,
	const uint32_t* __restrict__ prims, const float4* __restrict__ plo, const float4* __restrict__ phi)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= levelCount) return;
	const int ni = levelBegin + i;
	const BuildNode nd = nodes[ni];
	const int start = nd.leftFirst, cnt = nd.count;
	dec[i] = make_int4(-1, 0, 0, 0);
	if (cnt == 1) return;
	const int* d = scr + (int64_t)i * S_STRIDE;
	const v3 INF3 = mk3(__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()), NINF3 = mk3(-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf());
	int bestAxis = -1, bestSplit = -1;
	float bestCost = __builtin_huge_valf();
	for (int a = 0; a < 3; a++)
	{
		float cmin, scale;
		if (!axis_scale(d, a, cmin, scale)) continue;
		float la[NSPLIT], cost[NSPLIT];
		int lc[NSPLIT];
		v3 lo = INF3, hi = NINF3;
		int c = 0;
#pragma unroll
		for (int k = 0; k < NSPLIT; k++)
		{
			v3 bl, bh;
			bin_box(d, a * BINS + k, bl, bh);
			c += d[S_CNT + a * BINS + k], lo = kmin3(lo, bl), hi = kmax3(hi, bh);
			lc[k] = c, la[k] = half_area(lo, hi);
		}
		lo = INF3, hi = NINF3, c = 0;
#pragma unroll
		for (int k = NSPLIT; k > 0; k--)
		{
			v3 bl, bh;
			bin_box(d, a * BINS + k, bl, bh);
			c += d[S_CNT + a * BINS + k], lo = kmin3(lo, bl), hi = kmax3(hi, bh);
			const float ra = half_area(lo, hi);
			cost[k - 1] = (lc[k - 1] == 0 || c == 0) ? __builtin_huge_valf() : (float)lc[k - 1] * la[k - 1] + (float)c * ra;
		}
#pragma unroll
		for (int k = 0; k < NSPLIT; k++)
			if (cost[k] < bestCost) bestCost = cost[k], bestAxis = a, bestSplit = k;
	}
	v3 Llo = INF3, Lhi = NINF3, Rlo = INF3, Rhi = NINF3;
	int leftCount = 0, mode;
	if (bestAxis < 0)
	{
		if (cnt <= maxDegen) return;
		mode = 3;
		leftCount = cnt / 2;
		for (int k = start; k < start + cnt; k++)
		{
			const uint32_t p = prims[k];
			const v3 l = xyz(plo[p]), h = xyz(phi[p]);
			if (k < start + leftCount) Llo = kmin3(Llo, l), Lhi = kmax3(Lhi, h);
			else Rlo = kmin3(Rlo, l), Rhi = kmax3(Rhi, h);
		}
	}
	else
	{
		const float A = half_area(mk3(nd.lo[0], nd.lo[1], nd.lo[2]), mk3(nd.hi[0], nd.hi[1], nd.hi[2]));
		if (cnt <= maxLeaf && !(A + bestCost < (float)cnt * A)) return;
		mode = bestAxis;
		for (int k = 0; k < BINS; k++)
		{
			v3 bl, bh;
			bin_box(d, bestAxis * BINS + k, bl, bh);
			if (k <= bestSplit) Llo = kmin3(Llo, bl), Lhi = kmax3(Lhi, bh), leftCount += d[S_CNT + bestAxis * BINS + k];
			else Rlo = kmin3(Rlo, bl), Rhi = kmax3(Rhi, bh);
		}
	}
	const int child = atomicAdd(nodeCounter, 2);
	BuildNode L, R;
	L.lo[0] = Llo.x, L.lo[1] = Llo.y, L.lo[2] = Llo.z, L.hi[0] = Lhi.x, L.hi[1] = Lhi.y, L.hi[2] = Lhi.z;
	R.lo[0] = Rlo.x, R.lo[1] = Rlo.y, R.lo[2] = Rlo.z, R.hi[0] = Rhi.x, R.hi[1] = Rhi.y, R.hi[2] = Rhi.z;
	L.leftFirst = start, L.count = leftCount;
	R.leftFirst = start + leftCount, R.count = cnt - leftCount;
	nodes[child] = L, nodes[child + 1] = R;
	nodes[ni].leftFirst = child, nodes[ni].count = 0;
	dec[i] = make_int4(mode, bestSplit, leftCount, child);
	cursors[i] = make_int2(start, start + leftCount);
}

/* Partition of split nodes. Primitives of a node are contiguous, so most
 * waves hold one node: they take their left and right slots with one atomic
 * per side (ballot + mbcnt ranks the lanes); mixed waves use one atomic per
 * lane. Positions inside a child are therefore arbitrary; the child's set is
 * not. */
__global__ void __launch_bounds__(BLOCK) k_partition(int* __restrict__ posNode, const uint32_t* __restrict__ prims,
	const float4* __restrict__ lo, const float4* __restrict__ hi, int n, int levelBegin, int levelCount,
	const int* __restrict__ scr, const int4* __restrict__ dec, int2* __restrict__ cursors, const BuildNode* __restrict__ nodes,
	uint32_t* __restrict__ tmpPrims, int* __restrict__ tmpNode)
{
	const int p = blockIdx.x * BLOCK + threadIdx.x;
	const int lane = __lane_id();
	const int nd = p < n ? posNode[p] : -1;
	const int local = nd - levelBegin;
	const bool in = nd >= 0 && local >= 0 && local < levelCount;
	int4 dc = make_int4(-1, 0, 0, 0);
	if (in) dc = dec[local];
	if (in && dc.x < 0) posNode[p] = -1; /* became a leaf: finished */
	const bool part = in && dc.x >= 0;
	uint32_t prim = 0;
	bool left = false;
	if (part)
	{
		prim = prims[p];
		if (dc.x == 3) left = (p - nodes[dc.w].leftFirst) < dc.z;
		else
		{
			const int* d = scr + (int64_t)local * S_STRIDE;
			float cmin, scale;
			axis_scale(d, dc.x, cmin, scale);
			const v3 c = centroid(lo, hi, prim);
			left = bin_of(comp(c, dc.x), cmin, scale) <= dc.y;
		}
	}
	const bool binned = part && dc.x != 3;
	const uint64_t bm = __ballot(binned);
	int np = p;
	if (bm)
	{
		const int leader = __ffsll((unsigned long long)bm) - 1;
		const int l0 = __builtin_amdgcn_readlane(local, leader);
		const bool same = __ballot(binned && local == l0) == bm;
		if (same)
		{
			const uint64_t lm = __ballot(binned && left), rm = bm & ~lm;
			int bl = 0, br = 0;
			if (lane == leader)
			{
				if (lm) bl = atomicAdd(&cursors[l0].x, __popcll(lm));
				if (rm) br = atomicAdd(&cursors[l0].y, __popcll(rm));
			}
			bl = __builtin_amdgcn_readlane(bl, leader), br = __builtin_amdgcn_readlane(br, leader);
			const uint64_t mine = left ? lm : rm;
			const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mine >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mine, 0u));
			if (binned) np = (left ? bl : br) + rank;
		}
		else if (binned) np = left ? atomicAdd(&cursors[local].x, 1) : atomicAdd(&cursors[local].y, 1);
	}
	if (part)
	{
		tmpPrims[np] = prim;
		tmpNode[np] = left ? dc.w : dc.w + 1;
	}
}

__global__ void __launch_bounds__(BLOCK) k_scatter(int* __restrict__ posNode, uint32_t* __restrict__ prims, int n,
	int levelBegin, int levelCount, const int4* __restrict__ dec, const uint32_t* __restrict__ tmpPrims, const int* __restrict__ tmpNode)
{
	const int p = blockIdx.x * BLOCK + threadIdx.x;
	if (p >= n) return;
	const int nd = posNode[p];
	const int local = nd - levelBegin;
	if (nd < 0 || local < 0 || local >= levelCount) return;
	if (dec[local].x < 0) return;
	prims[p] = tmpPrims[p];
	posNode[p] = tmpNode[p];
}

__global__ void k_iota(uint32_t* a, int* b, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = (uint32_t)i, b[i] = 0;
}

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { err = e_; goto done; } } while (0)

hipError_t tri_bounds(const float4* dVerts, int n, float4* dLo, float4* dHi, hipStream_t stream)
{
	if (n <= 0) return hipSuccess;
	k_tri_bounds<<<(n + 255) / 256, 256, 0, stream>>>(dVerts, n, dLo, dHi);
	return hipGetLastError();
}

hipError_t gpu_build(const float4* dLo, const float4* dHi, int n, int maxLeaf, int maxDegen, DeviceBuild& out, hipStream_t stream)
{
	hipError_t err = hipSuccess;
	out.primCount = n, out.nodeCount = 0, out.depth = 0;
	out.levelStart.clear();
	if (n <= 0) return hipSuccess;
	const int chunkCap = std::min(n, 1 << 20);
	int* posNode = nullptr; int* tmpNode = nullptr; uint32_t* tmpPrims = nullptr;
	int* scr = nullptr; int4* dec = nullptr; int2* cursors = nullptr; int* counter = nullptr;
	const int grid = (n + BLOCK - 1) / BLOCK;
	int levelBegin = 0, levelEnd = 1, total = 1;
	CK(hipMalloc(&out.nodes, sizeof(BuildNode) * (size_t)(2 * n)));
	CK(hipMalloc(&out.prims, sizeof(uint32_t) * (size_t)n));
	CK(hipMalloc(&posNode, sizeof(int) * (size_t)n));
	CK(hipMalloc(&tmpNode, sizeof(int) * (size_t)n));
	CK(hipMalloc(&tmpPrims, sizeof(uint32_t) * (size_t)n));
	CK(hipMalloc(&scr, sizeof(int) * (size_t)chunkCap * S_STRIDE));
	CK(hipMalloc(&dec, sizeof(int4) * (size_t)chunkCap));
	CK(hipMalloc(&cursors, sizeof(int2) * (size_t)chunkCap));
	CK(hipMalloc(&counter, sizeof(int) * 8));
	{
		int init[8];
		for (int c = 0; c < 3; c++) init[c] = fkey(__builtin_huge_valf()), init[3 + c] = fkey(-__builtin_huge_valf());
		init[6] = 1, init[7] = 0;
		CK(hipMemcpyAsync(counter, init, sizeof(init), hipMemcpyHostToDevice, stream));
	}
	k_iota<<<grid, BLOCK, 0, stream>>>(out.prims, posNode, n);
	k_root_box<<<grid, BLOCK, 0, stream>>>(dLo, dHi, n, counter);
	k_root_node<<<1, 1, 0, stream>>>(out.nodes, counter, n);
	CK(hipGetLastError());
	while (levelEnd > levelBegin)
	{
		out.depth++;
		out.levelStart.push_back(levelBegin);
		for (int cb = levelBegin; cb < levelEnd; cb += chunkCap)
		{
			const int cc = std::min(chunkCap, levelEnd - cb);
			const int initGrid = (int)std::min<int64_t>(((int64_t)cc * S_STRIDE + 255) / 256, 65536);
			k_scratch_init<<<initGrid, 256, 0, stream>>>(scr, cc);
			k_centroid_bounds<<<grid, BLOCK, 0, stream>>>(posNode, out.prims, dLo, dHi, n, cb, cc, scr);
			k_bin<<<grid, BLOCK, 0, stream>>>(posNode, out.prims, dLo, dHi, n, cb, cc, scr);
			k_decide<<<(cc + 63) / 64, 64, 0, stream>>>(out.nodes, cb, cc, scr, dec, cursors, counter + 6, maxLeaf, maxDegen, out.prims, dLo, dHi);
			k_partition<<<grid, BLOCK, 0, stream>>>(posNode, out.prims, dLo, dHi, n, cb, cc, scr, dec, cursors, out.nodes, tmpPrims, tmpNode);
			k_scatter<<<grid, BLOCK, 0, stream>>>(posNode, out.prims, n, cb, cc, dec, tmpPrims, tmpNode);
			CK(hipGetLastError());
		}
		CK(hipMemcpyAsync(&total, counter + 6, sizeof(int), hipMemcpyDeviceToHost, stream));
		CK(hipStreamSynchronize(stream));
		levelBegin = levelEnd, levelEnd = total;
	}
	out.nodeCount = total;
	out.levelStart.push_back(total);
done:
	(void)hipFree(posNode), (void)hipFree(tmpNode), (void)hipFree(tmpPrims), (void)hipFree(scr), (void)hipFree(dec), (void)hipFree(cursors), (void)hipFree(counter);
	if (err != hipSuccess) free_build(out);
	return err;
}

void free_build(DeviceBuild& b)
{
	if (b.nodes) (void)hipFree(b.nodes);
	if (b.prims) (void)hipFree(b.prims);
	b.nodes = nullptr, b.prims = nullptr, b.nodeCount = 0;
}

/* min / max in std::min / std::max form, (b < a) ? b : a and (a < b) ? b : a:
 * a NaN operand resolves the same way as in the oracle's builder. */
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

__global__ void k_refit_level(BuildNode* __restrict__ nodes, const uint32_t* __restrict__ prims, const float4* __restrict__ v,
	int first, int count)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= count) return;
	const int x = first + i;
	BuildNode b = nodes[x];
	float lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
	if (b.count > 0)
		for (int k = 0; k < b.count; k++)
		{
			const float4* t = v + 3 * (size_t)prims[b.leftFirst + k];
			const float4 p0 = t[0], p1 = t[1], p2 = t[2];
			const float a0[3] = { p0.x, p0.y, p0.z }, a1[3] = { p1.x, p1.y, p1.z }, a2[3] = { p2.x, p2.y, p2.z };
			for (int c = 0; c < 3; c++)
				lo[c] = smin(lo[c], smin(smin(a0[c], a1[c]), a2[c])), hi[c] = smax(hi[c], smax(smax(a0[c], a1[c]), a2[c]));
		}
	else
		for (int k = 0; k < 2; k++)
		{
			const BuildNode ch = nodes[b.leftFirst + k];
			for (int c = 0; c < 3; c++) lo[c] = smin(lo[c], ch.lo[c]), hi[c] = smax(hi[c], ch.hi[c]);
		}
	for (int c = 0; c < 3; c++) nodes[x].lo[c] = lo[c], nodes[x].hi[c] = hi[c];
}

hipError_t refit_level(BuildNode* dNodes, const uint32_t* dPrims, const float4* dVerts, int first, int count, hipStream_t stream)
{
	if (count <= 0) return hipSuccess;
	k_refit_level<<<(count + 255) / 256, 256, 0, stream>>>(dNodes, dPrims, dVerts, first, count);
	return hipGetLastError();
}

} // namespace lh
