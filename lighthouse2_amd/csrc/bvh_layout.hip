/* bvh_layout.hip - lays a binned-SAH build tree out in the traversal formats,
 * on the device.
 *
 * Replaces the compaction and layout steps that OptiX runs inside its device
 * build (lib/rendercore_optix7/core_mesh.cpp:105-123, optixAccelBuild +
 * optixAccelCompact). Formats (bvh.h):
 *   NodeF32  binary nodes, slots in depth-first order;
 *   NodeQ8   4-wide nodes: a node adopts the two children of its interior
 *            child of largest surface area until it holds 4 children (first
 *            maximum wins), child boxes quantised to bytes against the node box
 *            in power-of-two steps, rounded outward;
 *   leaf triangles in depth-first (left first) leaf order, ascending primitive
 *   id inside a leaf, the last triangle of a leaf flagged.
 * Slot order of both node formats: a node, when it is expanded in depth-first
 * order, appends its interior children (in child order) to the node array. So
 * with A(x) = the array size when x is expanded and ik(x) its interior-child
 * count, x's interior children take slots A(x), A(x)+1, ...; the first child
 * is expanded next, A(c0) = A(x) + ik(x); a later child after the whole
 * subtree of the one before, A(c_j+1) = A(c_j) + size(c_j) - 1 (each node of a
 * subtree but its root is some node's interior child). The root sits in slot 0
 * with A = 1.
 *
 * The build tree is breadth-first: the nodes of depth d are the id range
 * [levelStart[d], levelStart[d+1]) and children lie one level below their
 * parent. Every pass is one launch per level, one thread per node:
 *   up1   subtree triangle count and interior-node count (deepest level first);
 *   down1 leaf bases, f32 slots and nodes, the leaf triangle order, and which
 *         binary nodes head a 4-wide node (with their children);
 *   up2   4-wide subtree sizes;
 *   down2 4-wide slots and quantised nodes.
 */
#include "bvh.h"
#include "lh_math.h"
#include <algorithm>

namespace lh {

static const int LBLOCK = 256;

/* scratch: one int per build node per field (+ one int4 of 4-wide children) */
enum { F_TRI = 0, F_N32, F_BASE, F_A32, F_S32, F_WLEV, F_NQ8, F_AQ8, F_SQ8, F_COUNT };

struct LayoutArgs
{
	const BuildNode* nodes;
	const uint32_t* prims;
	const uint32_t* leafRefMap;   /* non-null: a leaf (one primitive) refers to leafRefMap[prim] */
	int* f;                       /* F_COUNT x nodeCount ints */
	int4* kids;                   /* 4-wide children per binary node (-1 padded) */
	int nodeCount;
	NodeF32* f32;
	int32_t* f32Kids;
	NodeQ8* q8;                   /* null: no 4-wide layout */
	int32_t* q8Map;
	uint32_t* order;
	uint8_t* last;
	int* info;                    /* [0] = 4-wide depth */
};

__device__ __forceinline__ int& fld(const LayoutArgs& a, int field, int x) { return a.f[(size_t)field * a.nodeCount + x]; }

__device__ __forceinline__ float node_half_area(const BuildNode& b)
{
	const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
	return dx * dy + dy * dz + dz * dx;
}

__global__ void k_lay_root(LayoutArgs a)
{
	if (threadIdx.x != 0) return;
	fld(a, F_BASE, 0) = 0;
	fld(a, F_S32, 0) = 0, fld(a, F_A32, 0) = 1;
	const bool interior = a.nodes[0].count == 0;
	fld(a, F_WLEV, 0) = (a.q8 && interior) ? 0 : -1;
	fld(a, F_SQ8, 0) = 0, fld(a, F_AQ8, 0) = 1;
	a.info[0] = a.q8 ? 1 : 0;   /* a root leaf counts as one level, as the traversal stack sizing assumes */
}

/* WLEV = -1 for every node below the root (the root kernel sets node 0 after). */
__global__ void k_lay_clear(LayoutArgs a)
{
	const int x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x < a.nodeCount) fld(a, F_WLEV, x) = -1;
}

__global__ void __launch_bounds__(LBLOCK) k_lay_up1(LayoutArgs a, int b, int e)
{
	const int x = b + blockIdx.x * LBLOCK + threadIdx.x;
	if (x >= e) return;
	const BuildNode N = a.nodes[x];
	if (N.count > 0) fld(a, F_TRI, x) = N.count, fld(a, F_N32, x) = 0;
	else
	{
		const int L = N.leftFirst;
		fld(a, F_TRI, x) = fld(a, F_TRI, L) + fld(a, F_TRI, L + 1);
		fld(a, F_N32, x) = 1 + fld(a, F_N32, L) + fld(a, F_N32, L + 1);
	}
}

__device__ __forceinline__ uint32_t leaf_ref(const LayoutArgs& a, const BuildNode& leaf, int base)
{
	return a.leafRefMap ? (REF_LEAF | a.leafRefMap[a.prims[leaf.leftFirst]]) : (REF_LEAF | (uint32_t)base);
}

__global__ void __launch_bounds__(LBLOCK) k_lay_down1(LayoutArgs a, int b, int e)
{
	const int x = b + blockIdx.x * LBLOCK + threadIdx.x;
	if (x >= e) return;
	const BuildNode N = a.nodes[x];
	const int base = fld(a, F_BASE, x);
	if (N.count > 0)
	{	/* leaf: its triangles in ascending primitive order from its base slot */
		if (a.leafRefMap) return;
		const uint32_t* p = a.prims + N.leftFirst;
		for (int k = 0; k < N.count; k++)
		{
			const uint32_t pk = p[k];
			int rank = 0;
			for (int j = 0; j < N.count; j++) rank += p[j] < pk ? 1 : 0;
			a.order[base + rank] = pk;
			a.last[base + rank] = rank == N.count - 1 ? 1 : 0;
		}
		return;
	}
	const int L = N.leftFirst, R = L + 1;
	const BuildNode NL = a.nodes[L], NR = a.nodes[R];
	const int baseR = base + fld(a, F_TRI, L);
	fld(a, F_BASE, L) = base, fld(a, F_BASE, R) = baseR;
	const int lInt = NL.count == 0 ? 1 : 0, rInt = NR.count == 0 ? 1 : 0;
	const int ik = lInt + rInt, A = fld(a, F_A32, x), s = fld(a, F_S32, x);
	uint32_t refL, refR;
	if (lInt) fld(a, F_S32, L) = A, fld(a, F_A32, L) = A + ik, refL = (uint32_t)A;
	else refL = leaf_ref(a, NL, base);
	if (rInt)
	{
		fld(a, F_S32, R) = A + lInt;
		fld(a, F_A32, R) = lInt ? A + ik + fld(a, F_N32, L) - 1 : A + ik;
		refR = (uint32_t)(A + lInt);
	}
	else refR = leaf_ref(a, NR, baseR);
	NodeF32 o;
	o.lxy = make_float4(NL.lo[0], NL.hi[0], NL.lo[1], NL.hi[1]);
	o.rxy = make_float4(NR.lo[0], NR.hi[0], NR.lo[1], NR.hi[1]);
	o.lrz = make_float4(NL.lo[2], NL.hi[2], NR.lo[2], NR.hi[2]);
	o.refs = make_uint4(refL, refR, 0, 0);
	a.f32[s] = o;
	if (a.f32Kids) a.f32Kids[2 * s] = L, a.f32Kids[2 * s + 1] = R;
	/* 4-wide: a node heading a wide node collects its children and marks the interior ones */
	const int wl = fld(a, F_WLEV, x);
	if (!a.q8 || wl < 0) return;
	int k[4] = { L, R, -1, -1 };
	int cnt = 2;
	while (cnt < 4)
	{
		int best = -1;
		float ba = -1.0f;
		for (int j = 0; j < cnt; j++)
		{
			const BuildNode& C = a.nodes[k[j]];
			if (C.count == 0)
			{
				const float h = node_half_area(C);
				if (h > ba) ba = h, best = j;
			}
		}
		if (best < 0) break;
		const int c = a.nodes[k[best]].leftFirst;
		k[best] = c;
		k[cnt++] = c + 1;
	}
	a.kids[x] = make_int4(k[0], k[1], k[2], k[3]);
	bool any = false;
	for (int j = 0; j < cnt; j++)
		if (a.nodes[k[j]].count == 0) fld(a, F_WLEV, k[j]) = wl + 1, any = true;
	if (any) atomicMax(&a.info[0], wl + 2);
}

__global__ void __launch_bounds__(LBLOCK) k_lay_up2(LayoutArgs a, int b, int e)
{
	const int x = b + blockIdx.x * LBLOCK + threadIdx.x;
	if (x >= e || fld(a, F_WLEV, x) < 0) return;
	const int4 kv = a.kids[x];
	const int k[4] = { kv.x, kv.y, kv.z, kv.w };
	int size = 1;
	for (int j = 0; j < 4; j++)
		if (k[j] >= 0 && a.nodes[k[j]].count == 0) size += fld(a, F_NQ8, k[j]);
	fld(a, F_NQ8, x) = size;
}

/* Quantisation of one axis: the smallest power-of-two step from
 * ceil(log2(extent / 255)) up for which every child's lo / hi rounds outward
 * to a byte offset from the parent's lo (decoded lo = origin + q * step never
 * exceeds the true lo, decoded hi never falls below the true hi; q * step is
 * exact, so a fused multiply-add in the traversal decodes the same value). The step search ends by exponent 128
 * (an infinite step makes every test false, i.e. accepted). */
__device__ uint32_t quantise_axis_dev(float lo, float hi, const float* clo, const float* chi, int cnt, uint32_t* qlo, uint32_t* qhi)
{
	const float ext = hi - lo;
	int e = -126;
	if (ext > 0) e = max(-126, (int)ceil(log2((double)ext / 255.0)));
	for (; e <= 128; e++)
	{
		const float s = __uint_as_float((uint32_t)(e + 127) << 23);
		bool ok = true;
		for (int c = 0; c < cnt && ok; c++)
		{
			const double ql = floor(((double)clo[c] - (double)lo) / (double)s);
			const double qlc = (ql < 255.0) ? ql : 255.0;
			int a = (int)((0.0 < qlc) ? qlc : 0.0);
			while (a > 0 && __fadd_rn(lo, __fmul_rn((float)a, s)) > clo[c]) a--;
			const double qh = ceil(((double)chi[c] - (double)lo) / (double)s);
			const double qhc = (qh < 255.0) ? qh : 255.0;
			int b = (int)((0.0 < qhc) ? qhc : 0.0);
			while (b < 255 && __fadd_rn(lo, __fmul_rn((float)b, s)) < chi[c]) b++;
			if (__fadd_rn(lo, __fmul_rn((float)b, s)) < chi[c] || __fadd_rn(lo, __fmul_rn((float)a, s)) > clo[c]) { ok = false; break; }
			qlo[c] = (uint32_t)a, qhi[c] = (uint32_t)b;
		}
		if (ok) return (uint32_t)(e + 127);
	}
	return 255u;
}

__global__ void k_refit_f32(const BuildNode* __restrict__ nodes, NodeF32* __restrict__ out, const int32_t* __restrict__ kids, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const BuildNode L = nodes[kids[2 * i]], R = nodes[kids[2 * i + 1]];
	out[i].lxy = make_float4(L.lo[0], L.hi[0], L.lo[1], L.hi[1]);
	out[i].rxy = make_float4(R.lo[0], R.hi[0], R.lo[1], R.hi[1]);
	out[i].lrz = make_float4(L.lo[2], L.hi[2], R.lo[2], R.hi[2]);
}

__global__ void k_refit_q8(const BuildNode* __restrict__ nodes, NodeQ8* __restrict__ out, const int32_t* __restrict__ map, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const BuildNode B = nodes[map[5 * i]];
	int cnt = 0;
	int kid[4];
	for (int c = 0; c < 4; c++) { kid[c] = map[5 * i + 1 + c]; if (kid[c] >= 0) cnt = c + 1; }
	uint32_t ex[3], q[6][4] = {};
	for (int a = 0; a < 3; a++)
	{
		float clo[4], chi[4];
		for (int c = 0; c < cnt; c++) { const BuildNode& K = nodes[kid[c]]; clo[c] = K.lo[a], chi[c] = K.hi[a]; }
		ex[a] = quantise_axis_dev(B.lo[a], B.hi[a], clo, chi, cnt, q[a], q[3 + a]);
	}
	auto pack = [](const uint32_t* v) { return v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24); };
	out[i].origin = make_float4(B.lo[0], B.lo[1], B.lo[2], __uint_as_float(q8_step_bits(ex[2])));
	out[i].qa = make_uint4(pack(q[0]), pack(q[1]), pack(q[2]), pack(q[3]));
	out[i].qb = make_uint4(pack(q[4]), pack(q[5]), q8_step_bits(ex[0]), q8_step_bits(ex[1]));
}

hipError_t refit_layout(const BuildNode* dNodes, NodeF32* dF32, const int32_t* dF32Kids, int nF32, NodeQ8* dQ8, const int32_t* dQ8Map, int nQ8, hipStream_t stream)
{
	if (nF32 > 0) k_refit_f32<<<(nF32 + 255) / 256, 256, 0, stream>>>(dNodes, dF32, dF32Kids, nF32);
	if (nQ8 > 0) k_refit_q8<<<(nQ8 + 127) / 128, 128, 0, stream>>>(dNodes, dQ8, dQ8Map, nQ8);
	return hipGetLastError();
}

__global__ void __launch_bounds__(LBLOCK) k_lay_down2(LayoutArgs a, int b, int e)
{
	const int x = b + blockIdx.x * LBLOCK + threadIdx.x;
	if (x >= e || fld(a, F_WLEV, x) < 0) return;
	const int4 kv = a.kids[x];
	const int k[4] = { kv.x, kv.y, kv.z, kv.w };
	int cnt = 0, ik = 0;
	for (int j = 0; j < 4; j++)
		if (k[j] >= 0) cnt = j + 1, ik += a.nodes[k[j]].count == 0 ? 1 : 0;
	const int A = fld(a, F_AQ8, x), s = fld(a, F_SQ8, x);
	int run = A + ik, j2 = 0;
	uint32_t refs[4] = { Q8_EMPTY, Q8_EMPTY, Q8_EMPTY, Q8_EMPTY };
	for (int j = 0; j < cnt; j++)
	{
		const BuildNode& C = a.nodes[k[j]];
		if (C.count == 0)
		{
			fld(a, F_SQ8, k[j]) = A + j2, fld(a, F_AQ8, k[j]) = run;
			refs[j] = (uint32_t)(A + j2);
			run += fld(a, F_NQ8, k[j]) - 1;
			j2++;
		}
		else refs[j] = REF_LEAF | (uint32_t)fld(a, F_BASE, k[j]);
	}
	const BuildNode B = a.nodes[x];
	uint32_t ex[3], q[6][4] = {};
	for (int ax = 0; ax < 3; ax++)
	{
		float clo[4], chi[4];
		for (int c = 0; c < cnt; c++) { const BuildNode& K = a.nodes[k[c]]; clo[c] = K.lo[ax], chi[c] = K.hi[ax]; }
		ex[ax] = quantise_axis_dev(B.lo[ax], B.hi[ax], clo, chi, cnt, q[ax], q[3 + ax]);
	}
	auto pack = [](const uint32_t* v) { return v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24); };
	NodeQ8 o;
	o.origin = make_float4(B.lo[0], B.lo[1], B.lo[2], __uint_as_float(q8_step_bits(ex[2])));
	o.qa = make_uint4(pack(q[0]), pack(q[1]), pack(q[2]), pack(q[3]));
	o.qb = make_uint4(pack(q[4]), pack(q[5]), q8_step_bits(ex[0]), q8_step_bits(ex[1]));
	o.refs = make_uint4(refs[0], refs[1], refs[2], refs[3]);
	a.q8[s] = o;
	if (a.q8Map)
	{
		a.q8Map[5 * s] = x;
		for (int c = 0; c < 4; c++) a.q8Map[5 * s + 1 + c] = c < cnt ? k[c] : -1;
	}
}

__global__ void k_lay_counts(LayoutArgs a)
{
	if (threadIdx.x != 0) return;
	const bool interior = a.nodes[0].count == 0;
	a.info[1] = interior ? fld(a, F_N32, 0) : 0;
	a.info[2] = (interior && a.q8) ? fld(a, F_NQ8, 0) : 0;
}

#define LCK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { err = e_; goto done; } } while (0)

hipError_t layout_device(const BuildNode* dNodes, const uint32_t* dPrims, int nodeCount, const std::vector<int>& levelStart,
	const uint32_t* leafRefMap, const LayoutOut& out, int* countsOut, hipStream_t stream)
{
	hipError_t err = hipSuccess;
	int* f = nullptr;
	int4* kids = nullptr;
	int* info = nullptr;
	const int levels = (int)levelStart.size() - 1;
	if (nodeCount <= 0 || levels <= 0) return hipSuccess;
	LayoutArgs a;
	LCK(hipMalloc((void**)&f, sizeof(int) * (size_t)F_COUNT * nodeCount));
	LCK(hipMalloc((void**)&info, sizeof(int) * 4));
	if (out.q8) LCK(hipMalloc((void**)&kids, sizeof(int4) * (size_t)nodeCount));
	a.nodes = dNodes, a.prims = dPrims, a.leafRefMap = leafRefMap, a.f = f, a.kids = kids, a.nodeCount = nodeCount;
	a.f32 = out.f32, a.f32Kids = out.f32Kids, a.q8 = out.q8, a.q8Map = out.q8Map, a.order = out.order, a.last = out.last, a.info = info;
	k_lay_clear<<<(nodeCount + LBLOCK - 1) / LBLOCK, LBLOCK, 0, stream>>>(a);
	k_lay_root<<<1, 64, 0, stream>>>(a);
	for (int d = levels - 1; d >= 0; d--)
	{
		const int b = levelStart[d], e = levelStart[d + 1];
		k_lay_up1<<<(e - b + LBLOCK - 1) / LBLOCK, LBLOCK, 0, stream>>>(a, b, e);
	}
	for (int d = 0; d < levels; d++)
	{
		const int b = levelStart[d], e = levelStart[d + 1];
		k_lay_down1<<<(e - b + LBLOCK - 1) / LBLOCK, LBLOCK, 0, stream>>>(a, b, e);
	}
	if (out.q8)
	{
		for (int d = levels - 1; d >= 0; d--)
		{
			const int b = levelStart[d], e = levelStart[d + 1];
			k_lay_up2<<<(e - b + LBLOCK - 1) / LBLOCK, LBLOCK, 0, stream>>>(a, b, e);
		}
		for (int d = 0; d < levels; d++)
		{
			const int b = levelStart[d], e = levelStart[d + 1];
			k_lay_down2<<<(e - b + LBLOCK - 1) / LBLOCK, LBLOCK, 0, stream>>>(a, b, e);
		}
	}
	k_lay_counts<<<1, 64, 0, stream>>>(a);
	LCK(hipGetLastError());
	if (countsOut)
	{
		int h[4];
		LCK(hipMemcpyAsync(h, info, sizeof(h), hipMemcpyDeviceToHost, stream));
		LCK(hipStreamSynchronize(stream));
		countsOut[0] = h[1], countsOut[1] = h[2], countsOut[2] = h[0];
	}
done:
	if (f || kids || info)
	{
		const hipError_t se = hipStreamSynchronize(stream);
		if (err == hipSuccess) err = se;
	}
	if (f) (void)hipFree(f);
	if (kids) (void)hipFree(kids);
	if (info) (void)hipFree(info);
	return err;
}

__global__ void k_gather_tris(const float4* __restrict__ v, const uint32_t* __restrict__ order, const uint8_t* __restrict__ last, int n, float4* __restrict__ tris)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint32_t p = order[i];
	float4 a = v[3 * p];
	a.w = bitsf(p | (last[i] ? TRI_LAST : 0u));
	tris[3 * i] = a;
	tris[3 * i + 1] = v[3 * p + 1];
	tris[3 * i + 2] = v[3 * p + 2];
}

hipError_t gather_tris(const float4* dVerts, const uint32_t* dOrder, const uint8_t* dLast, int n, float4* dTris, hipStream_t stream)
{
	if (n <= 0) return hipSuccess;
	k_gather_tris<<<(n + 255) / 256, 256, 0, stream>>>(dVerts, dOrder, dLast, n, dTris);
	return hipGetLastError();
}

} // namespace lh
