/* oracle_trace.cpp - CPU oracle: binned-SAH BVH build and traversal.
 * TEST INFRASTRUCTURE (see oracle_math.h and oracle_trace.h for the contract). */
#include "oracle_trace.h"
#include <algorithm>

namespace orc {

static const int BINS = 16;
static const int MAXLEAF = 4;
static const int MAXDEGENERATELEAF = 255;

static inline float half_area(f3 lo, f3 hi)
{
	const f3 d = hi - lo;
	return d.x * d.y + d.y * d.z + d.z * d.x;
}

/* Breadth-first binned SAH over primitive boxes. The decision for a node
 * depends only on its primitive *set* (bins hold integer counts and min/max
 * bounds), so any builder following these rules, serial or parallel, produces
 * the same nodes in the same breadth-first order. Rules:
 *   centroid c = (lo + hi) * 0.5; per axis with centroid extent e > 0 and
 *   16/e finite, bin = min(15, int((c - cmin) * (16/e)));
 *   split cost after bin i = nL * halfArea(L) + nR * halfArea(R), first
 *   minimum over axis 0..2, i 0..14;
 *   n == 1 -> leaf; n <= 4 -> leaf unless A + cost < n * A (A = node
 *   halfArea); n > 4 -> split; no valid axis -> leaf if n <= 255, else the
 *   range is halved. */
static void build_generic(const std::vector<f3>& plo, const std::vector<f3>& phi,
	std::vector<BuildNode>& nodes, std::vector<uint>& prims, int& depth)
{
	const int n = (int)plo.size();
	std::vector<f3> cen(n);
	for (int i = 0; i < n; i++) cen[i] = (plo[i] + phi[i]) * 0.5f;
	prims.resize(n);
	for (int i = 0; i < n; i++) prims[i] = (uint)i;
	nodes.clear();
	depth = 0;
	if (n == 0) return;
	BuildNode root;
	root.lo = mk3(INFINITY, INFINITY, INFINITY), root.hi = mk3(-INFINITY, -INFINITY, -INFINITY);
	for (int i = 0; i < n; i++) root.lo = kmin3(root.lo, plo[i]), root.hi = kmax3(root.hi, phi[i]);
	root.leftFirst = 0, root.count = n;
	nodes.push_back(root);
	std::vector<int> level(1, 0), next;
	while (!level.empty())
	{
		depth++;
		next.clear();
		for (int ni : level)
		{
			BuildNode nd = nodes[ni];
			const int start = nd.leftFirst, cnt = nd.count;
			if (cnt == 1) continue; /* stays a leaf */
			f3 cmin = mk3(INFINITY, INFINITY, INFINITY), cmax = mk3(-INFINITY, -INFINITY, -INFINITY);
			for (int i = start; i < start + cnt; i++) cmin = kmin3(cmin, cen[prims[i]]), cmax = kmax3(cmax, cen[prims[i]]);
			int bestAxis = -1, bestSplit = -1;
			float bestCost = INFINITY, scales[3] = { 0, 0, 0 };
			f3 bestLlo = {}, bestLhi = {}, bestRlo = {}, bestRhi = {};
			for (int a = 0; a < 3; a++)
			{
				const float ext = fget(cmax, a) - fget(cmin, a);
				if (!(ext > 0.0f)) continue;
				const float scale = 16.0f / ext;
				if (!(scale < INFINITY)) continue;
				scales[a] = scale;
				int bc[BINS];
				f3 blo[BINS], bhi[BINS];
				for (int b = 0; b < BINS; b++) bc[b] = 0, blo[b] = mk3(INFINITY, INFINITY, INFINITY), bhi[b] = mk3(-INFINITY, -INFINITY, -INFINITY);
				const float ca = fget(cmin, a);
				for (int i = start; i < start + cnt; i++)
				{
					const uint p = prims[i];
					int b = (int)((fget(cen[p], a) - ca) * scale);
					if (b > BINS - 1) b = BINS - 1;
					bc[b]++, blo[b] = kmin3(blo[b], plo[p]), bhi[b] = kmax3(bhi[b], phi[p]);
				}
				int lc[BINS - 1], rc[BINS - 1];
				float la[BINS - 1], ra[BINS - 1];
				f3 llo[BINS - 1], lhi[BINS - 1], rlo[BINS - 1], rhi[BINS - 1];
				f3 lo = mk3(INFINITY, INFINITY, INFINITY), hi = mk3(-INFINITY, -INFINITY, -INFINITY);
				int c = 0;
				for (int i = 0; i < BINS - 1; i++)
				{
					c += bc[i], lo = kmin3(lo, blo[i]), hi = kmax3(hi, bhi[i]);
					lc[i] = c, la[i] = half_area(lo, hi), llo[i] = lo, lhi[i] = hi;
				}
				lo = mk3(INFINITY, INFINITY, INFINITY), hi = mk3(-INFINITY, -INFINITY, -INFINITY), c = 0;
				for (int i = BINS - 1; i > 0; i--)
				{
					c += bc[i], lo = kmin3(lo, blo[i]), hi = kmax3(hi, bhi[i]);
					rc[i - 1] = c, ra[i - 1] = half_area(lo, hi), rlo[i - 1] = lo, rhi[i - 1] = hi;
				}
				for (int i = 0; i < BINS - 1; i++)
				{
					if (lc[i] == 0 || rc[i] == 0) continue;
					const float cost = (float)lc[i] * la[i] + (float)rc[i] * ra[i];
					if (cost < bestCost)
						bestCost = cost, bestAxis = a, bestSplit = i,
						bestLlo = llo[i], bestLhi = lhi[i], bestRlo = rlo[i], bestRhi = rhi[i];
				}
			}
			int leftCount;
			if (bestAxis < 0)
			{
				if (cnt <= MAXDEGENERATELEAF) continue;
				leftCount = cnt / 2; /* degenerate: halve the current range */
				bestLlo = bestRlo = mk3(INFINITY, INFINITY, INFINITY), bestLhi = bestRhi = mk3(-INFINITY, -INFINITY, -INFINITY);
				for (int i = start; i < start + cnt; i++)
				{
					const uint p = prims[i];
					if (i < start + leftCount) bestLlo = kmin3(bestLlo, plo[p]), bestLhi = kmax3(bestLhi, phi[p]);
					else bestRlo = kmin3(bestRlo, plo[p]), bestRhi = kmax3(bestRhi, phi[p]);
				}
			}
			else
			{
				const float A = half_area(nd.lo, nd.hi);
				if (cnt <= MAXLEAF && !(A + bestCost < (float)cnt * A)) continue;
				const float ca = fget(cmin, bestAxis), scale = scales[bestAxis];
				uint* mid = std::partition(prims.data() + start, prims.data() + start + cnt, [&](uint p) {
					int b = (int)((fget(cen[p], bestAxis) - ca) * scale);
					if (b > BINS - 1) b = BINS - 1;
					return b <= bestSplit;
				});
				leftCount = (int)(mid - (prims.data() + start));
			}
			const int left = (int)nodes.size();
			BuildNode L, R;
			L.lo = bestLlo, L.hi = bestLhi, L.leftFirst = start, L.count = leftCount;
			R.lo = bestRlo, R.hi = bestRhi, R.leftFirst = start + leftCount, R.count = cnt - leftCount;
			nodes.push_back(L), nodes.push_back(R);
			nodes[ni].leftFirst = left, nodes[ni].count = 0;
			next.push_back(left), next.push_back(left + 1);
		}
		level.swap(next);
	}
}

void build_bvh(Mesh& m)
{
	const int n = (int)m.v.size() / 3;
	std::vector<f3> lo(n), hi(n);
	for (int i = 0; i < n; i++)
	{
		const f3 a = m.v[i * 3], b = m.v[i * 3 + 1], c = m.v[i * 3 + 2];
		lo[i] = kmin3(a, kmin3(b, c)), hi[i] = kmax3(a, kmax3(b, c));
	}
	build_generic(lo, hi, m.nodes, m.prims, m.depth);
}

/* mat4::Inverted, MESA cofactor expansion (lib/RenderSystem/common_types.h:559-606),
 * same term order; a singular matrix yields the identity. */
void mat4_inverse(const float* c, float* out)
{
	float inv[16];
	inv[0] = c[5] * c[10] * c[15] - c[5] * c[11] * c[14] - c[9] * c[6] * c[15] + c[9] * c[7] * c[14] + c[13] * c[6] * c[11] - c[13] * c[7] * c[10];
	inv[1] = -c[1] * c[10] * c[15] + c[1] * c[11] * c[14] + c[9] * c[2] * c[15] - c[9] * c[3] * c[14] - c[13] * c[2] * c[11] + c[13] * c[3] * c[10];
	inv[2] = c[1] * c[6] * c[15] - c[1] * c[7] * c[14] - c[5] * c[2] * c[15] + c[5] * c[3] * c[14] + c[13] * c[2] * c[7] - c[13] * c[3] * c[6];
	inv[3] = -c[1] * c[6] * c[11] + c[1] * c[7] * c[10] + c[5] * c[2] * c[11] - c[5] * c[3] * c[10] - c[9] * c[2] * c[7] + c[9] * c[3] * c[6];
	inv[4] = -c[4] * c[10] * c[15] + c[4] * c[11] * c[14] + c[8] * c[6] * c[15] - c[8] * c[7] * c[14] - c[12] * c[6] * c[11] + c[12] * c[7] * c[10];
	inv[5] = c[0] * c[10] * c[15] - c[0] * c[11] * c[14] - c[8] * c[2] * c[15] + c[8] * c[3] * c[14] + c[12] * c[2] * c[11] - c[12] * c[3] * c[10];
	inv[6] = -c[0] * c[6] * c[15] + c[0] * c[7] * c[14] + c[4] * c[2] * c[15] - c[4] * c[3] * c[14] - c[12] * c[2] * c[7] + c[12] * c[3] * c[6];
	inv[7] = c[0] * c[6] * c[11] - c[0] * c[7] * c[10] - c[4] * c[2] * c[11] + c[4] * c[3] * c[10] + c[8] * c[2] * c[7] - c[8] * c[3] * c[6];
	inv[8] = c[4] * c[9] * c[15] - c[4] * c[11] * c[13] - c[8] * c[5] * c[15] + c[8] * c[7] * c[13] + c[12] * c[5] * c[11] - c[12] * c[7] * c[9];
	inv[9] = -c[0] * c[9] * c[15] + c[0] * c[11] * c[13] + c[8] * c[1] * c[15] - c[8] * c[3] * c[13] - c[12] * c[1] * c[11] + c[12] * c[3] * c[9];
	inv[10] = c[0] * c[5] * c[15] - c[0] * c[7] * c[13] - c[4] * c[1] * c[15] + c[4] * c[3] * c[13] + c[12] * c[1] * c[7] - c[12] * c[3] * c[5];
	inv[11] = -c[0] * c[5] * c[11] + c[0] * c[7] * c[9] + c[4] * c[1] * c[11] - c[4] * c[3] * c[9] - c[8] * c[1] * c[7] + c[8] * c[3] * c[5];
	inv[12] = -c[4] * c[9] * c[14] + c[4] * c[10] * c[13] + c[8] * c[5] * c[14] - c[8] * c[6] * c[13] - c[12] * c[5] * c[10] + c[12] * c[6] * c[9];
	inv[13] = c[0] * c[9] * c[14] - c[0] * c[10] * c[13] - c[8] * c[1] * c[14] + c[8] * c[2] * c[13] + c[12] * c[1] * c[10] - c[12] * c[2] * c[9];
	inv[14] = -c[0] * c[5] * c[14] + c[0] * c[6] * c[13] + c[4] * c[1] * c[14] - c[4] * c[2] * c[13] - c[12] * c[1] * c[6] + c[12] * c[2] * c[5];
	inv[15] = c[0] * c[5] * c[10] - c[0] * c[6] * c[9] - c[4] * c[1] * c[10] + c[4] * c[2] * c[9] + c[8] * c[1] * c[6] - c[8] * c[2] * c[5];
	const float det = c[0] * inv[0] + c[1] * inv[4] + c[2] * inv[8] + c[3] * inv[12];
	for (int i = 0; i < 16; i++) out[i] = (i % 5 == 0) ? 1.0f : 0.0f;
	if (det != 0)
	{
		const float invdet = 1.0f / det;
		for (int i = 0; i < 16; i++) out[i] = inv[i] * invdet;
	}
}

static inline f3 xform_point(const float* m, f3 p)
{
	return mk3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
		m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
		m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
static inline f3 xform_vector(const float* m, f3 v)
{
	return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z,
		m[4] * v.x + m[5] * v.y + m[6] * v.z,
		m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

/* TLAS over instance world boxes (8 transformed corners of the mesh root box;
 * an identity instance keeps the root box). Same builder as the BLAS. */
void finalize_instances(Scene& s)
{
	const int n = (int)s.inst.size();
	std::vector<f3> lo(n), hi(n);
	for (int i = 0; i < n; i++)
	{
		Instance& I = s.inst[i];
		I.identity = true;
		for (int k = 0; k < 16; k++) if (I.T[k] != ((k % 5 == 0) ? 1.0f : 0.0f)) I.identity = false;
		mat4_inverse(I.T, I.inv);
		const Mesh& m = s.meshes[I.mesh];
		f3 blo = m.nodes.empty() ? mk3(0, 0, 0) : m.nodes[0].lo, bhi = m.nodes.empty() ? mk3(0, 0, 0) : m.nodes[0].hi;
		if (I.identity) { I.lo = blo, I.hi = bhi; }
		else
		{
			I.lo = mk3(INFINITY, INFINITY, INFINITY), I.hi = mk3(-INFINITY, -INFINITY, -INFINITY);
			for (int c = 0; c < 8; c++)
			{
				const f3 p = mk3((c & 1) ? bhi.x : blo.x, (c & 2) ? bhi.y : blo.y, (c & 4) ? bhi.z : blo.z);
				const f3 w = xform_point(I.T, p);
				I.lo = kmin3(I.lo, w), I.hi = kmax3(I.hi, w);
			}
		}
		lo[i] = I.lo, hi[i] = I.hi;
	}
	int depth;
	build_generic(lo, hi, s.tlas, s.tlasPrims, depth);
}

/* Conservative slab test: the far distance is widened by 2^-16 relative so a
 * box whose exact interval meets [tmin, tmax] is never rejected by rounding.
 * Box tests only prune; they never decide a hit. */
/* A direction component of zero gives an infinite reciprocal; a ray lying in
 * a face plane of the box then gives 0 * inf = NaN for that plane. Such a ray
 * stays inside the slab for all t, so the NaN bound becomes the opposite
 * infinity (the test stays conservative). */
static inline void slab_axis(float& t1, float& t2)
{
	if (t1 != t1 && t2 != t2) { t1 = -INFINITY, t2 = INFINITY; }
	else if (t1 != t1) t1 = -t2;
	else if (t2 != t2) t2 = -t1;
}
static inline bool slab(f3 lo, f3 hi, f3 O, f3 inv, float tmin, float tmax, float& tnear)
{
	float tx1 = (lo.x - O.x) * inv.x, tx2 = (hi.x - O.x) * inv.x;
	float ty1 = (lo.y - O.y) * inv.y, ty2 = (hi.y - O.y) * inv.y;
	float tz1 = (lo.z - O.z) * inv.z, tz2 = (hi.z - O.z) * inv.z;
	slab_axis(tx1, tx2), slab_axis(ty1, ty2), slab_axis(tz1, tz2);
	const float tn = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fmaxf(fminf(tz1, tz2), tmin));
	float tf = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fminf(fmaxf(tz1, tz2), tmax));
	tf = tf + fabsf(tf) * 1.52587890625e-5f;
	tnear = tn;
	return tn <= tf;
}

struct Best { float t; int inst, prim; float U, V, det; };

template <bool ANY>
static bool blas_trace(const Mesh& m, int instId, f3 O, f3 D, float tmin, Best& best, Counters* cnt)
{
	if (m.nodes.empty()) return false;
	const RayPre rp = ray_setup(O, D);
	const f3 inv = mk3(1.0f / D.x, 1.0f / D.y, 1.0f / D.z);
	int stack[128], sp = 0;
	int node = 0;
	float tn;
	if (!slab(m.nodes[0].lo, m.nodes[0].hi, O, inv, tmin, best.t, tn)) return false;
	while (true)
	{
		const BuildNode& nd = m.nodes[node];
		if (nd.count > 0)
		{
			if (cnt) cnt->leaves++;
			for (int i = nd.leftFirst; i < nd.leftFirst + nd.count; i++)
			{
				const uint p = m.prims[i];
				if (cnt) cnt->tris++;
				float t, U, V, det;
				if (!tri_intersect(rp, m.v[p * 3], m.v[p * 3 + 1], m.v[p * 3 + 2], t, U, V, det)) continue;
				if (!(t > tmin)) continue;
				if (ANY) { if (t < best.t) return true; continue; }
				if (better(t, instId, (int)p, best.t, best.inst, best.prim))
					best.t = t, best.inst = instId, best.prim = (int)p, best.U = U, best.V = V, best.det = det;
			}
			if (sp == 0) break;
			node = stack[--sp];
			continue;
		}
		if (cnt) cnt->nodes++;
		const int l = nd.leftFirst, r = l + 1;
		float tl, tr;
		const bool hl = slab(m.nodes[l].lo, m.nodes[l].hi, O, inv, tmin, best.t, tl);
		const bool hr = slab(m.nodes[r].lo, m.nodes[r].hi, O, inv, tmin, best.t, tr);
		if (hl && hr)
		{
			if (tr < tl) { stack[sp++] = l; node = r; }
			else { stack[sp++] = r; node = l; }
		}
		else if (hl) node = l;
		else if (hr) node = r;
		else
		{
			if (sp == 0) break;
			node = stack[--sp];
		}
	}
	return false;
}

template <bool ANY>
static bool scene_trace(const Scene& s, f3 O, float tmin, f3 D, float tmax, Best& best, Counters* cnt)
{
	best.t = tmax, best.inst = -1, best.prim = -1;
	if (s.tlas.empty()) return false;
	const f3 inv = mk3(1.0f / D.x, 1.0f / D.y, 1.0f / D.z);
	int stack[128], sp = 0, node = 0;
	float tn;
	if (!slab(s.tlas[0].lo, s.tlas[0].hi, O, inv, tmin, best.t, tn)) return false;
	while (true)
	{
		const BuildNode& nd = s.tlas[node];
		if (nd.count > 0)
		{
			for (int i = nd.leftFirst; i < nd.leftFirst + nd.count; i++)
			{
				const int id = (int)s.tlasPrims[i];
				const Instance& I = s.inst[id];
				f3 o = O, d = D;
				if (!I.identity) o = xform_point(I.inv, O), d = xform_vector(I.inv, D);
				if (blas_trace<ANY>(s.meshes[I.mesh], id, o, d, tmin, best, cnt)) return true;
			}
			if (sp == 0) break;
			node = stack[--sp];
			continue;
		}
		const int l = nd.leftFirst, r = l + 1;
		float tl, tr;
		const bool hl = slab(s.tlas[l].lo, s.tlas[l].hi, O, inv, tmin, best.t, tl);
		const bool hr = slab(s.tlas[r].lo, s.tlas[r].hi, O, inv, tmin, best.t, tr);
		if (hl && hr) { if (tr < tl) { stack[sp++] = l; node = r; } else { stack[sp++] = r; node = l; } }
		else if (hl) node = l;
		else if (hr) node = r;
		else { if (sp == 0) break; node = stack[--sp]; }
	}
	return false;
}

static Hit make_hit(const Best& b)
{
	Hit h;
	if (b.prim < 0) { h.uv = 0; h.inst = 0; h.prim = -1; h.t = b.t; return h; }
	h.uv = pack_uv(b.U, b.V, b.det), h.inst = b.inst, h.prim = b.prim, h.t = b.t;
	return h;
}

Hit brute_closest(const Scene& s, f3 O, float tmin, f3 D, float tmax);
bool brute_any(const Scene& s, f3 O, float tmin, f3 D, float tmax);

Hit trace_closest(const Scene& s, f3 O, float tmin, f3 D, float tmax, Counters* cnt)
{
	if (s.bruteForce) return brute_closest(s, O, tmin, D, tmax);
	Best b;
	scene_trace<false>(s, O, tmin, D, tmax, b, cnt);
	return make_hit(b);
}

bool trace_any(const Scene& s, f3 O, float tmin, f3 D, float tmax, Counters* cnt)
{
	if (s.bruteForce) return brute_any(s, O, tmin, D, tmax);
	Best b;
	return scene_trace<true>(s, O, tmin, D, tmax, b, cnt);
}

template <bool ANY>
static bool brute(const Scene& s, f3 O, float tmin, f3 D, float tmax, Best& b)
{
	b.t = tmax, b.inst = -1, b.prim = -1;
	for (int id = 0; id < (int)s.inst.size(); id++)
	{
		const Instance& I = s.inst[id];
		f3 o = O, d = D;
		if (!I.identity) o = xform_point(I.inv, O), d = xform_vector(I.inv, D);
		const RayPre rp = ray_setup(o, d);
		const Mesh& m = s.meshes[I.mesh];
		const int n = (int)m.v.size() / 3;
		for (int p = 0; p < n; p++)
		{
			float t, U, V, det;
			if (!tri_intersect(rp, m.v[p * 3], m.v[p * 3 + 1], m.v[p * 3 + 2], t, U, V, det)) continue;
			if (!(t > tmin)) continue;
			if (ANY) { if (t < tmax) return true; continue; }
			if (better(t, id, p, b.t, b.inst, b.prim)) b.t = t, b.inst = id, b.prim = p, b.U = U, b.V = V, b.det = det;
		}
	}
	return false;
}

Hit brute_closest(const Scene& s, f3 O, float tmin, f3 D, float tmax)
{
	Best b;
	brute<false>(s, O, tmin, D, tmax, b);
	return make_hit(b);
}

bool brute_any(const Scene& s, f3 O, float tmin, f3 D, float tmax)
{
	Best b;
	return brute<true>(s, O, tmin, D, tmax, b);
}

} // namespace orc
