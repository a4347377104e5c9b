/* oracle_layout.cpp - CPU restatement of RenderCore_HIP's BVH layout.
 * TEST INFRASTRUCTURE: loaded only by tests/. There is no reference
 * counterpart (OptiX lays its trees out inside its closed builder); this is
 * the checker for the device layout passes (bvh_layout.hip), which must
 * produce the same arrays from the same build tree.
 *
 * Converts a binned-SAH build tree (breadth-first BuildNode array, root 0,
 * children in pairs) into the traversal formats. Depth-first ordering puts a
 * node's subtree next to it and lays leaf triangles out in visiting order. The
 * 4-wide format collapses the binary tree (a node adopts the grandchildren of
 * its largest-area interior child until it has 4 children) and stores child
 * boxes as 8-bit offsets from the node's lower corner in power-of-two steps
 * per axis, rounded outward.
 */
#include <cmath>
#include <cstring>
#include <cstdint>
#include <vector>
#include <algorithm>

namespace orc_layout {

struct BuildNode { float lo[3]; int32_t leftFirst; float hi[3]; int32_t count; };
struct float4 { float x, y, z, w; };
struct uint4 { uint32_t x, y, z, w; };
static inline float4 make_float4(float x, float y, float z, float w) { return { x, y, z, w }; }
static inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return { x, y, z, w }; }
struct NodeF32 { float4 lxy, rxy, lrz; uint4 refs; };
struct NodeQ8 { float4 origin; uint4 qa, qb, refs; };
static const uint32_t REF_LEAF = 0x80000000u;
struct Layout
{
	std::vector<NodeF32> f32;
	std::vector<NodeQ8> q8;
	std::vector<uint32_t> order;
	std::vector<uint8_t> last;
	std::vector<int32_t> f32Kids, q8Map;
	uint32_t rootF32 = 0, rootQ8 = 0;
	int depthF32 = 0, depthQ8 = 0, stackF32 = 0, stackQ8 = 0;
};



static inline float bits_to_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t float_to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

struct Walker
{
	const std::vector<BuildNode>& n;
	const std::vector<uint32_t>& prims;
	Layout& out;
	std::vector<uint32_t> leafSlot; /* build node -> first slot (leaves) */
	Walker(const std::vector<BuildNode>& nodes, const std::vector<uint32_t>& p, Layout& o) : n(nodes), prims(p), out(o) {}

	/* Assign triangle slots to leaves in depth-first (left first) order. */
	void assign_leaves()
	{
		leafSlot.assign(n.size(), 0);
		std::vector<int> st;
		st.push_back(0);
		while (!st.empty())
		{
			const int x = st.back();
			st.pop_back();
			if (n[x].count > 0)
			{
				leafSlot[x] = (uint32_t)out.order.size();
				std::vector<uint32_t> ps(prims.begin() + n[x].leftFirst, prims.begin() + n[x].leftFirst + n[x].count);
				std::sort(ps.begin(), ps.end());
				for (size_t k = 0; k < ps.size(); k++) out.order.push_back(ps[k]), out.last.push_back(k + 1 == ps.size() ? 1 : 0);
				continue;
			}
			st.push_back(n[x].leftFirst + 1);
			st.push_back(n[x].leftFirst);
		}
	}

	int depth_of(int x) const
	{
		int best = 0;
		std::vector<std::pair<int, int>> st;
		st.push_back({ x, 1 });
		while (!st.empty())
		{
			auto [y, d] = st.back();
			st.pop_back();
			best = std::max(best, d);
			if (n[y].count == 0) st.push_back({ n[y].leftFirst, d + 1 }), st.push_back({ n[y].leftFirst + 1, d + 1 });
		}
		return best;
	}

	uint32_t ref_leaf(int x) const { return REF_LEAF | leafSlot[x]; }

	uint32_t emit_f32(int x)
	{
		if (n[x].count > 0) return ref_leaf(x);
		/* iterative preorder: reserve the slot, then fill children refs */
		struct Item { int node; uint32_t slot; };
		std::vector<Item> st;
		const uint32_t root = (uint32_t)out.f32.size();
		out.f32.push_back(NodeF32());
		out.f32Kids.push_back(0), out.f32Kids.push_back(0);
		st.push_back({ x, root });
		while (!st.empty())
		{
			Item it = st.back();
			st.pop_back();
			const BuildNode& L = n[n[it.node].leftFirst];
			const BuildNode& R = n[n[it.node].leftFirst + 1];
			NodeF32 nd;
			nd.lxy = make_float4(L.lo[0], L.hi[0], L.lo[1], L.hi[1]);
			nd.rxy = make_float4(R.lo[0], R.hi[0], R.lo[1], R.hi[1]);
			nd.lrz = make_float4(L.lo[2], L.hi[2], R.lo[2], R.hi[2]);
			uint32_t refs[2];
			int kids[2] = { n[it.node].leftFirst, n[it.node].leftFirst + 1 };
			Item pend[2];
			int np = 0;
			for (int c = 0; c < 2; c++)
			{
				if (n[kids[c]].count > 0) { refs[c] = ref_leaf(kids[c]); continue; }
				const uint32_t s = (uint32_t)out.f32.size();
				out.f32.push_back(NodeF32());
				out.f32Kids.push_back(0), out.f32Kids.push_back(0);
				refs[c] = s;
				pend[np++] = { kids[c], s };
			}
			nd.refs = make_uint4(refs[0], refs[1], 0, 0);
			out.f32[it.slot] = nd;
			out.f32Kids[2 * it.slot] = kids[0], out.f32Kids[2 * it.slot + 1] = kids[1];
			for (int k = np - 1; k >= 0; k--) st.push_back(pend[k]);
		}
		return root;
	}

	static float half_area(const BuildNode& b)
	{
		const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
		return dx * dy + dy * dz + dz * dx;
	}

	/* children of a 4-wide node rooted at binary node x */
	void wide_children(int x, std::vector<int>& kids) const
	{
		kids.clear();
		kids.push_back(n[x].leftFirst), kids.push_back(n[x].leftFirst + 1);
		while (kids.size() < 4)
		{
			int best = -1;
			float ba = -1;
			for (size_t k = 0; k < kids.size(); k++)
				if (n[kids[k]].count == 0 && half_area(n[kids[k]]) > ba) ba = half_area(n[kids[k]]), best = (int)k;
			if (best < 0) break;
			const int c = kids[best];
			kids[best] = n[c].leftFirst;
			kids.push_back(n[c].leftFirst + 1);
		}
	}

	static void quantise_axis(float lo, float hi, const float* clo, const float* chi, int cnt, uint8_t& exOut, uint8_t* qlo, uint8_t* qhi)
	{
		const float ext = hi - lo;
		int e = -126;
		if (ext > 0) e = std::max(-126, (int)std::ceil(std::log2((double)ext / 255.0)));
		for (;; e++)
		{
			const float s = bits_to_float((uint32_t)(e + 127) << 23);
			bool ok = true;
			for (int c = 0; c < cnt && ok; c++)
			{
				double ql = std::floor(((double)clo[c] - (double)lo) / (double)s);
				int a = (int)std::max(0.0, std::min(255.0, ql));
				while (a > 0 && lo + (float)a * s > clo[c]) a--;
				double qh = std::ceil(((double)chi[c] - (double)lo) / (double)s);
				int b = (int)std::max(0.0, std::min(255.0, qh));
				while (b < 255 && lo + (float)b * s < chi[c]) b++;
				if (lo + (float)b * s < chi[c] || lo + (float)a * s > clo[c]) { ok = false; break; }
				qlo[c] = (uint8_t)a, qhi[c] = (uint8_t)b;
			}
			if (ok) { exOut = (uint8_t)(e + 127); return; }
		}
	}

	uint32_t emit_q8(int x)
	{
		if (n[x].count > 0) return ref_leaf(x);
		struct Item { int node; uint32_t slot; };
		std::vector<Item> st;
		const uint32_t root = (uint32_t)out.q8.size();
		out.q8.push_back(NodeQ8());
		out.q8Map.resize(out.q8Map.size() + 5, -1);
		st.push_back({ x, root });
		std::vector<int> kids;
		while (!st.empty())
		{
			Item it = st.back();
			st.pop_back();
			wide_children(it.node, kids);
			const BuildNode& B = n[it.node];
			const int cnt = (int)kids.size();
			uint8_t ex[3], qlo[3][4] = {}, qhi[3][4] = {};
			for (int a = 0; a < 3; a++)
			{
				float clo[4], chi[4];
				for (int c = 0; c < cnt; c++) clo[c] = n[kids[c]].lo[a], chi[c] = n[kids[c]].hi[a];
				quantise_axis(B.lo[a], B.hi[a], clo, chi, cnt, ex[a], qlo[a], qhi[a]);
			}
			NodeQ8 nd;
			memset(&nd, 0, sizeof(nd));
			/* plane step 2^(e - 127) stored as the float with exponent field e */
			nd.origin = make_float4(B.lo[0], B.lo[1], B.lo[2], bits_to_float((uint32_t)ex[2] << 23));
			auto pack = [&](const uint8_t* q) { return (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24); };
			nd.qa = make_uint4(pack(qlo[0]), pack(qlo[1]), pack(qlo[2]), pack(qhi[0]));
			nd.qb = make_uint4(pack(qhi[1]), pack(qhi[2]), (uint32_t)ex[0] << 23, (uint32_t)ex[1] << 23);
			uint32_t refs[4] = { 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu };   /* unused slots */
			std::vector<Item> pend;
			for (int c = 0; c < cnt; c++)
			{
				if (n[kids[c]].count > 0) { refs[c] = ref_leaf(kids[c]); continue; }
				const uint32_t s = (uint32_t)out.q8.size();
				out.q8.push_back(NodeQ8());
				out.q8Map.resize(out.q8Map.size() + 5, -1);
				refs[c] = s;
				pend.push_back({ kids[c], s });
			}
			nd.refs = make_uint4(refs[0], refs[1], refs[2], refs[3]);
			out.q8[it.slot] = nd;
			out.q8Map[5 * it.slot] = it.node;
			for (int c = 0; c < cnt; c++) out.q8Map[5 * it.slot + 1 + c] = kids[c];
			for (int k = (int)pend.size() - 1; k >= 0; k--) st.push_back(pend[k]);
		}
		return root;
	}

	int wide_depth(int x)
	{
		int best = 0;
		std::vector<std::pair<int, int>> st;
		std::vector<int> kids;
		st.push_back({ x, 1 });
		while (!st.empty())
		{
			auto [y, d] = st.back();
			st.pop_back();
			best = std::max(best, d);
			if (n[y].count > 0) continue;
			wide_children(y, kids);
			for (int c : kids) if (n[c].count == 0) st.push_back({ c, d + 1 });
		}
		return best;
	}
};

void layout_bvh(const std::vector<BuildNode>& nodes, const std::vector<uint32_t>& prims, bool wantQ8, Layout& out)
{
	out = Layout();
	if (nodes.empty()) return;
	Walker w(nodes, prims, out);
	w.assign_leaves();
	out.rootF32 = w.emit_f32(0);
	out.depthF32 = w.depth_of(0);
	out.stackF32 = out.depthF32 + 1;
	if (wantQ8)
	{
		out.rootQ8 = w.emit_q8(0);
		out.depthQ8 = w.wide_depth(0);
		out.stackQ8 = 3 * out.depthQ8 + 1;
	}
}

} // namespace orc_layout

/* counts: [0] f32 nodes, [1] q8 nodes, [2] f32 root ref, [3] q8 root ref, [4] binary depth, [5] 4-wide depth.
 * Output arrays sized by the caller for (nodeCount - 1) / 2 nodes (q8MapOut: 5 ints per node) and n triangles. */
extern "C" __attribute__((visibility("default"))) void orc_bvh_layout(const void* nodes, int nodeCount, const uint32_t* prims, int n,
	void* f32Out, void* q8Out, int32_t* q8MapOut, uint32_t* orderOut, uint8_t* lastOut, int* counts)
{
	using namespace orc_layout;
	std::vector<BuildNode> nv((const BuildNode*)nodes, (const BuildNode*)nodes + nodeCount);
	std::vector<uint32_t> pv(prims, prims + n);
	Layout L;
	layout_bvh(nv, pv, true, L);
	memcpy(f32Out, L.f32.data(), L.f32.size() * sizeof(NodeF32));
	memcpy(q8Out, L.q8.data(), L.q8.size() * sizeof(NodeQ8));
	memcpy(q8MapOut, L.q8Map.data(), L.q8Map.size() * sizeof(int32_t));
	memcpy(orderOut, L.order.data(), L.order.size() * 4);
	memcpy(lastOut, L.last.data(), L.last.size());
	counts[0] = (int)L.f32.size(), counts[1] = (int)L.q8.size();
	counts[2] = (int)L.rootF32, counts[3] = (int)L.rootQ8, counts[4] = L.depthF32, counts[5] = L.depthQ8;
}
