/* bvh.h - acceleration-structure formats of RenderCore_HIP.
 *
 * Replaces the closed OptiX Prime / OptiX 7 acceleration structures
 * (lib/RenderCore_PrimeRef/core_mesh.cpp:53-64, rendercore.cpp:509-529;
 * lib/rendercore_optix7/core_mesh.cpp:67-129, rendercore.cpp:767-797) with a
 * two-level BVH: one BLAS per mesh over the SetGeometry vertex soup (primitive
 * id = soup index / 3) and a TLAS over instance boxes (instance id = position in
 * the SetInstance list).
 *
 * Build format (GPU builder output): breadth-first BuildNode array, root 0,
 * children allocated in pairs. Traversal formats (both 64 bytes per node):
 *   NodeF32  binary node holding both child boxes in fp32;
 *   NodeQ8   4-wide node, child boxes quantised to 8 bits per plane against
 *            the node box (lo rounded down, hi rounded up: conservative).
 * Child refs: bit 31 = leaf, low 30 bits = node index or first leaf triangle;
 * an unused q8 child slot holds Q8_EMPTY.
 * Leaf triangles: 3 float4 (vertex 0..2); v0.w = primitive id | bit 31 on the
 * last triangle of a leaf.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>

namespace lh {

struct BuildNode { float lo[3]; int32_t leftFirst; float hi[3]; int32_t count; }; /* count 0: interior */

struct NodeF32
{
	float4 lxy;   /* left  lo.x, hi.x, lo.y, hi.y */
	float4 rxy;   /* right lo.x, hi.x, lo.y, hi.y */
	float4 lrz;   /* left lo.z, hi.z, right lo.z, hi.z */
	uint4 refs;   /* x = left ref, y = right ref */
};

struct NodeQ8
{
	float4 origin;    /* xyz = quantisation origin; w = step of z (the float with exponent field ez) */
	uint4 qa;         /* lo.x[4], lo.y[4], lo.z[4], hi.x[4] (one byte per child) */
	uint4 qb;         /* hi.y[4], hi.z[4], step of x, step of y (float bits, exponent fields ex, ey) */
	uint4 refs;       /* child refs, Q8_EMPTY for unused slots */
};

static const uint32_t REF_LEAF = 0x80000000u;
static const uint32_t REF_TLAS = 0x40000000u;
static const uint32_t REF_MASK = 0x3fffffffu;
static const uint32_t TRI_LAST = 0x80000000u;
static const uint32_t Q8_EMPTY = 0xffffffffu;
/* A q8 plane lies at origin + q * step with step = 2^(e - 127) stored as the
 * float whose exponent field is e (e = 0 gives 0, e = 255 infinity). */
__host__ __device__ inline uint32_t q8_step_bits(uint32_t e) { return e << 23; }

/* Per-instance record for the two-level traversal (64 bytes). */
struct InstanceRec
{
	float4 inv0, inv1, inv2;   /* rows of the inverse transform */
	const void* nodes;         /* BLAS node array (NodeF32 or NodeQ8) */
	const float4* tris;        /* BLAS leaf triangles */
	uint32_t root;             /* BLAS root ref */
	uint32_t flags;            /* bit 0: identity transform */
};

/* Result of a GPU build: device arrays (owned by the caller). */
struct DeviceBuild
{
	BuildNode* nodes = nullptr;
	uint32_t* prims = nullptr;
	int nodeCount = 0, primCount = 0, depth = 0;
	std::vector<int> levelStart;   /* build ids of depth d: [levelStart[d], levelStart[d + 1]) */
};

/* Binned SAH (16 bins) over primitive boxes lo[i].xyz / hi[i].xyz on the GPU.
 * maxLeaf: largest SAH leaf; maxDegenerateLeaf: largest leaf when all
 * centroids coincide (beyond it the range is halved). Returns hipSuccess or an error. */
hipError_t gpu_build(const float4* dLo, const float4* dHi, int n, int maxLeaf, int maxDegenerateLeaf,
	DeviceBuild& out, hipStream_t stream);
void free_build(DeviceBuild& b);

/* Refit on the device, topology unchanged. Boxes of the build nodes
 * first .. first+count-1 (one tree level: leaves take the min / max of their
 * triangles' vertices, interior nodes the union of their two children, which
 * must already be refitted); then the laid-out f32 / q8 nodes are rewritten
 * from the refitted build boxes through the layout maps (refs unchanged). */
hipError_t refit_level(BuildNode* dNodes, const uint32_t* dPrims, const float4* dVerts, int first, int count, hipStream_t stream);
hipError_t refit_layout(const BuildNode* dNodes, NodeF32* dF32, const int32_t* dF32Kids, int nF32, NodeQ8* dQ8, const int32_t* dQ8Map, int nQ8, hipStream_t stream);

/* Triangle soup (3 float4 per triangle) -> primitive boxes. */
hipError_t tri_bounds(const float4* dVerts, int n, float4* dLo, float4* dHi, hipStream_t stream);

/* Device layout of a build tree into the traversal formats (bvh_layout.hip).
 * Arrays are caller-owned device memory sized for the worst case: a build
 * tree of N nodes has (N - 1) / 2 interior nodes, so at most that many f32
 * and q8 nodes; f32Kids holds 2 and q8Map 5 ints per node; order / last one
 * entry per primitive. q8 (with q8Map) null: binary layout only. f32Kids,
 * q8Map null: no refit maps. */
struct LayoutOut
{
	NodeF32* f32 = nullptr;
	int32_t* f32Kids = nullptr;       /* children (build ids) of f32 node i: [2i], [2i+1] */
	NodeQ8* q8 = nullptr;
	int32_t* q8Map = nullptr;         /* [5i] = build node of q8 node i, [5i+1..5i+4] = its children (-1 past the last) */
	uint32_t* order = nullptr;        /* primitive in leaf triangle slot i */
	uint8_t* last = nullptr;          /* 1 on the final slot of each leaf */
};
/* levelStart[d]..levelStart[d+1]: the build ids of depth d (gpu_build's breadth-first
 * order). leafRefMap non-null (top level, one primitive per leaf): leaf refs are
 * REF_LEAF | leafRefMap[primitive] and no triangle order is written.
 * countsOut (host, may be null; the call then waits for the stream): [0] f32
 * nodes, [1] q8 nodes, [2] q8 tree depth. */
hipError_t layout_device(const BuildNode* dNodes, const uint32_t* dPrims, int nodeCount, const std::vector<int>& levelStart,
	const uint32_t* leafRefMap, const LayoutOut& out, int* countsOut, hipStream_t stream);

/* Gathers leaf triangles: tris[3*i+k] = verts[3*order[i]+k], v0.w = prim | last bit. */
hipError_t gather_tris(const float4* dVerts, const uint32_t* dOrder, const uint8_t* dLast, int n, float4* dTris, hipStream_t stream);

} // namespace lh
