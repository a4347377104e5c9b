/* oracle_trace.h - CPU oracle: watertight intersection, binned-SAH BVH build,
 * two-level closest/any-hit traversal and brute force.
 *
 * TEST INFRASTRUCTURE (see oracle_math.h). The reference delegates all of this
 * to OptiX Prime (lib/RenderCore_PrimeRef/core_mesh.cpp:53-64,
 * lib/RenderCore_PrimeRef/rendercore.cpp:509-607), a closed binary; the oracle
 * restates the *contract* of its queries:
 *   closest hit  RTP_QUERY_TYPE_CLOSEST: nearest t with tmin < t < tmax, back
 *                faces included, instance id = position in the instance list
 *                (optix_prime_declarations.h:85-86), ties broken by the smaller
 *                (instance, primitive) pair so the answer does not depend on
 *                traversal order;
 *   any hit      RTP_QUERY_TYPE_ANY: some t with tmin < t < tmax.
 * The intersection test is Woop, Benthin & Wald, "Watertight Ray/Triangle
 * Intersection", JCGT 2(1), 2013, with u,v the weights of vertex 0 and vertex 1
 * (the convention PrimeRef's shading assumes, material_shared.h:62-66).
 */
#pragma once
#include "oracle_math.h"
#include <vector>

namespace orc {

struct RayPre
{
	int kx, ky, kz;
	float Sx, Sy, Sz;
	f3 O;
};

/* Per-ray setup (Woop et al. 2013, section 3): kz = dominant axis of D
 * (first of equal magnitudes), kx/ky swapped when D[kz] < 0 to keep winding. */
static inline RayPre ray_setup(f3 O, f3 D)
{
	RayPre p;
	const float ax = fabsf(D.x), ay = fabsf(D.y), az = fabsf(D.z);
	int kz = 0; float m = ax;
	if (ay > m) { kz = 1; m = ay; }
	if (az > m) kz = 2;
	int kx = kz + 1; if (kx == 3) kx = 0;
	int ky = kx + 1; if (ky == 3) ky = 0;
	const float dz = fget(D, kz);
	if (dz < 0.0f) { const int t = kx; kx = ky; ky = t; }
	p.kx = kx, p.ky = ky, p.kz = kz;
	p.Sx = fget(D, kx) / dz;
	p.Sy = fget(D, ky) / dz;
	p.Sz = 1.0f / dz;
	p.O = O;
	return p;
}

/* Watertight test. Returns true when the ray's line crosses the triangle
 * (edges and vertices included, no culling) with a non-zero determinant and
 * writes t = T/det and the scaled barycentrics; the caller applies the t range. */
static inline bool tri_intersect(const RayPre& r, f3 v0, f3 v1, f3 v2, float& t, float& U, float& V, float& det)
{
	const f3 A = v0 - r.O, B = v1 - r.O, C = v2 - r.O;
	const float Akz = fget(A, r.kz), Bkz = fget(B, r.kz), Ckz = fget(C, r.kz);
	const float Ax = fget(A, r.kx) - r.Sx * Akz, Ay = fget(A, r.ky) - r.Sy * Akz;
	const float Bx = fget(B, r.kx) - r.Sx * Bkz, By = fget(B, r.ky) - r.Sy * Bkz;
	const float Cx = fget(C, r.kx) - r.Sx * Ckz, Cy = fget(C, r.ky) - r.Sy * Ckz;
	float u = Cx * By - Cy * Bx;
	float v = Ax * Cy - Ay * Cx;
	float w = Bx * Ay - By * Ax;
	if (u == 0.0f || v == 0.0f || w == 0.0f)
	{	/* exact fallback on the edges: products of floats are exact in double */
		u = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
		v = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
		w = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
	}
	if ((u < 0.0f || v < 0.0f || w < 0.0f) && (u > 0.0f || v > 0.0f || w > 0.0f)) return false;
	const float d = u + v + w;
	if (d == 0.0f) return false;
	const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
	const float T = u * Az + v * Bz + w * Cz;
	t = T / d;
	U = u, V = v, det = d;
	return true;
}

/* Hit record (16 bytes): packed u16/u16 barycentrics, instance, primitive, t. */
struct Hit { uint uv; int inst; int prim; float t; };

static inline uint pack_uv(float U, float V, float det)
{
	const float u = U / det, v = V / det;
	return (uint)(65535.0f * u) + ((uint)(65535.0f * v) << 16);
}

/* Candidate acceptance with the order-independent tie rule. */
static inline bool better(float t, int inst, int prim, float bt, int binst, int bprim)
{
	if (t < bt) return true;
	if (t > bt) return false;
	return inst < binst || (inst == binst && prim < bprim);
}

/* ---- binned SAH build (16 bins, leaf <= 4) --------------------------------- */

struct BuildNode { f3 lo; int leftFirst; f3 hi; int count; }; /* count 0 = interior */

struct Mesh
{
	std::vector<f3> v;              /* 3 per triangle (vertex soup, primId = index / 3) */
	std::vector<BuildNode> nodes;   /* breadth-first order, root 0, children in pairs */
	std::vector<uint> prims;        /* leaf primitive lists */
	int depth = 0;
};

void build_bvh(Mesh& m);

struct Instance
{
	int mesh;
	float T[16];     /* object -> world, row-major */
	float inv[16];   /* MESA cofactor inverse (common_types.h:559-606) */
	bool identity;
	f3 lo, hi;       /* world bounds */
};

struct Counters { uint64_t nodes, tris, leaves; };

struct Scene
{
	std::vector<Mesh> meshes;
	std::vector<Instance> inst;
	std::vector<BuildNode> tlas;
	std::vector<uint> tlasPrims;
	bool bruteForce = false;   /* trace_closest / trace_any test every triangle (checks the traversal) */
};

void mat4_inverse(const float* c, float* out);
void finalize_instances(Scene& s);
Hit trace_closest(const Scene& s, f3 O, float tmin, f3 D, float tmax, Counters* cnt);
bool trace_any(const Scene& s, f3 O, float tmin, f3 D, float tmax, Counters* cnt);
Hit brute_closest(const Scene& s, f3 O, float tmin, f3 D, float tmax);
bool brute_any(const Scene& s, f3 O, float tmin, f3 D, float tmax);

} // namespace orc
