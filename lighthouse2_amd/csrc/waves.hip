/* waves.hip - ray-wave kernels of the incoherent traversal workload.
 *
 * k_primary follows generateEyeRaysKernel (lib/RenderCore_PrimeRef/kernels/camera.h:39-68):
 * jobIndex -> pixel, seed WangHash(jobIndex + R0), four xorshift floats for the
 * pixel jitter and the 9-blade lens sample. k_bounce turns a closest-hit wave into
 * the next diffuse wave: cosine-weighted direction (common_functions.h:104-110)
 * about the geometric normal flipped against the incoming ray, seed
 * WangHash(pathId * 17 + R0) as in shadeKernel (pathtracer.h:116); misses die and
 * survivors are compacted through an atomic counter.
 */
#include "waves.h"
#include "sample.h"

namespace lh {

__global__ void k_primary(CamParams c, uint32_t R0, float tmin, int n, float4* __restrict__ rays, uint32_t* __restrict__ pathId)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n) return;
	v3 O, D;
	eye_ray(c, j, R0, O, D);
	rays[2 * j] = make_float4(O.x, O.y, O.z, tmin);
	rays[2 * j + 1] = make_float4(D.x, D.y, D.z, 1e34f);
	if (pathId) pathId[j] = (uint32_t)j;
}

__global__ void k_bounce(const ShadeInst* __restrict__ inst, const float4* __restrict__ raysIn, const float4* __restrict__ hits,
	const uint32_t* __restrict__ pathIn, int n, uint32_t R0, float tmin, float4* __restrict__ raysOut, uint32_t* __restrict__ pathOut,
	uint32_t* __restrict__ counter)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n) return;
	const float4 h = hits[j];
	const uint32_t prim = fbits(h.z);
	if (prim == 0xffffffffu) return;
	const ShadeInst I = inst[fbits(h.y)];
	const v3 v0 = xyz(I.verts[3 * prim]), v1 = xyz(I.verts[3 * prim + 1]), v2 = xyz(I.verts[3 * prim + 2]);
	v3 N = cross(v1 - v0, v2 - v0);
	N = normalize(N.x * mk3(I.inv0.x, I.inv0.y, I.inv0.z) + N.y * mk3(I.inv1.x, I.inv1.y, I.inv1.z) + N.z * mk3(I.inv2.x, I.inv2.y, I.inv2.z));
	const float4 o4 = raysIn[2 * j], d4 = raysIn[2 * j + 1];
	const v3 O = xyz(o4), D = xyz(d4);
	if (dot(N, D) > 0) N = N * -1.0f;
	const v3 P = O + h.w * D;
	const uint32_t pid = pathIn[j];
	uint32_t seed = wang_hash(pid * 17u + R0);
	const float r0 = rand_float(seed), r1 = rand_float(seed);
	const v3 R = normalize(tangent2world(cosine_sample(r0, r1), N));
	const uint32_t k = atomicAdd(counter, 1u);
	raysOut[2 * k] = make_float4(P.x, P.y, P.z, tmin);
	raysOut[2 * k + 1] = make_float4(R.x, R.y, R.z, 1e34f);
	pathOut[k] = pid;
}

hipError_t launch_primary(const CamParams& c, uint32_t R0, float tmin, float4* rays, uint32_t* pathId, hipStream_t s)
{
	const int n = c.w * c.h;
	if (n <= 0) return hipSuccess;
	k_primary<<<(n + 255) / 256, 256, 0, s>>>(c, R0, tmin, n, rays, pathId);
	return hipGetLastError();
}

hipError_t launch_bounce(const ShadeInst* inst, const float4* raysIn, const float4* hits, const uint32_t* pathIn, int n,
	uint32_t R0, float tmin, float4* raysOut, uint32_t* pathOut, uint32_t* counter, hipStream_t s)
{
	hipError_t e = hipMemsetAsync(counter, 0, sizeof(uint32_t), s);
	if (e != hipSuccess || n <= 0) return e;
	k_bounce<<<(n + 255) / 256, 256, 0, s>>>(inst, raysIn, hits, pathIn, n, R0, tmin, raysOut, pathOut, counter);
	return hipGetLastError();
}

} // namespace lh
