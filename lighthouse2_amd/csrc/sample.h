/* sample.h - camera and direction sampling shared by the wave kernels. */
#pragma once
#include "lh_math.h"
#include "waves.h"

namespace lh {

/* RandomPointOnLens, 9-blade aperture (lib/RenderCore_PrimeRef/kernels/camera.h:22-33). */
__device__ __forceinline__ v3 point_on_lens(float r0, float r1, v3 pos, float aperture, v3 right, v3 up)
{
	const float PI = 3.14159265358979323846264f;
	const float blade = (float)(int)(r0 * 9);
	float r2 = (r0 - blade * (1.0f / 9.0f)) * 9.0f;
	float x1, y1, x2, y2;
	cw_sincos(blade * PI / 4.5f, &x1, &y1);
	cw_sincos((blade + 1.0f) * PI / 4.5f, &x2, &y2);
	if ((r1 + r2) > 1) r1 = 1.0f - r1, r2 = 1.0f - r2;
	const float xr = x1 * r1 + x2 * r2;
	const float yr = y1 * r1 + y2 * r2;
	return pos + aperture * (right * xr + up * yr);
}

/* Primary ray of job `job` (generateEyeRaysKernel, camera.h:39-68). */
__device__ __forceinline__ void eye_ray(const CamParams& c, int job, uint32_t R0, v3& O, v3& D)
{
	const uint32_t x = (uint32_t)job % (uint32_t)c.w;
	uint32_t y = (uint32_t)job / (uint32_t)c.w;
	y %= (uint32_t)c.h;
	uint32_t seed = wang_hash((uint32_t)job + R0);
	const float r0 = rand_float(seed), r1 = rand_float(seed);
	const float r2 = rand_float(seed), r3 = rand_float(seed);
	const v3 pos = mk3(c.pos[0], c.pos[1], c.pos[2]), p1 = mk3(c.p1[0], c.p1[1], c.p1[2]);
	const v3 right = mk3(c.p2[0], c.p2[1], c.p2[2]) - p1, up = mk3(c.p3[0], c.p3[1], c.p3[2]) - p1;
	const v3 posOnPixel = p1 + ((float)x + r0) * (right / (float)c.w) + ((float)y + r1) * (up / (float)c.h);
	const v3 posOnLens = point_on_lens(r2, r3, pos, c.aperture, right, up);
	O = posOnLens;
	D = normalize(posOnPixel - posOnLens);
}

/* SetupTangentSpace + Tangent2World, Duff et al. 2017 (common_functions.h:73-88). */
__device__ __forceinline__ v3 tangent2world(v3 V, v3 N)
{
	const float sign = copysignf(1.0f, N.z);
	const float a = -1.0f / (sign + N.z);
	const float b = N.x * N.y * a;
	const v3 B = mk3(1.0f + sign * N.x * N.x * a, sign * b, -sign * N.x);
	const v3 T = mk3(b, sign + N.y * N.y * a, -N.y);
	return V.x * T + V.y * B + V.z * N;
}

/* DiffuseReflectionCosWeighted (common_functions.h:104-110). */
__device__ __forceinline__ v3 cosine_sample(float r0, float r1)
{
	const float TWOPI = 6.28318530717958647692528f;
	const float term1 = TWOPI * r0, term2 = sqrtf(1 - r1);
	float s, c;
	cw_sincos(term1, &s, &c);
	return mk3(c * term2, s * term2, sqrtf(r1));
}

} // namespace lh
