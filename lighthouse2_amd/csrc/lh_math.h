/* lh_math.h - vector math, RNG and transcendental helpers used by the gfx950
 * kernels and the host side of RenderCore_HIP.
 *
 * All code is compiled with -ffp-contract=off: every expression is evaluated
 * left to right with separately rounded operations, and division and sqrt are
 * IEEE correctly rounded (hipcc's default for gfx950). Together with the
 * self-written sin/cos below this makes per-path arithmetic reproducible on
 * any IEEE-754 binary32 machine.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LH_HD __host__ __device__ __forceinline__

namespace lh {

struct v3 { float x, y, z; };

LH_HD v3 mk3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
LH_HD v3 operator+(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
LH_HD v3 operator-(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
LH_HD v3 operator*(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
LH_HD v3 operator*(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
LH_HD v3 operator*(float s, v3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
LH_HD v3 operator/(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
LH_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LH_HD v3 cross(v3 a, v3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
LH_HD float length(v3 a) { return sqrtf(dot(a, a)); }
LH_HD v3 normalize(v3 a) { const float inv = 1.0f / sqrtf(dot(a, a)); return a * inv; }
LH_HD float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
LH_HD v3 xyz(float4 a) { return mk3(a.x, a.y, a.z); }

LH_HD uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
LH_HD float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }

/* Order-preserving integer key of a float (-0 below +0); min/max of bounds
 * use it everywhere so that atomics and serial code agree. */
LH_HD int32_t fkey(float f) { const int32_t i = __builtin_bit_cast(int32_t, f); return i >= 0 ? i : (i ^ 0x7fffffff); }
LH_HD float keyf(int32_t k) { return __builtin_bit_cast(float, k >= 0 ? k : (k ^ 0x7fffffff)); }
LH_HD float kmin(float a, float b) { return fkey(b) < fkey(a) ? b : a; }
LH_HD float kmax(float a, float b) { return fkey(b) > fkey(a) ? b : a; }
LH_HD v3 kmin3(v3 a, v3 b) { return mk3(kmin(a.x, b.x), kmin(a.y, b.y), kmin(a.z, b.z)); }
LH_HD v3 kmax3(v3 a, v3 b) { return mk3(kmax(a.x, b.x), kmax(a.y, b.y), kmax(a.z, b.z)); }

/* Wang hash + Marsaglia xorshift32 (lib/CUDA/shared_kernel_code/tools_shared.h:60-62). */
LH_HD uint32_t wang_hash(uint32_t s) { s = (s ^ 61u) ^ (s >> 16); s *= 9u; s = s ^ (s >> 4); s *= 0x27d4eb2du; s = s ^ (s >> 15); return s; }
LH_HD uint32_t xor32(uint32_t& s) { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; }
LH_HD float rand_float(uint32_t& s) { return (float)xor32(s) * 2.3283064365387e-10f; }

/* sin and cos of x, |x| < 6000: Cody-Waite reduction by pi/2 with a three-part
 * constant (201/128, a 12-bit second part, a full third part: k*C1 and k*C2
 * are exact), then Taylor series to x^9 / x^10 on [-pi/4, pi/4] in Horner form. */
LH_HD void cw_sincos(float x, float* s, float* c)
{
	const float k = floorf(x * 0.63661975f + 0.5f);
	const float r = ((x - k * 1.5703125f) - k * 4.837512969970703e-4f) - k * 7.549790126404332e-8f;
	const float r2 = r * r;
	const float sr = r + r * r2 * (-0.16666667f + r2 * (8.3333333e-3f + r2 * (-1.9841270e-4f + r2 * 2.7557319e-6f)));
	const float cr = 1.0f + r2 * (-0.5f + r2 * (4.1666668e-2f + r2 * (-1.3888889e-3f + r2 * (2.4801587e-5f + r2 * -2.7557319e-7f))));
	const int q = ((int)k) & 3;
	if (q == 0) { *s = sr; *c = cr; }
	else if (q == 1) { *s = cr; *c = -sr; }
	else if (q == 2) { *s = -sr; *c = -cr; }
	else { *s = -cr; *c = sr; }
}

/* atan2(y, x): the octant reduction atan(min/max) (swap and reflect), then
 * atan(x) = pi/6 + atan((sqrt3 x - 1) / (sqrt3 + x)) for x > tan(pi/12), and
 * the Taylor series of atan to x^13 on |x| <= tan(pi/12) (truncation error
 * < 3e-10). atan2(0, 0) = 0; a negative zero y gives +0. */
LH_HD float lh_atan_unit(float x)
{
	const bool big = x > 0.26794919f;
	const float r = big ? (x * 1.7320508f - 1.0f) / (x + 1.7320508f) : x;
	const float z = r * r;
	const float p = r + r * z * (-0.33333333f + z * (0.2f + z * (-0.14285714f + z * (0.11111111f + z * (-0.09090909f + z * 0.07692308f)))));
	return big ? p + 0.52359878f : p;
}
LH_HD float lh_atan2f(float y, float x)
{
	const float ax = fabsf(x), ay = fabsf(y);
	const float mx = ay > ax ? ay : ax, mn = ay > ax ? ax : ay;
	float a = mx == 0.0f ? 0.0f : lh_atan_unit(mn / mx);
	if (ay > ax) a = 1.5707964f - a;
	if (x < 0.0f) a = 3.1415927f - a;
	return y < 0.0f ? -a : a;
}
/* acos(x), |x| <= 1: atan2(sqrt((1 - x)(1 + x)), x). */
LH_HD float lh_acosf(float x) { return lh_atan2f(sqrtf((1.0f - x) * (1.0f + x)), x); }
/* log2(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)), ln(m) = 2 atanh(s),
 * s = (m - 1) / (m + 1), |s| <= 0.1716, series to s^9 (truncation < 2e-9),
 * its multiply-adds fused. */
LH_HD float lh_log2f(float x)
{
	if (!(x > 0.0f)) return x == 0.0f ? -__builtin_huge_valf() : __builtin_nanf("");
	if (x == __builtin_huge_valf()) return x;
	uint32_t b = fbits(x);
	int e = 0;
	if (b < 0x00800000u) { x = x * 8388608.0f; b = fbits(x); e = -23; }
	e += (int)(b >> 23) - 127;
	float m = bitsf((b & 0x007fffffu) | 0x3f800000u);
	if (m > 1.4142135f) { m = m * 0.5f; e += 1; }
	const float s = (m - 1.0f) / (m + 1.0f), z = s * s;
	float q = fmaf(z, 0.11111111f, 0.14285714f);
	q = fmaf(q, z, 0.2f);
	q = fmaf(q, z, 0.33333333f);
	q = fmaf(q, z, 1.0f);
	return fmaf(2.0f * s * q, 1.4426950f, (float)e);
}
/* float -> uint32 with the saturation of the GPU's conversion (NaN and
 * negatives give 0). */
LH_HD uint32_t f2u_sat(float x) { return x > 0.0f ? (x >= 4294967040.0f ? 0xffffffffu : (uint32_t)x) : 0u; }
/* float -> int32 saturating (NaN gives 0). */
LH_HD int32_t f2i_sat(float x) { return x >= 2147483520.0f ? 0x7fffffff : (x <= -2147483648.0f ? (int32_t)0x80000000 : (x == x ? (int32_t)x : 0)); }

/* exp(x): k = round(x / ln2), Cody-Waite split of ln2 (0.693359375 exact in
 * 10 bits), Taylor series of exp(r) to r^7 on |r| <= ln2/2, scaled by 2^k.
 * Every multiply-add is one fused (singly rounded) fma. */
LH_HD float cw_expf(float x)
{
	if (x < -87.0f) return 0.0f;
	if (x > 88.0f) return __builtin_huge_valf();
	const float k = floorf(fmaf(x, 1.44269504f, 0.5f));
	const float r = fmaf(k, 2.12194440e-4f, fmaf(-k, 0.693359375f, x));
	float p = fmaf(r, 1.9841270e-4f, 1.3888889e-3f);
	p = fmaf(p, r, 8.3333333e-3f);
	p = fmaf(p, r, 4.1666668e-2f);
	p = fmaf(p, r, 0.16666667f);
	p = fmaf(p, r, 0.5f);
	p = fmaf(p, r, 1.0f);
	p = fmaf(p, r, 1.0f);
	return p * bitsf((uint32_t)((int)k + 127) << 23);
}

/* IEEE half -> float (exact). */
LH_HD float half_to_float(uint32_t h)
{
	const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
	if (e == 0)
	{
		if (m == 0) return bitsf(sign);
		const float f = (float)m * 5.9604644775390625e-8f;
		return sign ? -f : f;
	}
	if (e == 31) return bitsf(sign | 0x7f800000u | (m << 13));
	return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

/* ---- watertight ray/triangle test --------------------------------------------
 * Woop, Benthin & Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1),
 * 2013: shear the triangle into the ray's frame (kz = dominant axis, kx/ky
 * swapped for D[kz] < 0), 2D edge functions with an exact double fallback on
 * zero, no face culling. u, v are the weights of vertex 0 and vertex 1. */
struct RayPre { int kx, ky, kz; float Sx, Sy, Sz; };

LH_HD RayPre ray_setup(v3 D)
{
	RayPre p;
	const float ax = fabsf(D.x), ay = fabsf(D.y), az = fabsf(D.z);
	int kz = 0; float m = ax;
	if (ay > m) { kz = 1; m = ay; }
	if (az > m) kz = 2;
	int kx = kz + 1; if (kx == 3) kx = 0;
	int ky = kx + 1; if (ky == 3) ky = 0;
	const float dz = comp(D, kz);
	if (dz < 0.0f) { const int t = kx; kx = ky; ky = t; }
	p.kx = kx, p.ky = ky, p.kz = kz;
	p.Sx = comp(D, kx) / dz;
	p.Sy = comp(D, ky) / dz;
	p.Sz = 1.0f / dz;
	return p;
}

/* Returns false for a miss; else t = T/det and the scaled U, V, det. */
LH_HD bool tri_intersect(const RayPre& r, v3 O, v3 v0, v3 v1, v3 v2, float& t, float& U, float& V, float& det)
{
	const v3 A = v0 - O, B = v1 - O, C = v2 - O;
	const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
	const float Ax = comp(A, r.kx) - r.Sx * Akz, Ay = comp(A, r.ky) - r.Sy * Akz;
	const float Bx = comp(B, r.kx) - r.Sx * Bkz, By = comp(B, r.ky) - r.Sy * Bkz;
	const float Cx = comp(C, r.kx) - r.Sx * Ckz, Cy = comp(C, r.ky) - r.Sy * Ckz;
	float u = Cx * By - Cy * Bx;
	float v = Ax * Cy - Ay * Cx;
	float w = Bx * Ay - By * Ax;
	if (u == 0.0f || v == 0.0f || w == 0.0f)
	{
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

LH_HD uint32_t pack_uv(float U, float V, float det)
{
	const float u = U / det, v = V / det;
	return (uint32_t)(65535.0f * u) + ((uint32_t)(65535.0f * v) << 16);
}

/* Closest-hit order: smaller t, then smaller (instance, primitive). */
LH_HD bool better_hit(float t, int inst, int prim, float bt, int binst, int bprim)
{
	if (t < bt) return true;
	if (t > bt) return false;
	return inst < binst || (inst == binst && prim < bprim);
}

} // namespace lh
