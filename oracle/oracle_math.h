/* oracle_math.h - scalar/vector helpers of the CPU oracle.
 *
 * TEST INFRASTRUCTURE. The oracle is the checker for RenderCore_HIP: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it.
 * It never runs on the product path.
 *
 * Every expression below fixes its evaluation order (left to right, no fused
 * multiply-add: compiled with -ffp-contract=off) so that a float computed here
 * is bit-identical to the same quantity computed by the gfx950 kernels.
 */
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>

namespace orc {

typedef uint32_t uint;

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

static inline f3 mk3(float x, float y, float z) { f3 r = { x, y, z }; return r; }
static inline f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
static inline f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float length(f3 a) { return sqrtf(dot(a, a)); }
/* normalize = v * (1 / sqrt(v.v)); helper_math's rsqrtf is replaced by a correctly
 * rounded sqrt and divide so CPU and GPU agree bit for bit. */
static inline f3 normalize(f3 a) { const float inv = 1.0f / sqrtf(dot(a, a)); return a * inv; }
static inline float fget(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* Total order on floats used for every min/max of bounds: -0 sorts below +0,
 * which is also the order of the integer keys the GPU builder's atomics use. */
static inline int32_t fkey(float f) { int32_t i; memcpy(&i, &f, 4); return i >= 0 ? i : (i ^ 0x7fffffff); }
static inline float kmin(float a, float b) { return fkey(b) < fkey(a) ? b : a; }
static inline float kmax(float a, float b) { return fkey(b) > fkey(a) ? b : a; }
static inline f3 kmin3(f3 a, f3 b) { return mk3(kmin(a.x, b.x), kmin(a.y, b.y), kmin(a.z, b.z)); }
static inline f3 kmax3(f3 a, f3 b) { return mk3(kmax(a.x, b.x), kmax(a.y, b.y), kmax(a.z, b.z)); }

static inline uint f2u(float f) { uint u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint u) { float f; memcpy(&f, &u, 4); return f; }

/* RNG: Wang hash + Marsaglia xorshift32 (lib/CUDA/shared_kernel_code/tools_shared.h:60-62;
 * host variant lib/platform/system.cpp:46-52). */
static inline uint wang_hash(uint s) { s = (s ^ 61u) ^ (s >> 16); s *= 9u; s = s ^ (s >> 4); s *= 0x27d4eb2du; s = s ^ (s >> 15); return s; }
static inline uint xor32(uint& s) { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; }
static inline float rand_float(uint& s) { return (float)xor32(s) * 2.3283064365387e-10f; }

/* sin/cos: Cody & Waite (1980) three-part reduction by pi/2 (C1 = 201/128 and a
 * 12-bit C2 so k*C1 and k*C2 are exact for |k| < 4096), then the Taylor series
 * of sin to x^9 and cos to x^10 on [-pi/4, pi/4] in Horner form (truncation
 * error < 2e-9). Valid for |x| < 6000. */
static inline void cw_sincos(float x, float* s, float* c)
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
static inline float lh_atan_unit(float x)
{
	const bool big = x > 0.26794919f;
	const float r = big ? (x * 1.7320508f - 1.0f) / (x + 1.7320508f) : x;
	const float z = r * r;
	const float p = r + r * z * (-0.33333333f + z * (0.2f + z * (-0.14285714f + z * (0.11111111f + z * (-0.09090909f + z * 0.07692308f)))));
	return big ? p + 0.52359878f : p;
}
static inline float lh_atan2f(float y, float x)
{
	const float ax = fabsf(x), ay = fabsf(y);
	const float mx = ay > ax ? ay : ax, mn = ay > ax ? ax : ay;
	float a = mx == 0.0f ? 0.0f : lh_atan_unit(mn / mx);
	if (ay > ax) a = 1.5707964f - a;
	if (x < 0.0f) a = 3.1415927f - a;
	return y < 0.0f ? -a : a;
}
/* acos(x), |x| <= 1: atan2(sqrt((1 - x)(1 + x)), x). */
static inline float lh_acosf(float x) { return lh_atan2f(sqrtf((1.0f - x) * (1.0f + x)), x); }
/* log2(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)), ln(m) = 2 atanh(s),
 * s = (m - 1) / (m + 1), |s| <= 0.1716, series to s^9 (truncation < 2e-9),
 * its multiply-adds fused (fmaf, as the kernels' fma). */
static inline float lh_log2f(float x)
{
	if (!(x > 0.0f)) return x == 0.0f ? -HUGE_VALF : NAN;
	if (x == HUGE_VALF) return x;
	uint32_t b = f2u(x);
	int e = 0;
	if (b < 0x00800000u) { x = x * 8388608.0f; b = f2u(x); e = -23; }
	e += (int)(b >> 23) - 127;
	float m = u2f((b & 0x007fffffu) | 0x3f800000u);
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
static inline uint32_t f2u_sat(float x) { return x > 0.0f ? (x >= 4294967040.0f ? 0xffffffffu : (uint32_t)x) : 0u; }
/* float -> int32 saturating (NaN gives 0). */
static inline int32_t f2i_sat(float x) { return x >= 2147483520.0f ? 0x7fffffff : (x <= -2147483648.0f ? (int32_t)0x80000000 : (x == x ? (int32_t)x : 0)); }

/* exp(x) as the kernels compute it: k = round(x / ln2), Cody-Waite split
 * of ln2, Taylor series of exp(r) to r^7, scaled by 2^k; every multiply-add
 * fused (fmaf is the singly rounded IEEE operation, as the kernels' fma). */
static inline float cw_expf(float x)
{
	if (x < -87.0f) return 0.0f;
	if (x > 88.0f) return HUGE_VALF;
	const float k = floorf(fmaf(x, 1.44269504f, 0.5f));
	const float r = fmaf(k, 2.12194440e-4f, fmaf(-k, 0.693359375f, x));
	float p = fmaf(r, 1.9841270e-4f, 1.3888889e-3f);
	p = fmaf(p, r, 8.3333333e-3f);
	p = fmaf(p, r, 4.1666668e-2f);
	p = fmaf(p, r, 0.16666667f);
	p = fmaf(p, r, 0.5f);
	p = fmaf(p, r, 1.0f);
	p = fmaf(p, r, 1.0f);
	return p * u2f((uint)((int)k + 127) << 23);
}

/* Round-to-nearest-even float -> IEEE half bits and back (CUDA __float2half /
 * __half2float semantics), for the material colour fields. */
static inline uint16_t float_to_half(float f)
{
	uint x = f2u(f);
	const uint sign = (x >> 16) & 0x8000u;
	x &= 0x7fffffffu;
	if (x >= 0x7f800000u) return (uint16_t)(sign | (x > 0x7f800000u ? 0x7e00u : 0x7c00u));
	if (x >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to >= 65520: inf */
	if (x < 0x38800000u)
	{	/* subnormal half */
		if (x < 0x33000000u) return (uint16_t)sign;
		const uint e = x >> 23, m = (x & 0x7fffffu) | 0x800000u;
		const uint shift = 126u - e; /* 14..24 */
		uint h = m >> shift;
		const uint rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
		if (rem > half || (rem == half && (h & 1u))) h++;
		return (uint16_t)(sign | h);
	}
	uint h = ((x >> 13) - (112u << 10));
	const uint rem = x & 0x1fffu;
	if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
	return (uint16_t)(sign | h);
}
static inline float half_to_float(uint16_t h)
{
	const uint sign = (uint)(h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
	if (e == 0)
	{
		if (m == 0) return u2f(sign);
		float f = (float)m * 5.9604644775390625e-8f; /* 2^-24 */
		return sign ? -f : f;
	}
	if (e == 31) return u2f(sign | 0x7f800000u | (m << 13));
	return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

} // namespace orc
