/* svgf.hip - SVGF on gfx950 (see svgf.h).
 *
 * Restates lib/CUDA/shared_kernel_code/finalize_shared.h:169-484 with the
 * packers of tools_shared.h:113-136,162-183,282-304 and the history reads of
 * sampling_shared.h:111-115,136-220, in the reference's operation order.
 * Two library calls are replaced by functions of our own so that the CPU
 * oracle reproduces every value bit for bit: powf(x, 128) is x squared seven
 * times, and __expf is cw_expf (lh_math.h). One pixel per thread, 8x8 blocks.
 */
#include "svgf.h"
#include "lh_math.h"
#include "../../include/lh2hip.h"

namespace lh {

struct f2 { float x, y; };

__device__ __forceinline__ v3 unpack_normal2(uint32_t p)
{
	const uint32_t x = (p >> 2u) & 1023u, y = (p >> 12u) & 1023u, z = p >> 22u;
	return mk3(fmaf((float)x, 1.0f / 511.0f, -1.0f), fmaf((float)y, 1.0f / 511.0f, -1.0f), fmaf((float)z, 1.0f / 511.0f, -1.0f));
}
__device__ __forceinline__ v3 rgb32_to_hdr(uint32_t c)
{
	return mk3((float)(c >> 22) * (1.0f / 1023.0f), (float)((c >> 11) & 2047u) * (1.0f / 2047.0f), (float)(c & 2047u) * (1.0f / 2047.0f));
}
__device__ __forceinline__ v3 rgb32_to_hdr_min1(uint32_t c)
{
	const uint32_t r = c >> 22, g = (c >> 11) & 2047u, b = c & 2047u;
	return mk3((float)(r > 1u ? r : 1u) * (1.0f / 1023.0f), (float)(g > 1u ? g : 1u) * (1.0f / 2047.0f), (float)(b > 1u ? b : 1u) * (1.0f / 2047.0f));
}
__device__ __forceinline__ float luminance(v3 c)
{
	return fmaf(0.114f, fminf(c.z, 10.0f), fmaf(0.587f, fminf(c.y, 10.0f), 0.299f * fminf(c.x, 10.0f)));
}
/* dot product with its two adds fused into the products */
__device__ __forceinline__ float fdot(v3 a, v3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ float4 combine(v3 A, v3 B)
{
	const uint32_t Ar = f2u_sat(fminf(A.x, 31.999f) * 2048.0f), Ag = f2u_sat(fminf(A.y, 31.999f) * 2048.0f), Ab = f2u_sat(fminf(A.z, 31.999f) * 2048.0f);
	const uint32_t Br = f2u_sat(fminf(B.x, 31.999f) * 2048.0f), Bg = f2u_sat(fminf(B.y, 31.999f) * 2048.0f), Bb = f2u_sat(fminf(B.z, 31.999f) * 2048.0f);
	return make_float4(bitsf((Ar << 16) + Ag), bitsf(Ab), bitsf((Br << 16) + Bg), bitsf(Bb));
}
__device__ __forceinline__ v3 get_direct(float4 X)
{
	const uint32_t v0 = fbits(X.x), v1 = fbits(X.y);
	return mk3((float)(v0 >> 16) * (1.0f / 2048.0f), (float)(v0 & 65535u) * (1.0f / 2048.0f), (float)v1 * (1.0f / 2048.0f));
}
__device__ __forceinline__ v3 get_indirect(float4 X)
{
	const uint32_t v2 = fbits(X.z), v3b = fbits(X.w);
	return mk3((float)(v2 >> 16) * (1.0f / 2048.0f), (float)(v2 & 65535u) * (1.0f / 2048.0f), (float)v3b * (1.0f / 2048.0f));
}
__device__ __forceinline__ v3 rgb_to_ycocg(v3 RGB)
{
	const v3 rgb = mk3(fminf(4.0f, RGB.x), fminf(4.0f, RGB.y), fminf(4.0f, RGB.z));
	const float Y = (rgb.x * 1.0f + rgb.y * 2.0f + rgb.z * 1.0f) * 0.25f;
	const float Co = (rgb.x * 2.0f + rgb.y * 0.0f + rgb.z * -2.0f) * 0.25f + (0.5f * 256.0f / 255.0f);
	const float Cg = (rgb.x * -1.0f + rgb.y * 2.0f + rgb.z * -1.0f) * 0.25f + (0.5f * 256.0f / 255.0f);
	return mk3(Y, Co, Cg);
}
__device__ __forceinline__ v3 ycocg_to_rgb(v3 c)
{
	const float Y = c.x, Co = c.y - (0.5f * 256.0f / 255.0f), Cg = c.z - (0.5f * 256.0f / 255.0f);
	return mk3(Y + Co - Cg, Y + Cg, Y - Co - Cg);
}
__device__ __forceinline__ float dot4(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
__device__ __forceinline__ float pow128(float x)
{
	x = x * x, x = x * x, x = x * x, x = x * x, x = x * x, x = x * x, x = x * x;
	return x;
}
__device__ __forceinline__ float clampf(float v, float a, float b) { return fmaxf(a, fminf(v, b)); }

/* ReadWorldPos (sampling_shared.h:111-115) */
__device__ __forceinline__ float4 read_world_pos(const float4* buf, int x, int y, int w, int h)
{
	if (x >= 0 && y >= 0 && x < w && y < h) return buf[x + y * w];
	return make_float4(1e20f, 1e20f, 1e20f, bitsf(0u));
}

__device__ float world_distance(int px, int py, float4 cur, const float4* prev, int w, int h)
{
	const float4 p = read_world_pos(prev, px, py, w, h);
	const v3 ln = unpack_normal2(fbits(cur.w)), pn = unpack_normal2(fbits(p.w));
	if ((fbits(p.w) & 3u) != 1u) return 1e21f;
	if (dot(ln, pn) < 0.85f) return 1e21f;
	return length(mk3(cur.x - p.x, cur.y - p.y, cur.z - p.z));
}

__device__ float fine_world_distance(f2 pixel, float4 cur, const float4* prev, int w, int h)
{
	const int p0x = f2i_sat(pixel.x), p0y = f2i_sat(pixel.y);
	const float fx = pixel.x - floorf(pixel.x), fy = pixel.y - floorf(pixel.y);
	const float w0 = (1 - fx) * (1 - fy), w1 = fx * (1 - fy), w2 = (1 - fx) * fy, w3 = fx * fy;
	const float d0 = world_distance(p0x, p0y, cur, prev, w, h), d1 = world_distance(p0x + 1, p0y, cur, prev, w, h);
	const float d2 = world_distance(p0x, p0y + 1, cur, prev, w, h), d3 = world_distance(p0x + 1, p0y + 1, cur, prev, w, h);
	float tw = 0, td = 0;
	if (d0 < 1e20f) td += d0 * w0, tw += w0;
	if (d1 < 1e20f) td += d1 * w1, tw += w1;
	if (d2 < 1e20f) td += d2 * w2, tw += w2;
	if (d3 < 1e20f) td += d3 * w3, tw += w3;
	return tw == 0 ? 1e20f : td / tw;
}

__device__ int refine_history_pos(f2& pos, f2 off, float4 cur, float& best, float step, const float4* prev, int w, int h)
{
	int tap = 0;
	const f2 c[4] = { { pos.x - step, pos.y }, { pos.x + step, pos.y }, { pos.x, pos.y - step }, { pos.x, pos.y + step } };
	for (int i = 0; i < 4; i++)
	{
		const float d = fine_world_distance(f2{ c[i].x + off.x, c[i].y + off.y }, cur, prev, w, h);
		if (d < best) best = d, pos = c[i], tap = i + 1;
	}
	return tap;
}

/* ReadTexelConsistent (sampling_shared.h:136-178), active branch */
__device__ float4 read_texel_consistent(const float4* buf, const float4* prevWP, float4 localPos, v3 localNormal, float u, float v, int w, int h)
{
	const int iu1 = f2i_sat(floorf(u)), iv1 = f2i_sat(floorf(v)), iu0 = max(0, iu1 - 1), iv0 = max(0, iv1 - 1);
	if (iu1 >= w || iv1 >= h || iu1 < 0 || iv1 < 0) return make_float4(-1, -1, -1, -1);
	const float fu = u - floorf(u), fv = v - floorf(v);
	const float4 p0 = buf[iu0 + iv0 * w], pp0 = prevWP[iu0 + iv0 * w], p1 = buf[iu1 + iv0 * w], pp1 = prevWP[iu1 + iv0 * w];
	const float4 p2 = buf[iu0 + iv1 * w], pp2 = prevWP[iu0 + iv1 * w], p3 = buf[iu1 + iv1 * w], pp3 = prevWP[iu1 + iv1 * w];
	const uint32_t spec = fbits(localPos.w) & 3u;
	float w0 = (1 - fu) * (1 - fv), w1 = fu * (1 - fv), w2 = (1 - fu) * fv, w3 = 1 - (w0 + w1 + w2);
	const uint32_t n0 = fbits(pp0.w), n1 = fbits(pp1.w), n2 = fbits(pp2.w), n3 = fbits(pp3.w);
	if (dot(unpack_normal2(n0), localNormal) < 0.95f || (n0 & 3u) != spec) w0 = 0;
	if (dot(unpack_normal2(n1), localNormal) < 0.95f || (n1 & 3u) != spec) w1 = 0;
	if (dot(unpack_normal2(n2), localNormal) < 0.95f || (n2 & 3u) != spec) w2 = 0;
	if (dot(unpack_normal2(n3), localNormal) < 0.95f || (n3 & 3u) != spec) w3 = 0;
	const float sum = w0 + w1 + w2 + w3;
	if (sum == 0) return make_float4(-1, -1, -1, -1);
	const float r = 1.0f / sum;
	return make_float4((w0 * p0.x + w1 * p1.x + w2 * p2.x + w3 * p3.x) * r, (w0 * p0.y + w1 * p1.y + w2 * p2.y + w3 * p3.y) * r,
		(w0 * p0.z + w1 * p1.z + w2 * p2.z + w3 * p3.z) * r, (w0 * p0.w + w1 * p1.w + w2 * p2.w + w3 * p3.w) * r);
}

/* ReadTexelConsistent2 (sampling_shared.h:180-220), active branch; direct.x = -1 when rejected */
__device__ void read_texel_consistent2(const float4* buf, const float4* prevWP, float4 localPos, v3 localNormal, float u, float v, int w, int h,
	v3& direct, v3& indirect)
{
	direct.x = -1;
	const int iu1 = f2i_sat(floorf(u)), iv1 = f2i_sat(floorf(v)), iu0 = max(0, iu1 - 1), iv0 = max(0, iv1 - 1);
	if (iu1 >= w || iv1 >= h || iu1 < 0 || iv1 < 0) return;
	const float fu = u - floorf(u), fv = v - floorf(v);
	const int idx[4] = { iu0 + iv0 * w, iu1 + iv0 * w, iu0 + iv1 * w, iu1 + iv1 * w };
	float wt[4] = { (1 - fu) * (1 - fv), fu * (1 - fv), (1 - fu) * fv, 0 };
	wt[3] = 1 - (wt[0] + wt[1] + wt[2]);
	const uint32_t spec = fbits(localPos.w) & 3u;
	v3 pd[4], pi[4];
	for (int k = 0; k < 4; k++)
	{
		const float4 p = buf[idx[k]];
		pd[k] = get_direct(p), pi[k] = get_indirect(p);
		const uint32_t n = fbits(prevWP[idx[k]].w);
		if (dot(unpack_normal2(n), localNormal) < 0.975f || (n & 3u) != spec) wt[k] = 0;
	}
	const float sum = wt[0] + wt[1] + wt[2] + wt[3];
	if (sum == 0) return;
	const float r = 1.0f / sum;
	direct = mk3((wt[0] * pd[0].x + wt[1] * pd[1].x + wt[2] * pd[2].x + wt[3] * pd[3].x) * r,
		(wt[0] * pd[0].y + wt[1] * pd[1].y + wt[2] * pd[2].y + wt[3] * pd[3].y) * r,
		(wt[0] * pd[0].z + wt[1] * pd[1].z + wt[2] * pd[2].z + wt[3] * pd[3].z) * r);
	indirect = mk3((wt[0] * pi[0].x + wt[1] * pi[1].x + wt[2] * pi[2].x + wt[3] * pi[3].x) * r,
		(wt[0] * pi[0].y + wt[1] * pi[1].y + wt[2] * pi[2].y + wt[3] * pi[3].y) * r,
		(wt[0] * pi[0].z + wt[1] * pi[1].z + wt[2] * pi[2].z + wt[3] * pi[3].z) * r);
}


/* prepareFilterKernel (finalize_shared.h:226-294) */
/* Also keeps this frame's world positions (the next frame's history) in
 * worldPosOut, so the G-buffer needs no separate copy, and merges the frame's
 * G-buffer features (feat) into the filter's copy: x, y, z and the flags of w
 * from the frame, the history count (bits 0..3 of w) kept and updated. */
__global__ void k_prepare(const float4* __restrict__ acc, const uint4* __restrict__ feat, uint4* __restrict__ features, const float4* __restrict__ worldPos,
	float4* __restrict__ worldPosOut, const float4* __restrict__ prevWorldPos, float4* __restrict__ shading, float2* __restrict__ motion,
	float4* __restrict__ moments, const float4* __restrict__ prevMoments, SvgfCam cam, float j0, float j1, float pj0, float pj1,
	int w, int h, float pvs, float clampD, float clampI, int stationary)
{
	const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
	if (x >= w || y >= h) return;
	const int p = x + y * w;
	const float4 a0 = acc[p], a1 = acc[p + w * h];
	const v3 direct = mk3(a0.x, a0.y, a0.z) * pvs;
	const uint4 g = feat[p];
	const uint4 lf = make_uint4(g.x, g.y, g.z, (g.w & ~15u) | (features[p].w & 15u));
	uint32_t fw = lf.w;                                  /* w with the history count this frame leaves */
	const float4 lwp = worldPos[p];
	worldPosOut[p] = lwp;
	const v3 albedo = rgb32_to_hdr_min1(lf.x);
	const v3 indirect = mk3(a1.x, a1.y, a1.z) * pvs;
	const v3 reci = mk3(1.0f / albedo.x, 1.0f / albedo.y, 1.0f / albedo.z);
	const v3 dm = direct * reci, im = indirect * reci;
	const v3 dl = mk3(fminf(dm.x, clampD), fminf(dm.y, clampD), fminf(dm.z, clampD));
	const v3 il = mk3(fminf(im.x, clampI), fminf(im.y, clampI), fminf(im.z, clampI));
	shading[p] = combine(dl, il);
	float lumD = luminance(dl), lumD2 = lumD * lumD, lumI = luminance(il), lumI2 = lumI * lumI;
	f2 pp;
	if (((lf.w >> 4) & 3u) == 0)
	{
		const v3 dd = normalize(mk3(lwp.x - cam.prevPos.x, lwp.y - cam.prevPos.y, lwp.z - cam.prevPos.z));
		const float4 D = make_float4(dd.x, dd.y, dd.z, 0);
		const float s = cam.prevPos.w / dot4(cam.prevE, D);
		const float4 S = make_float4(cam.prevPos.x + D.x * s, cam.prevPos.y + D.y * s, cam.prevPos.z + D.z * s, 0);
		pp.x = dot4(S, cam.prevRight) - cam.prevRight.w - j0;
		pp.y = dot4(S, cam.prevUp) - cam.prevUp.w - j1;
	}
	else
	{
		pp.x = (float)x, pp.y = (float)y;
		if (!stationary)
		{
			const float4 q = prevWorldPos[p];
			float best = length(mk3(lwp.x - q.x, lwp.y - q.y, lwp.z - q.z)), step = 5.0f;
			const f2 off = { j0 - pj0, j1 - pj1 };
			int iter = 0;
			while (true)
			{
				const int tap = refine_history_pos(pp, off, lwp, best, step, prevWorldPos, w, h);
				if (tap == 0) { step *= 0.45f; if (step < 0.05f) break; }
				if (++iter == 25) break;
			}
		}
	}
	pp.x = pp.x + 0.5f, pp.y = pp.y + 0.5f;
	const float px = pp.x, py = pp.y;
	if (px >= 0 && px < (float)w && py >= 0 && py < (float)h)
	{
		const v3 ln = unpack_normal2(lf.y);
		const float4 hist = read_texel_consistent(prevMoments, prevWorldPos, lwp, ln, px, py, w, h);
		if (hist.x > -1)
		{
			lumD = 0.2f * lumD + 0.8f * hist.x;
			lumD2 = 0.2f * lumD2 + 0.8f * hist.y;
			lumI = 0.2f * lumI + 0.8f * hist.z;
			lumI2 = 0.2f * lumI2 + 0.8f * hist.w;
			if ((fw & 15u) < 15u) fw++;
		}
		else fw &= 0xfffffff0u;
	}
	else fw &= 0xfffffff0u;
	features[p] = make_uint4(lf.x, lf.y, lf.z, fw);
	motion[p] = make_float2(pp.x, pp.y);
	moments[p] = make_float4(lumD, lumD2, lumI, lumI2);
}

/* applyFilterKernel (finalize_shared.h:320-477) */
__global__ void k_apply(const uint4* __restrict__ features, const float4* __restrict__ prevWorldPos, const float4* __restrict__ worldPos,
	const float4* __restrict__ deltaDepth, const float2* __restrict__ motion, const float4* __restrict__ moments,
	const float4* __restrict__ A, const float4* __restrict__ B, float4* __restrict__ C, int w, int h, int phase, int lastPass)
{
	const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
	if (x >= w || y >= h) return;
	const int p = x + y * w;
	const uint4 lf = features[p];
	const float4 lpos = worldPos[p];
	const v3 ln = unpack_normal2(lf.y), lc = rgb32_to_hdr(lf.x);
	const uint32_t lmat = lf.w >> 4;
	float dws = 1, iws = 1;
	const float4 comb = A[p];
	v3 dsum = get_direct(comb) * dws, isum = get_indirect(comb) * iws;
	const float lD = luminance(get_direct(comb)), lI = luminance(get_indirect(comb));
	const float ldepth = bitsf(lf.z), ddx = deltaDepth[p].z, ddy = deltaDepth[p].w;
	const float sigma = 10.0f * bitsf((uint32_t)(127 - (phase - 1)) << 23);
	const float factor = (lf.w & 15u) == 0 ? 400.0f : 1.0f;
	const float4 m = moments[p];
	const float varD = m.y - m.x * m.x, varI = m.w - m.z * m.z;
	const float rD = -1.0f / (sigma * factor * sqrtf(varD + 0.00001f) + 0.00001f);
	const float rI = -1.0f / (sigma * factor * sqrtf(varI + 0.00001f) + 0.00001f);
	const int step = 1 << (phase - 1);
	for (int vv = -2; vv <= 2; vv++)
	{
		const int v = vv * step + y;
		const int r = (vv == 2 || vv == -2) ? 1 : 2;
		if (v < 0 || v >= h) continue;
		for (int uu = -r; uu <= r; uu++)
		{
			if (uu == 0 && vv == 0) continue;
			const int u = min(max(uu * step + x, 0), w - 1);
			const int q = u + v * w;
			const float4 c = A[q];
			const uint4 nf = features[q];
			const float wdist = (float)(uu * uu + vv * vv) * (-1.0f / 7.5f);
			const v3 nd = get_direct(c), ni = get_indirect(c), nn = unpack_normal2(nf.y);
			float wn = pow128(fmaxf(0.0f, fdot(nn, ln)));
			const float expected = fmaf(ddy, (float)(vv * step), fmaf(ddx, (float)(uu * step), ldepth));
			const float err = fabsf(expected - bitsf(nf.z));
			const float expDiff = fabsf(expected - ldepth);
			const float wdepth = err / fmaxf(0.00001f, (0.5f + (float)phase * 0.5f) * expDiff);
			if ((nf.w >> 4) != lmat) wn = (float)((double)wn * 0.0001);
			else wn = wn * fdot(lc, rgb32_to_hdr(nf.x));
			float wd = wn * cw_expf(fmaf(fabsf(lD - luminance(nd)), rD, wdist) - wdepth);
			float wi = wn * cw_expf(fmaf(fabsf(lI - luminance(ni)), rI, wdist) - wdepth);
			if (!__builtin_isfinite(wd)) wd = 0;
			if (!__builtin_isfinite(wi)) wi = 0;
			dsum = mk3(fmaf(nd.x, wd, dsum.x), fmaf(nd.y, wd, dsum.y), fmaf(nd.z, wd, dsum.z)), dws += wd;
			isum = mk3(fmaf(ni.x, wi, isum.x), fmaf(ni.y, wi, isum.y), fmaf(ni.z, wi, isum.z)), iws += wi;
		}
	}
	v3 df = dsum * (1.0f / fmaxf(0.0001f, dws));
	v3 inf = isum * (1.0f / fmaxf(0.0001f, iws));
	if (phase == 1)
	{
		const float2 mv = motion[p];
		const int px = f2i_sat(mv.x), py = f2i_sat(mv.y);
		if (px >= 0 && px < w && py >= 0 && py < h)
		{
			v3 pd, pi;
			read_texel_consistent2(B, prevWorldPos, lpos, ln, mv.x, mv.y, w, h, pd, pi);
			if (pd.x != -1)
			{
				pd = rgb_to_ycocg(pd), pi = rgb_to_ycocg(pi);
				v3 dAvg = rgb_to_ycocg(df), dVar = dAvg * dAvg;
				v3 iAvg = rgb_to_ycocg(inf), iVar = iAvg * iAvg;
				const auto add = [&](int q) {
					const float4 c4 = A[q];
					const v3 f = rgb_to_ycocg(get_direct(c4)), g = rgb_to_ycocg(get_indirect(c4));
					dAvg = dAvg + f, dVar = dVar + f * f, iAvg = iAvg + g, iVar = iVar + g * g;
				};
				if (x > 1)
				{
					if (y > 1) add(p - w - 1);
					add(p - 1);
					if (y < h - 1) add(p + w - 1);
				}
				if (y > 1) add(p - w);
				if (y < h - 1) add(p + w);
				if (x < w - 1)
				{
					if (y > 1) add(p + 1 - w);
					add(p + 1);
					if (y < h - 1) add(p + 1 + w);
				}
				dAvg = dAvg * (1.0f / 9.0f), dVar = dVar * (1.0f / 9.0f), iAvg = iAvg * (1.0f / 9.0f), iVar = iVar * (1.0f / 9.0f);
				const v3 sd0 = dVar - dAvg * dAvg, si0 = iVar - iAvg * iAvg;
				const v3 sD = mk3(sqrtf(fmaxf(0.0f, sd0.x)), sqrtf(fmaxf(0.0f, sd0.y)), sqrtf(fmaxf(0.0f, sd0.z)));
				const v3 sI = mk3(sqrtf(fmaxf(0.0f, si0.x)), sqrtf(fmaxf(0.0f, si0.y)), sqrtf(fmaxf(0.0f, si0.z)));
				const v3 mnD = dAvg - 0.75f * sD, mxD = dAvg + 0.75f * sD, mnI = iAvg - 0.75f * sI, mxI = iAvg + 0.75f * sI;
				pd = mk3(clampf(pd.x, mnD.x, mxD.x), clampf(pd.y, mnD.y, mxD.y), clampf(pd.z, mnD.z, mxD.z));
				pi = mk3(clampf(pi.x, mnI.x, mxI.x), clampf(pi.y, mnI.y, mxI.y), clampf(pi.z, mnI.z, mxI.z));
				df = df * 0.1f + ycocg_to_rgb(pd) * 0.9f;
				inf = inf * 0.1f + ycocg_to_rgb(pi) * 0.9f;
			}
		}
	}
	if (lastPass)
	{
		const v3 c = (df + inf) * rgb32_to_hdr(lf.x);
		C[p] = make_float4(sqrtf(c.x), sqrtf(c.y), sqrtf(c.z), 1);
	}
	else C[p] = combine(df, inf);
}

/* mitchellNetravali (sampling_shared.h:22-27), B = C = 1/3 */
__device__ __forceinline__ float mitchell_netravali(float v)
{
	const float B = 1.0f / 3.0f, C = 1.0f / 3.0f, x = fabsf(v), x2 = x * x, x3 = x2 * x;
	if (x < 1) return (1.0f / 6.0f) * ((12 - 9 * B - 6 * C) * x3 + (-18 + 12 * B + 6 * C) * x2 + (6 - 2 * B));
	if (x < 2) return 1.0f / 6.0f * ((-B - 6 * C) * x3 + (6 * B + 30 * C) * x2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C));
	return 0.0f;
}

/* ReadTexelBmitchellNetravali (sampling_shared.h:117-133): 4x4 taps from
 * (int)(u - 2), (int)(v - 2); row 0 is excluded (y > 0), as in the reference. */
__device__ v3 read_texel_mitchell(const float4* buf, float u, float v, int w, int h)
{
	const int x1 = f2i_sat(u - 2.0f), y1 = f2i_sat(v - 2.0f);
	float tw = 0;
	float4 t = make_float4(0, 0, 0, 0);
	for (int y = y1; y < y1 + 4; y++)
		for (int x = x1; x < x1 + 4; x++)
			if (x >= 0 && y > 0 && x < w && y < h)
			{
				const float dx = (float)x - u, dy = (float)y - v;
				const float wt = mitchell_netravali(dx) * mitchell_netravali(dy);
				const float4 b = buf[x + y * w];
				t = make_float4(t.x + b.x * wt, t.y + b.y * wt, t.z + b.z * wt, t.w + b.w * wt);
				tw += wt;
			}
	const float r = 1.0f / tw;
	return mk3(t.x * r, t.y * r, t.z * r);
}

/* TAApassKernel (finalize_shared.h:490-538): Salvi's variance clip of the
 * reprojected history (Mitchell-Netravali read of the previous output),
 * blend 0.1 new / 0.9 history in YCoCg, clamp to 10. The reference updates
 * its input in place while neighbours read it; here the neighbourhood is read
 * from the unmodified input `px` and the result goes to `out`. */
__global__ void k_taa(const float4* __restrict__ px, const float4* __restrict__ prevPixels, const float2* __restrict__ motion, int w, int h,
	float4* __restrict__ out)
{
	const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
	if (x >= w || y >= h) return;
	const int p = x + y * w;
	v3 pixel = mk3(px[p].x, px[p].y, px[p].z);
	const float2 mv = motion[p];
	const float ux = mv.x - 0.5f, uy = mv.y - 0.5f;
	if (ux >= 0 && ux < (float)w && uy >= 0 && uy < (float)h)
	{
		const v3 np = rgb_to_ycocg(pixel);
		v3 hist = rgb_to_ycocg(read_texel_mitchell(prevPixels, ux, uy, w, h));
		v3 avg = np, var = np * np;
		const auto add = [&](int q) {
			const v3 f = rgb_to_ycocg(mk3(px[q].x, px[q].y, px[q].z));
			avg = avg + f, var = var + f * f;
		};
		if (x > 1)
		{
			if (y > 1) add(p - w - 1);
			add(p - 1);
			if (y < h - 1) add(p + w - 1);
		}
		if (y > 1) add(p - w);
		if (y < h - 1) add(p + w);
		if (x < w - 1)
		{
			if (y > 1) add(p + 1 - w);
			add(p + 1);
			if (y < h - 1) add(p + 1 + w);
		}
		avg = avg * (1.0f / 9.0f), var = var * (1.0f / 9.0f);
		const v3 s0 = var - avg * avg;
		const v3 sg = mk3(sqrtf(fmaxf(0.0f, s0.x)), sqrtf(fmaxf(0.0f, s0.y)), sqrtf(fmaxf(0.0f, s0.z)));
		const v3 mn = avg - 1.25f * sg, mx = avg + 1.25f * sg;
		hist = mk3(clampf(hist.x, mn.x, mx.x), clampf(hist.y, mn.y, mx.y), clampf(hist.z, mn.z, mx.z));
		pixel = ycocg_to_rgb(np * 0.1f + hist * 0.9f);
		if (__builtin_isnan(pixel.x + pixel.y + pixel.z)) pixel = ycocg_to_rgb(np);
	}
	out[p] = make_float4(fminf(10.0f, pixel.x), fminf(10.0f, pixel.y), fminf(10.0f, pixel.z), 0);
}

/* unsharpenTAAKernel (finalize_shared.h:553-575): sharpen against the 8
 * neighbours, then square (the display path applies gamma again); border
 * pixels are not written. */
__global__ void k_unsharpen(const float4* __restrict__ px, int w, int h, float4* __restrict__ out)
{
	const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
	if (x == 0 || y == 0 || x >= w - 1 || y >= h - 1) return;
	const float4 q[8] = { px[x - 1 + (y - 1) * w], px[x + (y - 1) * w], px[x + 1 + (y - 1) * w], px[x + 1 + y * w],
		px[x + 1 + (y + 1) * w], px[x + (y + 1) * w], px[x - 1 + (y + 1) * w], px[x - 1 + y * w] };
	const float4 c = px[x + y * w];
	float4 s = make_float4(q[0].x * 0.35f, q[0].y * 0.35f, q[0].z * 0.35f, q[0].w * 0.35f);
	for (int k = 1; k < 8; k++)
	{
		const float wk = (k & 1) ? 0.5f : 0.35f;
		s = make_float4(s.x + q[k].x * wk, s.y + q[k].y * wk, s.z + q[k].z * wk, s.w + q[k].w * wk);
	}
	const float4 v = make_float4(fmaxf(c.x, c.x * 2.7f - s.x * 0.5f), fmaxf(c.y, c.y * 2.7f - s.y * 0.5f), fmaxf(c.z, c.z * 2.7f - s.z * 0.5f), 0);
	out[x + y * w] = make_float4(v.x * v.x, v.y * v.y, v.z * v.z, 0);
}

/* finalizeNoTAAKernel (finalize_shared.h:587-595): display value = sqrt of the last pass; border pixels are not written */
__global__ void k_finalize_no_taa(const float4* __restrict__ px, int w, int h, float4* __restrict__ out)
{
	const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
	if (x == 0 || y == 0 || x >= w - 1 || y >= h - 1) return;
	const float4 v = px[x + y * w];
	out[x + y * w] = make_float4(sqrtf(v.x), sqrtf(v.y), sqrtf(v.z), 0);
}

static inline float3 f3(const lighthouse2::float3& a) { return make_float3(a.x, a.y, a.z); }
static inline float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline float3 operator*(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }
static inline float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
static inline float hdot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float hlen(float3 a) { return sqrtf(hdot(a, a)); }
static inline float3 hnorm(float3 a) { const float inv = 1.0f / sqrtf(hdot(a, a)); return a * inv; }

SvgfCam svgf_camera(const lighthouse2::ViewPyramid& v, int h)
{
	const float3 pos = f3(v.pos), p1 = f3(v.p1), p2 = f3(v.p2), p3 = f3(v.p3);
	const float3 centre = 0.5f * (p2 + p3);
	const float3 direction = hnorm(centre - pos);
	const float3 right = hnorm(p2 - p1), up = hnorm(p3 - p1);
	const float screenSize = hlen(p3 - p1);
	const float lenReci = (float)h / screenSize;
	SvgfCam c;
	c.prevPos = make_float4(pos.x, pos.y, pos.z, -(hdot(pos, direction) - hdot(centre, direction)));
	c.prevE = make_float4(direction.x, direction.y, direction.z, 0);
	const float3 r = right * lenReci, u = up * lenReci;
	c.prevRight = make_float4(r.x, r.y, r.z, hdot(p1, right) * lenReci);
	c.prevUp = make_float4(u.x, u.y, u.z, hdot(p1, up) * lenReci);
	return c;
}

static inline void chk(hipError_t e, const char* what)
{
	if (e != hipSuccess) throw CoreError(e == hipErrorOutOfMemory ? LH2HIP_ERR_OOM : LH2HIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

Svgf::Svgf(int w_, int h_, hipStream_t s) : w(w_), h(h_)
{
	if (w < 3 || h < 3 || (int64_t)w * h > (1 << 26)) throw CoreError(LH2HIP_ERR_INVALID, "svgf: size 3x3 .. 2^26 pixels");
	const size_t n = (size_t)w * h;
	features.ensure(n * 16), worldPos.ensure(n * 16), prevWorldPos.ensure(n * 16);
	motion.ensure(n * 8), moments.ensure(n * 16), prevMoments.ensure(n * 16), shading.ensure(n * 16);
	filteredIn.ensure(n * 16), filteredOut.ensure(n * 16), lastPass.ensure(n * 16), prevPixels.ensure(n * 16), taaOut.ensure(n * 16);
	for (DevBuf* b : { &features, &worldPos, &prevWorldPos, &motion, &moments, &prevMoments, &shading, &filteredIn, &filteredOut, &lastPass,
		&prevPixels, &taaOut })
		chk(hipMemsetAsync(b->p, 0, b->bytes, s), "memset");
	if (!s) chk(hipStreamSynchronize(nullptr), "sync");
}

void Svgf::Reset(hipStream_t s)
{
	for (DevBuf* b : { &features, &worldPos, &prevWorldPos, &motion, &moments, &prevMoments, &shading, &filteredIn, &filteredOut, &lastPass,
		&prevPixels, &taaOut })
		chk(hipMemsetAsync(b->p, 0, b->bytes, s), "memset");
}

void Svgf::Frame(const float4* acc, const uint4* feat, const float4* wp, const float4* dd, const lighthouse2::ViewPyramid& prevView,
	float j0, float j1, float pj0, float pj1, int samplesTaken, float clampD, float clampI, int stationary, float4* out, hipStream_t s)
{
	if (samplesTaken <= 0) throw CoreError(LH2HIP_ERR_INVALID, "svgf frame: samplesTaken >= 1");
	const dim3 b(8, 8), g((w + 7) / 8, (h + 7) / 8);
	const SvgfCam cam = svgf_camera(prevView, h);
	/* wp and dd are read in place (stream order keeps them valid until the
	 * last pass); k_prepare keeps wp as the next frame's history */
	k_prepare<<<g, b, 0, s>>>(acc, feat, features.as<uint4>(), wp, worldPos.as<float4>(), prevWorldPos.as<float4>(), shading.as<float4>(),
		motion.as<float2>(), moments.as<float4>(), prevMoments.as<float4>(), cam, j0, j1, pj0, pj1, w, h, 1.0f / (float)samplesTaken, clampD,
		clampI, stationary);
	const auto apply = [&](int phase, DevBuf& A, const float4* B, DevBuf& C, int last) {
		k_apply<<<g, b, 0, s>>>(features.as<uint4>(), prevWorldPos.as<float4>(), wp, dd, motion.as<float2>(),
			moments.as<float4>(), A.as<float4>(), B, C.as<float4>(), w, h, phase, last);
	};
	apply(1, shading, filteredIn.as<float4>(), filteredOut, 0);
	apply(2, filteredOut, nullptr, filteredIn, 0);
	apply(3, filteredIn, nullptr, lastPass, 1);
	if (taa)
	{	/* TAApass + unsharpenTAA (rendercore.cpp:922-926); the TAA output is the next frame's history */
		k_taa<<<g, b, 0, s>>>(lastPass.as<float4>(), prevPixels.as<float4>(), motion.as<float2>(), w, h, taaOut.as<float4>());
		k_unsharpen<<<g, b, 0, s>>>(taaOut.as<float4>(), w, h, out);
		std::swap(taaOut, prevPixels);
	}
	else k_finalize_no_taa<<<g, b, 0, s>>>(lastPass.as<float4>(), w, h, out);
	chk(hipGetLastError(), "svgf kernels");
	std::swap(filteredIn, filteredOut);
	std::swap(prevWorldPos, worldPos);
	std::swap(moments, prevMoments);
}

} // namespace lh
