/* oracle_svgf.cpp - CPU oracle of SVGF. TEST INFRASTRUCTURE (see oracle_math.h).
 *
 * prepareFilterKernel, applyFilterKernel x 3, TAApassKernel + unsharpenTAAKernel
 * or finalizeNoTAAKernel of lib/CUDA/shared_kernel_code/finalize_shared.h:169-600
 * (with ReadTexelBmitchellNetravali, sampling_shared.h:22-27,117-133), driven as
 * RenderCore_Optix7Filter::FinalizeRender (rendercore.cpp:897-935), with the
 * packers of tools_shared.h:113-136,162-183,282-304 and the history reads of
 * sampling_shared.h:111-115,136-220. powf(x, 128) is seven squarings and
 * __expf is cw_expf, as in the kernels. */
#include "oracle_math.h"
#include "../include/lh2hip_types.h"
#include <algorithm>
#include <vector>

#define OAPI extern "C" __attribute__((visibility("default")))

namespace orc {

struct u4 { uint x, y, z, w; };
struct fl2 { float x, y; };

static f3 unpack_normal2(uint p)
{
	const uint x = (p >> 2u) & 1023u, y = (p >> 12u) & 1023u, z = p >> 22u;
	return mk3(fmaf((float)x, 1.0f / 511.0f, -1.0f), fmaf((float)y, 1.0f / 511.0f, -1.0f), fmaf((float)z, 1.0f / 511.0f, -1.0f));
}
static f3 rgb32_to_hdr(uint c) { return mk3((float)(c >> 22) * (1.0f / 1023.0f), (float)((c >> 11) & 2047u) * (1.0f / 2047.0f), (float)(c & 2047u) * (1.0f / 2047.0f)); }
static f3 rgb32_to_hdr_min1(uint c)
{
	return mk3((float)std::max(1u, c >> 22) * (1.0f / 1023.0f), (float)std::max(1u, (c >> 11) & 2047u) * (1.0f / 2047.0f),
		(float)std::max(1u, c & 2047u) * (1.0f / 2047.0f));
}
static float luminance(f3 c) { return fmaf(0.114f, std::min(c.z, 10.0f), fmaf(0.587f, std::min(c.y, 10.0f), 0.299f * std::min(c.x, 10.0f))); }
/* dot product with its two adds fused into the products (fmaf, as the kernels' fma) */
static float fdot(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
static f4 combine(f3 A, f3 B)
{
	const uint Ar = f2u_sat(std::min(A.x, 31.999f) * 2048.0f), Ag = f2u_sat(std::min(A.y, 31.999f) * 2048.0f), Ab = f2u_sat(std::min(A.z, 31.999f) * 2048.0f);
	const uint Br = f2u_sat(std::min(B.x, 31.999f) * 2048.0f), Bg = f2u_sat(std::min(B.y, 31.999f) * 2048.0f), Bb = f2u_sat(std::min(B.z, 31.999f) * 2048.0f);
	f4 r = { u2f((Ar << 16) + Ag), u2f(Ab), u2f((Br << 16) + Bg), u2f(Bb) };
	return r;
}
static f3 get_direct(f4 X) { const uint a = f2u(X.x), b = f2u(X.y); return mk3((float)(a >> 16) * (1.0f / 2048.0f), (float)(a & 65535u) * (1.0f / 2048.0f), (float)b * (1.0f / 2048.0f)); }
static f3 get_indirect(f4 X) { const uint a = f2u(X.z), b = f2u(X.w); return mk3((float)(a >> 16) * (1.0f / 2048.0f), (float)(a & 65535u) * (1.0f / 2048.0f), (float)b * (1.0f / 2048.0f)); }
static f3 rgb_to_ycocg(f3 RGB)
{
	const f3 rgb = mk3(std::min(4.0f, RGB.x), std::min(4.0f, RGB.y), std::min(4.0f, RGB.z));
	return mk3((rgb.x * 1.0f + rgb.y * 2.0f + rgb.z * 1.0f) * 0.25f, (rgb.x * 2.0f + rgb.y * 0.0f + rgb.z * -2.0f) * 0.25f + (0.5f * 256.0f / 255.0f),
		(rgb.x * -1.0f + rgb.y * 2.0f + rgb.z * -1.0f) * 0.25f + (0.5f * 256.0f / 255.0f));
}
static f3 ycocg_to_rgb(f3 c)
{
	const float Y = c.x, Co = c.y - (0.5f * 256.0f / 255.0f), Cg = c.z - (0.5f * 256.0f / 255.0f);
	return mk3(Y + Co - Cg, Y + Cg, Y - Co - Cg);
}
static float dot4(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
static float pow128(float x) { for (int i = 0; i < 7; i++) x = x * x; return x; }
static float clampf(float v, float a, float b) { return std::max(a, std::min(v, b)); }
static f3 xyz(f4 a) { return mk3(a.x, a.y, a.z); }

struct Cam { f4 pos, E, right, up; };

struct Svgf
{
	int w, h;
	std::vector<u4> features;
	std::vector<f4> worldPos, prevWorldPos, deltaDepth, moments, prevMoments, shading, filteredIn, filteredOut, lastPass;
	std::vector<f4> prevPixels, taaOut;   /* TAA history (the previous TAA output) and this frame's */
	std::vector<fl2> motion;
	Svgf(int w_, int h_) : w(w_), h(h_)
	{
		const size_t n = (size_t)w * h;
		const f4 z = { 0, 0, 0, 0 };
		features.assign(n, u4{ 0, 0, 0, 0 }), motion.assign(n, fl2{ 0, 0 });
		for (auto* v : { &worldPos, &prevWorldPos, &deltaDepth, &moments, &prevMoments, &shading, &filteredIn, &filteredOut, &lastPass,
			&prevPixels, &taaOut }) v->assign(n, z);
	}

	/* mitchellNetravali (sampling_shared.h:22-27), B = C = 1/3 */
	static float mitchell_netravali(float v)
	{
		const float B = 1.0f / 3.0f, C = 1.0f / 3.0f, x = fabsf(v), x2 = x * x, x3 = x2 * x;
		if (x < 1) return (1.0f / 6.0f) * ((12 - 9 * B - 6 * C) * x3 + (-18 + 12 * B + 6 * C) * x2 + (6 - 2 * B));
		if (x < 2) return 1.0f / 6.0f * ((-B - 6 * C) * x3 + (6 * B + 30 * C) * x2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C));
		return 0.0f;
	}
	/* ReadTexelBmitchellNetravali (sampling_shared.h:117-133); row 0 excluded (y > 0) as in the reference */
	f3 read_texel_mitchell(const std::vector<f4>& buf, float u, float v) const
	{
		const int x1 = f2i_sat(u - 2.0f), y1 = f2i_sat(v - 2.0f);
		float tw = 0;
		f4 t = { 0, 0, 0, 0 };
		for (int y = y1; y < y1 + 4; y++)
			for (int x = x1; x < x1 + 4; x++)
				if (x >= 0 && y > 0 && x < w && y < h)
				{
					const float dx = (float)x - u, dy = (float)y - v;
					const float wt = mitchell_netravali(dx) * mitchell_netravali(dy);
					const f4 b = buf[x + y * w];
					t = f4{ t.x + b.x * wt, t.y + b.y * wt, t.z + b.z * wt, t.w + b.w * wt };
					tw += wt;
				}
		const float r = 1.0f / tw;
		return mk3(t.x * r, t.y * r, t.z * r);
	}
	/* TAApassKernel (finalize_shared.h:490-538), reading the neighbourhood from the
	 * unmodified last pass; clamps with fminf/fmaxf (IEEE: a NaN operand loses). */
	void taa_pass()
	{
		for (int y = 0; y < h; y++) for (int x = 0; x < w; x++)
		{
			const int p = x + y * w;
			const std::vector<f4>& px = lastPass;
			f3 pixel = xyz(px[p]);
			const float ux = motion[p].x - 0.5f, uy = motion[p].y - 0.5f;
			if (ux >= 0 && ux < (float)w && uy >= 0 && uy < (float)h)
			{
				const f3 np = rgb_to_ycocg(pixel);
				f3 hist = rgb_to_ycocg(read_texel_mitchell(prevPixels, ux, uy));
				f3 avg = np, var = np * np;
				const auto add = [&](int q) { const f3 f = rgb_to_ycocg(xyz(px[q])); avg = avg + f, var = var + f * f; };
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
				const f3 s0 = var - avg * avg;
				const f3 sg = mk3(sqrtf(fmaxf(0.0f, s0.x)), sqrtf(fmaxf(0.0f, s0.y)), sqrtf(fmaxf(0.0f, s0.z)));
				const f3 mn = avg - 1.25f * sg, mx = avg + 1.25f * sg;
				hist = mk3(fmaxf(mn.x, fminf(hist.x, mx.x)), fmaxf(mn.y, fminf(hist.y, mx.y)), fmaxf(mn.z, fminf(hist.z, mx.z)));
				pixel = ycocg_to_rgb(np * 0.1f + hist * 0.9f);
				if (std::isnan(pixel.x + pixel.y + pixel.z)) pixel = ycocg_to_rgb(np);
			}
			taaOut[p] = f4{ fminf(10.0f, pixel.x), fminf(10.0f, pixel.y), fminf(10.0f, pixel.z), 0 };
		}
	}
	/* unsharpenTAAKernel (finalize_shared.h:553-575): interior pixels, squared */
	void unsharpen(float* out) const
	{
		for (int y = 1; y < h - 1; y++) for (int x = 1; x < w - 1; x++)
		{
			const std::vector<f4>& px = taaOut;
			const f4 q[8] = { px[x - 1 + (y - 1) * w], px[x + (y - 1) * w], px[x + 1 + (y - 1) * w], px[x + 1 + y * w],
				px[x + 1 + (y + 1) * w], px[x + (y + 1) * w], px[x - 1 + (y + 1) * w], px[x - 1 + y * w] };
			const f4 c = px[x + y * w];
			f4 s = { q[0].x * 0.35f, q[0].y * 0.35f, q[0].z * 0.35f, q[0].w * 0.35f };
			for (int k = 1; k < 8; k++)
			{
				const float wk = (k & 1) ? 0.5f : 0.35f;
				s = f4{ s.x + q[k].x * wk, s.y + q[k].y * wk, s.z + q[k].z * wk, s.w + q[k].w * wk };
			}
			const float vx = fmaxf(c.x, c.x * 2.7f - s.x * 0.5f), vy = fmaxf(c.y, c.y * 2.7f - s.y * 0.5f), vz = fmaxf(c.z, c.z * 2.7f - s.z * 0.5f);
			float* o = out + 4 * (size_t)(x + y * w);
			o[0] = vx * vx, o[1] = vy * vy, o[2] = vz * vz, o[3] = 0;
		}
	}

	f4 read_world_pos(int x, int y) const
	{
		if (x >= 0 && y >= 0 && x < w && y < h) return prevWorldPos[x + y * w];
		f4 r = { 1e20f, 1e20f, 1e20f, u2f(0u) };
		return r;
	}
	float world_distance(int px, int py, f4 cur) const
	{
		const f4 p = read_world_pos(px, py);
		if ((f2u(p.w) & 3u) != 1u) return 1e21f;
		if (dot(unpack_normal2(f2u(cur.w)), unpack_normal2(f2u(p.w))) < 0.85f) return 1e21f;
		return length(mk3(cur.x - p.x, cur.y - p.y, cur.z - p.z));
	}
	float fine_world_distance(fl2 pix, f4 cur) const
	{
		const int x0 = f2i_sat(pix.x), y0 = f2i_sat(pix.y);
		const float fx = pix.x - floorf(pix.x), fy = pix.y - floorf(pix.y);
		const float w0 = (1 - fx) * (1 - fy), w1 = fx * (1 - fy), w2 = (1 - fx) * fy, w3 = fx * fy;
		const float d[4] = { world_distance(x0, y0, cur), world_distance(x0 + 1, y0, cur), world_distance(x0, y0 + 1, cur), world_distance(x0 + 1, y0 + 1, cur) };
		const float wt[4] = { w0, w1, w2, w3 };
		float tw = 0, td = 0;
		for (int k = 0; k < 4; k++) if (d[k] < 1e20f) td += d[k] * wt[k], tw += wt[k];
		return tw == 0 ? 1e20f : td / tw;
	}
	/* ReadTexelConsistent: bilinear history read with normal / specularity rejection */
	f4 history(const std::vector<f4>& buf, f4 lp, f3 ln, float u, float v) const
	{
		const f4 none = { -1, -1, -1, -1 };
		const int iu1 = f2i_sat(floorf(u)), iv1 = f2i_sat(floorf(v)), iu0 = std::max(0, iu1 - 1), iv0 = std::max(0, iv1 - 1);
		if (iu1 >= w || iv1 >= h || iu1 < 0 || iv1 < 0) return none;
		const float fu = u - floorf(u), fv = v - floorf(v);
		const int idx[4] = { iu0 + iv0 * w, iu1 + iv0 * w, iu0 + iv1 * w, iu1 + iv1 * w };
		float wt[4] = { (1 - fu) * (1 - fv), fu * (1 - fv), (1 - fu) * fv, 0 };
		wt[3] = 1 - (wt[0] + wt[1] + wt[2]);
		const uint spec = f2u(lp.w) & 3u;
		for (int k = 0; k < 4; k++)
		{
			const uint n = f2u(prevWorldPos[idx[k]].w);
			if (dot(unpack_normal2(n), ln) < 0.95f || (n & 3u) != spec) wt[k] = 0;
		}
		const float sum = wt[0] + wt[1] + wt[2] + wt[3];
		if (sum == 0) return none;
		const float r = 1.0f / sum;
		const f4 &p0 = buf[idx[0]], &p1 = buf[idx[1]], &p2 = buf[idx[2]], &p3 = buf[idx[3]];
		f4 o = { (wt[0] * p0.x + wt[1] * p1.x + wt[2] * p2.x + wt[3] * p3.x) * r, (wt[0] * p0.y + wt[1] * p1.y + wt[2] * p2.y + wt[3] * p3.y) * r,
			(wt[0] * p0.z + wt[1] * p1.z + wt[2] * p2.z + wt[3] * p3.z) * r, (wt[0] * p0.w + wt[1] * p1.w + wt[2] * p2.w + wt[3] * p3.w) * r };
		return o;
	}
	/* ReadTexelConsistent2: the same for the packed direct/indirect history, threshold 0.975 */
	bool history2(const std::vector<f4>& buf, f4 lp, f3 ln, float u, float v, f3& dir, f3& ind) const
	{
		const int iu1 = f2i_sat(floorf(u)), iv1 = f2i_sat(floorf(v)), iu0 = std::max(0, iu1 - 1), iv0 = std::max(0, iv1 - 1);
		if (iu1 >= w || iv1 >= h || iu1 < 0 || iv1 < 0) return false;
		const float fu = u - floorf(u), fv = v - floorf(v);
		const int idx[4] = { iu0 + iv0 * w, iu1 + iv0 * w, iu0 + iv1 * w, iu1 + iv1 * w };
		float wt[4] = { (1 - fu) * (1 - fv), fu * (1 - fv), (1 - fu) * fv, 0 };
		wt[3] = 1 - (wt[0] + wt[1] + wt[2]);
		const uint spec = f2u(lp.w) & 3u;
		f3 pd[4], pi[4];
		for (int k = 0; k < 4; k++)
		{
			pd[k] = get_direct(buf[idx[k]]), pi[k] = get_indirect(buf[idx[k]]);
			const uint n = f2u(prevWorldPos[idx[k]].w);
			if (dot(unpack_normal2(n), ln) < 0.975f || (n & 3u) != spec) wt[k] = 0;
		}
		const float sum = wt[0] + wt[1] + wt[2] + wt[3];
		if (sum == 0) return false;
		const float r = 1.0f / sum;
		dir = mk3((wt[0] * pd[0].x + wt[1] * pd[1].x + wt[2] * pd[2].x + wt[3] * pd[3].x) * r, (wt[0] * pd[0].y + wt[1] * pd[1].y + wt[2] * pd[2].y + wt[3] * pd[3].y) * r,
			(wt[0] * pd[0].z + wt[1] * pd[1].z + wt[2] * pd[2].z + wt[3] * pd[3].z) * r);
		ind = mk3((wt[0] * pi[0].x + wt[1] * pi[1].x + wt[2] * pi[2].x + wt[3] * pi[3].x) * r, (wt[0] * pi[0].y + wt[1] * pi[1].y + wt[2] * pi[2].y + wt[3] * pi[3].y) * r,
			(wt[0] * pi[0].z + wt[1] * pi[1].z + wt[2] * pi[2].z + wt[3] * pi[3].z) * r);
		return true;
	}

	void prepare(const f4* acc, const Cam& cam, float j0, float j1, float pj0, float pj1, float pvs, float cD, float cI, int stationary)
	{
		for (int y = 0; y < h; y++) for (int x = 0; x < w; x++)
		{
			const int p = x + y * w;
			const f3 direct = xyz(acc[p]) * pvs;
			const u4 lf = features[p];
			const f4 lwp = worldPos[p];
			const f3 alb = rgb32_to_hdr_min1(lf.x);
			const f3 indirect = xyz(acc[p + w * h]) * pvs;
			const f3 reci = mk3(1.0f / alb.x, 1.0f / alb.y, 1.0f / alb.z);
			const f3 dm = direct * reci, im = indirect * reci;
			const f3 dl = mk3(std::min(dm.x, cD), std::min(dm.y, cD), std::min(dm.z, cD));
			const f3 il = mk3(std::min(im.x, cI), std::min(im.y, cI), std::min(im.z, cI));
			shading[p] = combine(dl, il);
			float lD = luminance(dl), lD2 = lD * lD, lI = luminance(il), lI2 = lI * lI;
			fl2 pp;
			if (((lf.w >> 4) & 3u) == 0)
			{
				const f3 dd = normalize(mk3(lwp.x - cam.pos.x, lwp.y - cam.pos.y, lwp.z - cam.pos.z));
				const f4 D = { dd.x, dd.y, dd.z, 0 };
				const float s = cam.pos.w / dot4(cam.E, D);
				const f4 S = { cam.pos.x + D.x * s, cam.pos.y + D.y * s, cam.pos.z + D.z * s, 0 };
				pp.x = dot4(S, cam.right) - cam.right.w - j0;
				pp.y = dot4(S, cam.up) - cam.up.w - j1;
			}
			else
			{
				pp.x = (float)x, pp.y = (float)y;
				if (!stationary)
				{
					const f4 q = prevWorldPos[p];
					float best = length(mk3(lwp.x - q.x, lwp.y - q.y, lwp.z - q.z)), step = 5.0f;
					const fl2 off = { j0 - pj0, j1 - pj1 };
					for (int iter = 0;;)
					{
						int tap = 0;
						const fl2 c[4] = { { pp.x - step, pp.y }, { pp.x + step, pp.y }, { pp.x, pp.y - step }, { pp.x, pp.y + step } };
						for (int i = 0; i < 4; i++)
						{
							const float d = fine_world_distance(fl2{ c[i].x + off.x, c[i].y + off.y }, lwp);
							if (d < best) best = d, pp = c[i], tap = i + 1;
						}
						if (tap == 0) { step *= 0.45f; if (step < 0.05f) break; }
						if (++iter == 25) break;
					}
				}
			}
			pp.x = pp.x + 0.5f, pp.y = pp.y + 0.5f;
			if (pp.x >= 0 && pp.x < (float)w && pp.y >= 0 && pp.y < (float)h)
			{
				const f4 hs = history(prevMoments, lwp, unpack_normal2(lf.y), pp.x, pp.y);
				if (hs.x > -1)
				{
					lD = 0.2f * lD + 0.8f * hs.x, lD2 = 0.2f * lD2 + 0.8f * hs.y;
					lI = 0.2f * lI + 0.8f * hs.z, lI2 = 0.2f * lI2 + 0.8f * hs.w;
					if ((features[p].w & 15u) < 15u) features[p].w++;
				}
				else features[p].w &= 0xfffffff0u;
			}
			else features[p].w &= 0xfffffff0u;
			motion[p] = pp;
			moments[p] = f4{ lD, lD2, lI, lI2 };
		}
	}

	void apply(int phase, const std::vector<f4>& A, const std::vector<f4>* B, std::vector<f4>& C, int last)
	{
		for (int y = 0; y < h; y++) for (int x = 0; x < w; x++)
		{
			const int p = x + y * w;
			const u4 lf = features[p];
			const f4 lpos = worldPos[p];
			const f3 ln = unpack_normal2(lf.y), lc = rgb32_to_hdr(lf.x);
			const uint lmat = lf.w >> 4;
			float dws = 1, iws = 1;
			const f4 comb = A[p];
			f3 dsum = get_direct(comb) * dws, isum = get_indirect(comb) * iws;
			const float lD = luminance(get_direct(comb)), lI = luminance(get_indirect(comb));
			const float ldepth = u2f(lf.z), ddx = deltaDepth[p].z, ddy = deltaDepth[p].w;
			const float sigma = 10.0f * u2f((uint)(127 - (phase - 1)) << 23);
			const float factor = (lf.w & 15u) == 0 ? 400.0f : 1.0f;
			const f4 m = moments[p];
			const float varD = m.y - m.x * m.x, varI = m.w - m.z * m.z;
			const float rD = -1.0f / (sigma * factor * sqrtf(varD + 0.00001f) + 0.00001f);
			const float rI = -1.0f / (sigma * factor * sqrtf(varI + 0.00001f) + 0.00001f);
			const int step = 1 << (phase - 1);
			for (int vv = -2; vv <= 2; vv++)
			{
				const int v = vv * step + y;
				const int r = std::abs(vv) == 2 ? 1 : 2;
				if (v < 0 || v >= h) continue;
				for (int uu = -r; uu <= r; uu++)
				{
					if (uu == 0 && vv == 0) continue;
					const int u = std::min(std::max(uu * step + x, 0), w - 1);
					const int q = u + v * w;
					const f4 c = A[q];
					const u4 nf = features[q];
					const float wdist = (float)(uu * uu + vv * vv) * (-1.0f / 7.5f);
					const f3 nd = get_direct(c), ni = get_indirect(c);
					float wn = pow128(std::max(0.0f, fdot(unpack_normal2(nf.y), ln)));
					const float expected = fmaf(ddy, (float)(vv * step), fmaf(ddx, (float)(uu * step), ldepth));
					const float err = fabsf(expected - u2f(nf.z));
					const float expDiff = fabsf(expected - ldepth);
					const float wdepth = err / std::max(0.00001f, (0.5f + (float)phase * 0.5f) * expDiff);
					if ((nf.w >> 4) != lmat) wn = (float)((double)wn * 0.0001);
					else wn = wn * fdot(lc, rgb32_to_hdr(nf.x));
					float wd = wn * cw_expf(fmaf(fabsf(lD - luminance(nd)), rD, wdist) - wdepth);
					float wi = wn * cw_expf(fmaf(fabsf(lI - luminance(ni)), rI, wdist) - wdepth);
					if (!std::isfinite(wd)) wd = 0;
					if (!std::isfinite(wi)) wi = 0;
					dsum = mk3(fmaf(nd.x, wd, dsum.x), fmaf(nd.y, wd, dsum.y), fmaf(nd.z, wd, dsum.z)), dws += wd;
					isum = mk3(fmaf(ni.x, wi, isum.x), fmaf(ni.y, wi, isum.y), fmaf(ni.z, wi, isum.z)), iws += wi;
				}
			}
			f3 df = dsum * (1.0f / std::max(0.0001f, dws));
			f3 inf = isum * (1.0f / std::max(0.0001f, iws));
			if (phase == 1)
			{
				const fl2 mv = motion[p];
				const int px = f2i_sat(mv.x), py = f2i_sat(mv.y);
				f3 pd, pi;
				if (px >= 0 && px < w && py >= 0 && py < h && history2(*B, lpos, ln, mv.x, mv.y, pd, pi))
				{
					pd = rgb_to_ycocg(pd), pi = rgb_to_ycocg(pi);
					f3 dAvg = rgb_to_ycocg(df), dVar = dAvg * dAvg, iAvg = rgb_to_ycocg(inf), iVar = iAvg * iAvg;
					const auto add = [&](int q) {
						const f3 f = rgb_to_ycocg(get_direct(A[q])), g = rgb_to_ycocg(get_indirect(A[q]));
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
					const f3 s0 = dVar - dAvg * dAvg, s1 = iVar - iAvg * iAvg;
					const f3 sD = mk3(sqrtf(std::max(0.0f, s0.x)), sqrtf(std::max(0.0f, s0.y)), sqrtf(std::max(0.0f, s0.z)));
					const f3 sI = mk3(sqrtf(std::max(0.0f, s1.x)), sqrtf(std::max(0.0f, s1.y)), sqrtf(std::max(0.0f, s1.z)));
					const f3 mnD = dAvg - 0.75f * sD, mxD = dAvg + 0.75f * sD, mnI = iAvg - 0.75f * sI, mxI = iAvg + 0.75f * sI;
					pd = mk3(clampf(pd.x, mnD.x, mxD.x), clampf(pd.y, mnD.y, mxD.y), clampf(pd.z, mnD.z, mxD.z));
					pi = mk3(clampf(pi.x, mnI.x, mxI.x), clampf(pi.y, mnI.y, mxI.y), clampf(pi.z, mnI.z, mxI.z));
					df = df * 0.1f + ycocg_to_rgb(pd) * 0.9f;
					inf = inf * 0.1f + ycocg_to_rgb(pi) * 0.9f;
				}
			}
			if (last)
			{
				const f3 c = (df + inf) * rgb32_to_hdr(lf.x);
				C[p] = f4{ sqrtf(c.x), sqrtf(c.y), sqrtf(c.z), 1 };
			}
			else C[p] = combine(df, inf);
		}
	}
};

/* prepareFilter's camera terms (finalize_shared.h:296-313), normalize = v * (1 / sqrt). */
static Cam camera(const lh2hip::ViewPyramid& v, int h)
{
	const f3 pos = mk3(v.pos.x, v.pos.y, v.pos.z), p1 = mk3(v.p1.x, v.p1.y, v.p1.z), p2 = mk3(v.p2.x, v.p2.y, v.p2.z), p3 = mk3(v.p3.x, v.p3.y, v.p3.z);
	const f3 centre = 0.5f * (p2 + p3);
	const f3 direction = normalize(centre - pos), right = normalize(p2 - p1), up = normalize(p3 - p1);
	const float lenReci = (float)h / length(p3 - p1);
	Cam c;
	c.pos = f4{ pos.x, pos.y, pos.z, -(dot(pos, direction) - dot(centre, direction)) };
	c.E = f4{ direction.x, direction.y, direction.z, 0 };
	const f3 r = right * lenReci, u = up * lenReci;
	c.right = f4{ r.x, r.y, r.z, dot(p1, right) * lenReci };
	c.up = f4{ u.x, u.y, u.z, dot(p1, up) * lenReci };
	return c;
}

} // namespace orc

using namespace orc;

OAPI void* orc_svgf_create(int w, int h) { return new Svgf(w, h); }
OAPI void orc_svgf_destroy(void* p) { delete (Svgf*)p; }

/* One frame with host arrays (layouts as lh2hip_svgf_frame); out: w*h float4, border untouched. */
OAPI void orc_svgf_frame(void* sp, const float* acc, const uint* feat, const float* wp, const float* dd, const lh2hip::ViewPyramid* prevView,
	float j0, float j1, float pj0, float pj1, int samplesTaken, float cD, float cI, int stationary, float* out, int taa)
{
	Svgf& s = *(Svgf*)sp;
	const int n = s.w * s.h;
	for (int i = 0; i < n; i++)
	{
		s.features[i] = u4{ feat[4 * i], feat[4 * i + 1], feat[4 * i + 2], (feat[4 * i + 3] & ~15u) | (s.features[i].w & 15u) };
		s.worldPos[i] = f4{ wp[4 * i], wp[4 * i + 1], wp[4 * i + 2], wp[4 * i + 3] };
		s.deltaDepth[i] = f4{ dd[4 * i], dd[4 * i + 1], dd[4 * i + 2], dd[4 * i + 3] };
	}
	s.prepare((const f4*)acc, camera(*prevView, s.h), j0, j1, pj0, pj1, 1.0f / (float)samplesTaken, cD, cI, stationary);
	s.apply(1, s.shading, &s.filteredIn, s.filteredOut, 0);
	s.apply(2, s.filteredOut, nullptr, s.filteredIn, 0);
	s.apply(3, s.filteredIn, nullptr, s.lastPass, 1);
	if (taa)
	{	/* TAApass + unsharpenTAA (RenderCore_Optix7Filter/rendercore.cpp:922-926) */
		s.taa_pass();
		s.unsharpen(out);
		std::swap(s.taaOut, s.prevPixels);
	}
	else for (int y = 1; y < s.h - 1; y++) for (int x = 1; x < s.w - 1; x++)
	{
		const f4 v = s.lastPass[x + y * s.w];
		float* o = out + 4 * (size_t)(x + y * s.w);
		o[0] = sqrtf(v.x), o[1] = sqrtf(v.y), o[2] = sqrtf(v.z), o[3] = 0;
	}
	std::swap(s.filteredIn, s.filteredOut);
	std::swap(s.prevWorldPos, s.worldPos);
	std::swap(s.moments, s.prevMoments);
}

/* One a-trous pass of applyFilterKernel (finalize_shared.h:320-484) on given
 * inputs, for closed-form checks of the edge-stopping weights. A, C: w*h
 * float4 in the 5:11 combined format (C: sqrt image with last = 1); feat,
 * wp, dd, moments as in orc_svgf_frame. Motion vectors point outside the
 * image, so phase 1 takes no history and reduces to the spatial pass. */
OAPI void orc_svgf_pass(void* sp, int phase, const float* A, const uint* feat, const float* wp, const float* dd, const float* mom, float* C, int last)
{
	Svgf& s = *(Svgf*)sp;
	const int n = s.w * s.h;
	std::vector<f4> a(n), c(n, f4{ 0, 0, 0, 0 });
	for (int i = 0; i < n; i++)
	{
		s.features[i] = u4{ feat[4 * i], feat[4 * i + 1], feat[4 * i + 2], feat[4 * i + 3] };
		s.worldPos[i] = f4{ wp[4 * i], wp[4 * i + 1], wp[4 * i + 2], wp[4 * i + 3] };
		s.deltaDepth[i] = f4{ dd[4 * i], dd[4 * i + 1], dd[4 * i + 2], dd[4 * i + 3] };
		s.moments[i] = f4{ mom[4 * i], mom[4 * i + 1], mom[4 * i + 2], mom[4 * i + 3] };
		s.motion[i] = fl2{ -1e9f, -1e9f };
		a[i] = f4{ A[4 * i], A[4 * i + 1], A[4 * i + 2], A[4 * i + 3] };
	}
	s.apply(phase, a, &s.filteredIn, c, last);
	memcpy(C, c.data(), (size_t)n * 16);
}

OAPI void orc_svgf_read(void* sp, int which, void* out)
{
	Svgf& s = *(Svgf*)sp;
	const size_t n = (size_t)s.w * s.h;
	if (which == 0) memcpy(out, s.features.data(), n * 16);
	else if (which == 1) memcpy(out, s.prevMoments.data(), n * 16);
	else if (which == 2) memcpy(out, s.motion.data(), n * 8);
	else if (which == 3) memcpy(out, s.filteredIn.data(), n * 16);
	else if (which == 4) memcpy(out, s.lastPass.data(), n * 16);
	else if (which == 5) memcpy(out, s.prevWorldPos.data(), n * 16);
	else if (which == 6) memcpy(out, s.prevPixels.data(), n * 16);
}
