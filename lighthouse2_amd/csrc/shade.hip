/* shade.hip - wavefront kernels of the reference estimator on gfx950.
 *
 * Follows RenderCore_PrimeRef ("REFERENCE, for image validation"):
 *   k_generate  generateEyeRaysKernel    lib/RenderCore_PrimeRef/kernels/camera.h:39-68
 *   k_shade     shadeKernel              lib/RenderCore_PrimeRef/kernels/pathtracer.h:45-168
 *               GetShadingData/SetupFrame lib/CUDA/shared_kernel_code/material_shared.h:42-206
 *               FetchTexel(Trilinear)    lib/CUDA/shared_kernel_code/sampling_shared.h:35-104
 *               SampleSkydome            lib/CUDA/shared_kernel_code/tools_shared.h:184-202
 *               RandomPointOnLight       lib/CUDA/shared_kernel_code/lights_shared.h:225-313
 *               Lambert+specular BSDF    lib/RenderCore_PrimeRef/kernels/bsdf.h:19-101
 *   k_connect   finalizeConnectionKernel lib/RenderCore_PrimeRef/kernels/connections.h:22-34
 *   k_finalize  finalizeRenderKernel     lib/CUDA/shared_kernel_code/finalize_shared.h:29-45
 * Differences of form: contributions go to a per-path slot (path slots are
 * owned by one path, so no two lanes ever add to the same address) and
 * k_resolve adds the slots to the pixel accumulators in sample order, giving
 * a deterministic image for any spp; the path index has its own 32-bit lane
 * (no 2^24 cap); light selection sums potentials in a first pass and
 * recomputes them while searching (no 64-entry stack array).
 */
#include "shade.h"
#include "sample.h"

namespace lh {

#define PI_F 3.14159265358979323846264f
#define INVPI_F 0.31830988618379067153777f
#define TWOPI_F 6.28318530717958647692528f
#define EPSILON_F 0.0001f
#define S_SPECULAR 1u

struct Shading
{
	v3 color, transmittance;
	uint4 params;
	float4 tint;   /* tint rgb and luminance of the untextured base colour (Disney BSDF) */
};
__device__ __forceinline__ float char2flt(uint32_t a, int s) { return ((float)((a >> s) & 255u)) * (1.0f / 255.0f); }
__device__ __forceinline__ float roughness(const Shading& sd) { const float r = char2flt(sd.params.x, 24); return r > 0.001f ? r : 0.001f; }
__device__ __forceinline__ float transmission(const Shading& sd) { return char2flt(sd.params.z, 16); }
__device__ __forceinline__ float eta(const Shading& sd) { return bitsf(sd.params.w); }

__device__ __forceinline__ v3 consistent_normal(v3 D, v3 iN, float alpha)
{
	const float t = PI_F - 2 * alpha, q = (t * t) / (PI_F * (PI_F + (2 * PI_F - 4) * alpha));
	const float b = dot(D, iN), g = 1 + q * (b - 1), rho = sqrtf(q * (1 + g) / (1 + b));
	const v3 Rc = (g + rho * b) * iN - (rho * D);
	return normalize(D + Rc);
}

/* __uchar4_to_float4 (sampling_shared.h:29-33): byte k of the texel word is
 * component k, scaled by 1/256. */
__device__ __forceinline__ float4 unpack_texel(uint32_t t)
{
	const float r = 1.0f / 256.0f;
	return make_float4((float)(t & 255u) * r, (float)((t >> 8) & 255u) * r, (float)((t >> 16) & 255u) * r, (float)(t >> 24) * r);
}

__device__ __forceinline__ float4 texel_at(const ShadeParams& P, int storage, int idx)
{
	if (storage == 1) return P.argb128[idx];
	return unpack_texel(storage == 2 ? P.nrm32[idx] : P.argb32[idx]);
}

/* FetchTexel, BILINEAR on (lib/CUDA/shared_kernel_code/sampling_shared.h:35-79). */
__device__ float4 fetch_texel(const ShadeParams& P, float tu, float tv, int o, int w, int h, int storage)
{
	const float tcx = (fmaxf(tu + 1000.0f, 0.0f) * (float)w) - 0.5f;
	const float tcy = (fmaxf(tv + 1000.0f, 0.0f) * (float)h) - 0.5f;
	const int iu = f2i_sat(tcx) % w, iv = f2i_sat(tcy) % h;
	const float fu = tcx - floorf(tcx), fv = tcy - floorf(tcy);
	const float w0 = (1 - fu) * (1 - fv), w1 = fu * (1 - fv), w2 = (1 - fu) * fv, w3 = 1 - (w0 + w1 + w2);
	const int iu1 = (iu + 1) % w, iv1 = (iv + 1) % h;
	const float4 p0 = texel_at(P, storage, o + iu + iv * w), p1 = texel_at(P, storage, o + iu1 + iv * w);
	const float4 p2 = texel_at(P, storage, o + iu + iv1 * w), p3 = texel_at(P, storage, o + iu1 + iv1 * w);
	return make_float4(p0.x * w0 + p1.x * w1 + p2.x * w2 + p3.x * w3, p0.y * w0 + p1.y * w1 + p2.y * w2 + p3.y * w3,
		p0.z * w0 + p1.z * w1 + p2.z * w2 + p3.z * w3, p0.w * w0 + p1.w * w1 + p2.w * w2 + p3.w * w3);
}

/* FetchTexelTrilinear (sampling_shared.h:81-104), 5 MIP levels. */
__device__ float4 fetch_texel_trilinear(const ShadeParams& P, float lambda, float tu, float tv, int offset, int width, int height)
{
	int level0 = 0, level1 = 0;
	float f = 0;
	if (lambda >= 0) level0 = min(4, f2i_sat(lambda)), level1 = min(4, level0 + 1), f = lambda - floorf(lambda);
	const float scale = (float)(width * height) * 1.3333333333f;
	const int o0 = offset + f2i_sat(scale * (1 - bitsf((uint32_t)(127 - 2 * level0) << 23)));
	const int o1 = offset + f2i_sat(scale * (1 - bitsf((uint32_t)(127 - 2 * level1) << 23)));
	const float4 p0 = fetch_texel(P, tu, tv, o0, width >> level0, height >> level0, 0);
	const float4 p1 = fetch_texel(P, tu, tv, o1, width >> level1, height >> level1, 0);
	return make_float4((1 - f) * p0.x + f * p1.x, (1 - f) * p0.y + f * p1.y, (1 - f) * p0.z + f * p1.z, (1 - f) * p0.w + f * p1.w);
}

/* SampleSkydome(-worldToSky.TransformVector(D)) (tools_shared.h:184-202;
 * common_types.h:608-613; kernels/pathtracer.h:73-78). */
__device__ v3 sample_sky(const ShadeParams& P, v3 D)
{
	if (P.skyW <= 0 || P.skyH <= 0) return mk3(0, 0, 0);
	const float4 a = P.w2s0, b = P.w2s1, c = P.w2s2;
	const v3 d = mk3(-(a.x * D.x + a.y * D.y + a.z * D.z), -(b.x * D.x + b.y * D.y + b.z * D.z), -(c.x * D.x + c.y * D.y + c.z * D.z));
	const float p = lh_atan2f(d.y, d.x);
	const float phi = p < 0 ? (p + 2 * PI_F) : p;
	const float theta = lh_acosf(fminf(fmaxf(d.z, -1.0f), 1.0f));
	const uint32_t u = f2u_sat((float)P.skyW * phi * 0.15915494309189533577f - 0.5f);
	const uint32_t v = f2u_sat((float)P.skyH * theta * 0.31830988618379067154f - 0.5f);
	const uint32_t idx = u + v * (uint32_t)P.skyW;
	if (idx >= (uint32_t)P.skyW * (uint32_t)P.skyH) return mk3(0, 0, 0);
	return xyz(P.sky[idx]);
}

__device__ __forceinline__ float h2f(uint32_t bits16) { return half_to_float(bits16 & 0xffffu); }

/* CoreTri as 13 float4 (common_classes.h:133-163); SetupFrame + GetShadingData
 * (material_shared.h:42-206) with the texture maps: diffuse (trilinear, LOD
 * TRI_LOD + log2(coneWidth / |D.N|); alpha below 0.5 ends the lookup with the
 * untextured colour), detail diffuse, normal maps (tangent frame T, B of the
 * triangle) and roughness; the normal-map scale's __expf is cw_expf. */
template <bool PRIME>
__device__ bool shading_data(const ShadeParams& P, v3 D, float u, float v, float coneWidth, const float4* tri, int instIdx,
	Shading& sd, v3& N, v3& iN, v3& fN, v3& T)
{
	const float4 t1 = tri[1], t2 = tri[2], t3 = tri[3], t4 = tri[4], t5 = tri[5], t7 = tri[7];
	const GpuMaterial& mat = P.mats[fbits(t1.w)];
	const uint32_t flags = mat.flags;
	sd.color = mk3(half_to_float(mat.diffuse_rg & 0xffffu), half_to_float(mat.diffuse_rg >> 16), half_to_float(mat.diffuse_b_medium_r & 0xffffu));
	if (!PRIME)
	{	/* tint4 (material_shared.h:19-33,118-119) */
		const v3 c = sd.color;
		const v3 X = mk3(fmaxf(0.0f, 0.412453f * c.x + 0.357580f * c.y + 0.180423f * c.z), fmaxf(0.0f, 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z),
			fmaxf(0.0f, 0.019334f * c.x + 0.119193f * c.y + 0.950227f * c.z));
		v3 t = mk3(1, 1, 1);
		if (X.y > 0)
		{
			const v3 n = X * (1.0f / X.y);
			t = mk3(fmaxf(0.0f, 3.240479f * n.x - 1.537150f * n.y - 0.498535f * n.z), fmaxf(0.0f, -0.969256f * n.x + 1.875992f * n.y + 0.041556f * n.z),
				fmaxf(0.0f, 0.055648f * n.x - 0.204043f * n.y + 1.057311f * n.z));
		}
		sd.tint = make_float4(t.x, t.y, t.z, X.y);
	}
	sd.transmittance = mk3(half_to_float(mat.diffuse_b_medium_r >> 16), half_to_float(mat.medium_gb & 0xffffu), half_to_float(mat.medium_gb >> 16));
	sd.params = mat.params;
	N = iN = fN = mk3(t2.w, t3.w, t4.w);
	T = xyz(t5);
	const float w = 1 - (u + v);
	if (flags & MF_HASSMOOTHNORMALS)
		iN = PRIME ? normalize(u * xyz(t2) + v * xyz(t3) + w * xyz(t4)) : normalize(w * xyz(t2) + u * xyz(t3) + v * xyz(t4));
	const ShadeInst& I = P.inst[instIdx];
	const v3 A = xyz(I.inv0), B = xyz(I.inv1), C = xyz(I.inv2);
	N = normalize(N.x * A + N.y * B + N.z * C);
	iN = normalize(iN.x * A + iN.y * B + iN.z * C);
	if (PRIME)
	{	/* CONSISTENTNORMALS, on in RenderCore_PrimeRef only */
		const bool backSide = dot(D, N) > 0;
		const float alpha = u * t7.x + v * t7.y + w * t7.z;
		iN = (backSide ? -1.0f : 1.0f) * consistent_normal(D * -1.0f, backSide ? (iN * -1.0f) : iN, alpha);
	}
	fN = iN;
	if (!(flags & (MF_HASDIFFUSEMAP | MF_HAS2NDDIFFUSEMAP | MF_HASSPECULARITYMAP | MF_HASNORMALMAP | MF_HAS2NDNORMALMAP | MF_HASROUGHNESSMAP))) return false;
	const float4 t0 = tri[0];
	const float ww = 1 - (u + v);
	const float tu = PRIME ? u * t0.x + v * t0.y + ww * t0.z : ww * t0.x + u * t0.y + v * t0.z;
	const float tv = PRIME ? u * t1.x + v * t1.y + ww * t1.z : ww * t1.x + u * t1.y + v * t1.z;
	if (flags & MF_HASDIFFUSEMAP)
	{
		const float lambda = t7.w + lh_log2f(coneWidth * (1.0f / fabsf(dot(D, N))));
		const uint4 m = mat.t0;
		const float x = h2f(m.y) * (h2f(m.z) + tu), y = h2f(m.y >> 16) * (h2f(m.z >> 16) + tv);
		const float4 texel = fetch_texel_trilinear(P, lambda, x, y, (int)m.w, (int)(m.x & 0xffffu), (int)(m.x >> 16));
		if (texel.w < 0.5f) return true;
		sd.color = sd.color * xyz(texel);
		if (flags & MF_HAS2NDDIFFUSEMAP)
		{
			const uint4 m1 = mat.t1;
			const float x1 = h2f(m1.y) * (h2f(m1.z) + tu), y1 = h2f(m1.y >> 16) * (h2f(m1.z >> 16) + tv);
			const float4 tx = fetch_texel(P, x1, y1, (int)m1.w, (int)(m1.x & 0xffffu), (int)(m1.x >> 16), 0);
			sd.color = sd.color + (xyz(tx) - mk3(0.5f, 0.5f, 0.5f));
		}
	}
	if (flags & MF_HASNORMALMAP)
	{
		const v3 Bt = xyz(tri[6]);
		const uint32_t part3 = mat.medium_gb;
		const float c0 = (float)((part3 >> 8) & 255u) - 128.0f;
		const float n0scale = copysignf(-0.0001f + 0.0001f * cw_expf(0.1f * fabsf(c0)), c0);
		const uint4 m = mat.n0;
		const float x = h2f(m.y) * (h2f(m.z) + tu), y = h2f(m.y >> 16) * (h2f(m.z >> 16) + tv);
		const float4 tx = fetch_texel(P, x, y, (int)m.w, (int)(m.x & 0xffffu), (int)(m.x >> 16), 2);
		v3 sn = (xyz(tx) - mk3(0.5f, 0.5f, 0.5f)) * 2.0f;
		sn.x = sn.x * n0scale, sn.y = sn.y * n0scale;
		if (flags & MF_HAS2NDNORMALMAP)
		{
			const float c1 = (float)((part3 >> 16) & 255u) - 128.0f;
			const float n1scale = copysignf(-0.0001f + 0.0001f * cw_expf(0.1f * c1), c1);
			const uint4 m1 = mat.n1;
			const float x1 = h2f(m1.y) * (h2f(m1.z) + tu), y1 = h2f(m1.y >> 16) * (h2f(m1.z >> 16) + tv);
			const float4 t1x = fetch_texel(P, x1, y1, (int)m1.w, (int)(m1.x & 0xffffu), (int)(m1.x >> 16), 2);
			v3 l1 = (xyz(t1x) - mk3(0.5f, 0.5f, 0.5f)) * 2.0f;
			l1.x = l1.x * n1scale, l1.y = l1.y * n1scale;
			sn = sn + l1;
		}
		sn = normalize(sn);
		fN = normalize(sn.x * T + sn.y * Bt + sn.z * iN);
	}
	if (flags & MF_HASROUGHNESSMAP)
	{
		const uint4 m = mat.r;
		const float x = h2f(m.y) * (h2f(m.z) + tu), y = h2f(m.y >> 16) * (h2f(m.z >> 16) + tv);
		const float4 tx = fetch_texel(P, x, y, (int)m.w, (int)(m.x & 0xffffu), (int)(m.x >> 16), 0);
		sd.params.x = (sd.params.x & 0x00ffffffu) + ((uint32_t)f2i_sat(tx.y * 255.0f) << 24);
		sd.params.x = (sd.params.x & 0xffffff00u) + (uint32_t)f2i_sat(tx.x * 255.0f);
	}
	return false;
}

__device__ v3 random_barycentrics(float r0)
{
	const uint32_t uf = (uint32_t)(uint64_t)(r0 * 4294967296.0f);
	float Ax = 1, Ay = 0, Bx = 0, By = 1, Cx = 0, Cy = 0;
	for (int i = 0; i < 16; ++i)
	{
		const int d = (uf >> (2 * (15 - i))) & 0x3;
		float Anx, Any, Bnx, Bny, Cnx, Cny;
		if (d == 0) { Anx = (Bx + Cx) * 0.5f, Any = (By + Cy) * 0.5f; Bnx = (Ax + Cx) * 0.5f, Bny = (Ay + Cy) * 0.5f; Cnx = (Ax + Bx) * 0.5f, Cny = (Ay + By) * 0.5f; }
		else if (d == 1) { Anx = Ax, Any = Ay; Bnx = (Ax + Bx) * 0.5f, Bny = (Ay + By) * 0.5f; Cnx = (Ax + Cx) * 0.5f, Cny = (Ay + Cy) * 0.5f; }
		else if (d == 2) { Anx = (Bx + Ax) * 0.5f, Any = (By + Ay) * 0.5f; Bnx = Bx, Bny = By; Cnx = (Bx + Cx) * 0.5f, Cny = (By + Cy) * 0.5f; }
		else { Anx = (Cx + Ax) * 0.5f, Any = (Cy + Ay) * 0.5f; Bnx = (Cx + Bx) * 0.5f, Bny = (Cy + By) * 0.5f; Cnx = Cx, Cny = Cy; }
		Ax = Anx, Ay = Any, Bx = Bnx, By = Bny, Cx = Cnx, Cy = Cny;
	}
	const float rx = (Ax + Bx + Cx) * 0.3333333f, ry = (Ay + By + Cy) * 0.3333333f;
	return mk3(rx, ry, 1 - rx - ry);
}

/* Light records do not change during a launch. Read through the constant
 * address space, the potential loops (every lane at the same light) fetch
 * them with scalar loads instead of one vector load per lane. */
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) float4* ConstF4;
#else
typedef const float4* ConstF4;   /* the host pass only parses the device functions */
#endif
__device__ __forceinline__ ConstF4 as_const(const float4* p) { return (ConstF4)p; }

/* dot product with its two adds fused into the products */
__device__ __forceinline__ float fdot3(v3 a, v3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
/* A triangle light's potential seen along Lv (light point - shading point):
 * energy max(0, -LN.L) max(0, N.L) / |Lv|^2 with L = Lv / |Lv|, its dot
 * products fused. */
__device__ __forceinline__ float tri_potential(float energy, v3 LN, v3 N, v3 Lv)
{
	const float d2 = fdot3(Lv, Lv);
	const float att = 1.0f / d2;
	Lv = Lv * (1.0f / sqrtf(d2));
	return energy * fmaxf(0.0f, -fdot3(LN, Lv)) * fmaxf(0.0f, fdot3(N, Lv)) * att;
}

__device__ __forceinline__ float potential(const ShadeParams& P, int i, v3 I, v3 N, v3 bary)
{
	if (i < P.nTri)
	{
		const ConstF4 L = as_const(P.triLights) + 6 * i;
		const float4 d0 = L[0], d1 = L[1];
		v3 Lv = I;
		if (bary.x >= 0)
		{	/* the barycentric point, fused */
			const float4 a = L[3], b = L[4], c = L[5];
			Lv = mk3(fmaf(bary.z, c.x, fmaf(bary.y, b.x, bary.x * a.x)), fmaf(bary.z, c.y, fmaf(bary.y, b.y, bary.x * a.y)),
				fmaf(bary.z, c.z, fmaf(bary.y, b.z, bary.x * a.z)));
		}
		return tri_potential(d0.w, xyz(d1), N, Lv - I);
	}
	i -= P.nTri;
	if (i < P.nPoint)
	{
		const float4 d0 = as_const(P.pointLights)[2 * i];
		const v3 Lv = xyz(d0) - I;
		const float NdotL = fmaxf(0.0f, dot(N, normalize(Lv)));
		const float att = 1.0f / dot(Lv, Lv);
		return d0.w * NdotL * att;
	}
	i -= P.nPoint;
	if (i < P.nSpot)
	{
		const ConstF4 S = as_const(P.spotLights) + 3 * i;
		const float4 d0 = S[0], d1 = S[1], d2 = S[2];
		v3 Lv = xyz(d0) - I;
		const float att = 1.0f / dot(Lv, Lv);
		Lv = normalize(Lv);
		const float d = (fmaxf(0.0f, -dot(Lv, xyz(d2))) - d1.w) / (d0.w - d1.w);
		const float NdotL = fmaxf(0.0f, dot(N, Lv));
		const float LNdotL = fmaxf(0.0f, fminf(1.0f, d));
		return (d1.x + d1.y + d1.z) * LNdotL * NdotL * att;
	}
	i -= P.nSpot;
	const float4 d0 = as_const(P.dirLights)[2 * i];
	const float LNdotL = fmaxf(0.0f, -(d0.x * N.x + d0.y * N.y + d0.z * N.z));
	return d0.w * LNdotL;
}

__device__ v3 random_point_on_light(const ShadeParams& P, float r0, float r1, v3 I, v3 N, float& pickProb, float& lightPdf, v3& lightColor)
{
	const int lightCount = P.nTri + P.nPoint + P.nSpot + P.nDir;
	const v3 bary = random_barycentrics(r0);
	/* One pass sums the potentials and keeps the running sum at the end of
	 * each of 8 equal segments of the light list. The search for the first
	 * light whose running sum reaches r1 then starts from the checkpoint
	 * before the segment that holds it: the sums are formed by the same
	 * additions in the same order as a full second pass (potentials are
	 * non-negative, so running sums never decrease), at most one segment of
	 * potentials is evaluated again. */
	const int seg = (lightCount + 7) >> 3;
	float ck[8];
	float sum = 0;
#pragma unroll
	for (int j = 0; j < 8; j++)
	{
		const int e = min((j + 1) * seg, lightCount);
		for (int i = j * seg; i < e; i++) sum += potential(P, i, I, N, bary);
		ck[j] = sum;
	}
	if (sum <= 0) { lightPdf = 0; return mk3(1, 1, 1); }
	r1 *= sum;
	int first = 7;
	float total = ck[6];
#pragma unroll
	for (int j = 6; j >= 0; j--)
		if (ck[j] >= r1) first = j, total = j > 0 ? ck[j - 1] : 0.0f;
	int lightIdx = 0;
	float pl = 0;
	const int e = min((first + 1) * seg, lightCount);
	for (int i = first * seg; i < e; i++)
	{
		const float c = potential(P, i, I, N, bary);
		total += c;
		if (total >= r1) { lightIdx = i; pl = c; break; }
	}
	if (lightIdx == 0) pl = potential(P, 0, I, N, bary);
	pickProb = pl / sum;
	if (lightIdx < P.nTri)
	{
		const float4* L = P.triLights + 6 * lightIdx;
		const float4 d1 = L[1];
		lightColor = xyz(L[2]);
		const v3 LN = xyz(d1);
		const v3 Pp = bary.x * xyz(L[3]) + bary.y * xyz(L[4]) + bary.z * xyz(L[5]);
		v3 Lv = I - Pp;
		const float sqDist = dot(Lv, Lv);
		Lv = normalize(Lv);
		const float LNdotL = Lv.x * LN.x + Lv.y * LN.y + Lv.z * LN.z;
		const float reciSolidAngle = sqDist / (d1.w * LNdotL);
		lightPdf = (LNdotL > 0 && dot(Lv, N) < 0) ? reciSolidAngle : 0;
		return Pp;
	}
	int k = lightIdx - P.nTri;
	if (k < P.nPoint)
	{
		const float4 d0 = P.pointLights[2 * k], d1 = P.pointLights[2 * k + 1];
		const v3 Pp = xyz(d0), Lv = Pp - I;
		const float sqDist = dot(Lv, Lv);
		lightColor = xyz(d1) / sqDist;
		lightPdf = dot(Lv, N) > 0 ? 1 : 0;
		return Pp;
	}
	k -= P.nPoint;
	if (k < P.nSpot)
	{
		const float4 d0 = P.spotLights[3 * k], d1 = P.spotLights[3 * k + 1], d2 = P.spotLights[3 * k + 2];
		const v3 Pp = xyz(d0);
		v3 Lv = I - Pp;
		const float sqDist = dot(Lv, Lv);
		Lv = normalize(Lv);
		const float d = (fmaxf(0.0f, Lv.x * d2.x + Lv.y * d2.y + Lv.z * d2.z) - d1.w) / (d0.w - d1.w);
		const float LNdotL = fminf(1.0f, d);
		lightPdf = (LNdotL > 0 && dot(Lv, N) < 0) ? (sqDist / LNdotL) : 0;
		lightColor = xyz(d1);
		return Pp;
	}
	k -= P.nSpot;
	const float4 d0 = P.dirLights[2 * k], d1 = P.dirLights[2 * k + 1];
	const v3 Ld = xyz(d0);
	lightColor = xyz(d1);
	const float NdotL = dot(Ld, N);
	lightPdf = NdotL < 0 ? 1 : 0;
	return I - 1000.0f * Ld;
}

__device__ __forceinline__ v3 reflect(v3 i, v3 n) { return i - 2.0f * n * dot(n, i); }

/* Fr_L; the shared Lambert BSDF of the optix7 core (sharedBSDFs/lambert.h:32-45)
 * also takes a view vector below the surface (inverted eta, |VDotN|). */
template <bool BELOW = false>
__device__ float fr_l(float VDotN, float eio)
{
	if (BELOW && VDotN < 0.0f) eio = 1.0f / eio, VDotN = fabsf(VDotN);
	const float SinThetaT2 = (eio * eio) * (1.0f - VDotN * VDotN);
	if (SinThetaT2 > 1.0f) return 1.0f;
	const float LDotN = sqrtf(1.0f - SinThetaT2);
	const float r1 = (VDotN - eio * LDotN) / (VDotN + eio * LDotN);
	const float r2 = (LDotN - eio * VDotN) / (LDotN + eio * VDotN);
	return 0.5f * (r1 * r1 + r2 * r2);
}

__device__ bool refract_l(v3 wi, v3 n, float et, v3& wt)
{
	const float cosThetaI = fabsf(dot(n, wi));
	const float sin2ThetaI = fmaxf(0.0f, 1.0f - cosThetaI * cosThetaI);
	const float sin2ThetaT = et * et * sin2ThetaI;
	if (sin2ThetaT >= 1) return false;
	const float cosThetaT = sqrtf(1.0f - sin2ThetaT);
	wt = et * (wi * -1.0f) + (et * cosThetaI - cosThetaT) * n;
	return true;
}

__device__ __forceinline__ v3 evaluate_bsdf(const Shading& sd, v3 iN, v3 wi, float& pdf)
{
	if (transmission(sd) > 0.999f || roughness(sd) <= 0.001f) { pdf = 0; return mk3(0, 0, 0); }
	pdf = fabsf(dot(wi, iN)) * INVPI_F;
	return sd.color * INVPI_F * roughness(sd);
}

template <bool SHARED = false>
__device__ v3 sample_bsdf(const Shading& sd, v3 iN, v3 N, v3 wo, float distance, float r3, float r4, v3& wi, float& pdf, bool& specular)
{
	const float flip = (dot(wo, N) < 0) ? -1 : 1;
	iN = iN * flip;
	specular = true, pdf = 1;
	v3 bsdf;
	const float TR = transmission(sd), RO = roughness(sd);
	if (r4 < TR)
	{
		const float eio = flip < 0 ? (1.0f / eta(sd)) : eta(sd), F = fr_l<SHARED>(dot(iN, wo), eio);
		v3 beer = mk3(1, 1, 1);
		beer.x = cw_expf(-sd.transmittance.x * distance * 2.0f);
		beer.y = cw_expf(-sd.transmittance.y * distance * 2.0f);
		beer.z = cw_expf(-sd.transmittance.z * distance * 2.0f);
		if (r3 < F)
		{
			wi = reflect(wo * -1.0f, iN);
			bsdf = sd.color * beer * (1 / fabsf(dot(iN, wi)));
		}
		else
		{
			if (!refract_l(wo, iN, eio, wi)) return mk3(0, 0, 0);
			return sd.color * beer * (1 / fabsf(dot(iN, wi)));
		}
	}
	else
	{
		const float pReflect = 1 - RO;
		if (r3 < pReflect)
		{
			wi = reflect(wo * -1.0f, iN);
			bsdf = sd.color * (1.0f / fabsf(dot(iN, wi)));
		}
		else
		{
			const float r5 = (r3 - pReflect) / (1 - pReflect);
			const float r6 = (r4 - TR) / (1 - TR);
			wi = normalize(tangent2world(cosine_sample(r5, r6), iN));
			pdf = fmaxf(0.0f, dot(wi, iN)) * INVPI_F;
			specular = false;
			bsdf = sd.color * INVPI_F;
		}
	}
	if (dot(N * flip, wi) <= 0) pdf = 0;
	return bsdf;
}

__device__ __forceinline__ v3 safe_origin(v3 O, v3 R, v3 N, float eps) { return O + N * (dot(N, R) > 0 ? eps : -eps); }

__device__ __forceinline__ void acc_add(float4* a, v3 c)
{
	float4 v = *a;
	v.x = v.x + c.x, v.y = v.y + c.y, v.z = v.z + c.z;
	*a = v;
}

__global__ void k_generate(CamParams c, uint32_t R0, uint32_t jobBase, int n, float eps, float4* __restrict__ rays,
	float4* __restrict__ state, uint32_t* __restrict__ pathId, float4* __restrict__ pathAcc, uint32_t* __restrict__ counters,
	uint32_t slotBase, uint32_t total, int storeSlot)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j == 0 && slotBase == 0) counters[0] = total, counters[1] = 0, counters[2] = 0;
	if (j >= n) return;
	const uint32_t job = jobBase + (uint32_t)j;
	const uint32_t o = slotBase + (uint32_t)j;
	v3 O, D;
	eye_ray(c, (int)job, R0, O, D);
	rays[2 * o] = make_float4(O.x, O.y, O.z, eps);
	rays[2 * o + 1] = make_float4(D.x, D.y, D.z, 1e34f);
	state[o] = make_float4(1.0f, 1.0f, 1.0f, bitsf(S_SPECULAR));
	pathId[o] = storeSlot ? o : job;
	pathAcc[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

/* The constants of the job a pass slot belongs to (passes shared by several
 * jobs): binary search of the segment table by slot, then the job's own
 * size, sample offset, seeds, camera and G-buffer in place of the pass's. */
template <bool MIS>
__device__ __forceinline__ void job_view(ShadeParams& P, uint32_t slot, uint32_t& pathIdx, uint32_t& slotOut)
{
	int g = 0;
	for (int step = JOB_SEGS_MAX / 2; step > 0; step >>= 1)
		if (g + step < P.nseg && P.segs[g + step].slotBase <= slot) g += step;
	const JobSeg& S = P.segs[g];
	pathIdx = S.jobBase + (slot - S.slotBase);
	slotOut = slot;
	P.w = S.w, P.h = S.h, P.pass = S.pass, P.maxDepth = S.maxDepth, P.spreadAngle = S.spreadAngle;
	P.R0 = MIS ? S.R0w[P.pathLength & 7] : P.R0 + S.R0add;
	P.shift = S.shift, P.blueSlot = S.blueSlot;
	P.gFeatures = S.gFeatures, P.gWorldPos = S.gWorldPos, P.gDeltaDepth = S.gDeltaDepth;
	P.camPos = S.camPos, P.camP1 = S.camP1, P.camRight = S.camRight, P.camUp = S.camUp;
	P.probe = nullptr;
}

__device__ __forceinline__ void shade_path(const ShadeParams& P, const int j, const uint32_t pathIdx, const uint32_t slot, const uint32_t stored,
	const float4* __restrict__ raysIn, const float4* __restrict__ stateIn,
	const uint32_t* __restrict__ pathIn, const float4* __restrict__ hits, float4* __restrict__ raysOut, float4* __restrict__ stateOut,
	uint32_t* __restrict__ pathOut, float4* __restrict__ shadowRays, float4* __restrict__ shadowPot, float4* __restrict__ pathAcc,
	uint32_t* __restrict__ counters)
{
	const float4 O4 = raysIn[2 * j], D4 = raysIn[2 * j + 1], T4 = stateIn[j];
	const float4 hd = hits[j];
	float4* acc = pathAcc + slot;
	const int wh = P.w * P.h;
	const uint32_t pixelIdx = pathIdx % (uint32_t)wh;
	const uint32_t sampleIdx = pathIdx / (uint32_t)wh + (uint32_t)P.pass;
	uint32_t flags = fbits(T4.w);
	const v3 D = xyz(D4);
	v3 throughput = xyz(T4);
	const uint32_t prim = fbits(hd.z);
	const int instIdx = (int)fbits(hd.y);
	const float t = hd.w;
	if (P.pathLength == 1) acc->w = acc->w + (prim == 0xffffffffu ? 10000.0f : t);
	if (prim == 0xffffffffu)
	{
		acc_add(acc, throughput * sample_sky(P, D));
		return;
	}
	if (P.probe && (int)pixelIdx == P.probePixel && P.pathLength == 1 && sampleIdx == 0)
		P.probe[0] = instIdx, P.probe[1] = (int)prim, P.probe[2] = (int)fbits(t);
	const uint32_t uvb = fbits(hd.x);
	const float u = (float)(uvb & 65535u) * (1.0f / 65535.0f), v = (float)(uvb >> 16) * (1.0f / 65535.0f);
	const ShadeInst& SI = P.inst[instIdx];
	const float4* tri = SI.coreTris + 13 * (size_t)prim;
	Shading sd;
	v3 N, iN, fN, T;
	const v3 I = xyz(O4) + t * D;
	shading_data<true>(P, D, u, v, P.spreadAngle * t, tri, instIdx, sd, N, iN, fN, T);
	if (sd.color.x > 1.0f || sd.color.y > 1.0f || sd.color.z > 1.0f)
	{
		const float DdotNL = -dot(D, N);
		if (DdotNL > 0)
		{
			const v3 c = throughput * sd.color;
			if (P.pathLength == 1 || (flags & S_SPECULAR)) acc_add(acc, c);
		}
		return;
	}
	if (roughness(sd) <= 0.001f || transmission(sd) > 0.999f) flags |= S_SPECULAR; else flags &= ~S_SPECULAR;
	uint32_t seed = wang_hash(pathIdx * 17u + P.R0);
	const float faceDir = (dot(D, N) > 0) ? -1 : 1;
	if (faceDir == 1) sd.transmittance = mk3(0, 0, 0);
	if ((flags & S_SPECULAR) == 0)
	{
		const float r0 = rand_float(seed), r1 = rand_float(seed);
		float pickProb = 0, lightPdf = 0;
		v3 lightColor = mk3(0, 0, 0);
		v3 L = random_point_on_light(P, r0, r1, I, fN * faceDir, pickProb, lightPdf, lightColor) - I;
		const float dist = length(L);
		L = L * (1.0f / dist);
		const float NdotL = dot(L, fN * faceDir);
		if (NdotL > 0 && lightPdf > 0 && P.pathLength < P.maxDepth)
		{
			float bsdfPdf;
			const v3 sampledBSDF = evaluate_bsdf(sd, fN, L, bsdfPdf);
			const v3 c = throughput * sampledBSDF * lightColor * (NdotL / (pickProb * lightPdf));
			const uint32_t k = atomicAdd(&counters[2], 1u);
			const v3 so = safe_origin(I, L, N, P.eps);
			shadowRays[2 * k] = make_float4(so.x, so.y, so.z, 0.0f);
			shadowRays[2 * k + 1] = make_float4(L.x, L.y, L.z, dist - 2 * P.eps);
			shadowPot[k] = make_float4(c.x, c.y, c.z, bitsf(slot));
		}
	}
	v3 R;
	float newBsdfPdf;
	bool specular = false;
	const float r3 = rand_float(seed), r4 = rand_float(seed), r5 = rand_float(seed);
	const v3 bsdf = sample_bsdf(sd, fN, N, D * -1.0f, t, r3, r4, R, newBsdfPdf, specular);
	if (newBsdfPdf < EPSILON_F || __builtin_isnan(newBsdfPdf)) return;
	if (specular) flags |= S_SPECULAR;
	const float pr = P.pathLength == P.maxDepth ? 0 : ((flags & S_SPECULAR) ? 1 : fminf(1.0f, fmaxf(fmaxf(bsdf.x, bsdf.y), bsdf.z)));
	if (pr <= r5) return;
	throughput = throughput * (bsdf * fabsf(dot(fN, R)) / (pr * newBsdfPdf));
	const uint32_t k = atomicAdd(&counters[1], 1u);
	const v3 no = safe_origin(I, R, N, P.eps);
	raysOut[2 * k] = make_float4(no.x, no.y, no.z, 0.0f);
	raysOut[2 * k + 1] = make_float4(R.x, R.y, R.z, 1e34f);
	stateOut[k] = make_float4(throughput.x, throughput.y, throughput.z, bitsf(flags));
	pathOut[k] = stored;
}

template <bool JOBS>
__global__ void __launch_bounds__(128) k_shade(const ShadeParams P, const float4* __restrict__ raysIn, const float4* __restrict__ stateIn,
	const uint32_t* __restrict__ pathIn, const float4* __restrict__ hits, float4* __restrict__ raysOut, float4* __restrict__ stateOut,
	uint32_t* __restrict__ pathOut, float4* __restrict__ shadowRays, float4* __restrict__ shadowPot, float4* __restrict__ pathAcc,
	uint32_t* __restrict__ counters)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= (int)counters[0]) return;
	const uint32_t stored = pathIn[j];
	if (JOBS)
	{	/* a path of one of several jobs sharing the pass: that job's constants */
		ShadeParams Q = P;
		uint32_t pathIdx, slot;
		job_view<false>(Q, stored, pathIdx, slot);
		shade_path(Q, j, pathIdx, slot, stored, raysIn, stateIn, pathIn, hits, raysOut, stateOut, pathOut, shadowRays, shadowPot, pathAcc, counters);
	}
	else shade_path(P, j, stored, stored - P.jobBase, stored, raysIn, stateIn, pathIn, hits, raysOut, stateOut, pathOut, shadowRays, shadowPot, pathAcc, counters);
}

__global__ void k_connect(const float4* __restrict__ shadowPot, const uint32_t* __restrict__ occluded, float4* __restrict__ pathAcc,
	const uint32_t* __restrict__ counters, float countW)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= (int)counters[2]) return;
	if (occluded[j]) return;
	const float4 p = shadowPot[j];
	float4* a = pathAcc + fbits(p.w);
	float4 v = *a;
	v.x = v.x + p.x, v.y = v.y + p.y, v.z = v.z + p.z, v.w = v.w + countW;
	*a = v;
}

/* Sample-ordered reduction of the path slots into the pixel accumulators.
 * fresh (a frame's first pass after a restart): the sums start from zero
 * instead of the stored values, the same additions as after clearing the
 * accumulators, without the clear and its reads. */
__global__ void k_resolve(const float4* __restrict__ pathAcc, int pixels, int samples, float4* __restrict__ acc32, double* __restrict__ acc64,
	int fresh)
{
	const int px = blockIdx.x * blockDim.x + threadIdx.x;
	if (px >= pixels) return;
	float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	double ax = 0.0, ay = 0.0, az = 0.0, aw = 0.0;
	if (!fresh) a = acc32[px], ax = acc64[4 * px], ay = acc64[4 * px + 1], az = acc64[4 * px + 2], aw = acc64[4 * px + 3];
	for (int s = 0; s < samples; s++)
	{
		const float4 c = pathAcc[(size_t)s * pixels + px];
		a.x = a.x + c.x, a.y = a.y + c.y, a.z = a.z + c.z, a.w = a.w + c.w;
		ax += (double)c.x, ay += (double)c.y, az += (double)c.z, aw += (double)c.w;
	}
	acc32[px] = a;
	acc64[4 * px] = ax, acc64[4 * px + 1] = ay, acc64[4 * px + 2] = az, acc64[4 * px + 3] = aw;
}

__global__ void k_finalize(const float4* __restrict__ acc32, int pixels, float scale, float4* __restrict__ image)
{
	const int px = blockIdx.x * blockDim.x + threadIdx.x;
	if (px >= pixels) return;
	const float4 a = acc32[px];
	image[px] = make_float4(a.x * scale, a.y * scale, a.z * scale, a.w * scale);
}

__global__ void k_set_u32(uint32_t* p, int n, uint32_t v)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src) { *dst = *src; }

__global__ void k_advance(uint32_t* c, uint32_t* logSlot, uint32_t* snapshot, uint32_t* live)
{
	const uint32_t next = c[1];
	if (snapshot) snapshot[0] = c[0], snapshot[1] = c[1], snapshot[2] = c[2], snapshot[3] = c[3];
	c[0] = next, c[1] = 0u, c[2] = 0u;
	if (logSlot) *logSlot = next;
	if (live) *live = next;
}

hipError_t launch_advance(uint32_t* counters, uint32_t* logSlot, hipStream_t s, uint32_t* snapshot, uint32_t* live)
{
	k_advance<<<1, 1, 0, s>>>(counters, logSlot, snapshot, live);
	return hipGetLastError();
}

#include "disney.h"

/* ==== optix7 estimator ======================================================
 * lib/rendercore_optix7: eye rays (optix/.optix.cu:85-129), shadeKernel
 * (kernels/pathtracer.h:54-238) with the Lambert BSDF variant of the core
 * (kernels/bsdf.h selects sharedBSDFs/lambert.h; BSDF_HAS_PURE_SPECULARS),
 * MAXPATHLENGTH 3, ENOUGH_BOUNCES = S_BOUNCED, CLAMPFIREFLIES and BLUENOISE on
 * (core_settings.h:24-33,185-197). Path state: state = (throughput, flags),
 * aux = (postponed bsdf pdf, packed normal of the last vertex). */
#define S7_SPECULAR 1u
#define S7_BOUNCED 2u
#define S7_VIASPECULAR 4u
#define S7_BOUNCEDTWICE 8u

/* blueNoiseSampler4 (tools_shared.h:323-337) */
__device__ __forceinline__ float4 blue_noise4(const uint32_t* bn, int x, int y, int sampleIndex, int dim)
{
	const uint32_t* rk = bn + dim + (x + y * 128) * 8 + 65536 * 3;
	const int rsi1 = (int)((sampleIndex ^ rk[0]) & 255u), rsi2 = (int)((sampleIndex ^ rk[1]) & 255u);
	const int rsi3 = (int)((sampleIndex ^ rk[2]) & 255u), rsi4 = (int)((sampleIndex ^ rk[3]) & 255u);
	const uint32_t v1 = bn[dim + 0 + rsi1 * 256], v2 = bn[dim + 1 + rsi2 * 256], v3 = bn[dim + 2 + rsi3 * 256], v4 = bn[dim + 3 + rsi4 * 256];
	const uint32_t* sc = bn + (dim & 7) + (x + y * 128) * 8 + 65536;
	return make_float4((0.5f + (float)(v1 ^ sc[0])) * (1.0f / 256.0f), (0.5f + (float)(v2 ^ sc[1])) * (1.0f / 256.0f),
		(0.5f + (float)(v3 ^ sc[2])) * (1.0f / 256.0f), (0.5f + (float)(v4 ^ sc[3])) * (1.0f / 256.0f));
}

/* SampleSmallSkydome (tools_shared.h:204-211): the 64x downsampled dome stored after the full one */
__device__ v3 sample_small_sky(const ShadeParams& P, v3 D)
{
	if (P.skyW <= 0 || P.skyH <= 0) return mk3(0, 0, 0);
	const float4 a = P.w2s0, b = P.w2s1, c = P.w2s2;
	const v3 d = mk3(-(a.x * D.x + a.y * D.y + a.z * D.z), -(b.x * D.x + b.y * D.y + b.z * D.z), -(c.x * D.x + c.y * D.y + c.z * D.z));
	const uint32_t w = (uint32_t)P.skyW >> 6, h = (uint32_t)P.skyH >> 6;
	const float p = lh_atan2f(d.y, d.x);
	const float phi = p < 0 ? (p + 2 * PI_F) : p;
	const float theta = lh_acosf(fminf(fmaxf(d.z, -1.0f), 1.0f));
	const uint32_t u = f2u_sat((float)w * phi * 0.15915494309189533577f - 0.5f);
	const uint32_t v = f2u_sat((float)h * theta * 0.31830988618379067154f - 0.5f);
	const uint32_t idx = u + v * w;
	if (idx >= w * h) return mk3(0, 0, 0);
	return xyz(P.sky[idx + (uint32_t)P.skyW * (uint32_t)P.skyH]);
}

/* LightPickProb with ISLIGHTS (lights_shared.h:174-190): potential of light idx
 * seen from the previous vertex O with normal N, the tri light's point being I. */
__device__ float light_pick_prob(const ShadeParams& P, int idx, v3 O, v3 N, v3 I)
{
	const int count = P.nTri + P.nPoint + P.nSpot + P.nDir;
	float sum = 0, mine = 0;
	for (int i = 0; i < count; i++)
	{
		float c;
		if (i < P.nTri)
		{
			const ConstF4 L = as_const(P.triLights) + 6 * i;
			c = tri_potential(L[0].w, xyz(L[1]), N, I - O);
		}
		else c = potential(P, i, O, N, mk3(-1, -1, -1));
		sum += c;
		if (i == idx) mine = c;
	}
	if (sum <= 0) return 0;
	return mine / sum;
}

/* UnpackNormal / PackNormal (tools_shared.h:101-120) */
__device__ __forceinline__ v3 unpack_normal(uint32_t p)
{
	float nx = (float)(p & 65535u) * (2.0f / 65535.0f), ny = (float)(p >> 16) * (2.0f / 65535.0f), nz = 0, nw = 0;
	nx += -1, ny += -1, nz += 1, nw += -1;
	float l = nx * -nx + ny * -ny + nz * -nw;
	nz = l, l = sqrtf(l), nx *= l, ny *= l;
	return mk3(nx * 2.0f + 0, ny * 2.0f + 0, nz * 2.0f + -1);
}
__device__ __forceinline__ uint32_t pack_normal(v3 N)
{
	const float f = 65535.0f / fmaxf(sqrtf(8.0f * N.z + 8.0f), 0.0001f);
	return f2u_sat(N.x * f + 32767.0f) + (f2u_sat(N.y * f + 32767.0f) << 16);
}

__device__ __forceinline__ v3 clamp_intensity(v3 c, float clampValue)
{
	const float v = fmaxf(c.x, fmaxf(c.y, c.z));
	if (v > clampValue) { const float m = clampValue / v; c.x *= m, c.y *= m, c.z *= m; }
	return c;
}
__device__ __forceinline__ v3 fixnan(v3 c)
{	/* FIXNAN_FLOAT3: a non-finite sum zeroes the vector */
	return __builtin_isfinite(c.x + c.y + c.z) ? c : mk3(0, 0, 0);
}

__global__ void k_generate7(CamParams c, ShadeParams P, uint32_t jobBase, int n, float eps, float4* __restrict__ rays,
	float4* __restrict__ state, uint32_t* __restrict__ pathId, float2* __restrict__ aux, float4* __restrict__ pathAcc, uint32_t* __restrict__ counters,
	uint32_t slotBase, uint32_t total, int storeSlot)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j == 0 && slotBase == 0) counters[0] = total, counters[1] = 0, counters[2] = 0;
	if (j >= n) return;
	const uint32_t pathIdx = jobBase + (uint32_t)j;
	const uint32_t o = slotBase + (uint32_t)j;
	const uint32_t wh = (uint32_t)(c.w * c.h);
	const uint32_t pixelIdx = pathIdx % wh, sampleIdx = pathIdx / wh + (uint32_t)P.pass;
	uint32_t seed = wang_hash(pathIdx * 16789u + (uint32_t)P.pass * 1791u);
	const int sx = (int)(pixelIdx % (uint32_t)c.w), sy = (int)(pixelIdx / (uint32_t)c.w);
	float4 r4;
	if (sampleIdx < 64)
		r4 = blue_noise4(P.blueNoise, (sx + (int)(P.shift & 127u)) & 127, (sy + (int)(P.shift >> 24)) & 127, (int)sampleIdx, 0);
	else
	{
		r4.x = rand_float(seed), r4.y = rand_float(seed);
		r4.z = rand_float(seed), r4.w = rand_float(seed);
	}
	const v3 pos = mk3(c.pos[0], c.pos[1], c.pos[2]), p1 = mk3(c.p1[0], c.p1[1], c.p1[2]);
	const v3 right = mk3(c.p2[0], c.p2[1], c.p2[2]) - p1, up = mk3(c.p3[0], c.p3[1], c.p3[2]) - p1;
	const v3 O = point_on_lens(r4.x, r4.z, pos, c.aperture, right, up);
	v3 target;
	if (c.distortion == 0)
	{	/* RayTarget (common_functions.h:29-50) */
		const float u = ((float)sx + r4.y) * (1.0f / (float)c.w), v = ((float)sy + r4.w) * (1.0f / (float)c.h);
		target = p1 + u * right + v * up;
	}
	else
	{
		const float tx = (float)sx / (float)c.w - 0.5f, ty = (float)sy / (float)c.h - 0.5f;
		const float rr = tx * tx + ty * ty;
		const float rq = sqrtf(rr) * (1.0f + c.distortion * rr + c.distortion * rr * rr);
		const float theta = lh_atan2f(tx, ty);
		float st, ct;
		cw_sincos(theta, &st, &ct);
		const float bx = (st * rq + 0.5f) * (float)c.w, by = (ct * rq + 0.5f) * (float)c.h;
		target = p1 + (bx + r4.y) * (right / (float)c.w) + (by + r4.w) * (up / (float)c.h);
	}
	const v3 D = normalize(target - O);
	rays[2 * o] = make_float4(O.x, O.y, O.z, eps);
	rays[2 * o + 1] = make_float4(D.x, D.y, D.z, 1e34f);
	state[o] = make_float4(1.0f, 1.0f, 1.0f, bitsf(S7_SPECULAR));
	aux[o] = make_float2(1.0f, bitsf(0u));
	pathId[o] = storeSlot ? o : pathIdx;
	pathAcc[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ void shade_path7(const ShadeParams& P, const int j, const uint32_t pathIdx, const uint32_t slot, const uint32_t stored,
	const float4* __restrict__ raysIn, const float4* __restrict__ stateIn,
	const float2* __restrict__ auxIn, const uint32_t* __restrict__ pathIn, const float4* __restrict__ hits, float4* __restrict__ raysOut,
	float4* __restrict__ stateOut, float2* __restrict__ auxOut, uint32_t* __restrict__ pathOut, float4* __restrict__ shadowRays,
	float4* __restrict__ shadowPot, float4* __restrict__ pathAcc, uint32_t* __restrict__ counters)
{
	const float4 O4 = raysIn[2 * j], D4 = raysIn[2 * j + 1], T4 = stateIn[j];
	const float2 A2 = auxIn[j];
	const float4 hd = hits[j];
	float4* acc = pathAcc + slot;
	const uint32_t wh = (uint32_t)(P.w * P.h);
	const uint32_t pixelIdx = pathIdx % wh;
	const uint32_t sampleIdx = pathIdx / wh + (uint32_t)P.pass;
	uint32_t flags = fbits(T4.w);
	const v3 D = xyz(D4);
	v3 throughput = xyz(T4);
	const float bsdfPdf = A2.x;
	const uint32_t prim = fbits(hd.z);
	const int instIdx = (int)fbits(hd.y);
	const float t = hd.w;
	if (prim == 0xffffffffu)
	{
		const v3 sky = (flags & S7_BOUNCED) ? sample_small_sky(P, D) : sample_sky(P, D);
		v3 c = throughput * sky * (1.0f / bsdfPdf);
		c = fixnan(clamp_intensity(c, P.clampValue));
		acc_add(acc, c);
		return;
	}
	if (P.probe && (int)pixelIdx == P.probePixel && P.pathLength == 1)
		P.probe[0] = instIdx, P.probe[1] = (int)prim, P.probe[2] = (int)fbits(t);
	const uint32_t uvb = fbits(hd.x);
	const float u = (float)(uvb & 65535u) * (1.0f / 65535.0f), v = (float)(uvb >> 16) * (1.0f / 65535.0f);
	const ShadeInst& SI = P.inst[instIdx];
	const float4* tri = SI.coreTris + 13 * (size_t)prim;
	Shading sd;
	v3 N, iN, fN, T;
	const v3 I = xyz(O4) + t * D;
	/* the hit record holds the weights of vertices 0 and 1; optix7's holds those of 1 and 2 */
	const bool alphaHit = shading_data<false>(P, D, v, 1.0f - (u + v), P.spreadAngle * t, tri, instIdx, sd, N, iN, fN, T);
	uint32_t seed = wang_hash(pathIdx * 17u + P.R0);
	if (alphaHit)
	{	/* continue through the cut-out, same path state */
		if (P.pathLength < MAXPATHLENGTH_OPTIX7)
		{
			const uint32_t k = atomicAdd(&counters[1], 1u);
			const v3 o = I + D * P.eps;
			raysOut[2 * k] = make_float4(o.x, o.y, o.z, P.eps);
			raysOut[2 * k + 1] = make_float4(D.x, D.y, D.z, 1e34f);
			float4 st = T4;
			if (!__builtin_isfinite(T4.x + T4.y + T4.z)) st = make_float4(0, 0, 0, T4.w);
			stateOut[k] = st;
			auxOut[k] = A2;
			pathOut[k] = stored;
		}
		return;
	}
	if (sd.color.x > 1.0f || sd.color.y > 1.0f || sd.color.z > 1.0f)
	{
		const float DdotNL = -dot(D, N);
		v3 c = mk3(0, 0, 0);
		if (DdotNL > 0)
		{
			if (P.pathLength == 1 || (flags & S7_SPECULAR) || !P.connections) c = sd.color;
			else
			{	/* MIS with the light sample of the previous vertex */
				const v3 lastN = unpack_normal(fbits(A2.y));
				const float area = tri[5].w;
				const int ltri = (int)fbits(tri[0].w);
				const float lightPdf = (t * t) / (fabsf(dot(D, N)) * area);
				const float pickProb = light_pick_prob(P, ltri, xyz(O4), lastN, I);
				if ((bsdfPdf + lightPdf * pickProb) > 0) c = throughput * sd.color * (1.0f / (bsdfPdf + lightPdf * pickProb));
			}
			c = fixnan(clamp_intensity(c, P.clampValue));
			acc_add(acc, c);
		}
		return;
	}
	if (flags & S7_BOUNCED) sd.params.x |= 255u << 24;   /* path regularisation: roughness 1 after a bounce */
	if (roughness(sd) <= 0.001f || transmission(sd) > 0.5f) flags |= S7_SPECULAR; else flags &= ~S7_SPECULAR;
	const float faceDir = (dot(D, N) > 0) ? -1 : 1;
	if (faceDir == 1) sd.transmittance = mk3(0, 0, 0);
	throughput = throughput * (1.0f / bsdfPdf);
	float4 r4;
	if (sampleIdx < 64)
	{
		const int x = (int)(((pathIdx % (uint32_t)P.w) + (P.shift & 127u)) & 127u);
		const int y = (int)(((pathIdx / (uint32_t)P.w) + (P.shift >> 24)) & 127u);
		r4 = blue_noise4(P.blueNoise, x, y, (int)sampleIdx, 4 * P.pathLength - 4);
	}
	else
	{
		r4.x = rand_float(seed), r4.y = rand_float(seed);
		r4.z = rand_float(seed), r4.w = rand_float(seed);
	}
	if ((flags & S7_SPECULAR) == 0 && P.connections)
	{
		float pickProb = 0, lightPdf = 0;
		v3 lightColor = mk3(0, 0, 0);
		v3 L = random_point_on_light(P, r4.x, r4.y, I, fN * faceDir, pickProb, lightPdf, lightColor) - I;
		const float dist = length(L);
		L = L * (1.0f / dist);
		const float NdotL = dot(L, fN * faceDir);
		if (NdotL > 0 && lightPdf > 0)
		{
			float pdf;
			/* Lambert: colour / pi times ROUGHNESS (BSDF_HAS_PURE_SPECULARS); Disney as evaluated */
			const v3 sampled = P.bsdf ? disney_evaluate(sd, fN, T, D * -1.0f, L, pdf) : evaluate_bsdf(sd, fN, L, pdf);
			if (pdf > 0)
			{
				v3 c = throughput * sampled * lightColor * (NdotL / (pickProb * lightPdf + pdf));
				c = clamp_intensity(fixnan(c), P.clampValue);
				const uint32_t k = atomicAdd(&counters[2], 1u);
				const v3 so = safe_origin(I, L, N, P.eps);
				shadowRays[2 * k] = make_float4(so.x, so.y, so.z, P.eps);
				shadowRays[2 * k + 1] = make_float4(L.x, L.y, L.z, dist - 2 * P.eps);
				shadowPot[k] = make_float4(c.x, c.y, c.z, bitsf(slot));
			}
		}
	}
	if ((flags & S7_BOUNCED) || P.pathLength == MAXPATHLENGTH_OPTIX7) return;
	v3 R;
	float newBsdfPdf;
	bool specular = false;
	const float r5 = rand_float(seed);              /* SampleBSDF's third random number, drawn at the call */
	const v3 bsdf = P.bsdf ? disney_sample(sd, fN, N, T, D * -1.0f, t, r4.z, r4.w, r5, R, newBsdfPdf, specular)
		: sample_bsdf<true>(sd, fN, N, D * -1.0f, t, r4.z, r4.w, R, newBsdfPdf, specular);
	if (newBsdfPdf < EPSILON_F || __builtin_isnan(newBsdfPdf)) return;
	if (specular) flags |= S7_SPECULAR;
	const float p = ((flags & S7_SPECULAR) || ((flags & S7_BOUNCED) == 0)) ? 1 : fminf(1.0f, fmaxf(fmaxf(bsdf.x, bsdf.y), bsdf.z));
	if (p < rand_float(seed)) return;
	throughput = throughput * (1 / p);
	const uint32_t k = atomicAdd(&counters[1], 1u);
	const uint32_t packedNormal = pack_normal(fN * faceDir);
	if (!(flags & S7_SPECULAR)) flags |= (flags & S7_BOUNCED) ? S7_BOUNCEDTWICE : S7_BOUNCED; else flags |= S7_VIASPECULAR;
	const v3 no = safe_origin(I, R, N, P.eps);
	raysOut[2 * k] = make_float4(no.x, no.y, no.z, P.eps);
	raysOut[2 * k + 1] = make_float4(R.x, R.y, R.z, 1e34f);
	const v3 tp = fixnan(throughput) * bsdf * fabsf(dot(fN, R));
	stateOut[k] = make_float4(tp.x, tp.y, tp.z, bitsf(flags));
	auxOut[k] = make_float2(newBsdfPdf, bitsf(packedNormal));
	pathOut[k] = stored;
}

template <bool JOBS>
__global__ void __launch_bounds__(128) k_shade7(const ShadeParams P, const float4* __restrict__ raysIn, const float4* __restrict__ stateIn,
	const float2* __restrict__ auxIn, const uint32_t* __restrict__ pathIn, const float4* __restrict__ hits, float4* __restrict__ raysOut,
	float4* __restrict__ stateOut, float2* __restrict__ auxOut, uint32_t* __restrict__ pathOut, float4* __restrict__ shadowRays,
	float4* __restrict__ shadowPot, float4* __restrict__ pathAcc, uint32_t* __restrict__ counters)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= (int)counters[0]) return;
	const uint32_t stored = pathIn[j];
	if (JOBS)
	{	/* a path of one of several jobs sharing the pass: that job's constants */
		ShadeParams Q = P;
		uint32_t pathIdx, slot;
		job_view<true>(Q, stored, pathIdx, slot);
		shade_path7(Q, j, pathIdx, slot, stored, raysIn, stateIn, auxIn, pathIn, hits, raysOut, stateOut, auxOut, pathOut, shadowRays, shadowPot, pathAcc, counters);
	}
	else shade_path7(P, j, stored, stored - P.jobBase, stored, raysIn, stateIn, auxIn, pathIn, hits, raysOut, stateOut, auxOut, pathOut, shadowRays, shadowPot, pathAcc, counters);
}

hipError_t launch_generate7(const CamParams& c, const ShadeParams& p, uint32_t jobBase, int n, float eps, const PathBuffers& b, int out, hipStream_t s,
	uint32_t slotBase, uint32_t total, bool storeSlot)
{
	if (n <= 0) return hipSuccess;
	k_generate7<<<(n + 255) / 256, 256, 0, s>>>(c, p, jobBase, n, eps, b.rays[out], b.state[out], b.pathId[out], b.aux[out], b.pathAcc, b.counters,
		slotBase, total ? total : (uint32_t)n, storeSlot ? 1 : 0);
	return hipGetLastError();
}

hipError_t launch_shade7(const ShadeParams& p, const PathBuffers& b, int in, int maxCount, hipStream_t s)
{
	if (maxCount <= 0) return hipSuccess;
	if (p.nseg > 0) k_shade7<true><<<(maxCount + 127) / 128, 128, 0, s>>>(p, b.rays[in], b.state[in], b.aux[in], b.pathId[in], b.hits, b.rays[in ^ 1], b.state[in ^ 1],
		b.aux[in ^ 1], b.pathId[in ^ 1], b.shadowRays, b.shadowPot, b.pathAcc, b.counters);
	else k_shade7<false><<<(maxCount + 127) / 128, 128, 0, s>>>(p, b.rays[in], b.state[in], b.aux[in], b.pathId[in], b.hits, b.rays[in ^ 1], b.state[in ^ 1],
		b.aux[in ^ 1], b.pathId[in ^ 1], b.shadowRays, b.shadowPot, b.pathAcc, b.counters);
	return hipGetLastError();
}

/* ==== filter estimator ======================================================
 * lib/RenderCore_Optix7Filter: generateEyeRay / setupPrimaryRay
 * (optix/.optix.cu:66-129), shadeKernel (kernels/pathtracer.h:76-306) with
 * PackFeatures / calculateDepthDerivatives (:44-58), the BSDF of the core
 * (Disney by default, kernels/bsdf.h), MAXPATHLENGTH 7, one diffuse bounce.
 * Radiance of paths that bounced off a diffuse vertex goes to the indirect
 * half (path slot + passPaths) when the G-buffer is on. */
#define SF_SPECULAR 1u
#define SF_BOUNCED 2u
#define SF_VIASPECULAR 4u

/* blueNoiseSampler (optix/.optix.cu:40-54) */
__device__ __forceinline__ float blue_noise1(const uint32_t* bn, int x, int y, int sampleIndex, int dim)
{
	x &= 127, y &= 127, sampleIndex &= 255, dim &= 255;
	const int ranked = (int)((uint32_t)(sampleIndex ^ (int)bn[dim + (x + y * 128) * 8 + 65536 * 3]) & 255u);
	uint32_t value = bn[dim + ranked * 256];
	value ^= bn[(dim & 7) + (x + y * 128) * 8 + 65536];
	return (0.5f + (float)value) * (1.0f / 256.0f);
}

/* PackNormal2, HDRtoRGB32, RGB32toHDR (tools_shared.h:122-129,162-175) */
__device__ __forceinline__ uint32_t pack_normal2(v3 N)
{
	const uint32_t x = min(f2u_sat((N.x + 1) * 511), 1023u), y = min(f2u_sat((N.y + 1) * 511), 1023u), z = min(f2u_sat((N.z + 1) * 511), 1023u);
	return (x << 2u) + (y << 12u) + (z << 22u);
}
__device__ __forceinline__ uint32_t hdr_to_rgb32(v3 c)
{
	const uint32_t r = f2u_sat(1023.0f * fminf(1.0f, c.x)), g = f2u_sat(2047.0f * fminf(1.0f, c.y)), b = f2u_sat(2047.0f * fminf(1.0f, c.z));
	return (r << 22) + (g << 11) + b;
}
__device__ __forceinline__ v3 rgb32_to_hdr(uint32_t c)
{
	return mk3((float)(c >> 22) * (1.0f / 1023.0f), (float)((c >> 11) & 2047u) * (1.0f / 2047.0f), (float)(c & 2047u) * (1.0f / 2047.0f));
}
__device__ __forceinline__ void pack_features(uint4* f, uint32_t albedo, uint32_t packedNormal, float t, uint32_t isSpecular, uint32_t matid)
{
	*f = make_uint4(albedo, packedNormal, fbits(t), (isSpecular << 4) + (matid << 6) + (f->w & 15u));
}
/* calculateDepthDerivatives: rays through (x+1, y) and (x, y+1) against the
 * object-space plane of the triangle */
__device__ void depth_derivatives(const ShadeParams& P, int x, int y, float depth, const float4* tri, float4* dd)
{
	const v3 triN = mk3(tri[2].w, tri[3].w, tri[4].w);
	const v3 p1 = xyz(P.camP1), p2 = xyz(P.camRight), p3 = xyz(P.camUp), pos = xyz(P.camPos), v0 = xyz(tri[8]);
	const float iw = 1.0f / (float)P.w, ih = 1.0f / (float)P.h;
	const v3 dxDir = normalize((p1 + ((float)x + 0.5f + 1) * iw * p2 + ((float)y + 0.5f) * ih * p3) - pos);
	const float dxDepth = dot(v0 - pos, triN) / dot(triN, dxDir);
	const v3 dyDir = normalize((p1 + ((float)x + 0.5f) * iw * p2 + ((float)y + 0.5f + 1) * ih * p3) - pos);
	const float dyDepth = dot(v0 - pos, triN) / dot(triN, dyDir);
	dd[x + y * P.w] = make_float4(0, 0, dxDepth - depth, dyDepth - depth);
}

__global__ void k_generateF(CamParams c, ShadeParams P, uint32_t jobBase, int n, float eps, float4* __restrict__ rays,
	float4* __restrict__ state, uint32_t* __restrict__ pathId, float2* __restrict__ aux, float4* __restrict__ pathAcc, uint32_t* __restrict__ counters,
	uint32_t slotBase, uint32_t total, int storeSlot)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j == 0 && slotBase == 0) counters[0] = total, counters[1] = 0, counters[2] = 0;
	if (j >= n) return;
	const uint32_t pathIdx = jobBase + (uint32_t)j;
	const uint32_t o = slotBase + (uint32_t)j;
	const uint32_t wh = (uint32_t)(c.w * c.h);
	const uint32_t pixelIdx = pathIdx % wh, sampleIdx = pathIdx / wh + (uint32_t)P.pass;
	const int sx = (int)(pixelIdx % (uint32_t)c.w), sy = (int)(pixelIdx / (uint32_t)c.w);
	const v3 pos = mk3(c.pos[0], c.pos[1], c.pos[2]), p1u = mk3(c.p1[0], c.p1[1], c.p1[2]);
	const v3 right = mk3(c.p2[0], c.p2[1], c.p2[2]) - p1u, up = mk3(c.p3[0], c.p3[1], c.p3[2]) - p1u;
	const v3 p1 = xyz(P.camP1);                     /* p1 moved by the TAA jitter (.optix.cu:85-129 with params.p1) */
	float r0 = 0, r1 = 0;
	v3 O = pos;
	if (!P.gFeatures)
	{	/* filter off: jittered pixel and lens */
		uint32_t seed = wang_hash(pathIdx * 16789u + (uint32_t)P.pass * 1791u);
		float r2, r3;
		if (sampleIdx < 256)
			r0 = blue_noise1(P.blueNoise, sx, sy, (int)sampleIdx, 0), r1 = blue_noise1(P.blueNoise, sx, sy, (int)sampleIdx, 1),
			r2 = blue_noise1(P.blueNoise, sx, sy, (int)sampleIdx, 2), r3 = blue_noise1(P.blueNoise, sx, sy, (int)sampleIdx, 3);
		else
			r0 = rand_float(seed), r1 = rand_float(seed), r2 = rand_float(seed), r3 = rand_float(seed);
		O = point_on_lens(r2, r3, pos, c.aperture, right, up);
	}
	const float u = ((float)sx + r0) * (1.0f / (float)c.w), v = ((float)sy + r1) * (1.0f / (float)c.h);
	const v3 target = p1 + u * right + v * up;
	const v3 D = normalize(target - O);
	rays[2 * o] = make_float4(O.x, O.y, O.z, eps);
	rays[2 * o + 1] = make_float4(D.x, D.y, D.z, 1e34f);
	state[o] = make_float4(1.0f, 1.0f, 1.0f, bitsf(SF_SPECULAR));
	aux[o] = make_float2(1.0f, bitsf(0u));
	pathId[o] = storeSlot ? o : pathIdx;
	pathAcc[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	pathAcc[o + total] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ void shade_pathF(const ShadeParams& P, const int j, const uint32_t pathIdx, const uint32_t slot0, const uint32_t stored,
	const float4* __restrict__ raysIn, const float4* __restrict__ stateIn,
	const float2* __restrict__ auxIn, const uint32_t* __restrict__ pathIn, const float4* __restrict__ hits, float4* __restrict__ raysOut,
	float4* __restrict__ stateOut, float2* __restrict__ auxOut, uint32_t* __restrict__ pathOut, float4* __restrict__ shadowRays,
	float4* __restrict__ shadowPot, float4* __restrict__ pathAcc, uint32_t* __restrict__ counters)
{
	const float4 O4 = raysIn[2 * j], D4 = raysIn[2 * j + 1], T4 = stateIn[j];
	const float2 A2 = auxIn[j];
	const float4 hd = hits[j];
	uint32_t flags = fbits(T4.w);
	const bool gbuf = P.gFeatures != nullptr;
	const uint32_t wh = (uint32_t)(P.w * P.h);
	const uint32_t pixelIdx = (pathIdx % wh) + ((gbuf && (flags & SF_BOUNCED)) ? wh : 0u);
	const uint32_t sampleIdx = pathIdx / wh + (uint32_t)P.pass;
	const uint32_t slot = slot0 + ((gbuf && (flags & SF_BOUNCED)) ? P.passPaths : 0u);
	float4* acc = pathAcc + slot;
	const v3 D = xyz(D4);
	v3 throughput = xyz(T4);
	const float bsdfPdf = A2.x;
	const uint32_t prim = fbits(hd.z);
	const int instIdx = (int)fbits(hd.y);
	const float t = hd.w;
	const bool firstHit = ((flags & SF_BOUNCED) == 0) && sampleIdx == 0 && gbuf;
	uint4* feat = gbuf ? P.gFeatures + pathIdx : nullptr;
	if (P.pathLength == 1 && firstHit)
	{
		pack_features(feat, 0, 0, 1e34f, 0, 0);
		P.gWorldPos[pathIdx] = make_float4(0, 0, 0, bitsf(0u));
		P.gDeltaDepth[pathIdx] = make_float4(0, 0, 0, 0);
	}
	if (P.pathLength == 1)
	{	/* depth for the DOF shader, in w of the (direct) accumulator */
		float4 a = *acc;
		a.w = a.w + (prim == 0xffffffffu ? 10000.0f : t);
		*acc = a;
	}
	if (prim == 0xffffffffu)
	{
		v3 c = throughput * sample_sky(P, D) * (1.0f / bsdfPdf);
		c = fixnan(clamp_intensity(c, P.clampValue));
		acc_add(acc, c);
		if (firstHit)
		{
			const uint32_t isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
			const uint32_t pn = pack_normal2(D * -1.0f) + isSpec;
			pack_features(feat, hdr_to_rgb32(c), pn, t, isSpec, 0);
			const v3 far = xyz(O4) + 50000.0f * D;
			P.gWorldPos[pathIdx] = make_float4(far.x, far.y, far.z, bitsf(pn));
			P.gDeltaDepth[pathIdx] = make_float4(0, 0, 0, 0);
		}
		return;
	}
	if (P.probe && (int)pixelIdx == P.probePixel && P.pathLength == 1 && sampleIdx == 0)
		P.probe[0] = instIdx, P.probe[1] = (int)prim, P.probe[2] = (int)fbits(t);
	const uint32_t uvb = fbits(hd.x);
	const float u = (float)(uvb & 65535u) * (1.0f / 65535.0f), v = (float)(uvb >> 16) * (1.0f / 65535.0f);
	const ShadeInst& SI = P.inst[instIdx];
	const float4* tri = SI.coreTris + 13 * (size_t)prim;
	Shading sd;
	v3 N, iN, fN, T;
	const v3 I = xyz(O4) + t * D;
	/* the hit record holds the weights of vertices 0 and 1; the core's holds those of 1 and 2 */
	const bool alphaHit = shading_data<false>(P, D, v, 1.0f - (u + v), P.spreadAngle * t, tri, instIdx, sd, N, iN, fN, T);
	if (alphaHit)
	{
		if (P.pathLength == MAXPATHLENGTH_FILTER)
		{
			if (gbuf && sampleIdx == 0 && ((flags & SF_BOUNCED) == 0))
			{
				const uint32_t isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
				const uint32_t pn = pack_normal2(N) + isSpec;
				pack_features(feat, 0, pn, t, isSpec, 0);
				P.gWorldPos[pathIdx] = make_float4(I.x, I.y, I.z, bitsf(pn));
			}
		}
		else
		{
			const uint32_t k = atomicAdd(&counters[1], 1u);
			const v3 o = I + D * P.eps;
			raysOut[2 * k] = make_float4(o.x, o.y, o.z, P.eps);
			raysOut[2 * k + 1] = make_float4(D.x, D.y, D.z, 1e34f);
			float4 st = T4;
			if (!__builtin_isfinite(T4.x + T4.y + T4.z)) st = make_float4(0, 0, 0, T4.w);
			stateOut[k] = st;
			auxOut[k] = A2;
			pathOut[k] = stored;
		}
		return;
	}
	if (sd.color.x > 1.0f || sd.color.y > 1.0f || sd.color.z > 1.0f)
	{
		const float DdotNL = -dot(D, N);
		if (DdotNL > 0)
		{
			v3 c = mk3(0, 0, 0);
			if (P.pathLength == 1 || (flags & SF_SPECULAR) || !P.connections) c = sd.color;
			else
			{
				const v3 lastN = unpack_normal(fbits(A2.y));
				const float area = tri[5].w;
				const int ltri = (int)fbits(tri[0].w);
				const float lightPdf = (t * t) / (fabsf(dot(D, N)) * area);
				const float pickProb = light_pick_prob(P, ltri, xyz(O4), lastN, I);
				if ((bsdfPdf + lightPdf * pickProb) > 0) c = throughput * sd.color * (1.0f / (bsdfPdf + lightPdf * pickProb));
			}
			c = fixnan(clamp_intensity(c, P.clampValue));
			acc_add(acc, c);
		}
		if (firstHit)
		{
			const uint32_t isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
			const uint32_t pn = pack_normal2(N) + isSpec;
			pack_features(feat, hdr_to_rgb32(sd.color), pn, t, isSpec, 0);
			P.gWorldPos[pathIdx] = make_float4(I.x, I.y, I.z, bitsf(pn));
			depth_derivatives(P, (int)(pathIdx % (uint32_t)P.w), (int)(pathIdx / (uint32_t)P.w), t, tri, P.gDeltaDepth);
		}
		return;
	}
	if (roughness(sd) <= 0.001f || transmission(sd) > 0.999f) flags |= SF_SPECULAR; else flags &= ~SF_SPECULAR;
	uint32_t seed = wang_hash(pathIdx * 17u + P.R0);
	if (firstHit)
	{
		if (flags & SF_SPECULAR) feat->x = hdr_to_rgb32(sd.color);
		else
		{
			v3 albedo = sd.color;
			if (flags & SF_VIASPECULAR) albedo = albedo * rgb32_to_hdr(feat->x);
			const uint32_t isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
			/* PackNormal2(flip ? -fN : fN) with flip = +-1 is always -fN */
			const uint32_t pn = pack_normal2(fN * -1.0f) + isSpec;
			pack_features(feat, hdr_to_rgb32(albedo), pn, t, isSpec, 0);
			P.gWorldPos[pathIdx] = make_float4(I.x, I.y, I.z, bitsf(pn));
			depth_derivatives(P, (int)(pathIdx % (uint32_t)P.w), (int)(pathIdx / (uint32_t)P.w), t, tri, P.gDeltaDepth);
		}
	}
	const float faceDir = (dot(D, N) > 0) ? -1 : 1;
	if (faceDir == 1) sd.transmittance = mk3(0, 0, 0);
	throughput = throughput * (1.0f / bsdfPdf);
	if ((flags & SF_SPECULAR) == 0 && P.connections)
	{
		float r0, r1;
		if (sampleIdx < 2)
		{
			const int x = (int)((pixelIdx % (uint32_t)P.w) & 127u), y = (int)((pixelIdx / (uint32_t)P.w) & 127u);
			r0 = blue_noise1(P.blueNoise, x, y, (int)sampleIdx + P.blueSlot, 4 + 4 * P.pathLength);
			r1 = blue_noise1(P.blueNoise, x, y, (int)sampleIdx + P.blueSlot, 5 + 4 * P.pathLength);
		}
		else r0 = rand_float(seed), r1 = rand_float(seed);
		float pickProb = 0, lightPdf = 0;
		v3 lightColor = mk3(0, 0, 0);
		v3 L = random_point_on_light(P, r0, r1, I, fN * faceDir, pickProb, lightPdf, lightColor) - I;
		const float dist = length(L);
		L = L * (1.0f / dist);
		const float NdotL = dot(L, fN * faceDir);
		if (NdotL > 0 && lightPdf > 0)
		{
			float pdf;
			const v3 sampled = P.bsdf ? disney_evaluate(sd, fN, T, D * -1.0f, L, pdf) : evaluate_bsdf(sd, fN, L, pdf);
			if (pdf > 0)
			{
				v3 c = throughput * sampled * lightColor * (NdotL / (pickProb * lightPdf + pdf));
				c = clamp_intensity(fixnan(c), P.clampValue);
				const uint32_t k = atomicAdd(&counters[2], 1u);
				const v3 so = safe_origin(I, L, N, P.eps);
				shadowRays[2 * k] = make_float4(so.x, so.y, so.z, P.eps);
				shadowRays[2 * k + 1] = make_float4(L.x, L.y, L.z, dist - 2 * P.eps);
				shadowPot[k] = make_float4(c.x, c.y, c.z, bitsf(slot));
			}
		}
	}
	if (flags & SF_BOUNCED) return;
	if (P.pathLength == MAXPATHLENGTH_FILTER)
	{
		if (firstHit)
		{
			const uint32_t isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
			const uint32_t pn = pack_normal2(N) + isSpec;
			pack_features(feat, 0, pn, t, isSpec, 0);
			P.gWorldPos[pathIdx] = make_float4(I.x, I.y, I.z, bitsf(pn));
		}
		return;
	}
	float r3, r4;
	if (sampleIdx < 256)
	{
		const int x = (int)((pixelIdx % (uint32_t)P.w) & 127u), y = (int)((pixelIdx / (uint32_t)P.w) & 127u);
		r3 = blue_noise1(P.blueNoise, x, y, (int)sampleIdx + P.blueSlot, 6 + 4 * P.pathLength);
		r4 = blue_noise1(P.blueNoise, x, y, (int)sampleIdx + P.blueSlot, 7 + 4 * P.pathLength);
	}
	else r3 = rand_float(seed), r4 = rand_float(seed);
	v3 R;
	float newBsdfPdf;
	bool specular = false;
	const float r5 = rand_float(seed);
	const v3 bsdf = P.bsdf ? disney_sample(sd, fN, N, T, D * -1.0f, t, r3, r4, r5, R, newBsdfPdf, specular)
		: sample_bsdf<true>(sd, fN, N, D * -1.0f, t, r3, r4, R, newBsdfPdf, specular);
	if (newBsdfPdf < EPSILON_F || __builtin_isnan(newBsdfPdf)) return;
	if (specular) flags |= SF_SPECULAR;
	const uint32_t k = atomicAdd(&counters[1], 1u);
	const uint32_t packedNormal = pack_normal(fN * faceDir);
	if (!(flags & SF_SPECULAR)) flags |= SF_BOUNCED; else flags |= SF_VIASPECULAR;
	const v3 no = safe_origin(I, R, N, P.eps);
	raysOut[2 * k] = make_float4(no.x, no.y, no.z, P.eps);
	raysOut[2 * k + 1] = make_float4(R.x, R.y, R.z, 1e34f);
	const v3 tp = fixnan(throughput) * bsdf * fabsf(dot(fN, R));
	stateOut[k] = make_float4(tp.x, tp.y, tp.z, bitsf(flags));
	auxOut[k] = make_float2(newBsdfPdf, bitsf(packedNormal));
	pathOut[k] = stored;
}

template <bool JOBS>
__global__ void __launch_bounds__(128) k_shadeF(const ShadeParams P, const float4* __restrict__ raysIn, const float4* __restrict__ stateIn,
	const float2* __restrict__ auxIn, const uint32_t* __restrict__ pathIn, const float4* __restrict__ hits, float4* __restrict__ raysOut,
	float4* __restrict__ stateOut, float2* __restrict__ auxOut, uint32_t* __restrict__ pathOut, float4* __restrict__ shadowRays,
	float4* __restrict__ shadowPot, float4* __restrict__ pathAcc, uint32_t* __restrict__ counters)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= (int)counters[0]) return;
	const uint32_t stored = pathIn[j];
	if (JOBS)
	{	/* a path of one of several jobs sharing the pass: that job's constants */
		ShadeParams Q = P;
		uint32_t pathIdx, slot;
		job_view<true>(Q, stored, pathIdx, slot);
		shade_pathF(Q, j, pathIdx, slot, stored, raysIn, stateIn, auxIn, pathIn, hits, raysOut, stateOut, auxOut, pathOut, shadowRays, shadowPot, pathAcc, counters);
	}
	else shade_pathF(P, j, stored, stored - P.jobBase, stored, raysIn, stateIn, auxIn, pathIn, hits, raysOut, stateOut, auxOut, pathOut, shadowRays, shadowPot, pathAcc, counters);
}

hipError_t launch_generateF(const CamParams& c, const ShadeParams& p, uint32_t jobBase, int n, float eps, const PathBuffers& b, int out, hipStream_t s,
	uint32_t slotBase, uint32_t total, bool storeSlot)
{
	if (n <= 0) return hipSuccess;
	k_generateF<<<(n + 255) / 256, 256, 0, s>>>(c, p, jobBase, n, eps, b.rays[out], b.state[out], b.pathId[out], b.aux[out], b.pathAcc, b.counters,
		slotBase, total ? total : (uint32_t)n, storeSlot ? 1 : 0);
	return hipGetLastError();
}

hipError_t launch_shadeF(const ShadeParams& p, const PathBuffers& b, int in, int maxCount, hipStream_t s)
{
	if (maxCount <= 0) return hipSuccess;
	if (p.nseg > 0) k_shadeF<true><<<(maxCount + 127) / 128, 128, 0, s>>>(p, b.rays[in], b.state[in], b.aux[in], b.pathId[in], b.hits, b.rays[in ^ 1], b.state[in ^ 1],
		b.aux[in ^ 1], b.pathId[in ^ 1], b.shadowRays, b.shadowPot, b.pathAcc, b.counters);
	else k_shadeF<false><<<(maxCount + 127) / 128, 128, 0, s>>>(p, b.rays[in], b.state[in], b.aux[in], b.pathId[in], b.hits, b.rays[in ^ 1], b.state[in ^ 1],
		b.aux[in ^ 1], b.pathId[in ^ 1], b.shadowRays, b.shadowPot, b.pathAcc, b.counters);
	return hipGetLastError();
}

/* Seeded tables in the sampler's layout (xorshift32 bytes); the sampler's
 * structure is the reference's, the table values are not its optimised ones. */
void blue_noise_tables(uint32_t* out, uint32_t seed)
{
	uint32_t s = seed ? seed : 1u;
	for (int i = 0; i < BLUE_NOISE_WORDS; i++) out[i] = 0;
	const auto next = [&]() { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; };
	for (int i = 0; i < 65536; i++) out[i] = next() & 255u;
	for (int i = 0; i < 128 * 128 * 8; i++) out[65536 + i] = next() & 255u;
	for (int i = 0; i < 128 * 128 * 8; i++) out[3 * 65536 + i] = next() & 255u;
}

hipError_t launch_generate(const CamParams& c, uint32_t R0, uint32_t jobBase, int n, float eps, const PathBuffers& b, int out, hipStream_t s,
	uint32_t slotBase, uint32_t total, bool storeSlot)
{
	k_generate<<<(n + 255) / 256, 256, 0, s>>>(c, R0, jobBase, n, eps, b.rays[out], b.state[out], b.pathId[out], b.pathAcc, b.counters,
		slotBase, total ? total : (uint32_t)n, storeSlot ? 1 : 0);
	return hipGetLastError();
}

hipError_t launch_shade(const ShadeParams& p, const PathBuffers& b, int in, int maxCount, hipStream_t s)
{
	if (maxCount <= 0) return hipSuccess;
	if (p.nseg > 0) k_shade<true><<<(maxCount + 127) / 128, 128, 0, s>>>(p, b.rays[in], b.state[in], b.pathId[in], b.hits, b.rays[in ^ 1], b.state[in ^ 1],
		b.pathId[in ^ 1], b.shadowRays, b.shadowPot, b.pathAcc, b.counters);
	else k_shade<false><<<(maxCount + 127) / 128, 128, 0, s>>>(p, b.rays[in], b.state[in], b.pathId[in], b.hits, b.rays[in ^ 1], b.state[in ^ 1],
		b.pathId[in ^ 1], b.shadowRays, b.shadowPot, b.pathAcc, b.counters);
	return hipGetLastError();
}

hipError_t launch_connect(const PathBuffers& b, int maxCount, hipStream_t s, float countW)
{
	if (maxCount <= 0) return hipSuccess;
	k_connect<<<(maxCount + 255) / 256, 256, 0, s>>>(b.shadowPot, b.occluded, b.pathAcc, b.counters, countW);
	return hipGetLastError();
}

hipError_t launch_resolve(const float4* pathAcc, int pixels, int samples, float4* acc32, double* acc64, hipStream_t s, bool fresh)
{
	k_resolve<<<(pixels + 255) / 256, 256, 0, s>>>(pathAcc, pixels, samples, acc32, acc64, fresh ? 1 : 0);
	return hipGetLastError();
}

hipError_t launch_finalize(const float4* acc32, int pixels, float scale, float4* image, hipStream_t s)
{
	k_finalize<<<(pixels + 255) / 256, 256, 0, s>>>(acc32, pixels, scale, image);
	return hipGetLastError();
}

hipError_t launch_set_u32(uint32_t* p, int n, uint32_t v, hipStream_t s)
{
	k_set_u32<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
	return hipGetLastError();
}

hipError_t launch_copy_u32(uint32_t* dst, const uint32_t* src, hipStream_t s)
{
	k_copy_u32<<<1, 1, 0, s>>>(dst, src);
	return hipGetLastError();
}

} // namespace lh
