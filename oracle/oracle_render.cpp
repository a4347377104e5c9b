/* oracle_render.cpp - CPU oracle of the reference estimator ("REFERENCE, for
 * image validation": RenderCore_PrimeRef). TEST INFRASTRUCTURE.
 *
 * Each function restates the reference lines it cites, in the same arithmetic
 * order, in fp32. Paths are traced depth-first: a path's result does not
 * depend on the wavefront's compaction order, and its contributions reach its
 * accumulator slot in the order the wavefront deposits them (connect of
 * bounce k before shade of bounce k+1). */
#include "oracle_render.h"
#include <thread>
#include <vector>
#include <algorithm>

namespace orc {

static const float PI = 3.14159265358979323846264f;
static const float INVPI = 0.31830988618379067153777f;
static const float TWOPI = 6.28318530717958647692528f;
static const float EPSILON = 0.0001f;

static inline f3 f3of(const lh2hip::lh2_float3& a) { return mk3(a.x, a.y, a.z); }

/* TOCHAR of rendercore.cpp:359, saturated to a byte (the reference overflows
 * into the next byte for values above 1). */
static inline uint tochar(float a)
{
	const float s = a * 255.0f;
	if (!(s > 0.0f)) return 0;
	if (s >= 255.0f) return 255;
	return (uint)s;
}
static inline uint touint4(float a, float b, float c, float d) { return tochar(a) + (tochar(b) << 8) + (tochar(c) << 16) + (tochar(d) << 24); }

/* RenderCore::SetMaterials (lib/RenderCore_PrimeRef/rendercore.cpp:360-402). */
OMat convert_material(const CoreMaterial& m, const std::vector<TexInfo>& tex)
{
	OMat g;
	memset(&g, 0, sizeof(g));
	g.diffuse[0] = float_to_half(m.color.value.x);
	g.diffuse[1] = float_to_half(m.color.value.y);
	g.diffuse[2] = float_to_half(m.color.value.z);
	g.transm[0] = float_to_half(1.0f - m.absorption.value.x);
	g.transm[1] = float_to_half(1.0f - m.absorption.value.y);
	g.transm[2] = float_to_half(1.0f - m.absorption.value.z);
	g.params[0] = touint4(m.metallic.value, m.subsurface.value, m.specular.value, m.roughness.value);
	g.params[1] = touint4(m.specularTint.value, m.anisotropic.value, m.sheen.value, m.sheenTint.value);
	g.params[2] = touint4(m.clearcoat.value, m.clearcoatGloss.value, m.transmission.value, 0.0f);
	g.params[3] = f2u(m.eta.value);
	const auto map = [&](uint* d, int id, decltype(m.color.uvscale) scale, decltype(m.color.uvoffset) offs) {
		if (id < 0 || id >= (int)tex.size()) return;
		const TexInfo& t = tex[id];
		d[0] = (t.width & 0xffffu) | ((t.height & 0xffffu) << 16);
		d[1] = (uint)float_to_half(scale.x) | ((uint)float_to_half(scale.y) << 16);
		d[2] = (uint)float_to_half(offs.x) | ((uint)float_to_half(offs.y) << 16);
		d[3] = t.firstPixel;
	};
	map(g.maps[0], m.color.textureID, m.color.uvscale, m.color.uvoffset);
	map(g.maps[1], m.detailColor.textureID, m.detailColor.uvscale, m.detailColor.uvoffset);
	map(g.maps[2], m.normals.textureID, m.normals.uvscale, m.normals.uvoffset);
	map(g.maps[3], m.detailNormals.textureID, m.detailNormals.uvscale, m.detailNormals.uvoffset);
	map(g.maps[4], m.specular.textureID, m.specular.uvscale, m.specular.uvoffset);
	map(g.maps[5], m.roughness.textureID, m.roughness.uvscale, m.roughness.uvoffset);
	const bool hdr = m.color.textureID >= 0 && m.color.textureID < (int)tex.size() && (tex[m.color.textureID].flags & 8u);
	g.flags = (m.eta.value < 1 ? M_ISDIELECTRIC : 0) + (hdr ? M_DIFFUSEMAPISHDR : 0) +
		(m.color.textureID != -1 ? M_HASDIFFUSEMAP : 0) + (m.normals.textureID != -1 ? M_HASNORMALMAP : 0) +
		(m.specular.textureID != -1 ? M_HASSPECULARITYMAP : 0) + (m.roughness.textureID != -1 ? M_HASROUGHNESSMAP : 0) +
		(m.detailNormals.textureID != -1 ? M_HAS2NDNORMALMAP : 0) + (m.detailColor.textureID != -1 ? M_HAS2NDDIFFUSEMAP : 0) +
		((m.flags & 1) ? M_HASSMOOTHNORMALS : 0) + ((m.flags & 2) ? M_HASALPHA : 0);
	return g;
}

static inline float char2flt(uint a, int s) { return ((float)((a >> s) & 255u)) * (1.0f / 255.0f); }

struct Shading
{
	f3 color, transmittance;
	uint params[4];
	float roughness() const { const float r = char2flt(params[0], 24); return r > 0.001f ? r : 0.001f; }
	float transmission() const { return char2flt(params[2], 16); }
	float eta() const { return u2f(params[3]); }
	bool emissive() const { return color.x > 1.0f || color.y > 1.0f || color.z > 1.0f; }
	f4 tint;   /* tint rgb and luminance of the untextured base colour (Disney BSDF) */
};

#include "oracle_disney.h"

/* ConsistentNormal, Reshetov et al. 2010, Eq. 1 approximation
 * (lib/CUDA/shared_kernel_code/tools_shared.h:266-280). */
static f3 consistent_normal(f3 D, f3 iN, float alpha)
{
	const float t = PI - 2 * alpha, q = (t * t) / (PI * (PI + (2 * PI - 4) * alpha));
	const float b = dot(D, iN), g = 1 + q * (b - 1), rho = sqrtf(q * (1 + g) / (1 + b));
	const f3 Rc = (g + rho * b) * iN - (rho * D);
	return normalize(D + Rc);
}

/* __uchar4_to_float4 (sampling_shared.h:29-33): byte k of the texel word is
 * component k, scaled by 1/256. */
static inline f4 unpack_texel(uint t)
{
	const float r = 1.0f / 256.0f;
	f4 o = { (float)(t & 255u) * r, (float)((t >> 8) & 255u) * r, (float)((t >> 16) & 255u) * r, (float)(t >> 24) * r };
	return o;
}

static inline f4 texel_at(const RenderScene& s, int storage, int idx)
{
	if (storage == 1) return s.argb128[idx];
	return unpack_texel(storage == 2 ? s.nrm32[idx] : s.argb32[idx]);
}

/* FetchTexel, BILINEAR on (lib/CUDA/shared_kernel_code/sampling_shared.h:35-79;
 * lib/RenderCore_PrimeRef/core_settings.h:31): wrap by +1000, bilinear
 * weights with w3 = 1 - (w0 + w1 + w2). */
f4 fetch_texel(const RenderScene& s, float tu, float tv, int o, int w, int h, int storage)
{
	const float tcx = (std::max(tu + 1000.0f, 0.0f) * (float)w) - 0.5f;
	const float tcy = (std::max(tv + 1000.0f, 0.0f) * (float)h) - 0.5f;
	const int iu = f2i_sat(tcx) % w, iv = f2i_sat(tcy) % h;
	const float fu = tcx - floorf(tcx), fv = tcy - floorf(tcy);
	const float w0 = (1 - fu) * (1 - fv), w1 = fu * (1 - fv), w2 = (1 - fu) * fv, w3 = 1 - (w0 + w1 + w2);
	const int iu1 = (iu + 1) % w, iv1 = (iv + 1) % h;
	const f4 p0 = texel_at(s, storage, o + iu + iv * w), p1 = texel_at(s, storage, o + iu1 + iv * w);
	const f4 p2 = texel_at(s, storage, o + iu + iv1 * w), p3 = texel_at(s, storage, o + iu1 + iv1 * w);
	f4 r;
	r.x = p0.x * w0 + p1.x * w1 + p2.x * w2 + p3.x * w3;
	r.y = p0.y * w0 + p1.y * w1 + p2.y * w2 + p3.y * w3;
	r.z = p0.z * w0 + p1.z * w1 + p2.z * w2 + p3.z * w3;
	r.w = p0.w * w0 + p1.w * w1 + p2.w * w2 + p3.w * w3;
	return r;
}

/* FetchTexelTrilinear (sampling_shared.h:81-104), MIPLEVELCOUNT 5
 * (common_settings.h:57); level offsets by the closed form of the reference,
 * scale * (1 - 4^-level). */
f4 fetch_texel_trilinear(const RenderScene& s, float lambda, float tu, float tv, int offset, int width, int height)
{
	int level0 = 0, level1 = 0;
	float f = 0;
	if (lambda >= 0) level0 = std::min(4, f2i_sat(lambda)), level1 = std::min(4, level0 + 1), f = lambda - floorf(lambda);
	const float scale = (float)(width * height) * 1.3333333333f;
	const int o0 = offset + f2i_sat(scale * (1 - u2f((uint)(127 - 2 * level0) << 23)));
	const int o1 = offset + f2i_sat(scale * (1 - u2f((uint)(127 - 2 * level1) << 23)));
	const f4 p0 = fetch_texel(s, tu, tv, o0, width >> level0, height >> level0, 0);
	const f4 p1 = fetch_texel(s, tu, tv, o1, width >> level1, height >> level1, 0);
	f4 r;
	r.x = (1 - f) * p0.x + f * p1.x, r.y = (1 - f) * p0.y + f * p1.y;
	r.z = (1 - f) * p0.z + f * p1.z, r.w = (1 - f) * p0.w + f * p1.w;
	return r;
}

/* SampleSkydome(-worldToSky.TransformVector(D)) (lib/CUDA/shared_kernel_code/
 * tools_shared.h:184-202; common_types.h:608-613; kernels/pathtracer.h:73-78):
 * equirect texel (uint)(w * phi / 2pi - 0.5), (uint)(h * theta / pi - 0.5). */
f3 sample_sky(const RenderScene& s, f3 D)
{
	if (s.skyW <= 0 || s.skyH <= 0) return mk3(0, 0, 0);
	const float* m = s.worldToSky;
	const f3 d = mk3(-(m[0] * D.x + m[1] * D.y + m[2] * D.z), -(m[4] * D.x + m[5] * D.y + m[6] * D.z), -(m[8] * D.x + m[9] * D.y + m[10] * D.z));
	const float p = lh_atan2f(d.y, d.x);
	const float phi = p < 0 ? (p + 2 * PI) : p;
	const float theta = lh_acosf(std::min(std::max(d.z, -1.0f), 1.0f));
	const uint u = f2u_sat((float)s.skyW * phi * 0.15915494309189533577f - 0.5f);
	const uint v = f2u_sat((float)s.skyH * theta * 0.31830988618379067154f - 0.5f);
	const uint idx = u + v * (uint)s.skyW;
	if (idx >= (uint)s.skyW * (uint)s.skyH) return mk3(0, 0, 0);
	const f4 t = s.sky[idx];
	return mk3(t.x, t.y, t.z);
}

static inline float h2f(uint bits16) { return half_to_float((uint16_t)(bits16 & 0xffffu)); }

/* SetupFrame + GetShadingData, OPTIXPRIMEBUILD barycentric convention
 * (lib/CUDA/shared_kernel_code/material_shared.h:42-94,96-206), CONSISTENTNORMALS
 * on (lib/RenderCore_PrimeRef/core_settings.h:26). Texture maps: diffuse
 * (trilinear, LOD TRI_LOD + log2(coneWidth / |D.N|); alpha below 0.5 ends the
 * lookup with the untextured colour), detail diffuse, normal maps (tangent
 * frame T, B of the triangle, not instance-transformed, as the reference) and
 * roughness. __expf of the normal-map scale is cw_expf. Returns the ALPHA flag. */
static bool shading_data(const RenderScene& s, f3 D, float u, float v, float coneWidth, const CoreTri& tri, int instIdx,
	Shading& sd, f3& N, f3& iN, f3& fN, f3& T, bool prime = true)
{
	const OMat& mat = s.mats[tri.material];
	sd.color = mk3(half_to_float(mat.diffuse[0]), half_to_float(mat.diffuse[1]), half_to_float(mat.diffuse[2]));
	sd.tint = { 1, 1, 1, 0 };
	if (!prime)
	{	/* tint4 = (ciexyz_to_linear_rgb(xyz / xyz.y), xyz.y) of the base colour (material_shared.h:19-33,118-119) */
		const f3 c = sd.color;
		const f3 X = mk3(fmaxf(0.0f, 0.412453f * c.x + 0.357580f * c.y + 0.180423f * c.z), fmaxf(0.0f, 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z),
			fmaxf(0.0f, 0.019334f * c.x + 0.119193f * c.y + 0.950227f * c.z));
		f3 t = mk3(1, 1, 1);
		if (X.y > 0)
		{
			const f3 n = X * (1.0f / X.y);
			t = mk3(fmaxf(0.0f, 3.240479f * n.x - 1.537150f * n.y - 0.498535f * n.z), fmaxf(0.0f, -0.969256f * n.x + 1.875992f * n.y + 0.041556f * n.z),
				fmaxf(0.0f, 0.055648f * n.x - 0.204043f * n.y + 1.057311f * n.z));
		}
		sd.tint = { t.x, t.y, t.z, X.y };
	}
	sd.transmittance = mk3(half_to_float(mat.transm[0]), half_to_float(mat.transm[1]), half_to_float(mat.transm[2]));
	for (int i = 0; i < 4; i++) sd.params[i] = mat.params[i];
	N = iN = fN = mk3(tri.Nx, tri.Ny, tri.Nz);
	T = f3of(tri.T);
	const float w = 1 - (u + v);
	/* !prime: the non-OPTIXPRIMEBUILD convention (w, u, v) and CONSISTENTNORMALS off,
	 * as rendercore_optix7 compiles material_shared.h (core_settings.h:26) */
	if (mat.flags & M_HASSMOOTHNORMALS)
		iN = prime ? normalize(u * f3of(tri.vN0) + v * f3of(tri.vN1) + w * f3of(tri.vN2)) : normalize(w * f3of(tri.vN0) + u * f3of(tri.vN1) + v * f3of(tri.vN2));
	const float* inv = s.inst[instIdx].inv;
	const f3 A = mk3(inv[0], inv[1], inv[2]), B = mk3(inv[4], inv[5], inv[6]), C = mk3(inv[8], inv[9], inv[10]);
	N = normalize(N.x * A + N.y * B + N.z * C);
	iN = normalize(iN.x * A + iN.y * B + iN.z * C);
	if (prime)
	{
		const bool backSide = dot(D, N) > 0;
		const float alpha = u * tri.alpha.x + v * tri.alpha.y + w * tri.alpha.z;
		iN = (backSide ? -1.0f : 1.0f) * consistent_normal(D * -1.0f, backSide ? (iN * -1.0f) : iN, alpha);
	}
	fN = iN;
	const uint fl = mat.flags;
	float tu = 0, tv = 0;
	if (fl & (M_HASDIFFUSEMAP | M_HAS2NDDIFFUSEMAP | M_HASSPECULARITYMAP | M_HASNORMALMAP | M_HAS2NDNORMALMAP | M_HASROUGHNESSMAP))
	{
		const float ww = 1 - (u + v);
		tu = prime ? u * tri.u0 + v * tri.u1 + ww * tri.u2 : ww * tri.u0 + u * tri.u1 + v * tri.u2;
		tv = prime ? u * tri.v0 + v * tri.v1 + ww * tri.v2 : ww * tri.v0 + u * tri.v1 + v * tri.v2;
	}
	const auto uvmap = [&](const uint* m, float& x, float& y) {
		x = h2f(m[1]) * (h2f(m[2]) + tu);
		y = h2f(m[1] >> 16) * (h2f(m[2] >> 16) + tv);
	};
	if (fl & M_HASDIFFUSEMAP)
	{
		const float lambda = tri.LOD + lh_log2f(coneWidth * (1.0f / fabsf(dot(D, N))));
		const uint* m = mat.maps[0];
		float x, y;
		uvmap(m, x, y);
		const f4 texel = fetch_texel_trilinear(s, lambda, x, y, (int)m[3], (int)(m[0] & 0xffffu), (int)(m[0] >> 16));
		if (texel.w < 0.5f) return true;
		sd.color = sd.color * mk3(texel.x, texel.y, texel.z);
		if (fl & M_HAS2NDDIFFUSEMAP)
		{
			const uint* m1 = mat.maps[1];
			uvmap(m1, x, y);
			const f4 t1 = fetch_texel(s, x, y, (int)m1[3], (int)(m1[0] & 0xffffu), (int)(m1[0] >> 16), 0);
			sd.color = sd.color + (mk3(t1.x, t1.y, t1.z) - mk3(0.5f, 0.5f, 0.5f));
		}
	}
	if (fl & M_HASNORMALMAP)
	{
		const f3 Bt = f3of(tri.B);
		const uint part3 = (uint)mat.transm[1] | ((uint)mat.transm[2] << 16);
		const float c0 = (float)((part3 >> 8) & 255) - 128.0f;
		const float n0scale = copysignf(-0.0001f + 0.0001f * cw_expf(0.1f * fabsf(c0)), c0);
		const uint* m = mat.maps[2];
		float x, y;
		uvmap(m, x, y);
		const f4 t = fetch_texel(s, x, y, (int)m[3], (int)(m[0] & 0xffffu), (int)(m[0] >> 16), 2);
		f3 sn = (mk3(t.x, t.y, t.z) - mk3(0.5f, 0.5f, 0.5f)) * 2.0f;
		sn.x = sn.x * n0scale, sn.y = sn.y * n0scale;
		if (fl & M_HAS2NDNORMALMAP)
		{
			const float c1 = (float)((part3 >> 16) & 255) - 128.0f;
			const float n1scale = copysignf(-0.0001f + 0.0001f * cw_expf(0.1f * c1), c1);
			const uint* m1 = mat.maps[3];
			uvmap(m1, x, y);
			const f4 t1 = fetch_texel(s, x, y, (int)m1[3], (int)(m1[0] & 0xffffu), (int)(m1[0] >> 16), 2);
			f3 l1 = (mk3(t1.x, t1.y, t1.z) - mk3(0.5f, 0.5f, 0.5f)) * 2.0f;
			l1.x = l1.x * n1scale, l1.y = l1.y * n1scale;
			sn = sn + l1;
		}
		sn = normalize(sn);
		fN = normalize(sn.x * T + sn.y * Bt + sn.z * iN);
	}
	if (fl & M_HASROUGHNESSMAP)
	{
		const uint* m = mat.maps[5];
		float x, y;
		uvmap(m, x, y);
		const f4 t = fetch_texel(s, x, y, (int)m[3], (int)(m[0] & 0xffffu), (int)(m[0] >> 16), 0);
		sd.params[0] = (sd.params[0] & 0x00ffffffu) + ((uint)f2i_sat(t.y * 255.0f) << 24);
		sd.params[0] = (sd.params[0] & 0xffffff00u) + (uint)f2i_sat(t.x * 255.0f);
	}
	return false;
}

/* RandomBarycentrics (lib/RenderSystem/common_functions.h:52-71); the 0:32
 * fixed-point conversion wraps 2^32 to 0 where the reference is undefined. */
static f3 random_barycentrics(float r0)
{
	const uint uf = (uint)(uint64_t)(r0 * 4294967296.0f);
	float Ax = 1, Ay = 0, Bx = 0, By = 1, Cx = 0, Cy = 0;
	for (int i = 0; i < 16; ++i)
	{
		const int d = (uf >> (2 * (15 - i))) & 0x3;
		float Anx, Any, Bnx, Bny, Cnx, Cny;
		switch (d)
		{
		case 0: Anx = (Bx + Cx) * 0.5f, Any = (By + Cy) * 0.5f; Bnx = (Ax + Cx) * 0.5f, Bny = (Ay + Cy) * 0.5f; Cnx = (Ax + Bx) * 0.5f, Cny = (Ay + By) * 0.5f; break;
		case 1: Anx = Ax, Any = Ay; Bnx = (Ax + Bx) * 0.5f, Bny = (Ay + By) * 0.5f; Cnx = (Ax + Cx) * 0.5f, Cny = (Ay + Cy) * 0.5f; break;
		case 2: Anx = (Bx + Ax) * 0.5f, Any = (By + Ay) * 0.5f; Bnx = Bx, Bny = By; Cnx = (Bx + Cx) * 0.5f, Cny = (By + Cy) * 0.5f; break;
		default: Anx = (Cx + Ax) * 0.5f, Any = (Cy + Ay) * 0.5f; Bnx = (Cx + Bx) * 0.5f, Bny = (Cy + By) * 0.5f; Cnx = Cx, Cny = Cy; break;
		}
		Ax = Anx, Ay = Any, Bx = Bnx, By = Bny, Cx = Cnx, Cy = Cny;
	}
	const float rx = (Ax + Bx + Cx) * 0.3333333f, ry = (Ay + By + Cy) * 0.3333333f;
	return mk3(rx, ry, 1 - rx - ry);
}

/* Potential contributions (lib/CUDA/shared_kernel_code/lights_shared.h:37-108).
 * The triangle light's barycentric point and dot products are fused (fmaf,
 * as the kernels' fma; nvcc contracts the same expressions by default). */
static float fdot3(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
static float potential_tri(const CoreLightTri& L, f3 O, f3 N, f3 I, f3 bary)
{
	f3 Lv = I;
	if (bary.x >= 0)
	{
		const f3 V0 = f3of(L.vertex0), V1 = f3of(L.vertex1), V2 = f3of(L.vertex2);
		Lv = mk3(fmaf(bary.z, V2.x, fmaf(bary.y, V1.x, bary.x * V0.x)), fmaf(bary.z, V2.y, fmaf(bary.y, V1.y, bary.x * V0.y)),
			fmaf(bary.z, V2.z, fmaf(bary.y, V1.z, bary.x * V0.z)));
	}
	Lv = Lv - O;
	const float d2 = fdot3(Lv, Lv);
	const float att = 1.0f / d2;
	Lv = Lv * (1.0f / sqrtf(d2));
	const float LNdotL = std::max(0.0f, -fdot3(f3of(L.N), Lv));
	const float NdotL = std::max(0.0f, fdot3(N, Lv));
	return L.energy * LNdotL * NdotL * att;
}
static float potential_point(const CorePointLight& L, f3 I, f3 N)
{
	const f3 Lv = f3of(L.position) - I;
	const float NdotL = std::max(0.0f, dot(N, normalize(Lv)));
	const float att = 1.0f / dot(Lv, Lv);
	return L.energy * NdotL * att;
}
static float potential_spot(const CoreSpotLight& L, f3 I, f3 N)
{
	f3 Lv = f3of(L.position) - I;
	const float att = 1.0f / dot(Lv, Lv);
	Lv = normalize(Lv);
	const float d = (std::max(0.0f, -dot(Lv, f3of(L.direction))) - L.cosOuter) / (L.cosInner - L.cosOuter);
	const float NdotL = std::max(0.0f, dot(N, Lv));
	const float LNdotL = std::max(0.0f, std::min(1.0f, d));
	return (L.radiance.x + L.radiance.y + L.radiance.z) * LNdotL * NdotL * att;
}
static float potential_dir(const CoreDirectionalLight& L, f3 N)
{
	const float LNdotL = std::max(0.0f, -(L.direction.x * N.x + L.direction.y * N.y + L.direction.z * N.z));
	return L.energy * LNdotL;
}

/* RandomPointOnLight with importance sampling (lights_shared.h:225-313). The
 * potentials are summed in one pass and recomputed (identically) while
 * searching, so there is no 64-light stack array. */
static f3 random_point_on_light(const RenderScene& s, float r0, float r1, f3 I, f3 N, float& pickProb, float& lightPdf, f3& lightColor)
{
	const int nt = (int)s.triLights.size(), np = (int)s.pointLights.size(), ns = (int)s.spotLights.size(), nd = (int)s.dirLights.size();
	const int lightCount = nt + np + ns + nd;
	const f3 bary = random_barycentrics(r0);
	auto pot = [&](int i) -> float {
		if (i < nt) return potential_tri(s.triLights[i], I, N, I, bary);
		if (i < nt + np) return potential_point(s.pointLights[i - nt], I, N);
		if (i < nt + np + ns) return potential_spot(s.spotLights[i - nt - np], I, N);
		return potential_dir(s.dirLights[i - nt - np - ns], N);
	};
	float sum = 0;
	for (int i = 0; i < lightCount; i++) sum += pot(i);
	if (sum <= 0) { lightPdf = 0; return mk3(1, 1, 1); }
	r1 *= sum;
	int lightIdx = 0;
	float total = 0, pl = 0;
	for (int i = 0; i < lightCount; i++)
	{
		const float c = pot(i);
		total += c;
		if (total >= r1) { lightIdx = i; pl = c; break; }
	}
	if (lightIdx == 0) pl = pot(0);
	pickProb = pl / sum;
	if (lightIdx < nt)
	{
		const CoreLightTri& L = s.triLights[lightIdx];
		const f3 V0 = f3of(L.vertex0), V1 = f3of(L.vertex1), V2 = f3of(L.vertex2);
		lightColor = f3of(L.radiance);
		const f3 LN = f3of(L.N);
		const f3 P = bary.x * V0 + bary.y * V1 + bary.z * V2;
		f3 Lv = I - P;
		const float sqDist = dot(Lv, Lv);
		Lv = normalize(Lv);
		const float LNdotL = Lv.x * LN.x + Lv.y * LN.y + Lv.z * LN.z;
		const float reciSolidAngle = sqDist / (L.area * LNdotL);
		lightPdf = (LNdotL > 0 && dot(Lv, N) < 0) ? reciSolidAngle : 0;
		return P;
	}
	else if (lightIdx < nt + np)
	{
		const CorePointLight& L = s.pointLights[lightIdx - nt];
		const f3 P = f3of(L.position), Lv = P - I;
		const float sqDist = dot(Lv, Lv);
		lightColor = f3of(L.radiance) / sqDist;
		lightPdf = dot(Lv, N) > 0 ? 1 : 0;
		return P;
	}
	else if (lightIdx < nt + np + ns)
	{
		const CoreSpotLight& L = s.spotLights[lightIdx - nt - np];
		const f3 P = f3of(L.position);
		f3 Lv = I - P;
		const float sqDist = dot(Lv, Lv);
		Lv = normalize(Lv);
		const float d = (std::max(0.0f, Lv.x * L.direction.x + Lv.y * L.direction.y + Lv.z * L.direction.z) - L.cosOuter) / (L.cosInner - L.cosOuter);
		const float LNdotL = std::min(1.0f, d);
		lightPdf = (LNdotL > 0 && dot(Lv, N) < 0) ? (sqDist / LNdotL) : 0;
		lightColor = f3of(L.radiance);
		return P;
	}
	const CoreDirectionalLight& L = s.dirLights[lightIdx - nt - np - ns];
	const f3 Ld = f3of(L.direction);
	lightColor = f3of(L.radiance);
	const float NdotL = dot(Ld, N);
	lightPdf = NdotL < 0 ? 1 : 0;
	return I - 1000.0f * Ld;
}

/* SetupTangentSpace / Tangent2World, Duff et al. 2017
 * (lib/RenderSystem/common_functions.h:73-88). */
static f3 tangent2world(f3 V, f3 N)
{
	const float sign = copysignf(1.0f, N.z);
	const float a = -1.0f / (sign + N.z);
	const float b = N.x * N.y * a;
	const f3 B = mk3(1.0f + sign * N.x * N.x * a, sign * b, -sign * N.x);
	const f3 T = mk3(b, sign + N.y * N.y * a, -N.y);
	return V.x * T + V.y * B + V.z * N;
}

/* DiffuseReflectionCosWeighted (common_functions.h:104-110). */
static f3 cosine_sample(float r0, float r1)
{
	const float term1 = TWOPI * r0, term2 = sqrtf(1 - r1);
	float sn, cs;
	cw_sincos(term1, &sn, &cs);
	return mk3(cs * term2, sn * term2, sqrtf(r1));
}

static inline f3 reflect(f3 i, f3 n) { return i - 2.0f * n * dot(n, i); }

/* Lambert + pure specular BSDF (lib/RenderCore_PrimeRef/kernels/bsdf.h:19-101).
 * below: the shared Lambert BSDF's Fr_L (sharedBSDFs/lambert.h:32-45), which
 * inverts eta for a view vector below the surface. */
static float fr_l(float VDotN, float eio, bool below = false)
{
	if (below && VDotN < 0.0f) eio = 1.0f / eio, VDotN = fabsf(VDotN);
	const float SinThetaT2 = (eio * eio) * (1.0f - VDotN * VDotN);
	if (SinThetaT2 > 1.0f) return 1.0f;
	const float LDotN = sqrtf(1.0f - SinThetaT2);
	const float r1 = (VDotN - eio * LDotN) / (VDotN + eio * LDotN);
	const float r2 = (LDotN - eio * VDotN) / (LDotN + eio * VDotN);
	return 0.5f * (r1 * r1 + r2 * r2);
}
static bool refract_l(f3 wi, f3 n, float eta, f3& wt)
{
	const float cosThetaI = fabsf(dot(n, wi));
	const float sin2ThetaI = std::max(0.0f, 1.0f - cosThetaI * cosThetaI);
	const float sin2ThetaT = eta * eta * sin2ThetaI;
	if (sin2ThetaT >= 1) return false;
	const float cosThetaT = sqrtf(1.0f - sin2ThetaT);
	wt = eta * (wi * -1.0f) + (eta * cosThetaI - cosThetaT) * n;
	return true;
}
static f3 evaluate_bsdf(const Shading& sd, f3 iN, f3 wo, f3 wi, float& pdf)
{
	(void)wo;
	if (sd.transmission() > 0.999f || sd.roughness() <= 0.001f) { pdf = 0; return mk3(0, 0, 0); }
	pdf = fabsf(dot(wi, iN)) * INVPI;
	return sd.color * INVPI * sd.roughness();
}
static f3 sample_bsdf(const Shading& sd, f3 iN, f3 N, f3 wo, float distance, float r3, float r4, f3& wi, float& pdf, bool& specular,
	bool shared = false)
{
	const float flip = (dot(wo, N) < 0) ? -1 : 1;
	iN = iN * flip;
	specular = true, pdf = 1;
	f3 bsdf;
	const float TR = sd.transmission(), RO = sd.roughness();
	if (r4 < TR)
	{
		const float eio = flip < 0 ? (1.0f / sd.eta()) : sd.eta(), F = fr_l(dot(iN, wo), eio, shared);
		f3 beer = mk3(1, 1, 1);
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
			pdf = std::max(0.0f, dot(wi, iN)) * INVPI;
			specular = false;
			bsdf = sd.color * INVPI;
		}
	}
	if (dot(N * flip, wi) <= 0) pdf = 0;
	return bsdf;
}

static inline f3 safe_origin(f3 O, f3 R, f3 N, float eps) { return O + N * (dot(N, R) > 0 ? eps : -eps); }

/* RandomPointOnLens, 9-blade aperture (lib/RenderCore_PrimeRef/kernels/camera.h:22-33). */
static f3 point_on_lens(float r0, float r1, f3 pos, float aperture, f3 right, f3 up)
{
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

/* generateEyeRaysKernel (lib/RenderCore_PrimeRef/kernels/camera.h:39-68). */
static void eye_ray(const ViewPyramid& v, int w, int h, uint R0, int job, f3& O, f3& D)
{
	const uint x = (uint)job % (uint)w;
	uint y = (uint)job / (uint)w;
	y %= (uint)h;
	uint seed = wang_hash((uint)job + R0);
	const float r0 = rand_float(seed), r1 = rand_float(seed);
	const float r2 = rand_float(seed), r3 = rand_float(seed);
	const f3 pos = f3of(v.pos), p1 = f3of(v.p1);
	const f3 right = f3of(v.p2) - p1, up = f3of(v.p3) - p1;
	const f3 posOnPixel = p1 + ((float)x + r0) * (right / (float)w) + ((float)y + r1) * (up / (float)h);
	const f3 posOnLens = point_on_lens(r2, r3, pos, v.aperture, right, up);
	O = posOnLens;
	D = normalize(posOnPixel - posOnLens);
}

/* One path, depth first: the wavefront loop of RenderCore::Render
 * (lib/RenderCore_PrimeRef/rendercore.cpp:564-611) with shadeKernel
 * (kernels/pathtracer.h:45-168) and finalizeConnectionKernel (connections.h:22-34).
 * NEE rays emitted at the last depth are not traced: the reference loop breaks
 * before the connect step (rendercore.cpp:586-591). */
static void trace_path(const RenderScene& s, const RenderParams& p, int job, uint R0gen, int pass, float* acc4)
{
	const int wh = p.w * p.h;
	f3 O, D;
	eye_ray(p.view, p.w, p.h, R0gen, job, O, D);
	float tmin = p.eps;
	f3 throughput = mk3(1, 1, 1);
	uint flags = S_SPECULAR;
	const uint pathIdx = (uint)job;
	const uint pixelIdx = pathIdx % (uint)wh;
	float ax = 0, ay = 0, az = 0, aw = 0;
	(void)pixelIdx;
	for (int pathLength = 1; pathLength <= p.maxDepth; pathLength++)
	{
		const Hit h = trace_closest(s, O, tmin, D, 1e34f, nullptr);
		const uint R0 = (uint)pass * 7907u + (uint)pathLength * 91771u;
		if (pathLength == 1) aw += h.prim < 0 ? 10000 : h.t;
		if (h.prim < 0)
		{
			const f3 c = throughput * sample_sky(s, D);
			ax += c.x, ay += c.y, az += c.z;
			break;
		}
		const float u = (float)(h.uv & 65535u) * (1.0f / 65535.0f), v = (float)(h.uv >> 16) * (1.0f / 65535.0f);
		const float t = h.t;
		const CoreTri& tri = s.tris[s.inst[h.inst].mesh][h.prim];
		Shading sd;
		f3 N, iN, fN, T;
		const f3 I = O + t * D;
		shading_data(s, D, u, v, p.view.spreadAngle * t, tri, h.inst, sd, N, iN, fN, T);
		if (sd.emissive())
		{
			const float DdotNL = -dot(D, N);
			if (DdotNL > 0)
			{
				const f3 c = throughput * sd.color;
				if (pathLength == 1 || (flags & S_SPECULAR)) ax += c.x, ay += c.y, az += c.z;
			}
			break;
		}
		if (sd.roughness() <= 0.001f || sd.transmission() > 0.999f) flags |= S_SPECULAR; else flags &= ~S_SPECULAR;
		uint seed = wang_hash(pathIdx * 17u + R0);
		const float faceDir = (dot(D, N) > 0) ? -1 : 1;
		if (faceDir == 1) sd.transmittance = mk3(0, 0, 0);
		if ((flags & S_SPECULAR) == 0)
		{
			const float r0 = rand_float(seed), r1 = rand_float(seed);
			float pickProb = 0, lightPdf = 0;
			f3 lightColor = mk3(0, 0, 0);
			f3 L = random_point_on_light(s, r0, r1, I, fN * faceDir, pickProb, lightPdf, lightColor) - I;
			const float dist = length(L);
			L = L * (1.0f / dist);
			const float NdotL = dot(L, fN * faceDir);
			if (NdotL > 0 && lightPdf > 0)
			{
				float bsdfPdf;
				const f3 sampledBSDF = evaluate_bsdf(sd, fN, D * -1.0f, L, bsdfPdf);
				const f3 c = throughput * sampledBSDF * lightColor * (NdotL / (pickProb * lightPdf));
				if (pathLength < p.maxDepth)
				{
					const f3 so = safe_origin(I, L, N, p.eps);
					if (!trace_any(s, so, 0.0f, L, dist - 2 * p.eps, nullptr)) ax += c.x, ay += c.y, az += c.z;
				}
			}
		}
		f3 R;
		float newBsdfPdf;
		bool specular = false;
		const float r3 = rand_float(seed), r4 = rand_float(seed), r5 = rand_float(seed);
		const f3 bsdf = sample_bsdf(sd, fN, N, D * -1.0f, t, r3, r4, R, newBsdfPdf, specular);
		if (newBsdfPdf < EPSILON || std::isnan(newBsdfPdf)) break;
		if (specular) flags |= S_SPECULAR;
		const float pr = pathLength == p.maxDepth ? 0 : ((flags & S_SPECULAR) ? 1 : std::min(1.0f, std::max(std::max(bsdf.x, bsdf.y), bsdf.z)));
		if (pr <= r5) break;
		throughput = throughput * (bsdf * fabsf(dot(fN, R)) / (pr * newBsdfPdf));
		O = safe_origin(I, R, N, p.eps);
		D = R;
		tmin = 0.0f;
	}
	acc4[0] = ax, acc4[1] = ay, acc4[2] = az, acc4[3] = aw;
}

void render_frame(const RenderScene& s, const RenderParams& p, uint R0gen, int pass, float* pathAcc, int threads)
{
	const int n = p.w * p.h * p.spp;
	if (threads < 1) threads = 1;
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; t++)
		pool.emplace_back([&, t]() {
			for (int j = t; j < n; j += threads) trace_path(s, p, j, R0gen, pass, pathAcc + (size_t)j * 4);
		});
	for (auto& th : pool) th.join();
}

/* ==== optix7 estimator ======================================================
 * rendercore_optix7: setupPrimaryRay/generateEyeRay (optix/.optix.cu:85-129),
 * shadeKernel (kernels/pathtracer.h:54-238) with the shared Lambert BSDF
 * (sharedBSDFs/lambert.h; BSDF_HAS_PURE_SPECULARS), connections traced after
 * each shade wave and added in that order (generateShadowRay, .optix.cu:142-154).
 * MAXPATHLENGTH 3, ENOUGH_BOUNCES S_BOUNCED, CLAMPFIREFLIES and BLUENOISE on
 * (core_settings.h:24-33,185-197). */
enum { S7_SPECULAR = 1, S7_BOUNCED = 2, S7_VIASPECULAR = 4, S7_BOUNCEDTWICE = 8 };
static const int MAXPATHLENGTH7 = 3;

/* blueNoiseSampler4 (optix/.optix.cu:56-69) */
static f4 blue_noise4(const std::vector<uint>& bn, int x, int y, int sampleIndex, int dim)
{
	const uint* rk = bn.data() + dim + (x + y * 128) * 8 + 65536 * 3;
	const int rsi1 = (int)((sampleIndex ^ rk[0]) & 255u), rsi2 = (int)((sampleIndex ^ rk[1]) & 255u);
	const int rsi3 = (int)((sampleIndex ^ rk[2]) & 255u), rsi4 = (int)((sampleIndex ^ rk[3]) & 255u);
	const uint v1 = bn[dim + 0 + rsi1 * 256], v2 = bn[dim + 1 + rsi2 * 256], v3 = bn[dim + 2 + rsi3 * 256], v4 = bn[dim + 3 + rsi4 * 256];
	const uint* sc = bn.data() + (dim & 7) + (x + y * 128) * 8 + 65536;
	return { (0.5f + (float)(v1 ^ sc[0])) * (1.0f / 256.0f), (0.5f + (float)(v2 ^ sc[1])) * (1.0f / 256.0f),
		(0.5f + (float)(v3 ^ sc[2])) * (1.0f / 256.0f), (0.5f + (float)(v4 ^ sc[3])) * (1.0f / 256.0f) };
}

/* SampleSmallSkydome (tools_shared.h:204-211) */
static f3 sample_small_sky(const RenderScene& s, f3 D)
{
	if (s.skyW <= 0 || s.skyH <= 0) return mk3(0, 0, 0);
	const float* m = s.worldToSky;
	const f3 d = mk3(-(m[0] * D.x + m[1] * D.y + m[2] * D.z), -(m[4] * D.x + m[5] * D.y + m[6] * D.z), -(m[8] * D.x + m[9] * D.y + m[10] * D.z));
	const uint w = (uint)s.skyW >> 6, h = (uint)s.skyH >> 6;
	const float p = lh_atan2f(d.y, d.x);
	const float phi = p < 0 ? (p + 2 * PI) : p;
	const float theta = lh_acosf(std::min(std::max(d.z, -1.0f), 1.0f));
	const uint u = f2u_sat((float)w * phi * 0.15915494309189533577f - 0.5f);
	const uint v = f2u_sat((float)h * theta * 0.31830988618379067154f - 0.5f);
	const uint idx = u + v * w;
	if (idx >= w * h) return mk3(0, 0, 0);
	const f4 t = s.sky[idx + (size_t)s.skyW * s.skyH];
	return mk3(t.x, t.y, t.z);
}

/* LightPickProb, ISLIGHTS (lights_shared.h:174-190) */
static float light_pick_prob(const RenderScene& s, int idx, f3 O, f3 N, f3 I)
{
	const int nt = (int)s.triLights.size(), np = (int)s.pointLights.size(), ns = (int)s.spotLights.size(), nd = (int)s.dirLights.size();
	float sum = 0, mine = 0;
	for (int i = 0; i < nt + np + ns + nd; i++)
	{
		float c;
		if (i < nt) c = potential_tri(s.triLights[i], O, N, I, mk3(-1, -1, -1));
		else if (i < nt + np) c = potential_point(s.pointLights[i - nt], O, N);
		else if (i < nt + np + ns) c = potential_spot(s.spotLights[i - nt - np], O, N);
		else c = potential_dir(s.dirLights[i - nt - np - ns], N);
		sum += c;
		if (i == idx) mine = c;
	}
	if (sum <= 0) return 0;
	return mine / sum;
}

/* UnpackNormal / PackNormal (tools_shared.h:101-120) */
static f3 unpack_normal(uint p)
{
	float nx = (float)(p & 65535u) * (2.0f / 65535.0f), ny = (float)(p >> 16) * (2.0f / 65535.0f), nz = 0, nw = 0;
	nx += -1, ny += -1, nz += 1, nw += -1;
	float l = nx * -nx + ny * -ny + nz * -nw;
	nz = l, l = sqrtf(l), nx *= l, ny *= l;
	return mk3(nx * 2.0f + 0, ny * 2.0f + 0, nz * 2.0f + -1);
}
static uint pack_normal(f3 N)
{
	const float f = 65535.0f / std::max(sqrtf(8.0f * N.z + 8.0f), 0.0001f);
	return f2u_sat(N.x * f + 32767.0f) + (f2u_sat(N.y * f + 32767.0f) << 16);
}

static f3 clamp_intensity(f3 c, float clampValue)
{
	const float v = std::max(c.x, std::max(c.y, c.z));
	if (v > clampValue) { const float m = clampValue / v; c.x *= m, c.y *= m, c.z *= m; }
	return c;
}
static f3 fixnan(f3 c) { return std::isfinite(c.x + c.y + c.z) ? c : mk3(0, 0, 0); }

/* generateEyeRay + RayTarget (optix/.optix.cu:85-104; common_functions.h:29-50) */
static void eye_ray7(const RenderScene& s, const Render7Params& p, uint pathIdx, f3& O, f3& D)
{
	const uint wh = (uint)(p.w * p.h);
	const uint pixelIdx = pathIdx % wh, sampleIdx = pathIdx / wh + (uint)p.pass;
	uint seed = wang_hash(pathIdx * 16789u + (uint)p.pass * 1791u);
	const int sx = (int)(pixelIdx % (uint)p.w), sy = (int)(pixelIdx / (uint)p.w);
	f4 r4;
	if (sampleIdx < 64) r4 = blue_noise4(s.blueNoise, (sx + (int)(p.shift & 127u)) & 127, (sy + (int)(p.shift >> 24)) & 127, (int)sampleIdx, 0);
	else
	{
		r4.x = rand_float(seed), r4.y = rand_float(seed);
		r4.z = rand_float(seed), r4.w = rand_float(seed);
	}
	const ViewPyramid& v = p.view;
	const f3 pos = f3of(v.pos), p1 = f3of(v.p1);
	const f3 right = f3of(v.p2) - p1, up = f3of(v.p3) - p1;
	O = point_on_lens(r4.x, r4.z, pos, v.aperture, right, up);
	f3 target;
	if (v.distortion == 0)
	{
		const float u = ((float)sx + r4.y) * (1.0f / (float)p.w), vv = ((float)sy + r4.w) * (1.0f / (float)p.h);
		target = p1 + u * right + vv * up;
	}
	else
	{
		const float tx = (float)sx / (float)p.w - 0.5f, ty = (float)sy / (float)p.h - 0.5f;
		const float rr = tx * tx + ty * ty;
		const float rq = sqrtf(rr) * (1.0f + v.distortion * rr + v.distortion * rr * rr);
		const float theta = lh_atan2f(tx, ty);
		float st, ct;
		cw_sincos(theta, &st, &ct);
		const float bx = (st * rq + 0.5f) * (float)p.w, by = (ct * rq + 0.5f) * (float)p.h;
		target = p1 + (bx + r4.y) * (right / (float)p.w) + (by + r4.w) * (up / (float)p.h);
	}
	D = normalize(target - O);
}

/* One optix7 path, depth first; returns the number of shade waves it took part in. */
static int trace_path7(const RenderScene& s, const Render7Params& p, const uint* R0wave, uint pathIdx, float* acc4)
{
	const uint wh = (uint)(p.w * p.h);
	const uint sampleIdx = pathIdx / wh + (uint)p.pass;
	const bool connections = (s.triLights.size() + s.pointLights.size() + s.spotLights.size() + s.dirLights.size()) > 0;
	f3 O, D;
	eye_ray7(s, p, pathIdx, O, D);
	f3 throughput = mk3(1, 1, 1);
	float bsdfPdf = 1;
	uint flags = S7_SPECULAR, lastNormal = 0;
	float ax = 0, ay = 0, az = 0, aw = 0;
	int waves = 0;
	for (int pathLength = 1; pathLength <= MAXPATHLENGTH7; pathLength++)
	{
		waves = pathLength;
		const Hit h = trace_closest(s, O, p.eps, D, 1e34f, nullptr);
		if (h.prim < 0)
		{
			const f3 sky = (flags & S7_BOUNCED) ? sample_small_sky(s, D) : sample_sky(s, D);
			f3 c = throughput * sky * (1.0f / bsdfPdf);
			c = fixnan(clamp_intensity(c, p.clampValue));
			ax += c.x, ay += c.y, az += c.z;
			break;
		}
		const float b0 = (float)(h.uv & 65535u) * (1.0f / 65535.0f), b1 = (float)(h.uv >> 16) * (1.0f / 65535.0f);
		const float t = h.t;
		const CoreTri& tri = s.tris[s.inst[h.inst].mesh][h.prim];
		Shading sd;
		f3 N, iN, fN, T;
		const f3 I = O + t * D;
		/* the hit record holds the weights of vertices 0 and 1 */
		const bool alphaHit = shading_data(s, D, b1, 1.0f - (b0 + b1), p.view.spreadAngle * t, tri, h.inst, sd, N, iN, fN, T, false);
		uint seed = wang_hash(pathIdx * 17u + R0wave[pathLength]);
		if (alphaHit)
		{
			if (pathLength == MAXPATHLENGTH7) break;
			if (!std::isfinite(throughput.x + throughput.y + throughput.z)) throughput = mk3(0, 0, 0);
			O = I + D * p.eps;
			continue;
		}
		if (sd.emissive())
		{
			const float DdotNL = -dot(D, N);
			if (DdotNL > 0)
			{
				f3 c = mk3(0, 0, 0);
				if (pathLength == 1 || (flags & S7_SPECULAR) || !connections) c = sd.color;
				else
				{
					const f3 lastN = unpack_normal(lastNormal);
					const float lightPdf = (t * t) / (fabsf(dot(D, N)) * tri.area);
					const float pickProb = light_pick_prob(s, tri.ltriIdx, O, lastN, I);
					if ((bsdfPdf + lightPdf * pickProb) > 0) c = throughput * sd.color * (1.0f / (bsdfPdf + lightPdf * pickProb));
				}
				c = fixnan(clamp_intensity(c, p.clampValue));
				ax += c.x, ay += c.y, az += c.z;
			}
			break;
		}
		if (flags & S7_BOUNCED) sd.params[0] |= 255u << 24;
		if (sd.roughness() <= 0.001f || sd.transmission() > 0.5f) flags |= S7_SPECULAR; else flags &= ~(uint)S7_SPECULAR;
		const float faceDir = (dot(D, N) > 0) ? -1 : 1;
		if (faceDir == 1) sd.transmittance = mk3(0, 0, 0);
		throughput = throughput * (1.0f / bsdfPdf);
		f4 r4;
		if (sampleIdx < 64)
		{
			const int x = (int)(((pathIdx % (uint)p.w) + (p.shift & 127u)) & 127u);
			const int y = (int)(((pathIdx / (uint)p.w) + (p.shift >> 24)) & 127u);
			r4 = blue_noise4(s.blueNoise, x, y, (int)sampleIdx, 4 * pathLength - 4);
		}
		else
		{
			r4.x = rand_float(seed), r4.y = rand_float(seed);
			r4.z = rand_float(seed), r4.w = rand_float(seed);
		}
		if ((flags & S7_SPECULAR) == 0 && connections)
		{
			float pickProb = 0, lightPdf = 0;
			f3 lightColor = mk3(0, 0, 0);
			f3 L = random_point_on_light(s, r4.x, r4.y, I, fN * faceDir, pickProb, lightPdf, lightColor) - I;
			const float dist = length(L);
			L = L * (1.0f / dist);
			const float NdotL = dot(L, fN * faceDir);
			if (NdotL > 0 && lightPdf > 0)
			{
				float pdf;
				/* Lambert: colour / pi times ROUGHNESS (BSDF_HAS_PURE_SPECULARS); Disney as evaluated */
				const f3 sampled = p.bsdf ? disney_evaluate(sd, fN, T, D * -1.0f, L, pdf) : evaluate_bsdf(sd, fN, D * -1.0f, L, pdf);
				if (pdf > 0)
				{
					f3 c = throughput * sampled * lightColor * (NdotL / (pickProb * lightPdf + pdf));
					c = clamp_intensity(fixnan(c), p.clampValue);
					const f3 so = safe_origin(I, L, N, p.eps);
					if (!trace_any(s, so, p.eps, L, dist - 2 * p.eps, nullptr)) ax += c.x, ay += c.y, az += c.z, aw += 1.0f;
				}
			}
		}
		if ((flags & S7_BOUNCED) || pathLength == MAXPATHLENGTH7) break;
		f3 R;
		float newBsdfPdf;
		bool specular = false;
		const float r5 = rand_float(seed);   /* SampleBSDF's third random number, drawn at the call */
		const f3 bsdf = p.bsdf ? disney_sample(sd, fN, N, T, D * -1.0f, t, r4.z, r4.w, r5, R, newBsdfPdf, specular)
			: sample_bsdf(sd, fN, N, D * -1.0f, t, r4.z, r4.w, R, newBsdfPdf, specular, true);
		if (newBsdfPdf < EPSILON || std::isnan(newBsdfPdf)) break;
		if (specular) flags |= S7_SPECULAR;
		const float pr = ((flags & S7_SPECULAR) || ((flags & S7_BOUNCED) == 0)) ? 1 : std::min(1.0f, std::max(std::max(bsdf.x, bsdf.y), bsdf.z));
		if (pr < rand_float(seed)) break;
		throughput = throughput * (1 / pr);
		lastNormal = pack_normal(fN * faceDir);
		if (!(flags & S7_SPECULAR)) flags |= (flags & S7_BOUNCED) ? S7_BOUNCEDTWICE : S7_BOUNCED; else flags |= S7_VIASPECULAR;
		O = safe_origin(I, R, N, p.eps);
		D = R;
		throughput = fixnan(throughput) * bsdf * fabsf(dot(fN, R));
		bsdfPdf = newBsdfPdf;
	}
	acc4[0] = ax, acc4[1] = ay, acc4[2] = az, acc4[3] = aw;
	return waves;
}

int render_frame7(const RenderScene& s, const Render7Params& p, uint camSeed, float* pathAcc, int threads)
{
	uint R0wave[MAXPATHLENGTH7 + 1] = { 0 };
	uint cs = camSeed;
	for (int k = 1; k <= MAXPATHLENGTH7; k++) R0wave[k] = xor32(cs) + (uint)k * 91771u;
	const int n = p.w * p.h * p.spp;
	if (threads < 1) threads = 1;
	std::vector<int> deepest(threads, 0);
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; t++)
		pool.emplace_back([&, t]() {
			for (int j = t; j < n; j += threads)
				deepest[t] = std::max(deepest[t], trace_path7(s, p, R0wave, (uint)j, pathAcc + (size_t)j * 4));
		});
	for (auto& th : pool) th.join();
	return *std::max_element(deepest.begin(), deepest.end());
}

void blue_noise_sample4(const RenderScene& s, int x, int y, int sampleIndex, int dim, float* out4)
{
	const f4 r = blue_noise4(s.blueNoise, x, y, sampleIndex, dim);
	out4[0] = r.x, out4[1] = r.y, out4[2] = r.z, out4[3] = r.w;
}
f3 sample_sky_small(const RenderScene& s, f3 D) { return sample_small_sky(s, D); }

/* Disney BSDF probes for the closed-form tests: shading data from a colour,
 * the packed parameter words and a transmittance; tint as shading_data computes it. */
static Shading probe_shading(const float* color3, const uint* params4, const float* transm3)
{
	Shading sd;
	sd.color = mk3(color3[0], color3[1], color3[2]);
	sd.transmittance = mk3(transm3[0], transm3[1], transm3[2]);
	for (int i = 0; i < 4; i++) sd.params[i] = params4[i];
	const f3 c = sd.color;
	const f3 X = mk3(fmaxf(0.0f, 0.412453f * c.x + 0.357580f * c.y + 0.180423f * c.z), fmaxf(0.0f, 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z),
		fmaxf(0.0f, 0.019334f * c.x + 0.119193f * c.y + 0.950227f * c.z));
	f3 t = mk3(1, 1, 1);
	if (X.y > 0)
	{
		const f3 n = X * (1.0f / X.y);
		t = mk3(fmaxf(0.0f, 3.240479f * n.x - 1.537150f * n.y - 0.498535f * n.z), fmaxf(0.0f, -0.969256f * n.x + 1.875992f * n.y + 0.041556f * n.z),
			fmaxf(0.0f, 0.055648f * n.x - 0.204043f * n.y + 1.057311f * n.z));
	}
	sd.tint = { t.x, t.y, t.z, X.y };
	return sd;
}

void disney_probe_eval(const float* color3, const uint* params4, const float* transm3, const float* N3, const float* T3,
	const float* wo3, const float* wi3, float* out4)
{
	const Shading sd = probe_shading(color3, params4, transm3);
	float pdf = 0;
	const f3 v = disney_evaluate(sd, mk3(N3[0], N3[1], N3[2]), mk3(T3[0], T3[1], T3[2]), mk3(wo3[0], wo3[1], wo3[2]), mk3(wi3[0], wi3[1], wi3[2]), pdf);
	out4[0] = v.x, out4[1] = v.y, out4[2] = v.z, out4[3] = pdf;
}

void disney_probe_sample(const float* color3, const uint* params4, const float* transm3, const float* N3, const float* T3,
	const float* wo3, const float* r3, float* out8)
{
	const Shading sd = probe_shading(color3, params4, transm3);
	const f3 N = mk3(N3[0], N3[1], N3[2]);
	f3 wi = mk3(0, 0, 0);
	float pdf = 0;
	bool specular = false;
	const f3 v = disney_sample(sd, N, N, mk3(T3[0], T3[1], T3[2]), mk3(wo3[0], wo3[1], wo3[2]), 1.0f, r3[0], r3[1], r3[2], wi, pdf, specular);
	out8[0] = v.x, out8[1] = v.y, out8[2] = v.z, out8[3] = pdf, out8[4] = wi.x, out8[5] = wi.y, out8[6] = wi.z, out8[7] = specular ? 1.0f : 0.0f;
}

void ggx_probe(int which, const float* v3in, const float* m3, float ax, float ay, float r0, float r1, float* out3)
{
	const f3 v = mk3(v3in[0], v3in[1], v3in[2]), m = mk3(m3[0], m3[1], m3[2]);
	if (which == 0) out3[0] = ggx_D(m, ax, ay);
	else if (which == 1) out3[0] = ggx_pdf(v, m, ax, ay);
	else if (which == 2) { const f3 h = ggx_sample(v, r0, r1, ax, ay); out3[0] = h.x, out3[1] = h.y, out3[2] = h.z; }
	else if (which == 3) out3[0] = gtr1_D(m, ax);
	else if (which == 4) { const f3 h = gtr1_sample(r0, r1, ax); out3[0] = h.x, out3[1] = h.y, out3[2] = h.z; }
	else if (which == 5) out3[0] = ggx_G1(v, ax, ay);
}
uint normal_pack(const float* n3) { return pack_normal(mk3(n3[0], n3[1], n3[2])); }
void normal_unpack(uint p, float* n3) { const f3 n = unpack_normal(p); n3[0] = n.x, n3[1] = n.y, n3[2] = n.z; }

/* ==== filter estimator ======================================================
 * RenderCore_Optix7Filter: generateEyeRay / setupPrimaryRay
 * (optix/.optix.cu:66-129), shadeKernel (kernels/pathtracer.h:76-306) with
 * PackFeatures / calculateDepthDerivatives (:44-58); MAXPATHLENGTH 7, one
 * diffuse bounce; radiance after a diffuse bounce goes to the indirect sum
 * when the G-buffer is written. Connections are traced after each shade wave
 * and added in that order. */
enum { SF_SPECULAR = 1, SF_BOUNCED = 2, SF_VIASPECULAR = 4 };
static const int MAXPATHLENGTH_F = 7;

/* blueNoiseSampler (optix/.optix.cu:40-54) */
static float blue_noise1(const std::vector<uint>& bn, int x, int y, int sampleIndex, int dim)
{
	x &= 127, y &= 127, sampleIndex &= 255, dim &= 255;
	const int ranked = (int)((uint)(sampleIndex ^ (int)bn[dim + (x + y * 128) * 8 + 65536 * 3]) & 255u);
	uint value = bn[dim + ranked * 256];
	value ^= bn[(dim & 7) + (x + y * 128) * 8 + 65536];
	return (0.5f + (float)value) * (1.0f / 256.0f);
}
/* PackNormal2, HDRtoRGB32, RGB32toHDR (tools_shared.h:122-129,162-175) */
static uint pack_normal2(f3 N)
{
	const uint x = std::min(f2u_sat((N.x + 1) * 511), 1023u), y = std::min(f2u_sat((N.y + 1) * 511), 1023u), z = std::min(f2u_sat((N.z + 1) * 511), 1023u);
	return (x << 2u) + (y << 12u) + (z << 22u);
}
static uint hdr_to_rgb32(f3 c)
{
	const uint r = f2u_sat(1023.0f * fminf(1.0f, c.x)), g = f2u_sat(2047.0f * fminf(1.0f, c.y)), b = f2u_sat(2047.0f * fminf(1.0f, c.z));
	return (r << 22) + (g << 11) + b;
}
static f3 rgb32_to_hdr(uint c)
{
	return mk3((float)(c >> 22) * (1.0f / 1023.0f), (float)((c >> 11) & 2047u) * (1.0f / 2047.0f), (float)(c & 2047u) * (1.0f / 2047.0f));
}
static void pack_features(uint* f, uint albedo, uint packedNormal, float t, uint isSpecular, uint matid)
{
	f[0] = albedo, f[1] = packedNormal, f[2] = f2u(t), f[3] = (isSpecular << 4) + (matid << 6) + (f[3] & 15u);
}
/* The view moved by the TAA jitter (RenderCore_Optix7Filter/rendercore.cpp:781-786):
 * p1, p2, p3 += (j0 / w) (p2 - p1) + (j1 / h) (p3 - p1). */
static void jittered_corners(const RenderFParams& p, f3& p1, f3& p2, f3& p3)
{
	const ViewPyramid& v = p.view;
	p1 = f3of(v.p1), p2 = f3of(v.p2), p3 = f3of(v.p3);
	if (p.j0 == 0 && p.j1 == 0) return;
	const f3 right = p2 - p1, up = p3 - p1;
	const float sx = p.j0 / (float)p.w, sy = p.j1 / (float)p.h;
	const f3 j = mk3(sx * right.x + sy * up.x, sx * right.y + sy * up.y, sx * right.z + sy * up.z);
	p1 = p1 + j, p2 = p2 + j, p3 = p3 + j;
}

static void depth_derivatives(const RenderFParams& p, int x, int y, float depth, const CoreTri& tri, float* dd)
{
	const f3 triN = mk3(tri.Nx, tri.Ny, tri.Nz);
	const ViewPyramid& v = p.view;
	f3 jp1, jp2, jp3;
	jittered_corners(p, jp1, jp2, jp3);   /* the shade kernel receives the jittered view (rendercore.cpp:836) */
	const f3 p1 = jp1, p2 = jp2 - jp1, p3 = jp3 - jp1, pos = f3of(v.pos), v0 = f3of(tri.vertex0);
	const float iw = 1.0f / (float)p.w, ih = 1.0f / (float)p.h;
	const f3 dxDir = normalize((p1 + ((float)x + 0.5f + 1) * iw * p2 + ((float)y + 0.5f) * ih * p3) - pos);
	const float dxDepth = dot(v0 - pos, triN) / dot(triN, dxDir);
	const f3 dyDir = normalize((p1 + ((float)x + 0.5f) * iw * p2 + ((float)y + 0.5f + 1) * ih * p3) - pos);
	const float dyDepth = dot(v0 - pos, triN) / dot(triN, dyDir);
	float* o = dd + 4 * (size_t)(x + y * p.w);
	o[0] = 0, o[1] = 0, o[2] = dxDepth - depth, o[3] = dyDepth - depth;
}

/* One path of the filter estimator, depth first; returns the number of shade waves it took part in. */
static int trace_pathF(const RenderScene& s, const RenderFParams& p, const uint* R0wave, uint pathIdx, float* accD, float* accI,
	uint* features, float* worldPos, float* deltaDepth)
{
	const uint wh = (uint)(p.w * p.h);
	const uint sampleIdx = pathIdx / wh + (uint)p.pass;
	const bool gbuf = features != nullptr;
	const bool connections = (s.triLights.size() + s.pointLights.size() + s.spotLights.size() + s.dirLights.size()) > 0;
	const ViewPyramid& v = p.view;
	/* eye ray */
	const uint pix0 = pathIdx % wh;
	const int sx = (int)(pix0 % (uint)p.w), sy = (int)(pix0 / (uint)p.w);
	const f3 pos = f3of(v.pos), right = f3of(v.p2) - f3of(v.p1), up = f3of(v.p3) - f3of(v.p1);
	f3 p1, jp2, jp3;
	jittered_corners(p, p1, jp2, jp3);      /* eye rays: jittered p1, unjittered right/up (rendercore.cpp:782,791-793) */
	float r0 = 0, r1 = 0;
	f3 O = pos;
	if (!gbuf)
	{
		uint seed = wang_hash(pathIdx * 16789u + (uint)p.pass * 1791u);
		float r2, r3;
		if (sampleIdx < 256)
			r0 = blue_noise1(s.blueNoise, sx, sy, (int)sampleIdx, 0), r1 = blue_noise1(s.blueNoise, sx, sy, (int)sampleIdx, 1),
			r2 = blue_noise1(s.blueNoise, sx, sy, (int)sampleIdx, 2), r3 = blue_noise1(s.blueNoise, sx, sy, (int)sampleIdx, 3);
		else
			r0 = rand_float(seed), r1 = rand_float(seed), r2 = rand_float(seed), r3 = rand_float(seed);
		O = point_on_lens(r2, r3, pos, v.aperture, right, up);
	}
	const float uu = ((float)sx + r0) * (1.0f / (float)p.w), vv = ((float)sy + r1) * (1.0f / (float)p.h);
	f3 D = normalize((p1 + uu * right + vv * up) - O);
	f3 throughput = mk3(1, 1, 1);
	float bsdfPdf = 1;
	uint flags = SF_SPECULAR, lastNormal = 0;
	float d4[4] = { 0, 0, 0, 0 }, i4[4] = { 0, 0, 0, 0 };
	int waves = 0;
	uint* feat = gbuf ? features + 4 * (size_t)pathIdx : nullptr;
	auto acc = [&]() -> float* { return (gbuf && (flags & SF_BOUNCED)) ? i4 : d4; };
	auto add = [&](f3 c) { float* a = acc(); a[0] += c.x, a[1] += c.y, a[2] += c.z; };
	auto setwp = [&](f3 P, uint pn) { float* o = worldPos + 4 * (size_t)pathIdx; o[0] = P.x, o[1] = P.y, o[2] = P.z, o[3] = u2f(pn); };
	for (int pathLength = 1; pathLength <= MAXPATHLENGTH_F; pathLength++)
	{
		waves = pathLength;
		const Hit h = trace_closest(s, O, p.eps, D, 1e34f, nullptr);
		const uint pixelIdx = (pathIdx % wh) + ((gbuf && (flags & SF_BOUNCED)) ? wh : 0u);
		const bool firstHit = ((flags & SF_BOUNCED) == 0) && sampleIdx == 0 && gbuf;
		if (pathLength == 1 && firstHit)
		{
			pack_features(feat, 0, 0, 1e34f, 0, 0);
			setwp(mk3(0, 0, 0), 0);
			float* dd = deltaDepth + 4 * (size_t)pathIdx;
			dd[0] = dd[1] = dd[2] = dd[3] = 0;
		}
		if (pathLength == 1) acc()[3] += h.prim < 0 ? 10000.0f : h.t;
		if (h.prim < 0)
		{
			f3 c = throughput * sample_sky(s, D) * (1.0f / bsdfPdf);
			c = fixnan(clamp_intensity(c, p.clampValue));
			add(c);
			if (firstHit)
			{
				const uint isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
				const uint pn = pack_normal2(D * -1.0f) + isSpec;
				pack_features(feat, hdr_to_rgb32(c), pn, h.t, isSpec, 0);
				setwp(O + 50000.0f * D, pn);
				float* dd = deltaDepth + 4 * (size_t)pathIdx;
				dd[0] = dd[1] = dd[2] = dd[3] = 0;
			}
			break;
		}
		const float b0 = (float)(h.uv & 65535u) * (1.0f / 65535.0f), b1 = (float)(h.uv >> 16) * (1.0f / 65535.0f);
		const float t = h.t;
		const CoreTri& tri = s.tris[s.inst[h.inst].mesh][h.prim];
		Shading sd;
		f3 N, iN, fN, T;
		const f3 I = O + t * D;
		const bool alphaHit = shading_data(s, D, b1, 1.0f - (b0 + b1), v.spreadAngle * t, tri, h.inst, sd, N, iN, fN, T, false);
		if (alphaHit)
		{
			if (pathLength == MAXPATHLENGTH_F)
			{
				if (gbuf && sampleIdx == 0 && ((flags & SF_BOUNCED) == 0))
				{
					const uint isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
					const uint pn = pack_normal2(N) + isSpec;
					pack_features(feat, 0, pn, t, isSpec, 0);
					setwp(I, pn);
				}
				break;
			}
			if (!std::isfinite(throughput.x + throughput.y + throughput.z)) throughput = mk3(0, 0, 0);
			O = I + D * p.eps;
			continue;
		}
		if (sd.emissive())
		{
			const float DdotNL = -dot(D, N);
			if (DdotNL > 0)
			{
				f3 c = mk3(0, 0, 0);
				if (pathLength == 1 || (flags & SF_SPECULAR) || !connections) c = sd.color;
				else
				{
					const f3 lastN = unpack_normal(lastNormal);
					const float lightPdf = (t * t) / (fabsf(dot(D, N)) * tri.area);
					const float pickProb = light_pick_prob(s, tri.ltriIdx, O, lastN, I);
					if ((bsdfPdf + lightPdf * pickProb) > 0) c = throughput * sd.color * (1.0f / (bsdfPdf + lightPdf * pickProb));
				}
				c = fixnan(clamp_intensity(c, p.clampValue));
				add(c);
			}
			if (firstHit)
			{
				const uint isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
				const uint pn = pack_normal2(N) + isSpec;
				pack_features(feat, hdr_to_rgb32(sd.color), pn, t, isSpec, 0);
				setwp(I, pn);
				depth_derivatives(p, (int)(pathIdx % (uint)p.w), (int)(pathIdx / (uint)p.w), t, tri, deltaDepth);
			}
			break;
		}
		if (sd.roughness() <= 0.001f || sd.transmission() > 0.999f) flags |= SF_SPECULAR; else flags &= ~(uint)SF_SPECULAR;
		uint seed = wang_hash(pathIdx * 17u + R0wave[pathLength]);
		if (firstHit)
		{
			if (flags & SF_SPECULAR) feat[0] = hdr_to_rgb32(sd.color);
			else
			{
				f3 albedo = sd.color;
				if (flags & SF_VIASPECULAR) albedo = albedo * rgb32_to_hdr(feat[0]);
				const uint isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
				const uint pn = pack_normal2(fN * -1.0f) + isSpec;   /* PackNormal2(flip ? -fN : fN), flip = +-1 */
				pack_features(feat, hdr_to_rgb32(albedo), pn, t, isSpec, 0);
				setwp(I, pn);
				depth_derivatives(p, (int)(pathIdx % (uint)p.w), (int)(pathIdx / (uint)p.w), t, tri, deltaDepth);
			}
		}
		const float faceDir = (dot(D, N) > 0) ? -1 : 1;
		if (faceDir == 1) sd.transmittance = mk3(0, 0, 0);
		throughput = throughput * (1.0f / bsdfPdf);
		if ((flags & SF_SPECULAR) == 0 && connections)
		{
			float q0, q1;
			if (sampleIdx < 2)
			{
				const int x = (int)((pixelIdx % (uint)p.w) & 127u), y = (int)((pixelIdx / (uint)p.w) & 127u);
				q0 = blue_noise1(s.blueNoise, x, y, (int)sampleIdx + p.blueSlot, 4 + 4 * pathLength);
				q1 = blue_noise1(s.blueNoise, x, y, (int)sampleIdx + p.blueSlot, 5 + 4 * pathLength);
			}
			else q0 = rand_float(seed), q1 = rand_float(seed);
			float pickProb = 0, lightPdf = 0;
			f3 lightColor = mk3(0, 0, 0);
			f3 L = random_point_on_light(s, q0, q1, I, fN * faceDir, pickProb, lightPdf, lightColor) - I;
			const float dist = length(L);
			L = L * (1.0f / dist);
			const float NdotL = dot(L, fN * faceDir);
			if (NdotL > 0 && lightPdf > 0)
			{
				float pdf;
				const f3 sampled = p.bsdf ? disney_evaluate(sd, fN, T, D * -1.0f, L, pdf) : evaluate_bsdf(sd, fN, D * -1.0f, L, pdf);
				if (pdf > 0)
				{
					f3 c = throughput * sampled * lightColor * (NdotL / (pickProb * lightPdf + pdf));
					c = clamp_intensity(fixnan(c), p.clampValue);
					const f3 so = safe_origin(I, L, N, p.eps);
					if (!trace_any(s, so, p.eps, L, dist - 2 * p.eps, nullptr)) { float* a = acc(); a[0] += c.x, a[1] += c.y, a[2] += c.z, a[3] += 1.0f; }
				}
			}
		}
		if (flags & SF_BOUNCED) break;
		if (pathLength == MAXPATHLENGTH_F)
		{
			if (firstHit)
			{
				const uint isSpec = (flags & SF_VIASPECULAR) ? 1u : 0u;
				const uint pn = pack_normal2(N) + isSpec;
				pack_features(feat, 0, pn, t, isSpec, 0);
				setwp(I, pn);
			}
			break;
		}
		float q3, q4;
		if (sampleIdx < 256)
		{
			const int x = (int)((pixelIdx % (uint)p.w) & 127u), y = (int)((pixelIdx / (uint)p.w) & 127u);
			q3 = blue_noise1(s.blueNoise, x, y, (int)sampleIdx + p.blueSlot, 6 + 4 * pathLength);
			q4 = blue_noise1(s.blueNoise, x, y, (int)sampleIdx + p.blueSlot, 7 + 4 * pathLength);
		}
		else q3 = rand_float(seed), q4 = rand_float(seed);
		f3 R;
		float newBsdfPdf;
		bool specular = false;
		const float q5 = rand_float(seed);
		const f3 bsdf = p.bsdf ? disney_sample(sd, fN, N, T, D * -1.0f, t, q3, q4, q5, R, newBsdfPdf, specular)
			: sample_bsdf(sd, fN, N, D * -1.0f, t, q3, q4, R, newBsdfPdf, specular, true);
		if (newBsdfPdf < EPSILON || std::isnan(newBsdfPdf)) break;
		if (specular) flags |= SF_SPECULAR;
		lastNormal = pack_normal(fN * faceDir);
		if (!(flags & SF_SPECULAR)) flags |= SF_BOUNCED; else flags |= SF_VIASPECULAR;
		O = safe_origin(I, R, N, p.eps);
		D = R;
		throughput = fixnan(throughput) * bsdf * fabsf(dot(fN, R));
		bsdfPdf = newBsdfPdf;
	}
	for (int k = 0; k < 4; k++) accD[k] = d4[k], accI[k] = i4[k];
	return waves;
}

int render_frameF(const RenderScene& s, const RenderFParams& p, uint camSeed, float* pathAcc, uint* features, float* worldPos,
	float* deltaDepth, int threads)
{
	uint R0wave[MAXPATHLENGTH_F + 1] = { 0 };
	uint cs = camSeed;
	for (int k = 1; k <= MAXPATHLENGTH_F; k++) R0wave[k] = xor32(cs) + (uint)k * 91771u;
	const int n = p.w * p.h * p.spp;
	if (threads < 1) threads = 1;
	std::vector<int> deepest(threads, 0);
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; t++)
		pool.emplace_back([&, t]() {
			for (int j = t; j < n; j += threads)
				deepest[t] = std::max(deepest[t], trace_pathF(s, p, R0wave, (uint)j, pathAcc + (size_t)j * 4, pathAcc + ((size_t)n + j) * 4,
					features, worldPos, deltaDepth));
		});
	for (auto& th : pool) th.join();
	return *std::max_element(deepest.begin(), deepest.end());
}

/* Seeded blue-noise tables of the core (xorshift32 bytes in the sampler's layout). */
void blue_noise_default(uint* out, uint seed)
{
	uint x = seed ? seed : 1u;
	const int words = 65536 * 5 + 64;
	for (int i = 0; i < words; i++) out[i] = 0;
	for (int i = 0; i < 65536; i++) out[i] = xor32(x) & 255u;
	for (int i = 0; i < 128 * 128 * 8; i++) out[65536 + i] = xor32(x) & 255u;
	for (int i = 0; i < 128 * 128 * 8; i++) out[3 * 65536 + i] = xor32(x) & 255u;
}

void primary_ray(const ViewPyramid& v, int w, int h, uint R0, int job, float tmin, float* ray8)
{
	f3 O, D;
	eye_ray(v, w, h, R0, job, O, D);
	ray8[0] = O.x, ray8[1] = O.y, ray8[2] = O.z, ray8[3] = tmin;
	ray8[4] = D.x, ray8[5] = D.y, ray8[6] = D.z, ray8[7] = 1e34f;
}

/* Diffuse bounce of the traversal microbenchmark: geometric normal (object
 * space cross product, instance inverse-transpose applied), flipped against
 * the ray; cosine-weighted direction; origin at the hit point. */
bool bounce_ray(const RenderScene& s, const float* in8, const Hit& h, uint pathId, uint R0, float tmin, float* out8)
{
	if (h.prim < 0) return false;
	const Instance& I = s.inst[h.inst];
	const Mesh& m = s.meshes[I.mesh];
	const f3 v0 = m.v[h.prim * 3], v1 = m.v[h.prim * 3 + 1], v2 = m.v[h.prim * 3 + 2];
	f3 N = cross(v1 - v0, v2 - v0);
	const float* inv = I.inv;
	N = normalize(N.x * mk3(inv[0], inv[1], inv[2]) + N.y * mk3(inv[4], inv[5], inv[6]) + N.z * mk3(inv[8], inv[9], inv[10]));
	const f3 O = mk3(in8[0], in8[1], in8[2]), D = mk3(in8[4], in8[5], in8[6]);
	if (dot(N, D) > 0) N = N * -1.0f;
	const f3 P = O + h.t * D;
	uint seed = wang_hash(pathId * 17u + R0);
	const float r0 = rand_float(seed), r1 = rand_float(seed);
	const f3 R = normalize(tangent2world(cosine_sample(r0, r1), N));
	out8[0] = P.x, out8[1] = P.y, out8[2] = P.z, out8[3] = tmin;
	out8[4] = R.x, out8[5] = R.y, out8[6] = R.z, out8[7] = 1e34f;
	return true;
}

} // namespace orc
