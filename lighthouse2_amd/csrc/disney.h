/* disney.h - the Disney principled BSDF of the optix7 cores, device side.
 *
 * Restates lib/sharedBSDFs/disney.h:33-360 (diffuse, sheen, GGX specular,
 * GTR1 clearcoat, rough dielectric), ggxmdf.h:39-223 (anisotropic GGX and
 * GTR1 distributions, visible-normal sampling) and frosted.h:24-103 (Fresnel,
 * refraction helpers), in the reference's operation order. Included by
 * shade.hip after the Shading struct. Differences of form, each deliberate:
 *   - sincosf/expf/logf/powf are this project's own (cw_sincos, cw_expf,
 *     lh_log2f based), identical in the oracle;
 *   - outputs the reference leaves unset on early returns (pdf, value, the
 *     sampled direction, the running contribution) start at zero, and a
 *     component evaluation that returns early leaves the running contribution
 *     as it was;
 *   - normalize is 1/sqrt, not rsqrtf. */
#pragma once
/* (included inside namespace lh) */

#define DISNEY_GGX 0
#define DISNEY_GTR1 1

__device__ __forceinline__ float dn_sqr(float a) { return a * a; }
__device__ __forceinline__ float dn_saturate(float x) { return fmaxf(0.0f, fminf(1.0f, x)); }
__device__ __forceinline__ float dn_lerp(float a, float b, float t) { return a + t * (b - a); }
/* mix of tools_shared.h:98: clamps t to the ends */
__device__ __forceinline__ float dn_mix(float a, float b, float x) { return x <= 0 ? a : x >= 1 ? b : dn_lerp(a, b, x); }
__device__ __forceinline__ float dn_logf(float x) { return lh_log2f(x) * 0.693147180559945309f; }
__device__ __forceinline__ float dn_powf(float a, float b) { return cw_expf(b * dn_logf(a)); }
__device__ __forceinline__ v3 dn_reflect(v3 i, v3 n) { return i - 2.0f * n * dot(n, i); }
__device__ __forceinline__ v3 w2t(v3 V, v3 N, v3 T, v3 B) { return mk3(dot(V, T), dot(V, B), dot(V, N)); }
__device__ __forceinline__ v3 t2w(v3 V, v3 N, v3 T, v3 B) { return V.x * T + V.y * B + V.z * N; }

/* material parameters (tools_shared.h:40-54) */
__device__ __forceinline__ float p_metallic(const Shading& s) { return char2flt(s.params.x, 0); }
__device__ __forceinline__ float p_subsurface(const Shading& s) { return char2flt(s.params.x, 8); }
__device__ __forceinline__ float p_specular(const Shading& s) { return char2flt(s.params.x, 16); }
__device__ __forceinline__ float p_spectint(const Shading& s) { return char2flt(s.params.y, 0); }
__device__ __forceinline__ float p_anisotropic(const Shading& s) { return char2flt(s.params.y, 8); }
__device__ __forceinline__ float p_sheen(const Shading& s) { return char2flt(s.params.y, 16); }
__device__ __forceinline__ float p_sheentint(const Shading& s) { return char2flt(s.params.y, 24); }
__device__ __forceinline__ float p_clearcoat(const Shading& s) { return char2flt(s.params.z, 0); }
__device__ __forceinline__ float p_clearcoatgloss(const Shading& s) { return char2flt(s.params.z, 8); }

/* ---- ggxmdf.h ---- */
__device__ float ggx_D(v3 m, float ax, float ay)
{
	if (m.z == 0) return dn_sqr(ax) * INVPI_F;
	const float c2 = dn_sqr(m.z);
	const float s = sqrtf(fmaxf(0.0f, 1 - c2));
	const float t2 = (1.0f - c2) / c2;
	float stretched;
	if (ax == ay || s == 0.0f) stretched = 1.0f / dn_sqr(ax);
	else stretched = dn_sqr(m.x / (s * ax)) + dn_sqr(m.y / (s * ay));
	return 1.0f / (PI_F * ax * ay * dn_sqr(c2) * dn_sqr(1.0f + t2 * stretched));
}
__device__ float ggx_lambda(v3 v, float ax, float ay)
{
	if (v.z == 0) return 0;
	const float c2 = v.z * v.z;
	const float s = sqrtf(fmaxf(0.0f, 1 - c2));
	float pr;
	if (ax == ay || s == 0.0f) pr = ax;
	else pr = sqrtf(dn_sqr((v.x * ax) / s) + dn_sqr((v.y * ay) / s));
	const float t2 = dn_sqr(s) / c2;
	const float a2rcp = dn_sqr(pr) * t2;
	return (-1.0f + sqrtf(1.0f + a2rcp)) * 0.5f;
}
__device__ __forceinline__ float ggx_G(v3 wi, v3 wo, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(wo, ax, ay) + ggx_lambda(wi, ax, ay)); }
__device__ __forceinline__ float ggx_G1(v3 v, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(v, ax, ay)); }
__device__ v3 ggx_sample(v3 v, float r0, float r1, float ax, float ay)
{
	const float sgn = v.z < 0.0f ? -1.0f : 1.0f;
	const v3 st = normalize(mk3(sgn * v.x * ax, sgn * v.y * ay, sgn * v.z));
	const v3 t1 = v.z < 0.9999f ? normalize(cross(st, mk3(0, 0, 1))) : mk3(1, 0, 0);
	const v3 t2 = cross(t1, st);
	const float a = 1.0f / (1.0f + st.z);
	const float r = sqrtf(r0);
	const float phi = r1 < a ? (r1 / a * PI_F) : (PI_F + (r1 - a) / (1.0f - a) * PI_F);
	float p1, p2;
	cw_sincos(phi, &p2, &p1);
	p1 *= r;
	p2 *= r * (r1 < a ? 1.0f : st.z);
	const v3 h = p1 * t1 + p2 * t2 + sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2)) * st;
	return normalize(mk3(h.x * ax, h.y * ay, fmaxf(0.0f, h.z)));
}
__device__ float ggx_pdf(v3 v, v3 m, float ax, float ay)
{	/* pdf_visible_normals */
	const float c = v.z;
	if (c == 0.0f) return 0;
	return ggx_G1(v, ax, ay) * fabsf(dot(v, m)) * ggx_D(m, ax, ay) / fabsf(c);
}
__device__ float gtr1_D(v3 m, float ax)
{
	const float alpha = fminf(fmaxf(ax, 0.001f), 0.999f);
	const float a2 = dn_sqr(alpha);
	const float a = (a2 - 1.0f) / (PI_F * dn_logf(a2));
	const float b = (1 / (1 + (a2 - 1) * dn_sqr(m.z)));
	return a * b;
}
__device__ float gtr1_lambda(v3 v, float ax)
{
	if (v.z == 0) return 0;
	const float c2 = dn_sqr(v.z);
	const float s = sqrtf(fmaxf(0.0f, 1.0f - c2));
	if (s == 0) return 0;
	const float cot2 = c2 / dn_sqr(s);
	const float cot = sqrtf(cot2);
	const float alpha2 = dn_sqr(fminf(fmaxf(ax, 0.001f), 0.999f));
	const float a = sqrtf(cot2 + alpha2);
	const float b = sqrtf(cot2 + 1.0f);
	const float c = dn_logf(cot + b);
	const float d = dn_logf(cot + a);
	return (a - b + cot * (c - d)) / (cot * dn_logf(alpha2));
}
__device__ __forceinline__ float gtr1_G(v3 wi, v3 wo, float ax) { return 1.0f / (1.0f + gtr1_lambda(wo, ax) + gtr1_lambda(wi, ax)); }
__device__ v3 gtr1_sample(float r0, float r1, float ax)
{
	const float alpha = fminf(fmaxf(ax, 0.001f), 0.999f);
	const float a2 = dn_sqr(alpha);
	const float c2 = (1.0f - dn_powf(a2, 1.0f - r0)) / (1.0f - a2);
	const float s = sqrtf(fmaxf(0.0f, 1.0f - c2));
	float cp, sp;
	cw_sincos(TWOPI_F * r1, &sp, &cp);
	return mk3(cp * s, sp * s, sqrtf(c2));
}
__device__ __forceinline__ float gtr1_pdf(v3 m, float ax) { return gtr1_D(m, ax) * fabsf(m.z); }
__device__ __forceinline__ void alpha_from_roughness(float roughness, float aniso, float& ax, float& ay)
{
	const float sq = roughness * roughness;
	const float aspect = sqrtf(1.0f + aniso * (aniso < 0 ? 0.9f : -0.9f));
	ax = fmaxf(0.001f, sq / aspect);
	ay = fmaxf(0.001f, sq * aspect);
}

/* ---- frosted.h helpers ---- */
__device__ __forceinline__ float fresnel_dielectric(float eta, float ci, float ct)
{
	if (ci == 0 && ct == 0) return 1;
	const float k0 = eta * ct, k1 = eta * ci;
	return 0.5f * (dn_sqr((ci - k0) / (ci + k0)) + dn_sqr((ct - k1) / (ct + k1)));
}
__device__ __forceinline__ float fresnel_reflectance(float ci, float eta, float& ct)
{
	const float st2 = (1 - dn_sqr(ci)) * dn_sqr(eta);
	if (st2 > 1) { ct = 0; return 1; }
	ct = fminf(sqrtf(fmaxf(1 - st2, 0.0f)), 1.0f);
	return fresnel_dielectric(eta, fabsf(ci), ct);
}
__device__ __forceinline__ v3 refracted_direction(v3 wo, v3 m, float cos_wom, float ct, float rcp_eta)
{
	const v3 wi = cos_wom > 0 ? (rcp_eta * cos_wom - ct) * m - rcp_eta * wo : (rcp_eta * cos_wom + ct) * m - rcp_eta * wo;
	return wi * ((3 - dot(wi, wi)) * 0.5f);
}
__device__ __forceinline__ float choose_reflection_probability(float F)
{
	const float r = F * 1.0f, t = (1 - F) * 1.0f, s = r + t;
	return s != 0 ? r / s : 1;
}
__device__ __forceinline__ v3 half_vector_up(v3 h) { return h.z < 0 ? (h * -1.0f) : h; }
__device__ v3 evaluate_reflection(v3 color, v3 wo, v3 wi, v3 m, float ax, float ay, float F)
{
	const float denom = fabsf(4 * wo.z * wi.z);
	if (denom == 0) return mk3(0, 0, 0);
	const float D = ggx_D(m, ax, ay), G = ggx_G(wi, wo, ax, ay);
	return color * (F * D * G / denom);
}
__device__ v3 evaluate_refraction(float eta, v3 color, v3 wo, v3 wi, v3 m, float ax, float ay, float T)
{
	if (wo.z == 0 || wi.z == 0) return mk3(0, 0, 0);
	const float cih = dot(m, wi), coh = dot(m, wo);
	const float dots = (cih * coh) / (wi.z * wo.z);
	const float sd = coh + eta * cih;
	if (fabsf(sd) < 1.0e-6f) return mk3(0, 0, 0);
	const float D = ggx_D(m, ax, ay), G = ggx_G(wi, wo, ax, ay);
	float mult = fabsf(dots) * T * D * G / dn_sqr(sd);
	mult *= dn_sqr(eta);   /* not adjoint */
	return color * mult;
}
__device__ __forceinline__ float reflection_jacobian(float coh) { return coh == 0 ? 0 : 1 / (4 * fabsf(coh)); }
__device__ __forceinline__ float refraction_jacobian(v3 wo, v3 wi, v3 m, float eta)
{
	const float cih = dot(m, wi), coh = dot(m, wo);
	const float sd = coh + eta * cih;
	if (fabsf(sd) < 1.0e-6f) return 0;
	return fabsf(cih) * dn_sqr(eta / sd);
}

/* ---- disney.h ---- */
__device__ __forceinline__ float schlick_fresnel(float u) { const float m = dn_saturate(1.0f - u), m2 = dn_sqr(m), m4 = dn_sqr(m2); return m4 * m; }
__device__ __forceinline__ v3 mix_one_with_spectra(v3 b, float t) { return mk3((1.0f - t) + t * b.x, (1.0f - t) + t * b.y, (1.0f - t) + t * b.z); }
__device__ __forceinline__ v3 mix_spectra(v3 a, v3 b, float t) { return (1.0f - t) * a + t * b; }
__device__ __forceinline__ v3 mix_spectra_with_one(v3 a, float t) { const v3 r = (1.0f - t) * a; return mk3(r.x + t, r.y + t, r.z + t); }
__device__ __forceinline__ float clearcoat_roughness(const Shading& s) { return dn_mix(0.1f, 0.001f, p_clearcoatgloss(s)); }
__device__ v3 specular_fresnel(const Shading& s, v3 o, v3 h)
{
	v3 value = mix_one_with_spectra(xyz(s.tint), p_spectint(s));
	value = value * (p_specular(s) * 0.08f);
	value = mix_spectra(value, s.color, p_metallic(s));
	const float coh = fabsf(dot(o, h));
	return mix_spectra_with_one(value, schlick_fresnel(coh));
}
__device__ __forceinline__ v3 clearcoat_fresnel(const Shading& s, v3 o, v3 h)
{
	const float coh = fabsf(dot(o, h));
	const float v = dn_mix(0.04f, 1.0f, schlick_fresnel(coh)) * 0.25f * p_clearcoat(s);
	return mk3(v, v, v);
}

template <int MDF>
__device__ void sample_mf(const Shading& s, float r0, float r1, float ax, float ay, v3 wol, v3& wil, float& pdf, v3& value)
{
	if (wol.z == 0) { value = mk3(0, 0, 0); pdf = 0; return; }
	const v3 m = MDF == DISNEY_GGX ? ggx_sample(wol, r0, r1, ax, ay) : gtr1_sample(r0, r1, ax);
	wil = dn_reflect(wol * -1.0f, m);
	if (wil.z == 0) return;
	const float coh = dot(wol, m);
	pdf = (MDF == DISNEY_GGX ? ggx_pdf(wol, m, ax, ay) : gtr1_pdf(m, ax)) / fabsf(4.0f * coh);
	if (pdf < 1.0e-6f) return;
	const float D = MDF == DISNEY_GGX ? ggx_D(m, ax, ay) : gtr1_D(m, ax);
	const float G = MDF == DISNEY_GGX ? ggx_G(wil, wol, ax, ay) : gtr1_G(wil, wol, ax);
	value = MDF == DISNEY_GGX ? specular_fresnel(s, wol, m) : clearcoat_fresnel(s, wol, m);
	value = value * (D * G);
}

/* returns the pdf; `bsdf` keeps its value when the evaluation returns early */
template <int MDF>
__device__ float evaluate_mf(const Shading& s, float ax, float ay, v3 wol, v3 wil, v3 m, v3& bsdf)
{
	if (wol.z == 0 || wil.z == 0) return 0;
	const float coh = dot(wol, m);
	if (coh == 0) return 0;
	const float D = MDF == DISNEY_GGX ? ggx_D(m, ax, ay) : gtr1_D(m, ax);
	const float G = MDF == DISNEY_GGX ? ggx_G(wil, wol, ax, ay) : gtr1_G(wil, wol, ax);
	bsdf = MDF == DISNEY_GGX ? specular_fresnel(s, wol, m) : clearcoat_fresnel(s, wol, m);
	bsdf = bsdf * (D * G / fabsf(4.0f * wol.z * wil.z));
	return (MDF == DISNEY_GGX ? ggx_pdf(wol, m, ax, ay) : gtr1_pdf(m, ax)) / fabsf(4.0f * coh);
}

__device__ float evaluate_diffuse(const Shading& s, v3 iN, v3 wow, v3 wiw, v3 m, v3& value)
{
	const float con = dot(iN, wow), cin = dot(iN, wiw), cih = dot(wiw, m);
	const float fl = schlick_fresnel(cin), fv = schlick_fresnel(con);
	const float ss_ = p_subsurface(s), rough = roughness(s);
	float fd = 0;
	if (ss_ != 1.0f)
	{
		const float fd90 = 0.5f + 2.0f * dn_sqr(cih) * rough;
		fd = dn_mix(1.0f, fd90, fl) * dn_mix(1.0f, fd90, fv);
	}
	if (ss_ > 0)
	{
		const float fss90 = dn_sqr(cih) * rough;
		const float fss = dn_mix(1.0f, fss90, fl) * dn_mix(1.0f, fss90, fv);
		const float ss = 1.25f * (fss * (1.0f / (fabsf(con) + fabsf(cin)) - 0.5f) + 0.5f);
		fd = dn_mix(fd, ss, ss_);
	}
	value = s.color * fd * INVPI_F * (1.0f - p_metallic(s));
	return fabsf(cin) * INVPI_F;
}

__device__ float evaluate_sheen(const Shading& s, v3 wiw, v3 m, v3& value)
{
	const float cih = dot(wiw, m);
	const float fh = schlick_fresnel(cih);
	value = mix_one_with_spectra(xyz(s.tint), p_sheentint(s));
	value = value * (fh * p_sheen(s) * (1.0f - p_metallic(s)));
	return 1.0f / (2 * PI_F);
}

/* SampleBSDF (disney.h:150-268); r0..r2 are SampleBSDF's three random numbers */
__device__ v3 disney_sample(const Shading& s, v3 iN, v3 N, v3 iT, v3 wow, float distance, float r0, float r1, float r2,
	v3& wiw, float& pdf, bool& specular)
{
	pdf = 0;
	const float flip = (dot(wow, N) < 0) ? -1 : 1;
	iN = iN * flip;
	const v3 B = normalize(cross(iN, iT));
	const v3 T = normalize(cross(iN, B));
	const float TR = transmission(s);
	if (r0 < TR)
	{
		specular = true;
		const float r3 = r0 / TR;
		const v3 wol = w2t(wow, iN, T, B);
		const float e = flip < 0 ? (1 / eta(s)) : eta(s);
		if (e == 1) return mk3(0, 0, 0);
		const v3 beer = mk3(cw_expf(-s.transmittance.x * distance * 2.0f), cw_expf(-s.transmittance.y * distance * 2.0f),
			cw_expf(-s.transmittance.z * distance * 2.0f));
		float ax, ay;
		alpha_from_roughness(roughness(s), p_anisotropic(s), ax, ay);
		const v3 m = ggx_sample(wol, r1, r3, ax, ay);
		const float rcp_eta = 1 / e, cos_wom = fminf(fmaxf(dot(wol, m), -1.0f), 1.0f);
		(void)rcp_eta;
		float ct, jacobian;
		const float F = fresnel_reflectance(cos_wom, e, ct);
		v3 wil, ret;
		if (r2 < F)
		{
			wil = dn_reflect(wol * -1.0f, m);
			if (wil.z * wol.z <= 0) return mk3(0, 0, 0);
			ret = evaluate_reflection(s.color, wol, wil, m, ax, ay, F);
			pdf = F, jacobian = reflection_jacobian(cos_wom);
		}
		else
		{	/* refracted_direction takes eta in its rcp_eta slot, as the reference passes it */
			wil = refracted_direction(wol, m, cos_wom, ct, e);
			if (wil.z * wol.z > 0) return mk3(0, 0, 0);
			ret = evaluate_refraction(1 / e, s.color, wol, wil, m, ax, ay, 1 - F);
			pdf = 1 - F, jacobian = refraction_jacobian(wol, wil, m, 1 / e);
		}
		pdf *= jacobian * ggx_pdf(wol, m, ax, ay);
		if (pdf > 1.0e-6f) wiw = t2w(wil, iN, T, B);
		return ret * beer;
	}
	const float r3 = (r0 - TR) / (1 - TR);
	const float met = p_metallic(s);
	float w0 = dn_lerp(s.tint.w, 0, met), w1 = dn_lerp(p_sheen(s), 0, met), w2 = dn_lerp(p_specular(s), 1, met), w3 = p_clearcoat(s) * 0.25f;
	const float inv = 1.0f / (w0 + w1 + w2 + w3);
	w0 *= inv, w1 *= inv, w2 *= inv, w3 *= inv;
	const float cdfx = w0, cdfy = w0 + w1, cdfz = w0 + w1 + w2;
	float probability, cpdf = 0;
	v3 contrib = mk3(0, 0, 0), value = mk3(0, 0, 0);
	if (r3 < cdfy)
	{
		const float rr = r3 / cdfy;
		{	/* DiffuseReflectionCosWeighted (common_functions.h:118-124) */
			const float term1 = TWOPI_F * rr, term2 = sqrtf(1 - r1);
			float sn, cs;
			cw_sincos(term1, &sn, &cs);
			wiw = (cs * term2 * T) + (sn * term2) * B + sqrtf(r1) * iN;
		}
		const v3 m = normalize(wiw + wow);
		if (r3 < cdfx) { cpdf = evaluate_diffuse(s, iN, wow, wiw, m, value); probability = w0 * cpdf, w0 = 0; }
		else { cpdf = evaluate_sheen(s, wiw, m, value); probability = w1 * cpdf, w1 = 0; }
	}
	else
	{
		const v3 wol = w2t(wow, iN, T, B);
		v3 wil = mk3(0, 0, 0);
		if (r3 < cdfz)
		{
			const float rr = (r3 - cdfy) / (cdfz - cdfy);
			float ax, ay;
			alpha_from_roughness(roughness(s), p_anisotropic(s), ax, ay);
			sample_mf<DISNEY_GGX>(s, rr, r1, ax, ay, wol, wil, cpdf, value);
			probability = w2 * cpdf, w2 = 0;
		}
		else
		{
			const float rr = (r3 - cdfz) / (1 - cdfz);
			const float a = clearcoat_roughness(s);
			sample_mf<DISNEY_GTR1>(s, rr, r1, a, a, wol, wil, cpdf, value);
			probability = w3 * cpdf, w3 = 0;
		}
		value = value * (1.0f / fabsf(4.0f * wol.z * wil.z));
		wiw = t2w(wil, iN, T, B);
	}
	if (w0 + w1 > 0)
	{
		const v3 m = normalize(wiw + wow);
		if (w0 > 0) probability += w0 * evaluate_diffuse(s, iN, wow, wiw, m, contrib), value = value + contrib;
		if (w1 > 0) probability += w1 * evaluate_sheen(s, wiw, m, contrib), value = value + contrib;
	}
	if (w2 + w3 > 0)
	{
		const v3 wol = w2t(wow, iN, T, B), wil = w2t(wiw, iN, T, B);
		const v3 m = normalize(wol + wil);
		if (w2 > 0)
		{
			float ax, ay;
			alpha_from_roughness(roughness(s), p_anisotropic(s), ax, ay);
			probability += w2 * evaluate_mf<DISNEY_GGX>(s, ax, ay, wol, wil, m, contrib);
			value = value + contrib;
		}
		if (w3 > 0)
		{
			const float a = clearcoat_roughness(s);
			probability += w3 * evaluate_mf<DISNEY_GTR1>(s, a, a, wol, wil, m, contrib);
			value = value + contrib;
		}
	}
	pdf = probability > 1.0e-6f ? probability : 0;
	return value;
}

/* EvaluateBSDF (disney.h:270-358) */
__device__ v3 disney_evaluate(const Shading& s, v3 iN, v3 iT, v3 wow, v3 wiw, float& pdf)
{
	if (transmission(s) > 0.5f)
	{
		const v3 B = normalize(cross(iN, iT));
		const v3 T = normalize(cross(iN, B));
		const v3 wol = w2t(wow, iN, T, B), wil = w2t(wiw, iN, T, B);
		const float e = wol.z > 0 ? eta(s) : (1.0f / eta(s));
		if (e == 1) { pdf = 0; return mk3(0, 0, 0); }
		float ax, ay, jacobian;
		alpha_from_roughness(roughness(s), p_anisotropic(s), ax, ay);
		v3 ret, m;
		if (wil.z * wol.z >= 0)
		{
			m = half_vector_up(normalize(wil + wol));
			const float cos_wom = dot(wol, m);
			float ct;
			const float F = fresnel_reflectance(cos_wom, 1 / e, ct);
			ret = evaluate_reflection(s.color, wol, wil, m, ax, ay, F);
			pdf = choose_reflection_probability(F), jacobian = reflection_jacobian(cos_wom);
		}
		else
		{
			m = half_vector_up(normalize(wol + e * wil));
			const float cos_wom = dot(wol, m);
			float ct;
			const float F = fresnel_reflectance(cos_wom, 1 / e, ct);
			ret = evaluate_refraction(e, s.color, wol, wil, m, ax, ay, 1 - F);
			pdf = 1 - choose_reflection_probability(F), jacobian = refraction_jacobian(wol, wil, m, e);
		}
		pdf *= jacobian * ggx_pdf(wol, m, ax, ay);
		return ret;
	}
	if (roughness(s) <= 0.001f) { pdf = 0; return mk3(0, 0, 0); }
	const v3 B = normalize(cross(iN, iT));
	const v3 T = normalize(cross(iN, B));
	const float met = p_metallic(s);
	float w0 = dn_lerp(s.tint.w, 0, met), w1 = dn_lerp(p_sheen(s), 0, met), w2 = dn_lerp(p_specular(s), 1, met), w3 = p_clearcoat(s) * 0.25f;
	const float inv = 1.0f / (w0 + w1 + w2 + w3);
	w0 *= inv, w1 *= inv, w2 *= inv, w3 *= inv;
	pdf = 0;
	v3 value = mk3(0, 0, 0);
	if (w0 + w1 > 0)
	{	/* the sheen term replaces the diffuse value, as in the reference */
		const v3 m = normalize(wiw + wow);
		if (w0 > 0) pdf += w0 * evaluate_diffuse(s, iN, wow, wiw, m, value);
		if (w1 > 0) pdf += w1 * evaluate_sheen(s, wiw, m, value);
	}
	if (w2 + w3 > 0)
	{
		const v3 wol = w2t(wow, iN, T, B), wil = w2t(wiw, iN, T, B);
		const v3 m = normalize(wol + wil);
		if (w2 > 0)
		{
			float ax, ay;
			alpha_from_roughness(roughness(s), p_anisotropic(s), ax, ay);
			v3 c = mk3(0, 0, 0);
			const float sp = evaluate_mf<DISNEY_GGX>(s, ax, ay, wol, wil, m, c);
			if (sp > 0) pdf += w2 * sp, value = value + c;
		}
		if (w3 > 0)
		{
			const float a = clearcoat_roughness(s);
			v3 c = mk3(0, 0, 0);
			const float cp = evaluate_mf<DISNEY_GTR1>(s, a, a, wol, wil, m, c);
			if (cp > 0) pdf += w3 * cp, value = value + c;
		}
	}
	return value;
}
