/* oracle_disney.h - CPU oracle of the Disney principled BSDF of the optix7
 * cores. TEST INFRASTRUCTURE (see oracle_math.h).
 *
 * Restates lib/sharedBSDFs/disney.h:33-360, ggxmdf.h:39-223 and
 * frosted.h:24-103 in the reference's operation order, in fp32, with the
 * oracle's own sincos/exp/log2 (log = log2 * ln 2, pow(a, b) = exp(b log a)).
 * Outputs the reference leaves unset on early returns start at zero; a
 * component evaluation that returns early leaves the running contribution as
 * it was. Included by oracle_render.cpp after its Shading struct. */
#pragma once

#define DISNEY_GGX 0
#define DISNEY_GTR1 1

static inline float dn_sqr(float a) { return a * a; }
static inline float dn_saturate(float x) { return fmaxf(0.0f, fminf(1.0f, x)); }
static inline float dn_lerp(float a, float b, float t) { return a + t * (b - a); }
/* mix of tools_shared.h:98: clamps t to the ends */
static inline float dn_mix(float a, float b, float x) { return x <= 0 ? a : x >= 1 ? b : dn_lerp(a, b, x); }
static inline float dn_logf(float x) { return lh_log2f(x) * 0.693147180559945309f; }
static inline float dn_powf(float a, float b) { return cw_expf(b * dn_logf(a)); }
static inline f3 dn_reflect(f3 i, f3 n) { return i - 2.0f * n * dot(n, i); }
static inline f3 w2t(f3 V, f3 N, f3 T, f3 B) { return mk3(dot(V, T), dot(V, B), dot(V, N)); }
static inline f3 t2w(f3 V, f3 N, f3 T, f3 B) { return V.x * T + V.y * B + V.z * N; }

/* material parameters (tools_shared.h:40-54) */
static inline float p_metallic(const Shading& s) { return char2flt(s.params[0], 0); }
static inline float p_subsurface(const Shading& s) { return char2flt(s.params[0], 8); }
static inline float p_specular(const Shading& s) { return char2flt(s.params[0], 16); }
static inline float p_spectint(const Shading& s) { return char2flt(s.params[1], 0); }
static inline float p_anisotropic(const Shading& s) { return char2flt(s.params[1], 8); }
static inline float p_sheen(const Shading& s) { return char2flt(s.params[1], 16); }
static inline float p_sheentint(const Shading& s) { return char2flt(s.params[1], 24); }
static inline float p_clearcoat(const Shading& s) { return char2flt(s.params[2], 0); }
static inline float p_clearcoatgloss(const Shading& s) { return char2flt(s.params[2], 8); }

/* ---- ggxmdf.h ---- */
static float ggx_D(f3 m, float ax, float ay)
{
	if (m.z == 0) return dn_sqr(ax) * INVPI;
	const float c2 = dn_sqr(m.z);
	const float s = sqrtf(fmaxf(0.0f, 1 - c2));
	const float t2 = (1.0f - c2) / c2;
	float stretched;
	if (ax == ay || s == 0.0f) stretched = 1.0f / dn_sqr(ax);
	else stretched = dn_sqr(m.x / (s * ax)) + dn_sqr(m.y / (s * ay));
	return 1.0f / (PI * ax * ay * dn_sqr(c2) * dn_sqr(1.0f + t2 * stretched));
}
static float ggx_lambda(f3 v, float ax, float ay)
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
static inline float ggx_G(f3 wi, f3 wo, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(wo, ax, ay) + ggx_lambda(wi, ax, ay)); }
static inline float ggx_G1(f3 v, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(v, ax, ay)); }
static f3 ggx_sample(f3 v, float r0, float r1, float ax, float ay)
{
	const float sgn = v.z < 0.0f ? -1.0f : 1.0f;
	const f3 st = normalize(mk3(sgn * v.x * ax, sgn * v.y * ay, sgn * v.z));
	const f3 t1 = v.z < 0.9999f ? normalize(cross(st, mk3(0, 0, 1))) : mk3(1, 0, 0);
	const f3 t2 = cross(t1, st);
	const float a = 1.0f / (1.0f + st.z);
	const float r = sqrtf(r0);
	const float phi = r1 < a ? (r1 / a * PI) : (PI + (r1 - a) / (1.0f - a) * PI);
	float p1, p2;
	cw_sincos(phi, &p2, &p1);
	p1 *= r;
	p2 *= r * (r1 < a ? 1.0f : st.z);
	const f3 h = p1 * t1 + p2 * t2 + sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2)) * st;
	return normalize(mk3(h.x * ax, h.y * ay, fmaxf(0.0f, h.z)));
}
static float ggx_pdf(f3 v, f3 m, float ax, float ay)
{	/* pdf_visible_normals */
	const float c = v.z;
	if (c == 0.0f) return 0;
	return ggx_G1(v, ax, ay) * fabsf(dot(v, m)) * ggx_D(m, ax, ay) / fabsf(c);
}
static float gtr1_D(f3 m, float ax)
{
	const float alpha = fminf(fmaxf(ax, 0.001f), 0.999f);
	const float a2 = dn_sqr(alpha);
	const float a = (a2 - 1.0f) / (PI * dn_logf(a2));
	const float b = (1 / (1 + (a2 - 1) * dn_sqr(m.z)));
	return a * b;
}
static float gtr1_lambda(f3 v, float ax)
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
static inline float gtr1_G(f3 wi, f3 wo, float ax) { return 1.0f / (1.0f + gtr1_lambda(wo, ax) + gtr1_lambda(wi, ax)); }
static f3 gtr1_sample(float r0, float r1, float ax)
{
	const float alpha = fminf(fmaxf(ax, 0.001f), 0.999f);
	const float a2 = dn_sqr(alpha);
	const float c2 = (1.0f - dn_powf(a2, 1.0f - r0)) / (1.0f - a2);
	const float s = sqrtf(fmaxf(0.0f, 1.0f - c2));
	float cp, sp;
	cw_sincos(TWOPI * r1, &sp, &cp);
	return mk3(cp * s, sp * s, sqrtf(c2));
}
static inline float gtr1_pdf(f3 m, float ax) { return gtr1_D(m, ax) * fabsf(m.z); }
static inline void alpha_from_roughness(float roughness, float aniso, float& ax, float& ay)
{
	const float sq = roughness * roughness;
	const float aspect = sqrtf(1.0f + aniso * (aniso < 0 ? 0.9f : -0.9f));
	ax = fmaxf(0.001f, sq / aspect);
	ay = fmaxf(0.001f, sq * aspect);
}

/* ---- frosted.h helpers ---- */
static inline float fresnel_dielectric(float eta, float ci, float ct)
{
	if (ci == 0 && ct == 0) return 1;
	const float k0 = eta * ct, k1 = eta * ci;
	return 0.5f * (dn_sqr((ci - k0) / (ci + k0)) + dn_sqr((ct - k1) / (ct + k1)));
}
static inline float fresnel_reflectance(float ci, float eta, float& ct)
{
	const float st2 = (1 - dn_sqr(ci)) * dn_sqr(eta);
	if (st2 > 1) { ct = 0; return 1; }
	ct = fminf(sqrtf(fmaxf(1 - st2, 0.0f)), 1.0f);
	return fresnel_dielectric(eta, fabsf(ci), ct);
}
static inline f3 refracted_direction(f3 wo, f3 m, float cos_wom, float ct, float rcp_eta)
{
	const f3 wi = cos_wom > 0 ? (rcp_eta * cos_wom - ct) * m - rcp_eta * wo : (rcp_eta * cos_wom + ct) * m - rcp_eta * wo;
	return wi * ((3 - dot(wi, wi)) * 0.5f);
}
static inline float choose_reflection_probability(float F)
{
	const float r = F * 1.0f, t = (1 - F) * 1.0f, s = r + t;
	return s != 0 ? r / s : 1;
}
static inline f3 half_vector_up(f3 h) { return h.z < 0 ? (h * -1.0f) : h; }
static f3 evaluate_reflection(f3 color, f3 wo, f3 wi, f3 m, float ax, float ay, float F)
{
	const float denom = fabsf(4 * wo.z * wi.z);
	if (denom == 0) return mk3(0, 0, 0);
	const float D = ggx_D(m, ax, ay), G = ggx_G(wi, wo, ax, ay);
	return color * (F * D * G / denom);
}
static f3 evaluate_refraction(float eta, f3 color, f3 wo, f3 wi, f3 m, float ax, float ay, float T)
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
static inline float reflection_jacobian(float coh) { return coh == 0 ? 0 : 1 / (4 * fabsf(coh)); }
static inline float refraction_jacobian(f3 wo, f3 wi, f3 m, float eta)
{
	const float cih = dot(m, wi), coh = dot(m, wo);
	const float sd = coh + eta * cih;
	if (fabsf(sd) < 1.0e-6f) return 0;
	return fabsf(cih) * dn_sqr(eta / sd);
}

/* ---- disney.h ---- */
static inline float schlick_fresnel(float u) { const float m = dn_saturate(1.0f - u), m2 = dn_sqr(m), m4 = dn_sqr(m2); return m4 * m; }
static inline f3 mix_one_with_spectra(f3 b, float t) { return mk3((1.0f - t) + t * b.x, (1.0f - t) + t * b.y, (1.0f - t) + t * b.z); }
static inline f3 mix_spectra(f3 a, f3 b, float t) { return (1.0f - t) * a + t * b; }
static inline f3 mix_spectra_with_one(f3 a, float t) { const f3 r = (1.0f - t) * a; return mk3(r.x + t, r.y + t, r.z + t); }
static inline float clearcoat_roughness(const Shading& s) { return dn_mix(0.1f, 0.001f, p_clearcoatgloss(s)); }
static f3 specular_fresnel(const Shading& s, f3 o, f3 h)
{
	f3 value = mix_one_with_spectra(mk3(s.tint.x, s.tint.y, s.tint.z), p_spectint(s));
	value = value * (p_specular(s) * 0.08f);
	value = mix_spectra(value, s.color, p_metallic(s));
	const float coh = fabsf(dot(o, h));
	return mix_spectra_with_one(value, schlick_fresnel(coh));
}
static inline f3 clearcoat_fresnel(const Shading& s, f3 o, f3 h)
{
	const float coh = fabsf(dot(o, h));
	const float v = dn_mix(0.04f, 1.0f, schlick_fresnel(coh)) * 0.25f * p_clearcoat(s);
	return mk3(v, v, v);
}

template <int MDF>
static void sample_mf(const Shading& s, float r0, float r1, float ax, float ay, f3 wol, f3& wil, float& pdf, f3& value)
{
	if (wol.z == 0) { value = mk3(0, 0, 0); pdf = 0; return; }
	const f3 m = MDF == DISNEY_GGX ? ggx_sample(wol, r0, r1, ax, ay) : gtr1_sample(r0, r1, ax);
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
static float evaluate_mf(const Shading& s, float ax, float ay, f3 wol, f3 wil, f3 m, f3& bsdf)
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

static float evaluate_diffuse(const Shading& s, f3 iN, f3 wow, f3 wiw, f3 m, f3& value)
{
	const float con = dot(iN, wow), cin = dot(iN, wiw), cih = dot(wiw, m);
	const float fl = schlick_fresnel(cin), fv = schlick_fresnel(con);
	const float ss_ = p_subsurface(s), rough = s.roughness();
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
	value = s.color * fd * INVPI * (1.0f - p_metallic(s));
	return fabsf(cin) * INVPI;
}

static float evaluate_sheen(const Shading& s, f3 wiw, f3 m, f3& value)
{
	const float cih = dot(wiw, m);
	const float fh = schlick_fresnel(cih);
	value = mix_one_with_spectra(mk3(s.tint.x, s.tint.y, s.tint.z), p_sheentint(s));
	value = value * (fh * p_sheen(s) * (1.0f - p_metallic(s)));
	return 1.0f / (2 * PI);
}

/* SampleBSDF (disney.h:150-268); r0..r2 are SampleBSDF's three random numbers */
static f3 disney_sample(const Shading& s, f3 iN, f3 N, f3 iT, f3 wow, float distance, float r0, float r1, float r2,
	f3& wiw, float& pdf, bool& specular)
{
	pdf = 0;
	const float flip = (dot(wow, N) < 0) ? -1 : 1;
	iN = iN * flip;
	const f3 B = normalize(cross(iN, iT));
	const f3 T = normalize(cross(iN, B));
	const float TR = s.transmission();
	if (r0 < TR)
	{
		specular = true;
		const float r3 = r0 / TR;
		const f3 wol = w2t(wow, iN, T, B);
		const float e = flip < 0 ? (1 / s.eta()) : s.eta();
		if (e == 1) return mk3(0, 0, 0);
		const f3 beer = mk3(cw_expf(-s.transmittance.x * distance * 2.0f), cw_expf(-s.transmittance.y * distance * 2.0f),
			cw_expf(-s.transmittance.z * distance * 2.0f));
		float ax, ay;
		alpha_from_roughness(s.roughness(), p_anisotropic(s), ax, ay);
		const f3 m = ggx_sample(wol, r1, r3, ax, ay);
		const float rcp_eta = 1 / e, cos_wom = fminf(fmaxf(dot(wol, m), -1.0f), 1.0f);
		(void)rcp_eta;
		float ct, jacobian;
		const float F = fresnel_reflectance(cos_wom, e, ct);
		f3 wil, ret;
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
	f3 contrib = mk3(0, 0, 0), value = mk3(0, 0, 0);
	if (r3 < cdfy)
	{
		const float rr = r3 / cdfy;
		{	/* DiffuseReflectionCosWeighted (common_functions.h:118-124) */
			const float term1 = TWOPI * rr, term2 = sqrtf(1 - r1);
			float sn, cs;
			cw_sincos(term1, &sn, &cs);
			wiw = (cs * term2 * T) + (sn * term2) * B + sqrtf(r1) * iN;
		}
		const f3 m = normalize(wiw + wow);
		if (r3 < cdfx) { cpdf = evaluate_diffuse(s, iN, wow, wiw, m, value); probability = w0 * cpdf, w0 = 0; }
		else { cpdf = evaluate_sheen(s, wiw, m, value); probability = w1 * cpdf, w1 = 0; }
	}
	else
	{
		const f3 wol = w2t(wow, iN, T, B);
		f3 wil = mk3(0, 0, 0);
		if (r3 < cdfz)
		{
			const float rr = (r3 - cdfy) / (cdfz - cdfy);
			float ax, ay;
			alpha_from_roughness(s.roughness(), p_anisotropic(s), ax, ay);
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
		const f3 m = normalize(wiw + wow);
		if (w0 > 0) probability += w0 * evaluate_diffuse(s, iN, wow, wiw, m, contrib), value = value + contrib;
		if (w1 > 0) probability += w1 * evaluate_sheen(s, wiw, m, contrib), value = value + contrib;
	}
	if (w2 + w3 > 0)
	{
		const f3 wol = w2t(wow, iN, T, B), wil = w2t(wiw, iN, T, B);
		const f3 m = normalize(wol + wil);
		if (w2 > 0)
		{
			float ax, ay;
			alpha_from_roughness(s.roughness(), p_anisotropic(s), ax, ay);
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
static f3 disney_evaluate(const Shading& s, f3 iN, f3 iT, f3 wow, f3 wiw, float& pdf)
{
	if (s.transmission() > 0.5f)
	{
		const f3 B = normalize(cross(iN, iT));
		const f3 T = normalize(cross(iN, B));
		const f3 wol = w2t(wow, iN, T, B), wil = w2t(wiw, iN, T, B);
		const float e = wol.z > 0 ? s.eta() : (1.0f / s.eta());
		if (e == 1) { pdf = 0; return mk3(0, 0, 0); }
		float ax, ay, jacobian;
		alpha_from_roughness(s.roughness(), p_anisotropic(s), ax, ay);
		f3 ret, m;
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
	if (s.roughness() <= 0.001f) { pdf = 0; return mk3(0, 0, 0); }
	const f3 B = normalize(cross(iN, iT));
	const f3 T = normalize(cross(iN, B));
	const float met = p_metallic(s);
	float w0 = dn_lerp(s.tint.w, 0, met), w1 = dn_lerp(p_sheen(s), 0, met), w2 = dn_lerp(p_specular(s), 1, met), w3 = p_clearcoat(s) * 0.25f;
	const float inv = 1.0f / (w0 + w1 + w2 + w3);
	w0 *= inv, w1 *= inv, w2 *= inv, w3 *= inv;
	pdf = 0;
	f3 value = mk3(0, 0, 0);
	if (w0 + w1 > 0)
	{	/* the sheen term replaces the diffuse value, as in the reference */
		const f3 m = normalize(wiw + wow);
		if (w0 > 0) pdf += w0 * evaluate_diffuse(s, iN, wow, wiw, m, value);
		if (w1 > 0) pdf += w1 * evaluate_sheen(s, wiw, m, value);
	}
	if (w2 + w3 > 0)
	{
		const f3 wol = w2t(wow, iN, T, B), wil = w2t(wiw, iN, T, B);
		const f3 m = normalize(wol + wil);
		if (w2 > 0)
		{
			float ax, ay;
			alpha_from_roughness(s.roughness(), p_anisotropic(s), ax, ay);
			f3 c = mk3(0, 0, 0);
			const float sp = evaluate_mf<DISNEY_GGX>(s, ax, ay, wol, wil, m, c);
			if (sp > 0) pdf += w2 * sp, value = value + c;
		}
		if (w3 > 0)
		{
			const float a = clearcoat_roughness(s);
			f3 c = mk3(0, 0, 0);
			const float cp = evaluate_mf<DISNEY_GTR1>(s, a, a, wol, wil, m, c);
			if (cp > 0) pdf += w3 * cp, value = value + c;
		}
	}
	return value;
}
