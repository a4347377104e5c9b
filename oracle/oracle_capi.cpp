/* oracle_capi.cpp - ctypes surface of the CPU oracle (liboracle.so).
 * TEST INFRASTRUCTURE: loaded only by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py. */
#include "oracle_render.h"
#include <thread>
#include <vector>
#include <atomic>

using namespace orc;

#define OAPI extern "C" __attribute__((visibility("default")))

OAPI void* orc_create() { return new RenderScene(); }
OAPI void orc_destroy(void* s) { delete (RenderScene*)s; }

/* vert4: 3 float4 per triangle (w ignored); tris may be NULL (traversal only). */
OAPI int orc_add_mesh(void* sp, const float* vert4, int ntri, const CoreTri* tris)
{
	RenderScene& s = *(RenderScene*)sp;
	Mesh m;
	m.v.resize((size_t)ntri * 3);
	for (size_t i = 0; i < (size_t)ntri * 3; i++) m.v[i] = mk3(vert4[i * 4], vert4[i * 4 + 1], vert4[i * 4 + 2]);
	build_bvh(m);
	s.meshes.push_back(std::move(m));
	s.tris.emplace_back();
	if (tris) s.tris.back().assign(tris, tris + ntri);
	return (int)s.meshes.size() - 1;
}

OAPI int orc_add_instance(void* sp, int mesh, const float* T16)
{
	RenderScene& s = *(RenderScene*)sp;
	Instance I;
	I.mesh = mesh;
	for (int i = 0; i < 16; i++) I.T[i] = T16[i];
	s.inst.push_back(I);
	return (int)s.inst.size() - 1;
}

OAPI void orc_finalize(void* sp) { finalize_instances(*(RenderScene*)sp); }

OAPI void orc_set_materials(void* sp, const CoreMaterial* m, int n)
{
	RenderScene& s = *(RenderScene*)sp;
	s.mats.resize(n);
	for (int i = 0; i < n; i++) s.mats[i] = convert_material(m[i], s.tex);
}

/* RenderCore::SetTextures + SyncStorageType (lib/RenderCore_PrimeRef/
 * rendercore.cpp:283-355): texels of each storage kind go to one continuous
 * array in texture order; firstPixel is the texture's start in its array.
 * ARGB128 texels are copied whole (float4 each). */
OAPI void orc_set_textures(void* sp, const lh2hip::CoreTexDesc* t, int n)
{
	RenderScene& s = *(RenderScene*)sp;
	s.tex.assign(n, TexInfo());
	s.argb32.clear(), s.nrm32.clear(), s.argb128.clear();
	for (int i = 0; i < n; i++)
	{
		TexInfo& d = s.tex[i];
		d.width = t[i].width, d.height = t[i].height, d.flags = t[i].flags, d.storage = t[i].storage;
		if (t[i].storage == 1)
		{
			d.firstPixel = (uint)s.argb128.size();
			const f4* src = (const f4*)t[i].fdata;
			s.argb128.insert(s.argb128.end(), src, src + t[i].pixelCount);
		}
		else
		{
			std::vector<uint>& dst = t[i].storage == 2 ? s.nrm32 : s.argb32;
			d.firstPixel = (uint)dst.size();
			dst.insert(dst.end(), t[i].idata, t[i].idata + t[i].pixelCount);
		}
	}
}

/* RenderCore::SetSkyData (rendercore.cpp:466-478): float3 texels stored as float4, w = 0. */
OAPI void orc_set_sky(void* sp, const float* pixels3, int w, int h, const float* worldToSky16)
{
	RenderScene& s = *(RenderScene*)sp;
	s.skyW = w, s.skyH = h;
	const size_t n = (size_t)w * h;
	const int sw = w >> 6, sh = h >> 6;
	s.sky.resize(n + (size_t)sw * sh);
	for (size_t i = 0; i < n; i++) s.sky[i] = { pixels3[3 * i], pixels3[3 * i + 1], pixels3[3 * i + 2], 0.0f };
	/* 64x64 box average (rendercore_optix7/rendercore.cpp:726-738) */
	for (int y = 0; y < sh; y++) for (int x = 0; x < sw; x++)
	{
		float t0 = 0, t1 = 0, t2 = 0, t3 = 0;
		const f4* tile = s.sky.data() + (size_t)x * 64 + (size_t)y * 64 * w;
		for (int v = 0; v < 64; v++) for (int u = 0; u < 64; u++)
		{
			const f4 q = tile[u + (size_t)v * w];
			t0 += q.x, t1 += q.y, t2 += q.z, t3 += q.w;
		}
		const float k = 1.0f / (64 * 64);
		s.sky[n + x + (size_t)y * sw] = { t0 * k, t1 * k, t2 * k, t3 * k };
	}
	for (int i = 0; i < 16; i++) s.worldToSky[i] = worldToSky16 ? worldToSky16[i] : (i % 5 == 0 ? 1.0f : 0.0f);
}

OAPI void orc_sample_sky(void* sp, const float* dirs3, int n, float* out3, int small)
{
	const RenderScene& s = *(const RenderScene*)sp;
	for (int i = 0; i < n; i++)
	{
		const f3 d = mk3(dirs3[3 * i], dirs3[3 * i + 1], dirs3[3 * i + 2]);
		const f3 c = small ? sample_sky_small(s, d) : sample_sky(s, d);
		out3[3 * i] = c.x, out3[3 * i + 1] = c.y, out3[3 * i + 2] = c.z;
	}
}

/* uv (n, 2) -> float4 texels of texture `id` (bilinear, or trilinear at `lambda` when trilinear != 0). */
OAPI void orc_fetch_texels(void* sp, int id, const float* uv, int n, int trilinear, float lambda, float* out4)
{
	const RenderScene& s = *(const RenderScene*)sp;
	const TexInfo& t = s.tex[id];
	for (int i = 0; i < n; i++)
	{
		const f4 r = trilinear ? fetch_texel_trilinear(s, lambda, uv[2 * i], uv[2 * i + 1], (int)t.firstPixel, (int)t.width, (int)t.height)
			: fetch_texel(s, uv[2 * i], uv[2 * i + 1], (int)t.firstPixel, (int)t.width, (int)t.height, t.storage);
		out4[4 * i] = r.x, out4[4 * i + 1] = r.y, out4[4 * i + 2] = r.z, out4[4 * i + 3] = r.w;
	}
}

/* The self-written transcendental functions, elementwise (which: 0 atan2(a, b), 1 acos(a), 2 log2(a), 3 exp(a)). */
OAPI void orc_math(int which, const float* a, const float* b, int n, float* out)
{
	for (int i = 0; i < n; i++)
		out[i] = which == 0 ? lh_atan2f(a[i], b[i]) : which == 1 ? lh_acosf(a[i]) : which == 2 ? lh_log2f(a[i]) : cw_expf(a[i]);
}

OAPI void orc_set_lights(void* sp, const CoreLightTri* t, int nt, const CorePointLight* p, int np,
	const CoreSpotLight* sl, int ns, const CoreDirectionalLight* d, int nd)
{
	RenderScene& s = *(RenderScene*)sp;
	s.triLights.assign(t, t + nt);
	s.pointLights.assign(p, p + np);
	s.spotLights.assign(sl, sl + ns);
	s.dirLights.assign(d, d + nd);
}

template <typename F>
static void parallel_for(int n, int threads, F f)
{
	if (threads < 1) threads = 1;
	std::atomic<int> next(0);
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; t++)
		pool.emplace_back([&, t]() {
			while (true)
			{
				const int b = next.fetch_add(1024);
				if (b >= n) break;
				const int e = b + 1024 < n ? b + 1024 : n;
				for (int i = b; i < e; i++) f(i, t);
			}
		});
	for (auto& th : pool) th.join();
}

/* rays8: {O.xyz, tmin, D.xyz, tmax} per ray. hits4 (closest): {uv bits, inst,
 * prim (-1 miss), t}; for any-hit, hits4[i*4] = 1 if occluded. mode 0 = BVH,
 * 1 = brute force. counters3 (optional) += {interior nodes, triangles, leaves}. */
OAPI void orc_trace(void* sp, const float* rays8, int n, float* hits4, int any, int mode, uint64_t* counters3, int threads)
{
	const RenderScene& s = *(RenderScene*)sp;
	std::vector<Counters> per(threads < 1 ? 1 : threads);
	for (auto& c : per) c.nodes = c.tris = c.leaves = 0;
	parallel_for(n, threads, [&](int i, int t) {
		const float* r = rays8 + (size_t)i * 8;
		const f3 O = mk3(r[0], r[1], r[2]), D = mk3(r[4], r[5], r[6]);
		Counters* c = counters3 ? &per[t] : nullptr;
		float* h = hits4 + (size_t)i * 4;
		if (any)
		{
			const bool occ = mode ? brute_any(s, O, r[3], D, r[7]) : trace_any(s, O, r[3], D, r[7], c);
			h[0] = occ ? 1.0f : 0.0f;
		}
		else
		{
			const Hit hh = mode ? brute_closest(s, O, r[3], D, r[7]) : trace_closest(s, O, r[3], D, r[7], c);
			memcpy(h + 0, &hh.uv, 4), memcpy(h + 1, &hh.inst, 4), memcpy(h + 2, &hh.prim, 4), h[3] = hh.t;
		}
	});
	if (counters3) for (auto& c : per) counters3[0] += c.nodes, counters3[1] += c.tris, counters3[2] += c.leaves;
}

/* Canonical breadth-first build tree: nodes as {lo.xyz, leftFirst, hi.xyz, count}. */
OAPI int orc_bvh_nodes(void* sp, int mesh, float* nodes8, uint* prims, int* depth)
{
	const RenderScene& s = *(RenderScene*)sp;
	const Mesh& m = s.meshes[mesh];
	if (nodes8) memcpy(nodes8, m.nodes.data(), m.nodes.size() * sizeof(BuildNode));
	if (prims) memcpy(prims, m.prims.data(), m.prims.size() * 4);
	if (depth) *depth = m.depth;
	return (int)m.nodes.size();
}

OAPI int orc_tlas_nodes(void* sp, float* nodes8, uint* prims)
{
	const RenderScene& s = *(RenderScene*)sp;
	if (nodes8) memcpy(nodes8, s.tlas.data(), s.tlas.size() * sizeof(BuildNode));
	if (prims) memcpy(prims, s.tlasPrims.data(), s.tlasPrims.size() * 4);
	return (int)s.tlas.size();
}

OAPI void orc_render_frame(void* sp, const ViewPyramid* view, int w, int h, int spp, int maxDepth, float eps,
	uint R0gen, int pass, float* pathAcc, int threads)
{
	RenderParams p;
	p.view = *view, p.w = w, p.h = h, p.spp = spp, p.maxDepth = maxDepth, p.eps = eps;
	render_frame(*(RenderScene*)sp, p, R0gen, pass, pathAcc, threads);
}

/* Blue-noise tables from 8-bit arrays (rendercore_optix7/rendercore.cpp:246-255);
 * null arrays select the core's seeded default tables. */
OAPI void orc_set_blue_noise(void* sp, const uint8_t* seq, const uint8_t* scr, const uint8_t* rnk, uint seed)
{
	RenderScene& s = *(RenderScene*)sp;
	s.blueNoise.assign(65536 * 5 + 64, 0u);
	if (!seq || !scr || !rnk) { blue_noise_default(s.blueNoise.data(), seed); return; }
	for (int i = 0; i < 65536; i++) s.blueNoise[i] = seq[i];
	for (int i = 0; i < 128 * 128 * 8; i++) s.blueNoise[65536 + i] = scr[i], s.blueNoise[3 * 65536 + i] = rnk[i];
}

OAPI void orc_blue_noise4(void* sp, const int* xysd, int n, float* out4)
{
	const RenderScene& s = *(const RenderScene*)sp;
	for (int i = 0; i < n; i++) blue_noise_sample4(s, xysd[4 * i], xysd[4 * i + 1], xysd[4 * i + 2], xysd[4 * i + 3], out4 + 4 * i);
}

OAPI void orc_normals(int pack, float* n3, uint* packed, int n)
{
	for (int i = 0; i < n; i++)
		if (pack) packed[i] = normal_pack(n3 + 3 * i); else normal_unpack(packed[i], n3 + 3 * i);
}

/* Batched Disney / GGX probes (rows of inputs). */
OAPI void orc_disney_eval(const float* color3, const uint* params4, const float* transm3, const float* N3, const float* T3,
	const float* wo3, const float* wi3, int n, float* out4)
{
	for (int i = 0; i < n; i++) disney_probe_eval(color3, params4, transm3, N3, T3, wo3 + 3 * i, wi3 + 3 * i, out4 + 4 * i);
}
OAPI void orc_disney_sample(const float* color3, const uint* params4, const float* transm3, const float* N3, const float* T3,
	const float* wo3, const float* r3, int n, float* out8)
{
	for (int i = 0; i < n; i++) disney_probe_sample(color3, params4, transm3, N3, T3, wo3 + 3 * i, r3 + 3 * i, out8 + 8 * i);
}
OAPI void orc_ggx(int which, const float* v3, const float* m3, const float* r2, int n, float ax, float ay, float* out3)
{
	for (int i = 0; i < n; i++) ggx_probe(which, v3 + 3 * i, m3 + 3 * i, ax, ay, r2[2 * i], r2[2 * i + 1], out3 + 3 * i);
}

OAPI int orc_render_frame7(void* sp, const ViewPyramid* view, int w, int h, int spp, int pass, float eps, float clampValue,
	uint shift, uint camSeed, float* pathAcc, int threads, int bsdf)
{
	RenderScene& s = *(RenderScene*)sp;
	if (s.blueNoise.empty()) return -1;
	Render7Params p;
	p.view = *view, p.w = w, p.h = h, p.spp = spp, p.pass = pass, p.eps = eps, p.clampValue = clampValue, p.shift = shift, p.bsdf = bsdf;
	return render_frame7(s, p, camSeed, pathAcc, threads);
}

OAPI void orc_set_brute(void* sp, int on) { ((RenderScene*)sp)->bruteForce = on != 0; }

OAPI int orc_render_frameF(void* sp, const ViewPyramid* view, int w, int h, int spp, int pass, int blueSlot, float eps, float clampValue,
	uint camSeed, float* pathAcc, uint* features, float* worldPos, float* deltaDepth, int threads, int bsdf, float j0, float j1)
{
	RenderScene& s = *(RenderScene*)sp;
	if (s.blueNoise.empty()) return -1;
	RenderFParams p;
	p.view = *view, p.w = w, p.h = h, p.spp = spp, p.pass = pass, p.blueSlot = blueSlot, p.bsdf = bsdf, p.eps = eps, p.clampValue = clampValue;
	p.j0 = j0, p.j1 = j1;
	return render_frameF(s, p, camSeed, pathAcc, features, worldPos, deltaDepth, threads);
}

OAPI void orc_primary_rays(const ViewPyramid* view, int w, int h, uint R0, float tmin, float* rays8)
{
	const int n = w * h;
	for (int j = 0; j < n; j++) primary_ray(*view, w, h, R0, j, tmin, rays8 + (size_t)j * 8);
}

/* Compacts survivors in input order; returns their count. */
OAPI int orc_bounce(void* sp, const float* raysIn8, const float* hits4, const uint* pathIn, int n, uint R0,
	float tmin, float* raysOut8, uint* pathOut)
{
	const RenderScene& s = *(RenderScene*)sp;
	int k = 0;
	for (int i = 0; i < n; i++)
	{
		Hit h;
		memcpy(&h.uv, hits4 + i * 4, 4), memcpy(&h.inst, hits4 + i * 4 + 1, 4), memcpy(&h.prim, hits4 + i * 4 + 2, 4);
		h.t = hits4[i * 4 + 3];
		if (bounce_ray(s, raysIn8 + (size_t)i * 8, h, pathIn[i], R0, tmin, raysOut8 + (size_t)k * 8)) pathOut[k++] = pathIn[i];
	}
	return k;
}

/* Scalar kernels exposed for golden-vector and closed-form tests. */
OAPI uint orc_wang_hash(uint s) { return wang_hash(s); }
OAPI void orc_xor32_seq(uint seed, int n, uint* out) { for (int i = 0; i < n; i++) out[i] = xor32(seed); }
OAPI void orc_sincos(float x, float* s, float* c) { cw_sincos(x, s, c); }
OAPI uint16_t orc_float_to_half(float f) { return float_to_half(f); }
OAPI float orc_half_to_float(uint16_t h) { return half_to_float(h); }
OAPI void orc_mat4_inverse(const float* m, float* out) { mat4_inverse(m, out); }
OAPI void orc_convert_material(const CoreMaterial* m, OMat* out) { *out = convert_material(*m, std::vector<TexInfo>()); }
OAPI int orc_sizeof_omat() { return (int)sizeof(OMat); }
OAPI int orc_tri_intersect(const float* O3, const float* D3, const float* v9, float* out4)
{
	const RayPre r = ray_setup(mk3(O3[0], O3[1], O3[2]), mk3(D3[0], D3[1], D3[2]));
	float t, U, V, det;
	if (!tri_intersect(r, mk3(v9[0], v9[1], v9[2]), mk3(v9[3], v9[4], v9[5]), mk3(v9[6], v9[7], v9[8]), t, U, V, det)) return 0;
	out4[0] = t, out4[1] = U, out4[2] = V, out4[3] = det;
	return 1;
}
