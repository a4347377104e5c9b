"""The oracle's optix7 estimator (rendercore_optix7 shadeKernel with the
shared Lambert BSDF, setting bsdf 0) pinned by what the mathematics fixes: normal packing
inverts, the blue-noise sampler reduces to its sequence table and ranks by
permutation, the small skydome is a 64x64 box average, a white-furnace-style
scene (uniform sky over a plane) returns albedo x sky radiance for every
path, and next-event estimation with MIS under a square emitter converges
to the closed-form form factor."""
import numpy as np
import pytest

from lighthouse2_amd import formats as F
from lighthouse2_amd import scenes
import oracle as orc

f32 = np.float32


def test_normal_pack_round_trip():
    rng = np.random.default_rng(0)
    n = rng.normal(size=(4000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n = n[n[:, 2] > -0.9].astype(f32)
    back = orc.unpack_normals(orc.pack_normals(n))
    # 16-bit spheremap coordinates: error ~1e-4, growing towards z = -1
    assert np.abs(back - n).max() < 5e-4
    assert np.abs(back - n)[n[:, 2] > 0].max() < 2e-4
    # the two 16-bit halves: +z maps to the centre of the octahedral square
    p = int(orc.pack_normals(np.array([[0, 0, 1]], f32))[0])
    assert abs((p & 0xffff) - 32767) <= 1 and abs((p >> 16) - 32767) <= 1


def _tables(seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 256, 65536, dtype=np.uint8), rng.integers(0, 256, 131072, dtype=np.uint8),
            rng.integers(0, 256, 131072, dtype=np.uint8))


def test_blue_noise_reduces_to_sequence():
    seq, scr, rnk = _tables(1)
    o = orc.OracleScene(scenes.sky_scene(4))
    o.set_blue_noise((seq, np.zeros_like(scr), np.zeros_like(rnk)))
    rng = np.random.default_rng(2)
    q = np.stack([rng.integers(0, 128, 500), rng.integers(0, 128, 500), rng.integers(0, 256, 500),
                  rng.integers(0, 3, 500) * 4], 1)
    got = o.blue_noise4(q)
    for (x, y, s, d), g in zip(q, got):
        want = (0.5 + seq[d + np.arange(4) + s * 256].astype(np.float64)) / 256.0
        assert np.array_equal(g, want.astype(f32))


def test_blue_noise_ranking_permutes_samples():
    """Over the 256 sample indices of one texel and dimension, ranking only
    permutes which sequence entry is used and scrambling XORs all of them
    with one byte: the multiset of values is fixed by the texel's scramble."""
    seq, scr, rnk = _tables(3)
    o = orc.OracleScene(scenes.sky_scene(4))
    o.set_blue_noise((seq, scr, rnk))
    for x, y, d in [(0, 0, 0), (17, 90, 4), (127, 127, 8), (64, 3, 4)]:
        q = np.array([[x, y, s, d] for s in range(256)])
        got = o.blue_noise4(q)
        for k in range(4):
            sc = int(scr[(d & 7) + k + (x + y * 128) * 8])
            want = np.sort((0.5 + (seq[d + k + np.arange(256) * 256].astype(np.int64) ^ sc)) / 256.0).astype(f32)
            assert np.array_equal(np.sort(got[:, k]), want)


def test_small_sky_is_box_average():
    rng = np.random.default_rng(4)
    W, H = 256, 128
    px = rng.uniform(0, 2, (H, W, 3)).astype(f32)
    s = scenes.sky_scene(4)
    s.sky = (px, F.IDENTITY.copy())
    o = orc.OracleScene(s)
    d = rng.normal(size=(3000, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f32)
    got = o.sample_sky(d, small=True)
    # direction -> texel of the 4 x 2 dome; sampled at -D (tools_shared.h:184-211, pathtracer.h:87)
    t = -d.astype(np.float64)
    phi = np.arctan2(t[:, 1], t[:, 0])
    phi = np.where(phi < 0, phi + 2 * np.pi, phi)
    theta = np.arccos(np.clip(t[:, 2], -1, 1))
    u = np.floor(np.maximum(4 * phi / (2 * np.pi) - 0.5, 0)).astype(int)
    v = np.floor(np.maximum(2 * theta / np.pi - 0.5, 0)).astype(int)
    blocks = px.reshape(2, 64, 4, 64, 3).astype(np.float64).mean((1, 3))
    want = blocks[v, u]
    # skip directions within float rounding of a block edge
    fu, fv = 4 * phi / (2 * np.pi) - 0.5, 2 * theta / np.pi - 0.5
    ok = (np.abs(fu - np.round(fu)) > 1e-4) & (np.abs(fv - np.round(fv)) > 1e-4)
    assert ok.sum() > 2500
    assert np.allclose(got[ok], want[ok], rtol=2e-6, atol=1e-6)


def _floor_scene(albedo, roughness, sky_value=None):
    s = scenes.Scene()
    m = s.add_material(F.lambert_material(albedo, roughness=roughness))
    v, t = scenes.add_quad((0, 1, 0), (0, 0, 0), 400.0, 400.0, m)
    s.add_instance(s.add_mesh(v, t))
    if sky_value is not None:
        s.sky = (np.full((128, 256, 3), sky_value, f32), F.IDENTITY.copy())
    return s


@pytest.mark.parametrize("roughness", [1.0, 0.4, 0.0])
def test_uniform_sky_over_plane(roughness):
    """Every path: camera -> plane (albedo a) -> sky of radiance L after one
    bounce. bsdf * cos / pdf = a for the diffuse and the specular lobe, so each
    sample returns a * L up to fp32 rounding (no lights: no NEE)."""
    a = np.array([0.25, 0.5, 0.75], f32)
    L = 0.75
    s = _floor_scene(a, roughness, L)
    view = scenes.camera_view((0.0, 1.0, -2.0), (0.0, 0.0, 0.0), 40.0, 32, 32)
    r32, r64, n, _, _ = orc.render7(s, view, 32, 32, spp=2, frames=2, threads=8, bsdf=0)
    img = r64[..., :3] / n
    assert np.allclose(img, a * L, rtol=2e-6, atol=0)
    assert np.all(r64[..., 3] == 0)


def _form_factor_square(half, d):
    """Differential area to a parallel, centred square of side 2*half at
    distance d: four corner rectangles (Howell's catalogue, B-3)."""
    A = B = half / d
    sa, sb = np.sqrt(1 + A * A), np.sqrt(1 + B * B)
    corner = (A / sa * np.arctan(B / sa) + B / sb * np.arctan(A / sb)) / (2 * np.pi)
    return 4 * corner


def _form_factor_polygon(P, verts, n):
    """Lambert's contour formula: differential area at P (normal n) to a polygon."""
    tot = 0.0
    for i in range(len(verts)):
        a, b = verts[i] - P, verts[(i + 1) % len(verts)] - P
        a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
        c = np.cross(a, b)
        tot += np.arccos(np.clip(a @ b, -1, 1)) * (c / np.linalg.norm(c)) @ n
    return abs(tot) / (2 * np.pi)


def test_form_factor_formulas_agree():
    sq = np.array([[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]], float) * [0.7, 1.3, 0.7]
    assert abs(_form_factor_polygon(np.zeros(3), sq, np.array([0, 1.0, 0])) - _form_factor_square(0.7, 1.3)) < 1e-12


def test_triangle_emitter_form_factor():
    """Floor of albedo a under one downward triangle emitter of radiance Le:
    NEE plus MIS-weighted implicit hits converge to a * Le * F (F from
    Lambert's formula), 16384 samples at the point below the light.
    (With several emitting triangles the estimator is biased by about 1% for
    a two-triangle quad: LightPickProb, lights_shared.h:174-190, evaluates
    every light's potential at the hit point instead of a point on that light,
    so the implicit and explicit pick probabilities disagree.)"""
    a, Le = 0.5, 2.0
    s = _floor_scene((a, a, a), 1.0)
    lm = s.add_material(F.lambert_material((Le, Le, Le)))
    tri = np.array([[-1.5, 1.0, -1.0], [1.5, 1.0, -1.0], [0.0, 1.0, 1.5]], f32)
    v, t = scenes.add_triangles(tri[:1], tri[1:2], tri[2:], lm)
    assert t["Ny"][0] == -1.0
    li = s.add_instance(s.add_mesh(v, t))
    scenes.add_emitters(s, t, (Le, Le, Le), li)
    # narrow view from below the emitter straight down at the origin
    view = scenes.camera_view((0.0, 0.5, 0.0001), (0.0, 0.0, 0.0), 1.0, 32, 32)
    r32, r64, n, _, _ = orc.render7(s, view, 32, 32, spp=16, frames=1, threads=8, bsdf=0)
    est = r64[..., 0] / n
    want = a * Le * _form_factor_polygon(np.zeros(3), tri.astype(float), np.array([0, 1.0, 0]))
    err = abs(est.mean() - want) / want
    print(f"mean {est.mean():.5f} closed form {want:.5f} rel {err:.2e}")
    assert err < 0.005
    assert np.all(r64[..., 3] > 0)          # every pixel had unoccluded connections (w counts them)


@pytest.mark.parametrize("taa", [False, True])
def test_filter_gbuffer_closed_form(taa):
    """Filter estimator G-buffer on the Lambert Cornell box: eye rays through
    pixel corners hit the walls at t = plane distance / D.n; the world
    position is O + t D; the depth derivatives are those of the neighbouring
    pixel corners' rays against the same plane; albedo is the wall colour.
    With TAA the first frame's view is shifted by (0.3, 0.2) pixels
    (RenderCore_Optix7Filter/rendercore.cpp:755-757,781-786)."""
    s = scenes.cornell_box(seed=5, sphere_tris=200)
    w, h = 48, 40
    view = scenes.cornell_view(w, h)
    r = orc.RenderF(s, w, h, spp=1, bsdf=0, taa=taa)
    r.frame(view, orc.RESTART)
    jx, jy = (0.3, 0.2) if taa else (0.0, 0.0)
    v = view[0]
    pos, p1 = v["pos"].astype(np.float64), v["p1"].astype(np.float64)
    right, up = (v["p2"] - v["p1"]).astype(np.float64), (v["p3"] - v["p1"]).astype(np.float64)
    back = s.materials[2]["color"]["value"].astype(np.float64)
    checked = 0
    for y in range(h):
        for x in range(w):
            D = p1 + ((x + jx) / w) * right + ((y + jy) / h) * up - pos
            D /= np.linalg.norm(D)
            t = (1.0 - pos[2]) / D[2]
            P = pos + t * D
            if not (0.02 < P[0] < 0.98 and 0.02 < P[1] < 0.98) or np.hypot(P[0] - 0.5, P[1] - 0.5) < 0.42:
                continue                                  # back wall only, away from edges and the sphere (its shadow from the eye has radius 0.39 there)
            f = r.features[y, x]
            assert abs(f[2:3].view(np.float32)[0] - t) < 1e-5 * t
            assert np.allclose(r.world_pos[y, x, :3], P, atol=2e-6)
            alb = np.array([f[0] >> 22, (f[0] >> 11) & 2047, f[0] & 2047], np.float64) / [1023, 2047, 2047]
            assert np.allclose(alb, back, atol=1.5e-3)
            def depth(px, py):
                d = p1 + px / w * right + py / h * up - pos
                d /= np.linalg.norm(d)
                return (1.0 - pos[2]) / d[2]
            dd = r.delta_depth[y, x]
            assert abs(dd[2] - (depth(x + 1.5 + jx, y + 0.5 + jy) - t)) < 2e-5
            assert abs(dd[3] - (depth(x + 0.5 + jx, y + 1.5 + jy) - t)) < 2e-5
            checked += 1
    assert checked > 150
    # direct and indirect halves: the indirect sum holds radiance of paths after a diffuse bounce
    assert r.acc64[0, :, :3].sum() > 0 and r.acc64[1, :, :3].sum() > 0
