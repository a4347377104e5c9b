"""Pins the CPU oracle (oracle/) against golden vectors, closed forms, brute
force and analytic integrals, so that the GPU parity tests compare against a
checked reference. CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as orc
from lighthouse2_amd import formats as F
from lighthouse2_amd import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rng.json")


def test_rng_golden_vectors():
    g = json.load(open(GOLDEN))
    for x, y in zip(g["wang_hash"]["in"], g["wang_hash"]["out"]):
        assert orc.wang_hash(x) == y
    for seed, seq in g["xor32"].items():
        assert list(orc.xor32_seq(int(seed), 64)) == seq


def test_sincos_accuracy():
    x = np.concatenate([np.linspace(0, 2 * np.pi, 20001), np.linspace(-100, 100, 5001)]).astype(np.float32)
    err = 0.0
    for v in x:
        s, c = orc.sincos(float(v))
        err = max(err, abs(s - math.sin(float(v))), abs(c - math.cos(float(v))))
    assert err < 2.5e-7


def test_half_conversion_is_round_to_nearest_even():
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.uniform(-70000, 70000, 2000), rng.uniform(-1, 1, 2000), rng.uniform(-1e-4, 1e-4, 2000),
                           [0.0, -0.0, 1.0, 15.0, 65504.0, 65520.0, 5.96e-8, 2.98e-8, 1e-9, np.inf, -np.inf]]).astype(np.float32)
    for v in vals:
        h = orc.lib().orc_float_to_half(float(v))
        assert h == int(np.float32(v).astype(np.float16).view(np.uint16)), v
        assert orc.lib().orc_half_to_float(h) == np.float32(np.float16(v))


def test_mat4_inverse():
    import ctypes
    rng = np.random.default_rng(2)
    out = np.zeros(16, np.float32)
    for _ in range(50):
        A = np.eye(4, dtype=np.float32)
        A[:3, :3] = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        A[:3, 3] = rng.normal(size=3)
        a = np.ascontiguousarray(A.reshape(16))
        orc.lib().orc_mat4_inverse(a.ctypes.data, out.ctypes.data)
        assert np.allclose(out.reshape(4, 4), np.linalg.inv(A.astype(np.float64)), atol=1e-5)
    eye = np.eye(4, dtype=np.float32).reshape(16)
    orc.lib().orc_mat4_inverse(eye.ctypes.data, out.ctypes.data)
    assert np.array_equal(out, eye)
    sing = np.zeros(16, np.float32)
    orc.lib().orc_mat4_inverse(sing.ctypes.data, out.ctypes.data)
    assert np.array_equal(out, eye)  # singular -> identity, as mat4::Inverted


def _isect(O, D, v):
    out = np.zeros(4, np.float32)
    o, d, vv = (np.ascontiguousarray(x, np.float32) for x in (O, D, v))
    hit = orc.lib().orc_tri_intersect(o.ctypes.data, d.ctypes.data, vv.ctypes.data, out.ctypes.data)
    return bool(hit), out


def test_watertight_closed_form_plane():
    rng = np.random.default_rng(7)
    tri = np.array([[-1, -1, 2], [3, -1, 2], [-1, 3, 2]], np.float32).reshape(9)
    for _ in range(500):
        O = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(0, 1.5)], np.float32)
        D = np.array([rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2), 1.0], np.float32)
        hit, out = _isect(O, D, tri)
        assert hit
        t = out[0]
        t_exact = (2.0 - float(O[2])) / float(D[2])
        assert abs(t - t_exact) <= 4 * np.spacing(np.float32(abs(t_exact)) + 1e-30)
        u, v, det = out[1] / out[3], out[2] / out[3], out[3]
        P = O.astype(np.float64) + t_exact * D.astype(np.float64)
        V = tri.reshape(3, 3).astype(np.float64)
        w = 1 - u - v
        assert np.allclose(u * V[0] + v * V[1] + w * V[2], P, atol=1e-5)  # u, v weight vertices 0 and 1


def test_watertight_shared_edge_has_no_crack():
    # quad in z = 0 split along its diagonal x = y: rays through points of the diagonal hit at least one triangle
    a = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], np.float32).reshape(9)
    b = np.array([[0, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32).reshape(9)
    rng = np.random.default_rng(3)
    for k in range(1, 512):
        x = np.float32(k / 512.0)
        O = np.array([x, x, -1], np.float32)
        D = np.array([rng.uniform(-1e-3, 1e-3), rng.uniform(-1e-3, 1e-3), 1], np.float32)
        D[0] = D[1]  # keep the ray in the diagonal plane
        assert _isect(O, D, a)[0] or _isect(O, D, b)[0]


def test_parallel_and_degenerate_miss():
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32).reshape(9)
    assert not _isect([0.2, 0.2, 1], [1, 0, 0], tri)[0]          # parallel, off-plane
    line = np.array([[0, 0, 0], [1, 1, 0], [2, 2, 0]], np.float32).reshape(9)
    assert not _isect([0.5, 0.5, -1], [0, 0, 1], line)[0]       # collinear vertices: det = 0


def _random_rays(rng, n, lo=-0.2, hi=1.2):
    O = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    D = rng.normal(size=(n, 3)).astype(np.float32)
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7], r[:, 7] = O, 0.0, D, 1e34
    return r


@pytest.mark.parametrize("seed,n", [(1, 2000), (2, 5000), (3, 700)])
def test_bvh_traversal_equals_brute_force(seed, n):
    s = scenes.random_triangle_scene(n, seed=seed)
    o = orc.OracleScene(s)
    rng = np.random.default_rng(seed)
    rays = np.concatenate([_random_rays(rng, 3000), orc.primary_rays(scenes.microbench_view(48, 48), 48, 48, 77, 1e-5)])
    hb, hr = o.trace(rays), o.trace(rays, brute=True)
    assert np.array_equal(hb.view(np.uint32), hr.view(np.uint32))
    assert (hb.view(np.int32)[:, 2] >= 0).sum() > 20
    ab, ar = o.trace(rays, any_hit=True), o.trace(rays, any_hit=True, brute=True)
    assert np.array_equal(ab, ar)
    assert np.array_equal(ab, hb.view(np.int32)[:, 2] >= 0)  # any-hit occluded == closest hit exists in (tmin, tmax)


def test_tie_break_duplicate_triangles():
    v = scenes.random_triangles(50, seed=4)
    dup = np.concatenate([v[30:33], v, v[30:33]])  # triangle 10 of v appears as prims 0, 11 and 51
    s = scenes.Scene()
    s.add_instance(s.add_mesh(dup, None))
    o = orc.OracleScene(s)
    c = (dup[0, :3] + dup[1, :3] + dup[2, :3]) / 3
    r = np.array([[c[0], c[1], -1, 0, 0, 0, 1, 1e34]], np.float32)
    h = o.trace(r).view(np.int32)
    assert h[0, 2] == 0 and np.array_equal(h, o.trace(r, brute=True).view(np.int32))


def test_tmin_tmax_are_exclusive():
    s = scenes.Scene()
    v = np.array([[-1, -1, 2, 1], [3, -1, 2, 1], [-1, 3, 2, 1]], np.float32)
    s.add_instance(s.add_mesh(v, None))
    o = orc.OracleScene(s)
    r = np.array([[0, 0, 0, 0, 0, 0, 1, 1e34], [0, 0, 0, 0, 0, 0, 1, 2.0], [0, 0, 0, 2.0, 0, 0, 1, 1e34],
                  [0, 0, 0, 0, 0, 0, 1, np.nextafter(np.float32(2), np.float32(3))]], np.float32)
    h = o.trace(r).view(np.int32)
    assert list(h[:, 2]) == [0, -1, -1, 0]


def test_bvh_structure_invariants():
    s = scenes.random_triangle_scene(3000, seed=9)
    o = orc.OracleScene(s)
    nodes, prims, depth = o.bvh_nodes(0)
    assert sorted(prims.tolist()) == list(range(3000))
    leaves = nodes[nodes["count"] > 0]
    assert leaves["count"].max() <= 4 and len(nodes) == 2 * len(leaves) - 1
    v = s.meshes[0][0][:, :3].reshape(-1, 3, 3)
    for nd in leaves:
        tv = v[prims[nd["leftFirst"]:nd["leftFirst"] + nd["count"]]].reshape(-1, 3)
        assert (tv >= nd["lo"]).all() and (tv <= nd["hi"]).all()
    for nd in nodes[nodes["count"] == 0]:
        for c in (nd["leftFirst"], nd["leftFirst"] + 1):
            assert (nodes[c]["lo"] >= nd["lo"]).all() and (nodes[c]["hi"] <= nd["hi"]).all()


def _affine(rng):
    from scipy.spatial.transform import Rotation
    M = np.eye(4, dtype=np.float32)
    M[:3, :3] = Rotation.from_quat(rng.normal(size=4)).as_matrix() * rng.uniform(0.5, 2.0)
    M[:3, 3] = rng.uniform(-2, 2, 3)
    return M.reshape(16)


def instanced_scene(seed=11, meshes=3, instances=12, tris=400):
    rng = np.random.default_rng(seed)
    s = scenes.Scene()
    for m in range(meshes):
        s.add_mesh(scenes.random_triangles(tris, seed=seed + m, sigma=0.05), None)
    for i in range(instances):
        s.add_instance(int(rng.integers(meshes)), None if i == 0 else _affine(rng))
    return s


def test_two_level_traversal_equals_brute_force():
    s = instanced_scene()
    o = orc.OracleScene(s)
    rng = np.random.default_rng(1)
    rays = _random_rays(rng, 4000, -3, 3)
    hb, hr = o.trace(rays), o.trace(rays, brute=True)
    assert np.array_equal(hb.view(np.uint32), hr.view(np.uint32))
    assert len(set(hb.view(np.int32)[hb.view(np.int32)[:, 2] >= 0, 1])) > 5
    assert np.array_equal(o.trace(rays, any_hit=True), o.trace(rays, any_hit=True, brute=True))


def test_cosine_bounce_distribution():
    # diffuse bounces off the plane z = 0: E[cos] = 2/3, E[cos^2] = 1/2 for a cosine-weighted lobe
    s = scenes.Scene()
    s.add_instance(s.add_mesh(np.array([[-10, -10, 0, 1], [10, -10, 0, 1], [-10, 10, 0, 1]], np.float32), None))
    o = orc.OracleScene(s)
    n = 40000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = [0, 0, 1]
    rays[:, 6], rays[:, 7] = -1, 1e34
    hits = o.trace(rays)
    out, _ = o.bounce(rays, hits, np.arange(n, dtype=np.uint32), 91771, 1e-5)
    cz = out[:, 6].astype(np.float64)
    assert len(out) == n and (cz > 0).all()
    assert abs(cz.mean() - 2 / 3) < 0.01 and abs((cz ** 2).mean() - 0.5) < 0.01
    assert np.allclose(np.linalg.norm(out[:, 4:7], axis=1), 1, atol=1e-6)


def test_direct_lighting_matches_integral():
    """Lambert floor (albedo 0.5) under a 0.2 x 0.2 light of radiance 15 at
    height 1: the pixel seeing the point below the light equals
    rho/pi * integral(Le cos cos' / r^2 dA) (max depth 2: direct light only)."""
    s = scenes.Scene()
    m = s.add_material(F.lambert_material((0.5, 0.5, 0.5)))
    v, t = scenes.add_quad((0, 1, 0), (0, 0, 0), 8.0, 8.0, m)
    s.add_instance(s.add_mesh(v, t))
    lm = s.add_material(F.lambert_material((15.0, 15.0, 15.0)))
    v, t = scenes.add_quad((0, -1, 0), (0, 1, 0), 0.2, 0.2, lm)
    li = s.add_instance(s.add_mesh(v, t))
    s.tri_lights = scenes.concat([scenes.light_triangle(t[k], (15, 15, 15), k, li) for k in range(2)], F.CoreLightTri)
    view = scenes.camera_view((0.0, 0.5, -0.5), (0, 0, 0), 0.05, 1, 1)
    _, acc, n = orc.render(s, view, 1, 1, spp=8192, frames=1, max_depth=2, eps=1e-4)
    L = acc[0, 0, :3] / n
    # numerical integral over the light (midpoint rule, 400 x 400)
    g = (np.arange(400) + 0.5) / 400 * 0.2 - 0.1
    X, Z = np.meshgrid(g, g)
    r2 = X ** 2 + Z ** 2 + 1.0
    E = 15.0 * np.sum((1 / np.sqrt(r2)) * (1 / np.sqrt(r2)) / r2) * (0.2 / 400) ** 2
    expect = 0.5 / np.pi * E
    half = float(np.float16(0.5))
    assert np.allclose(L, expect * half / 0.5, rtol=0.02), (L, expect)


def test_oracle_render_is_deterministic_and_accumulates():
    s = scenes.cornell_box(sphere_tris=2000)
    v = scenes.cornell_view(24, 24)
    a32, a64, n = orc.render(s, v, 24, 24, spp=2, frames=2, max_depth=3)
    b32, b64, _ = orc.render(s, v, 24, 24, spp=2, frames=2, max_depth=3)
    assert n == 4 and np.array_equal(a32, b32) and np.array_equal(a64, b64)
    rel = np.abs(a32[..., :3] - a64[..., :3]) / np.maximum(a64[..., :3], 1e-6)
    assert rel.max() < 1e-5 and a64[..., :3].mean() > 0.01


def test_rays_in_box_planes_match_brute_force():
    """Rays with zero direction components lying in the face planes of BVH
    boxes (0 * inf in the slab test) must not be culled: pixel-corner rays of
    the centred camera through the Cornell sphere, and axis-parallel rays."""
    s = scenes.cornell_box(sphere_tris=5000)
    o = orc.OracleScene(s)
    w, h = 72, 56
    v = scenes.cornell_view(w, h)[0]
    pos, p1 = v["pos"].astype(np.float32), v["p1"].astype(np.float32)
    right, up = (v["p2"] - v["p1"]).astype(np.float32), (v["p3"] - v["p1"]).astype(np.float32)
    rays = []
    for y in range(h):
        D = p1 + np.float32(0.5) * right + np.float32(y / h) * up - pos      # the centre column: D.x == 0
        rays.append(np.concatenate([pos, [1e-3], D / np.linalg.norm(D), [1e34]]))
    for k in range(64):
        c = np.float32(0.2 + 0.6 * k / 63)
        rays.append([0.5, c, -1.0, 1e-3, 0.0, 0.0, 1.0, 1e34])            # along z in the plane x = 0.5
        rays.append([c, 0.5, -1.0, 1e-3, 0.0, 0.0, 1.0, 1e34])
    r = np.array(rays, np.float32)
    assert np.array_equal(o.trace(r).view(np.uint32), o.trace(r, brute=True).view(np.uint32))
    assert (o.trace(r).view(np.int32)[:, 2] >= 0).all()
