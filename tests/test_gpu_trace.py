"""GPU traversal parity: the gfx950 extend / any-hit kernels (fp32 BVH2 and
8-bit quantised BVH4 nodes, single- and two-level) against the CPU oracle,
bit-exact on (t, u|v, instance, primitive), through the C ABI."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import scenes
import oracle as orc
from test_oracle import _random_rays, instanced_scene

pytestmark = pytest.mark.gpu


def gpu_closest(torch, core, rays):
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    h = torch.zeros((len(rays), 4), dtype=torch.float32, device="cuda")
    core.trace_closest(r.data_ptr(), h.data_ptr(), len(rays))
    torch.cuda.synchronize()
    return h.cpu().numpy()


def gpu_any(torch, core, rays):
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    m = torch.zeros(((len(rays) + 31) // 32,), dtype=torch.int32, device="cuda")
    core.trace_any(r.data_ptr(), m.data_ptr(), len(rays))
    torch.cuda.synchronize()
    bits = m.cpu().numpy().view(np.uint32)
    idx = np.arange(len(rays))
    return ((bits[idx >> 5] >> (idx & 31)) & 1).astype(bool)


def make_core(scene, fmt, loop=0):
    core = lh.RenderCore()
    core.Setting("nodeFormat", fmt)
    core.Setting("traceLoop", loop)
    lh.load_scene(core, scene)
    return core


def workload_rays(o, seed, n_random=20000, res=96):
    rng = np.random.default_rng(seed)
    prim = orc.primary_rays(scenes.microbench_view(res, res), res, res, 1234, 1e-5)
    h = o.trace(prim)
    b1, _ = o.bounce(prim, h, np.arange(len(prim), dtype=np.uint32), 91771, 1e-5)
    return np.concatenate([prim, b1, _random_rays(rng, n_random)])


@pytest.mark.parametrize("loop", [0, 1, 2])
@pytest.mark.parametrize("fmt", [0, 1])
def test_closest_and_any_match_oracle(torch_cuda, fmt, loop):
    s = scenes.random_triangle_scene(20000, seed=5)
    o = orc.OracleScene(s)
    rays = workload_rays(o, 5)
    core = make_core(s, fmt, loop)
    got = gpu_closest(torch_cuda, core, rays)
    ref = o.trace(rays)
    print(f"fmt {fmt}: {len(rays)} rays, {(ref.view(np.int32)[:, 2] >= 0).sum()} hits")
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    sub = rays[::7]
    assert np.array_equal(o.trace(sub, brute=True).view(np.uint32), got[::7].view(np.uint32))
    assert np.array_equal(gpu_any(torch_cuda, core, rays), o.trace(rays, any_hit=True))
    core.Shutdown()


@pytest.mark.parametrize("loop", [0, 1, 2])
@pytest.mark.parametrize("fmt", [0, 1])
def test_two_level_matches_brute_force(torch_cuda, fmt, loop):
    s = instanced_scene(seed=21, meshes=4, instances=40, tris=600)
    o = orc.OracleScene(s)
    rays = _random_rays(np.random.default_rng(2), 30000, -3, 3)
    core = make_core(s, fmt, loop)
    got = gpu_closest(torch_cuda, core, rays)
    ref = o.trace(rays, brute=True)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(gpu_any(torch_cuda, core, rays), o.trace(rays, any_hit=True, brute=True))
    core.Shutdown()


def canonical(nodes, prims):
    """Depth-first serialisation (left first) of a build tree: independent of
    node numbering and of the primitive order inside leaves."""
    out, st = [], [0]
    while st:
        i = st.pop()
        n = nodes[i]
        if n["count"] > 0:
            out.append(("L", tuple(n["lo"]), tuple(n["hi"]), tuple(sorted(prims[n["leftFirst"]:n["leftFirst"] + n["count"]].tolist()))))
        else:
            out.append(("N", tuple(n["lo"]), tuple(n["hi"])))
            st.append(n["leftFirst"] + 1), st.append(n["leftFirst"])
    return out


def test_gpu_bvh_build_equals_oracle_build(torch_cuda):
    for s in (scenes.random_triangle_scene(30000, seed=8), scenes.cornell_box()):
        o = orc.OracleScene(s)
        core = make_core(s, 1)
        for m in range(len(s.meshes)):
            gn, gp = core.bvh_download(m)
            on, op, _ = o.bvh_nodes(m)
            assert len(gn) == len(on)
            assert canonical(gn, gp) == canonical(on, op)
        core.Shutdown()


def test_wave_kernels_match_oracle(torch_cuda):
    torch = torch_cuda
    s = scenes.random_triangle_scene(20000, seed=6)
    o = orc.OracleScene(s)
    core = make_core(s, 1)
    W, H = 80, 60
    view = scenes.microbench_view(W, H)
    n = W * H
    rays = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
    pid = torch.zeros(n, dtype=torch.int32, device="cuda")
    lh.core.generate_primary(view, W, H, 4321, 1e-5, rays.data_ptr(), pid.data_ptr())
    torch.cuda.synchronize()
    ref = orc.primary_rays(view, W, H, 4321, 1e-5)
    assert np.array_equal(rays.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    hits = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    core.trace_closest(rays.data_ptr(), hits.data_ptr(), n)
    out = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
    pout = torch.zeros(n, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    core.bounce_diffuse(rays.data_ptr(), hits.data_ptr(), pid.data_ptr(), n, 91771, 1e-5, out.data_ptr(), pout.data_ptr(), cnt.data_ptr())
    torch.cuda.synchronize()
    k = int(cnt.item())
    rb, pb = o.bounce(ref, o.trace(ref), np.arange(n, dtype=np.uint32), 91771, 1e-5)
    assert k == len(rb)
    got_p = pout.cpu().numpy()[:k].view(np.uint32)
    order = np.argsort(got_p)
    assert np.array_equal(got_p[order], pb)
    assert np.array_equal(out.cpu().numpy()[:k][order].view(np.uint32), rb.view(np.uint32))
    core.Shutdown()


@pytest.mark.parametrize("loop", [0, 1, 2])
def test_edge_cases(torch_cuda, loop):
    s = scenes.Scene()
    v = np.array([[-1, -1, 2, 1], [3, -1, 2, 1], [-1, 3, 2, 1]], np.float32)
    s.add_instance(s.add_mesh(np.concatenate([v, v]), None))  # duplicate triangle: tie -> prim 0
    core = make_core(s, 1, loop)
    r = np.array([[0, 0, 0, 0, 0, 0, 1, 1e34], [0, 0, 0, 0, 0, 0, 1, 2.0], [0, 0, 0, 2.0, 0, 0, 1, 1e34],
                  [0, 0, 0, 0, 0, 0, -1, 1e34], [0, 0, 0, 0, 0, 0, 1, np.nextafter(np.float32(2), np.float32(3))]], np.float32)
    h = gpu_closest(torch_cuda, core, r).view(np.int32)
    assert list(h[:, 2]) == [0, -1, -1, -1, 0]
    ref = orc.OracleScene(s).trace(r)
    assert np.array_equal(h.view(np.uint32), ref.view(np.uint32))
    core.trace_closest(0, 0, 0)  # empty wave
    with pytest.raises(lh.LH2HIPError):
        core.trace_closest(0, 0, 5)
    core.Shutdown()
    empty = lh.RenderCore()
    with pytest.raises(lh.LH2HIPError):
        empty.trace_closest(1, 1, 1)  # no scene yet
    empty.Shutdown()


def test_large_scene_against_oracle_bvh(torch_cuda):
    s = scenes.random_triangle_scene(300000, seed=1)
    o = orc.OracleScene(s)
    rays = workload_rays(o, 9, n_random=100000, res=256)
    core = make_core(s, 1)
    got = gpu_closest(torch_cuda, core, rays)
    ref = o.trace(rays, threads=16)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    core.Setting("traceLoop", 1)
    assert np.array_equal(gpu_closest(torch_cuda, core, rays).view(np.uint32), ref.view(np.uint32))
    core.Setting("nodeFormat", 0)
    assert np.array_equal(gpu_closest(torch_cuda, core, rays).view(np.uint32), ref.view(np.uint32))
    core.Setting("traceLoop", 0)
    assert np.array_equal(gpu_closest(torch_cuda, core, rays).view(np.uint32), ref.view(np.uint32))
    core.Setting("traceLoop", 2)
    assert np.array_equal(gpu_closest(torch_cuda, core, rays).view(np.uint32), ref.view(np.uint32))
    core.Shutdown()
