"""Animation refit ("refit" setting): SetGeometry with the same triangle count
keeps the mesh's build tree and recomputes its boxes from the new vertices
(on the device, level by level, then the laid-out nodes from the refitted
boxes; on the host copy when the setting was turned on after the build).
Traversal must still return exactly the oracle's hits on the moved geometry
(boxes only prune; the watertight test decides), for both node formats, and
a render of the moved scene must equal the oracle's render of it. Refitting
unchanged vertices reproduces the built tree, so the hits equal those of a
fresh build ray for ray."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import scenes
import oracle as orc
from test_gpu_trace import gpu_any, gpu_closest, workload_rays

pytestmark = pytest.mark.gpu


def moved(scene, seed, amount):
    """The scene with every vertex of mesh 0 displaced (same triangle count)."""
    rng = np.random.default_rng(seed)
    v, t = scene.meshes[0]
    v2 = v.copy()
    v2[:, :3] += rng.normal(0, amount, (len(v), 3)).astype(np.float32)
    s2 = scenes.random_triangle_scene(len(v) // 3, seed=5)
    s2.meshes[0] = (v2, t)
    return s2


@pytest.mark.parametrize("fmt", [0, 1])
def test_refit_hits_equal_oracle(torch_cuda, fmt):
    torch = torch_cuda
    s = scenes.random_triangle_scene(20000, seed=5)
    core = lh.RenderCore()
    core.Setting("nodeFormat", fmt)
    core.Setting("refit", 1)
    lh.load_scene(core, s)
    build_ms = float(core.GetCoreStats()["bvhBuildTime"]) * 1e3
    o0 = orc.OracleScene(s)
    rays = workload_rays(o0, 5)
    first = gpu_closest(torch, core, rays)
    # refit with the same vertices: same boxes, same hits
    core.SetGeometry(0, *s.meshes[0])
    core.FinalizeInstances()
    assert np.array_equal(gpu_closest(torch, core, rays).view(np.uint32), first.view(np.uint32))
    for k, amount in enumerate([0.002, 0.02]):
        s2 = moved(s, 10 + k, amount)
        core.SetGeometry(0, *s2.meshes[0])
        core.FinalizeInstances()
        refit_ms = float(core.GetCoreStats()["bvhBuildTime"]) * 1e3
        o = orc.OracleScene(s2)
        got = gpu_closest(torch, core, rays)
        ref = o.trace(rays)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), amount
        assert np.array_equal(o.trace(rays[::9], brute=True).view(np.uint32), got[::9].view(np.uint32))
        sub = rays[::5]
        assert np.array_equal(gpu_any(torch, core, sub), o.trace(sub, any_hit=True)), amount
        print(f"fmt {fmt}, displacement {amount}: build {build_ms:.2f} ms, refit {refit_ms:.2f} ms")
    core.Shutdown()


def test_refit_render_matches_oracle(torch_cuda):
    """Cornell box whose sphere moves between frames (refit), rendered by the
    reference estimator: fp64 accumulators equal the oracle's render of the
    moved scene."""
    s = scenes.cornell_box(sphere_tris=3000)
    sphere = [i for i, (v, t) in enumerate(s.meshes) if len(t) > 100][0]
    view = scenes.cornell_view(48, 48)
    core = lh.RenderCore()
    core.Setting("refit", 1)
    lh.load_scene(core, s)
    core.SetTarget(48, 48, 1)
    core.Setting("maxDepth", 3)
    core.Render(view, lh.LH2_RESTART)
    v, t = s.meshes[sphere]
    v2 = v.copy()
    v2[:, 0] += np.float32(0.08)
    v2[:, 1] += np.float32(0.03) * np.sin(v[:, 2] * 20).astype(np.float32)
    t2 = t.copy()
    for k in range(3):
        t2[f"vertex{k}"] = v2[k::3, :3]
    core.SetGeometry(sphere, v2, t2)
    core.FinalizeInstances()
    core.Render(view, lh.LH2_RESTART)
    acc = core.read_accumulator_f64()
    core.Shutdown()
    s.meshes[sphere] = (v2, t2)
    r32, r64, n = orc.render(s, view, 48, 48, 1, 0, max_depth=3, threads=16)
    assert np.array_equal(acc, r64)


@pytest.mark.parametrize("fmt", [0, 1])
def test_refit_enabled_after_the_build(torch_cuda, fmt):
    """The refit setting turned on after the mesh was built: the first refit
    runs on the host copy of the build tree and lays the mesh out again, the
    following ones on the device; hits equal the oracle's after each."""
    torch = torch_cuda
    s = scenes.random_triangle_scene(6000, seed=5)
    core = lh.RenderCore()
    core.Setting("nodeFormat", fmt)
    lh.load_scene(core, s)
    core.Setting("refit", 1)
    rays = workload_rays(orc.OracleScene(s), 3)
    for k, amount in enumerate([0.01, 0.03, 0.005]):
        s2 = moved(s, 20 + k, amount)
        core.SetGeometry(0, *s2.meshes[0])
        core.FinalizeInstances()
        ms = float(core.GetCoreStats()["bvhBuildTime"]) * 1e3
        o = orc.OracleScene(s2)
        got = gpu_closest(torch, core, rays)
        assert np.array_equal(got.view(np.uint32), o.trace(rays).view(np.uint32)), k
        sub = rays[::4]
        assert np.array_equal(gpu_any(torch, core, sub), o.trace(sub, any_hit=True)), k
        print(f"fmt {fmt}, refit {k} ({'host' if k == 0 else 'device'}): {ms:.3f} ms")
    core.Shutdown()
