"""The drop-in path a Lighthouse 2 RenderSystem takes: tests/plugin_host.cpp
loads librendercore_hip.so with dlopen, takes CreateCore with dlsym, calls
Init and drives the core through the CoreAPI_Base vtable only
(lib/RenderSystem/core_api_base.cpp:97-130, rendersystem.cpp:46-211).

CPU: the host program is built and resolves every symbol it needs from the
library. GPU: its render of a small Cornell box equals the C-ABI render bit
for bit and the CPU oracle within 1e-4 relative, and GetCoreStats comes back
by value."""
import os
import subprocess

import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import formats as F, scenes
import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "bin", "plugin_host")


def write_scene(path, scene, view, w, h, spp, max_depth, eps, seed):
    with open(path, "wb") as f:
        f.write(np.array([w, h, spp, max_depth], np.int32).tobytes())
        f.write(np.array([eps], np.float32).tobytes())
        f.write(np.array([seed], np.uint32).tobytes())
        f.write(np.ascontiguousarray(view, F.ViewPyramid).tobytes())
        mats = np.ascontiguousarray(scene.materials, F.CoreMaterial)
        f.write(np.array([len(mats)], np.int32).tobytes() + mats.tobytes())
        f.write(np.array([len(scene.meshes)], np.int32).tobytes())
        for v, t in scene.meshes:
            f.write(np.array([len(v) // 3], np.int32).tobytes())
            f.write(np.ascontiguousarray(v, np.float32).tobytes())
            f.write(np.ascontiguousarray(t, F.CoreTri).tobytes())
        tl = np.ascontiguousarray(scene.tri_lights, F.CoreLightTri)
        f.write(np.array([len(tl)], np.int32).tobytes() + tl.tobytes())
        f.write(np.array([len(scene.instances)], np.int32).tobytes())
        for m, T in scene.instances:
            f.write(np.array([m], np.int32).tobytes() + np.ascontiguousarray(T, np.float32).reshape(16).tobytes())


def test_plugin_host_builds_and_resolves_symbols(tmp_path):
    assert os.path.exists(HOST), "tests/bin/plugin_host missing: build with `make -C lighthouse2_amd/csrc`"
    # a missing scene file fails after dlopen + dlsym succeeded, before any HIP call
    r = subprocess.run([HOST, lh.core.LIB_PATH, str(tmp_path / "none.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert r.returncode == 2, (r.returncode, r.stderr)
    r = subprocess.run([HOST, "/nonexistent/librendercore_hip.so", "a", "b"], capture_output=True, text=True)
    assert r.returncode == 3


@pytest.mark.gpu
def test_render_through_the_vtable_matches_oracle(tmp_path, torch_cuda):
    s = scenes.cornell_box(sphere_tris=2000)
    w, h, spp, depth, eps, seed = 64, 48, 2, 3, 1e-3, 1234
    view = scenes.cornell_view(w, h)
    scene_file, out_file = tmp_path / "scene.bin", tmp_path / "out.bin"
    write_scene(scene_file, s, view, w, h, spp, depth, eps, seed)
    print("plugin host: cornell box 64x48, 2 spp", flush=True)
    r = subprocess.run([HOST, lh.core.LIB_PATH, str(scene_file), str(out_file)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    raw = open(out_file, "rb").read()
    n = w * h
    img = np.frombuffer(raw, np.float32, 4 * n).reshape(h, w, 4)
    acc = np.frombuffer(raw, np.float64, 4 * n, offset=16 * n).reshape(h, w, 4)
    stats = np.frombuffer(raw, F.CoreStats, 1, offset=48 * n)[0]
    assert stats["primaryRayCount"] == n * spp and stats["renderTime"] > 0
    # the same frame through the C ABI
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(w, h, spp)
    core.Setting("maxDepth", depth)
    core.Setting("epsilon", eps)
    core.Setting("seed", seed)
    core.Render(view, lh.LH2_RESTART)
    assert np.array_equal(acc, core.read_accumulator_f64()) and np.array_equal(img, core.read_image())
    core.Shutdown()
    # and the oracle
    _, ref64, _ = orc.render(s, view, w, h, spp, 0, depth, eps, seed=seed, modes=[orc.RESTART])
    err = np.abs(acc[..., :3] - ref64[..., :3]) / np.maximum(np.abs(ref64[..., :3]), 1e-6)
    print(f"plugin path vs oracle: mean rel {err.mean():.3g}, max rel {err.max():.3g}")
    assert err.max() <= 1e-4
