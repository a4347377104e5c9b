"""An animated sequence through the headless RenderSystem on the GPU: the
sphere mesh moves (re-sent, refitted on the device) and a second instance of
it moves by transform, frame by frame; each restarted frame's fp64
accumulator equals the oracle's render of that frame's scene, and each
synchronize() sent only what changed."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import rendersystem as RS
from lighthouse2_amd import scenes
import oracle as orc

pytestmark = pytest.mark.gpu


def test_animated_sequence_matches_oracle(torch_cuda):
    s = scenes.cornell_box(sphere_tris=2000)
    sphere = [i for i, (v, t) in enumerate(s.meshes) if len(t) > 100][0]
    T = np.eye(4, dtype=np.float32)
    T[:3, :3] *= 0.5
    T[:3, 3] = (0.25, 0.2, 0.3)
    s.add_instance(sphere, T.reshape(16))
    v0, t0 = s.meshes[sphere]
    w, h = 40, 40
    view = scenes.cornell_view(w, h)
    core = lh.RenderCore()
    core.Setting("refit", 1)
    core.SetTarget(w, h, 1)
    core.Setting("maxDepth", 3)
    rs = RS.RenderSystem(core, s)
    for k in range(3):
        d = np.float32(0.04 * k)
        v = v0.copy()
        v[:, 0] += d
        t = t0.copy()
        for c in range(3):
            t[f"vertex{c}"] = v[c::3, :3]
        s.meshes[sphere] = (v, t)
        T2 = T.copy()
        T2[1, 3] += d
        s.instances[-1] = (sphere, T2.reshape(16))
        sent = [c[0] for c in rs.synchronize()]
        if k > 0:
            assert sent == ["SetGeometry"] + ["SetInstance"] * (len(s.instances) + 1) + ["FinalizeInstances"]
        rs.render(view, lh.LH2_RESTART)
        acc = core.read_accumulator_f64()
        ref = orc.render(s, view, w, h, 1, 0, max_depth=3, threads=16)[1]
        assert np.array_equal(acc, ref), k
        print(f"frame {k}: bvh time {float(core.GetCoreStats()['bvhBuildTime']) * 1e3:.3f} ms")
    core.Shutdown()
