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


def test_animated_gltf_through_render_system(torch_cuda, tmp_path):
    """A glTF file whose lamp slides and whose sphere turns (node animations,
    LINEAR translation and rotation): after each update(dt) the instances and
    the lamp's world-space lights are re-sent, and the frame equals the
    oracle's render of the animated scene."""
    from gltf_build import GltfBuilder
    from lighthouse2_amd import loaders as L
    g = GltfBuilder()
    grey = g.material(pbrMetallicRoughness={"baseColorFactor": [0.6, 0.6, 0.6, 1], "roughnessFactor": 1.0})
    lamp = g.material(pbrMetallicRoughness={"baseColorFactor": [12, 12, 10, 1], "roughnessFactor": 1.0})
    P, I = [], []
    n = 12
    for i in range(n + 1):
        th = np.pi * i / n
        for j in range(2 * n + 1):
            ph = np.pi * j / n
            P.append((0.3 * np.sin(th) * np.cos(ph), 0.3 * np.cos(th), 0.3 * np.sin(th) * np.sin(ph)))
    for i in range(n):
        for j in range(2 * n):
            a, b = i * (2 * n + 1) + j, (i + 1) * (2 * n + 1) + j
            I += [a, a + 1, b + 1, a, b + 1, b]
    sphere = g.mesh([g.primitive(np.float32(P), I, None, None, grey)])
    quad = lambda y0, s: np.float32([[-s, y0, -s], [s, y0, -s], [s, y0, s], [-s, y0, s]])
    floor = g.mesh([g.primitive(quad(0, 2), [0, 3, 2, 0, 2, 1], None, None, grey)])
    light = g.mesh([g.primitive(quad(0, 0.15), [0, 1, 2, 0, 2, 3], None, None, lamp)])   # faces -y
    g.node(root=True, mesh=sphere, translation=[0, 0.35, 0])
    g.node(root=True, mesh=floor)
    g.node(root=True, mesh=light, translation=[0, 1.2, 0])
    tin = g.accessor(np.float32([0, 1]))
    move = g.accessor(np.float32([[-0.4, 1.2, 0], [0.4, 1.2, 0.2]]))
    h = np.sqrt(0.5)
    turn = g.accessor(np.float32([[0, 0, 0, 1], [0, h, 0, h]]))
    g.doc["animations"] = [{"samplers": [{"input": tin, "output": move}, {"input": tin, "output": turn}],
                            "channels": [{"sampler": 0, "target": {"node": 2, "path": "translation"}},
                                         {"sampler": 1, "target": {"node": 0, "path": "rotation"}}]}]
    s = scenes.Scene()
    anim = L.load_gltf(s, str(g.write_glb(tmp_path / "anim.glb")), animated=True)
    w, hgt = 40, 32
    view = scenes.camera_view((0, 0.8, -2.0), (0, 0.4, 0), 50.0, w, hgt)
    core = lh.RenderCore()
    core.SetTarget(w, hgt, 1)
    core.Setting("maxDepth", 3)
    rs = RS.RenderSystem(core, s)
    for k in range(3):
        if k:
            anim.update(0.3)
        sent = [c[0] for c in rs.synchronize()]
        if k:
            assert sent == ["SetLights"] + ["SetInstance"] * 4 + ["FinalizeInstances"]
        rs.render(view, lh.LH2_RESTART)
        acc = core.read_accumulator_f64()
        ref = orc.render(s, view, w, hgt, 1, 0, max_depth=3, threads=16)[1]
        assert np.array_equal(acc, ref), k
        assert ref[..., :3].sum() > 0
    core.Shutdown()
