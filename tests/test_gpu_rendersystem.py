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


def test_skinned_gltf_refits_and_matches_oracle(torch_cuda, tmp_path):
    """A skinned strip bending at its upper joint (LINEAR rotation): each
    update re-poses the mesh, the RenderSystem re-sends it, the core refits
    it on the device, and the frame equals the oracle's render."""
    from gltf_build import GltfBuilder
    from lighthouse2_amd import loaders as L
    f32 = np.float32
    g = GltfBuilder()
    grey = g.material(pbrMetallicRoughness={"baseColorFactor": [0.7, 0.6, 0.5, 1], "roughnessFactor": 1.0})
    lamp = g.material(pbrMetallicRoughness={"baseColorFactor": [14, 12, 10, 1], "roughnessFactor": 1.0})
    rows = 9
    P = np.float32([[x, 0.2 * r, 0] for r in range(rows) for x in (-0.2, 0.2)])
    idx = []
    for r in range(rows - 1):
        a = 2 * r
        idx += [a, a + 1, a + 3, a, a + 3, a + 2]
    wt = np.float32([[1 - min(max((0.2 * r - 0.6) / 0.6, 0), 1), min(max((0.2 * r - 0.6) / 0.6, 0), 1), 0, 0] for r in range(rows) for _ in (0, 1)])
    prim = g.primitive(P, idx, material=grey)
    prim["attributes"]["JOINTS_0"] = g.accessor(np.array([[0, 1, 0, 0]] * (2 * rows), "u2"))
    prim["attributes"]["WEIGHTS_0"] = g.accessor(wt)
    strip = g.mesh([prim])
    quad = lambda y0, s: np.float32([[-s, y0, -s], [s, y0, -s], [s, y0, s], [-s, y0, s]])
    floor = g.mesh([g.primitive(quad(0, 2), [0, 3, 2, 0, 2, 1], None, None, grey)])
    light = g.mesh([g.primitive(quad(0, 0.2), [0, 1, 2, 0, 2, 3], None, None, lamp)])
    ibm = np.stack([np.eye(4, dtype=f32), np.eye(4, dtype=f32)])
    ibm[1, 1, 3] = -0.8
    view_i = g.view(ibm.transpose(0, 2, 1).astype(f32).tobytes())
    g.doc["accessors"].append({"bufferView": view_i, "componentType": 5126, "count": 2, "type": "MAT4"})
    g.doc["skins"] = [{"joints": [3, 4], "inverseBindMatrices": len(g.doc["accessors"]) - 1}]
    g.node(root=True, mesh=strip, skin=0)
    g.node(root=True, mesh=floor)
    g.node(root=True, mesh=light, translation=[0, 2.2, 0])
    g.node(root=True, children=[4])
    g.node(translation=[0, 0.8, 0])
    tin = g.accessor(np.float32([0, 1]))
    h = np.sqrt(0.5)
    rot = g.accessor(np.float32([[0, 0, 0, 1], [0, 0, h, h]]))
    g.doc["animations"] = [{"samplers": [{"input": tin, "output": rot}], "channels": [{"sampler": 0, "target": {"node": 4, "path": "rotation"}}]}]
    s = scenes.Scene()
    anim = L.load_gltf(s, str(g.write_glb(tmp_path / "skin.glb")), animated=True)
    w, hgt = 40, 40
    view = scenes.camera_view((0, 1.0, -2.5), (0, 0.8, 0), 50.0, w, hgt)
    core = lh.RenderCore()
    core.Setting("refit", 1)
    core.SetTarget(w, hgt, 1)
    core.Setting("maxDepth", 3)
    rs = RS.RenderSystem(core, s)
    for k in range(3):
        if k:
            anim.update(0.3)
        sent = [c[0] for c in rs.synchronize()]
        if k:
            assert sent[0] == "SetGeometry" and "SetInstance" in sent
        rs.render(view, lh.LH2_RESTART)
        acc = core.read_accumulator_f64()
        ref = orc.render(s, view, w, hgt, 1, 0, max_depth=3, threads=16)[1]
        assert np.array_equal(acc, ref), k
        assert ref[..., :3].sum() > 0
    core.Shutdown()
