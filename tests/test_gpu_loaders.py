"""A scene read from OBJ + MTL + PNG files (lighthouse2_amd/loaders.py) and
handed to the core through the boundary renders on the GPU exactly as the
oracle renders the same arrays: a smooth-shaded sphere (vertex normals, the
consistent-normal alphas), a textured floor (FLIPPED + LINEARIZED texture,
triangle LOD) and an emissive quad that becomes two lights; fp64
accumulators identical for the reference and the optix7 estimators."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import loaders as L
from lighthouse2_amd import scenes
import oracle as orc
from test_loaders import encode_png

pytestmark = pytest.mark.gpu


def write_obj_scene(d, slices=24, stacks=12):
    lines = ["mtllib scene.mtl"]
    vs, vts, vns = [], [], []
    for i in range(stacks + 1):
        th = np.pi * i / stacks
        for j in range(slices + 1):
            ph = 2 * np.pi * j / slices
            n = (np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph))
            vs.append((0.5 + 0.25 * n[0], 0.3 + 0.25 * n[1], 0.5 + 0.25 * n[2]))
            vns.append(n)
            vts.append((j / slices, i / stacks))
    lines += [f"v {x:.6f} {y:.6f} {z:.6f}" for x, y, z in vs]
    floor = len(vs)
    lines += ["v -1 0 -1", "v 2 0 -1", "v 2 0 2", "v -1 0 2", "v 0.3 1.2 0.3", "v 0.7 1.2 0.3", "v 0.7 1.2 0.7", "v 0.3 1.2 0.7"]
    lines += [f"vt {u:.6f} {v:.6f}" for u, v in vts] + ["vt 0 0", "vt 3 0", "vt 3 3", "vt 0 3"]
    lines += [f"vn {x:.6f} {y:.6f} {z:.6f}" for x, y, z in vns]
    lines.append("usemtl plastic")
    for i in range(stacks):
        for j in range(slices):
            a, b = i * (slices + 1) + j + 1, (i + 1) * (slices + 1) + j + 1
            lines.append(f"f {a}/{a}/{a} {a + 1}/{a + 1}/{a + 1} {b + 1}/{b + 1}/{b + 1} {b}/{b}/{b}")
    t0 = len(vts)
    lines.append("usemtl floor")
    lines.append(f"f {floor + 1}/{t0 + 1} {floor + 4}/{t0 + 4} {floor + 3}/{t0 + 3} {floor + 2}/{t0 + 2}")
    lines.append("usemtl lamp")
    lines.append(f"f {floor + 5} {floor + 6} {floor + 7} {floor + 8}")
    (d / "scene.obj").write_text("\n".join(lines) + "\n")
    (d / "scene.mtl").write_text("newmtl plastic\nKd 0.7 0.3 0.2\nnewmtl floor\nKd 1 1 1\nmap_Kd floor.png\nnewmtl lamp\nKd 12 11 9\n")
    y, x = np.mgrid[0:32, 0:32]
    img = np.zeros((32, 32, 4), np.uint8)
    img[..., 0] = np.where(((x // 4) + (y // 4)) % 2, 230, 60)
    img[..., 1] = 90 + 4 * x
    img[..., 2] = 200 - 5 * y
    img[..., 3] = 255
    (d / "floor.png").write_bytes(encode_png(img, 6, [0, 1, 2, 3, 4]))
    return d / "scene.obj"


@pytest.mark.parametrize("estimator", [0, 1])
def test_obj_scene_matches_oracle(torch_cuda, tmp_path, estimator):
    s = scenes.Scene()
    mesh = L.load_obj(s, str(write_obj_scene(tmp_path)))
    L.add_instance_with_lights(s, mesh)
    assert len(s.tri_lights) == 2
    w, h = 48, 40
    view = scenes.camera_view((0.5, 0.6, -1.0), (0.5, 0.35, 0.5), 50.0, w, h)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(w, h, 2)
    core.Setting("estimator", estimator)
    core.Setting("maxDepth", 3)
    if estimator == 1:
        core.Setting("bsdf", 0)
    for m in [lh.LH2_RESTART, lh.LH2_CONVERGE, lh.LH2_CONVERGE]:
        core.Render(view, m)
    acc = core.read_accumulator_f64()
    core.Shutdown()
    if estimator == 0:
        ref = orc.render(s, view, w, h, 2, 2, max_depth=3, threads=16)
        r64 = ref[1]
    else:
        _, r64, n, _, _ = orc.render7(s, view, w, h, 2, 2, threads=16, bsdf=0)
    assert np.array_equal(acc, r64)
    assert r64[..., :3].sum() > 0


def write_gltf_scene(d, slices=20, stacks=10):
    """A glTF scene: one sphere mesh (u16 indices, normals, uvs) instanced by a
    rotated, scaled root node and by its child (column-major matrix), a
    textured floor (u8 indices) and an emissive quad under a translated node."""
    from gltf_build import GltfBuilder
    g = GltfBuilder()
    y, x = np.mgrid[0:16, 0:16]
    img = np.zeros((16, 16, 4), np.uint8)
    img[..., 0] = np.where(((x // 2) + (y // 2)) % 2, 220, 50)
    img[..., 1] = 60 + 8 * x
    img[..., 2] = 190 - 6 * y
    img[..., 3] = 255
    tex = g.image_png(encode_png(img, 6, [1, 2, 3, 4, 0]))
    plastic = g.material(pbrMetallicRoughness={"baseColorFactor": [0.7, 0.3, 0.2, 1], "roughnessFactor": 1.0, "metallicFactor": 0.0})
    floor = g.material(pbrMetallicRoughness={"baseColorFactor": [1, 1, 1, 1], "roughnessFactor": 1.0, "metallicFactor": 0.0,
                                             "baseColorTexture": {"index": tex}})
    lamp = g.material(pbrMetallicRoughness={"baseColorFactor": [14, 12, 10, 1], "roughnessFactor": 1.0})
    P, N, UV, I = [], [], [], []
    for i in range(stacks + 1):
        th = np.pi * i / stacks
        for j in range(slices + 1):
            ph = 2 * np.pi * j / slices
            n = (np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph))
            P.append(n), N.append(n), UV.append((j / slices, i / stacks))
    for i in range(stacks):
        for j in range(slices):
            a, b = i * (slices + 1) + j, (i + 1) * (slices + 1) + j
            I += [a, a + 1, b + 1, a, b + 1, b]
    sphere = g.mesh([g.primitive(np.float32(P), I, np.float32(N), np.float32(UV), plastic)])
    quad = lambda y0, s: np.float32([[-s, y0, -s], [s, y0, -s], [s, y0, s], [-s, y0, s]])
    ground = g.mesh([g.primitive(quad(0, 1.5) + [0.5, 0, 0.5], [0, 3, 2, 0, 2, 1], None,
                                 np.float32([[0, 0], [0, 3], [3, 3], [3, 0]]), floor, index_dtype="u1")])
    light = g.mesh([g.primitive(quad(0, 0.2), [0, 2, 1, 0, 3, 2], None, None, lamp)])     # faces +y
    child = g.node(mesh=sphere, matrix=[0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0, 2.2, 0.1, 0, 1])
    g.node(root=True, mesh=sphere, children=[child], translation=[0.35, 0.22, 0.5],
           rotation=[0, np.sin(0.3), 0, np.cos(0.3)], scale=[0.22, 0.22, 0.22])
    g.node(root=True, mesh=ground)
    g.node(root=True, mesh=light, translation=[0.5, 1.1, 0.5], rotation=[1, 0, 0, 0])   # flipped to face down
    return g.write_glb(d / "scene.glb")


@pytest.mark.parametrize("estimator", [0, 1])
def test_gltf_scene_matches_oracle(torch_cuda, tmp_path, estimator):
    s = scenes.Scene()
    L.load_gltf(s, str(write_gltf_scene(tmp_path)))
    assert len(s.instances) == 4 and len(s.tri_lights) == 2
    w, h = 48, 40
    view = scenes.camera_view((0.5, 0.6, -1.0), (0.5, 0.3, 0.5), 55.0, w, h)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(w, h, 2)
    core.Setting("estimator", estimator)
    core.Setting("maxDepth", 3)
    if estimator == 1:
        core.Setting("bsdf", 0)
    for m in [lh.LH2_RESTART, lh.LH2_CONVERGE]:
        core.Render(view, m)
    acc = core.read_accumulator_f64()
    core.Shutdown()
    if estimator == 0:
        r64 = orc.render(s, view, w, h, 2, 1, max_depth=3, threads=16)[1]
    else:
        r64 = orc.render7(s, view, w, h, 2, 1, threads=16, bsdf=0)[1]
    assert np.array_equal(acc, r64)
    assert r64[..., :3].sum() > 0
