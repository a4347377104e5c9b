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
