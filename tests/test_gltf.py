"""glTF producer (lighthouse2_amd/loaders.py) against closed forms of the
reference's glTF path: HostMesh::ConvertFromGTLFMesh (host_mesh.cpp:327-513:
index widths, strided attributes, strip / fan expansion, skipped modes),
BuildFromIndexedData (:520-681: unclamped consistent-normal alphas, the flip
rule, unflipped face normals, uv / PBRT / edge tangents, single-texel
material copies), HostMaterial::ConvertFrom (host_material.cpp:76-170),
HostNode's TRS / matrix transforms and the post-order instance list
(host_node.cpp:77-199), and world-space lights of emissive triangles. The
.glb, .gltf + .bin and .gltf + data: URI containers must give identical
arrays. CPU only."""
import numpy as np
import pytest

from lighthouse2_amd import formats as F
from lighthouse2_amd import loaders as L
from lighthouse2_amd import scenes
from gltf_build import GltfBuilder
from test_loaders import encode_png

f32 = np.float32
P6 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 2, 0], [-1, 1, 0]], f32)


def modes_file(tmp_path):
    g = GltfBuilder()
    inter = np.concatenate([P6, np.tile([[0, 0, 1]], (6, 1)).astype(f32)], 1)     # position + normal, stride 24
    v = g.view(inter.tobytes(), stride=24)
    pos = g.accessor(P6, view=v, offset=0, count=6, width=3, dtype="f4")
    nrm = g.accessor(None, view=v, offset=12, count=6, width=3, dtype="f4")
    at = {"POSITION": pos, "NORMAL": nrm}
    prims = [
        {"attributes": at, "indices": g.accessor(np.array([0, 1, 2], "u1"))},                       # TRIANGLES, u8
        {"attributes": at, "indices": g.accessor(np.array([0, 1, 2, 3], "u2")), "mode": 5},          # strip
        {"attributes": at, "indices": g.accessor(np.array([0, 2, 3, 4, 5], "u4")), "mode": 6},       # fan
        {"attributes": at, "indices": g.accessor(np.array([0, 1], "u2")), "mode": 1},                # lines: skipped
        {"attributes": at},                                                                          # no indices
        {"attributes": at, "indices": g.accessor(np.array([5, 4, 3], "i1"))},                        # signed bytes
    ]
    g.mesh(prims)
    g.node(root=True, mesh=0)
    return g


def test_primitive_modes_and_index_widths(tmp_path):
    s = scenes.Scene()
    L.load_gltf(s, str(modes_file(tmp_path).write_glb(tmp_path / "m.glb")))
    verts, tris = s.meshes[0]
    want = [(0, 1, 2), (0, 1, 2), (1, 2, 3), (0, 2, 3), (0, 3, 4), (0, 4, 5), (0, 1, 2), (3, 4, 5), (5, 4, 3)]
    assert len(tris) == len(want) and verts.shape == (3 * len(want), 4)
    assert np.array_equal(verts[:, :3], P6[np.array(want).reshape(-1)]) and np.all(verts[:, 3] == 1)
    assert np.array_equal(tris["vertex1"], P6[np.array(want)[:, 1]])
    # no materials in the file: every triangle uses scene material 0 (added as grey Lambert here)
    assert len(s.materials) == 1 and np.all(tris["material"] == 0)


@pytest.mark.parametrize("kind", ["glb", "gltf", "uri"])
def test_containers_agree(tmp_path, kind):
    g = modes_file(tmp_path)
    ref = scenes.Scene()
    L.load_gltf(ref, str(g.write_glb(tmp_path / "a.glb")))
    s = scenes.Scene()
    path = {"glb": lambda: g.write_glb(tmp_path / "b.glb"), "gltf": lambda: g.write_gltf(tmp_path / "b.gltf"),
            "uri": lambda: g.write_gltf(tmp_path / "c.gltf", data_uri=True)}[kind]()
    L.load_gltf(s, str(path))
    assert np.array_equal(s.meshes[0][0], ref.meshes[0][0])
    assert s.meshes[0][1].tobytes() == ref.meshes[0][1].tobytes()


def test_alphas_flip_rule_and_tangents(tmp_path):
    g = GltfBuilder()
    m = g.material(pbrMetallicRoughness={"baseColorFactor": [0.5, 0.5, 0.5, 1]})
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32)                      # face normal +z
    c, s_ = f32(np.cos(np.pi / 3)), f32(np.sin(np.pi / 3))
    n_tilt = np.array([[0, 0, 1], [s_, 0, c], [0, 0, 1]], f32)                  # dot 0.5 at vertex 1
    n_away = np.tile([[0, 0, -1]], (3, 1)).astype(f32)                          # all three face away: flip
    n_mixed = np.array([[0, 0, -1], [0, 0, -1], [0, 0, 1]], f32)                # two away: no flip
    uv = np.array([[0, 0], [1, 0], [0, 1]], f32)
    uv_deg = np.array([[0, 0], [0, 0], [0, 1]], f32)                            # |uv01| = 0: PBRT frame
    prims = [g.primitive(tri, [0, 1, 2], n_tilt, uv, m), g.primitive(tri, [0, 1, 2], n_away, uv_deg, m),
             g.primitive(tri, [0, 1, 2], n_mixed, None, m), g.primitive(tri, [0, 1, 2], None, None, m)]
    g.mesh(prims)
    g.node(root=True, mesh=0)
    s = scenes.Scene()
    L.load_gltf(s, str(g.write_glb(tmp_path / "a.glb")))
    _, t = s.meshes[0]
    w = 0.03632
    al = lambda d: f32(np.arccos(f32(d)) * (1 + w * (1 - d) ** 2))
    # tilted normal: acos(0.5)(1 + 0.03632 * 0.25), no clamp at 0.7 (the OBJ path clamps)
    assert abs(t["alpha"][0][1] - al(0.5)) < 1e-6 and t["alpha"][0][0] == 0 and t["alpha"][0][2] == 0
    # all normals away: N flips for the alpha (dot 1 -> 0), the stored face normal stays +z
    assert np.all(t["alpha"][1] == 0) and (t["Nx"][1], t["Ny"][1], t["Nz"][1]) == (0, 0, 1)
    assert np.array_equal(t["vN0"][1], [0, 0, -1])
    # two away, one toward: no flip; dot -1 -> acos(-1)(1 + 0.03632 * 4)
    assert abs(t["alpha"][2][0] - al(-1.0)) < 1e-5 and t["alpha"][2][2] == 0
    # uv tangents, PBRT frame for N = +z: T = (0, 1, 0), B = N x T = (-1, 0, 0)
    assert np.allclose(t["T"][0], [1, 0, 0]) and np.allclose(t["B"][0], [0, 1, 0])
    assert np.array_equal(t["T"][1], [0, 1, 0]) and np.allclose(t["B"][1], [-1, 0, 0])
    # no uvs: edge tangent; no normals: vN = N and a tiny alpha of acos(min(1, N.N))
    assert np.allclose(t["T"][2], [1, 0, 0]) and np.allclose(t["B"][2], [0, 1, 0])
    assert np.array_equal(t["vN1"][3], [t["Nx"][3], t["Ny"][3], t["Nz"][3]]) and np.all((t["alpha"][3] >= 0) & (t["alpha"][3] < 1e-3))
    assert np.all(t["LOD"] == 0) and np.all(t["area"] == 0) and np.all(t["ltriIdx"] == -1)


def _png(rgb):
    h, w = rgb.shape[:2]
    img = np.concatenate([rgb, np.full((h, w, 1), 255, np.uint8)], 2)
    return encode_png(img, 6, [0, 1, 4])


def test_materials_textures_and_single_texel_copies(tmp_path):
    y, x = np.mgrid[0:8, 0:8]
    col = np.stack([20 + 20 * x, 10 + 25 * y, (x * y) % 256], -1).astype(np.uint8)
    nrm = np.stack([128 + 0 * x, 128 + 0 * x, 255 + 0 * x], -1).astype(np.uint8)
    g = GltfBuilder()
    t0 = g.image_png(_png(col))
    t1 = g.image_png(_png(nrm), as_uri=True)
    m0 = g.material(pbrMetallicRoughness={"baseColorFactor": [0.2, 0.4, 0.6, 1], "metallicFactor": 0.3,
                                          "roughnessFactor": 0.7, "baseColorTexture": {"index": t0},
                                          "metallicRoughnessTexture": {"index": t0}},
                    normalTexture={"index": t1, "scale": 0.5})
    m1 = g.material(pbrMetallicRoughness={"baseColorFactor": [6, 5, 4, 1]})
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32)
    one = np.full((3, 2), [0.25, 0.75], f32)                                   # a single texel: (2, 6)
    prims = [g.primitive(tri, [0, 1, 2], None, np.array([[0, 0], [1, 0], [0, 1]], f32), m0),
             g.primitive(tri, [0, 1, 2], None, one, m0), g.primitive(tri, [0, 1, 2], None, one, m0),
             g.primitive(tri, [0, 1, 2], None, None, m1)]
    g.mesh(prims)
    g.node(root=True, mesh=0)
    s = scenes.Scene()
    L.load_gltf(s, str(g.write_glb(tmp_path / "a.glb")))
    mats, (_, t) = s.materials, s.meshes[0]
    assert len(s.textures) == 2 and len(mats) == 3                             # two file materials + one copy
    a = mats[0]
    assert np.allclose(a["color"]["value"], [0.2, 0.4, 0.6]) and a["color"]["textureID"] == 0
    assert abs(a["metallic"]["value"] - 0.3) < 1e-7 and abs(a["roughness"]["value"] - 0.7) < 1e-7
    assert a["roughness"]["textureID"] == 0 and a["metallic"]["textureID"] == 0
    assert a["normals"]["textureID"] == 1 and a["normals"]["scale"] == 0.5 and a["flags"] == 1 | L.MAT_FROM_MTL
    b = mats[1]
    assert np.allclose(b["color"]["value"], [6, 5, 4]) and b["metallic"]["value"] == f32(1e-32) and b["roughness"]["value"] == f32(1e-32)
    # textures: top row first, not linearised, byte 0 = red; the normal map as NRM32
    px = s.textures[0]["pixels"][:64].reshape(8, 8)
    assert np.array_equal(px & 255, col[..., 0]) and np.array_equal((px >> 8) & 255, col[..., 1])
    assert np.array_equal((px >> 16) & 255, col[..., 2])
    assert s.textures[0]["storage"] == F.ARGB32 and s.textures[1]["storage"] == F.NRM32
    assert s.textures[1]["flags"] == L.TEX_LDR | L.TEX_NORMALMAP
    # single-texel triangles: one shared untextured copy of material 0 coloured by texel (2, 6)
    assert t["material"][0] == 0 and t["material"][1] == 2 and t["material"][2] == 2 and t["material"][3] == 1
    cp = mats[2]
    assert cp["color"]["textureID"] == -1 and cp["flags"] == 1 | L.MAT_FROM_MTL | L.MAT_SINGLE_COLOR_COPY
    assert np.array_equal(cp["color"]["value"], col[6, 2].astype(f32) * f32(1 / 255))
    assert cp["roughness"]["value"] == a["roughness"]["value"] and cp["normals"]["textureID"] == 1


def test_node_transforms_instances_and_lights(tmp_path):
    g = GltfBuilder()
    glow = g.material(pbrMetallicRoughness={"baseColorFactor": [5, 4, 3, 1]})
    grey = g.material(pbrMetallicRoughness={"baseColorFactor": [0.5, 0.5, 0.5, 1]})
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32)
    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], f32)
    body = g.mesh([g.primitive(tri, [0, 1, 2], material=grey)])
    lamp = g.mesh([g.primitive(quad, [0, 2, 1, 0, 3, 2], material=glow)])
    q = [0, np.sin(np.pi / 4), 0, np.cos(np.pi / 4)]                           # 90 degrees about y
    child = g.node(mesh=body, matrix=[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 5, 1])   # column-major: translate z 5
    g.node(root=True, mesh=body, children=[child], translation=[1, 2, 3], rotation=q, scale=[2, 2, 2])
    g.node(root=True, mesh=lamp, translation=[0, 4, 0])
    g.node(mesh=lamp)                                                           # not in the scene: no instance
    s = scenes.Scene()
    root = np.eye(4, dtype=f32)
    root[0, 3] = 10
    inst = L.load_gltf(s, str(g.write_gltf(tmp_path / "n.gltf")), transform=root.reshape(16))
    assert inst == [0, 1, 2]
    meshes = [m for m, _ in s.instances]
    assert meshes == [body, body, lamp]                                        # child before its parent
    Ry = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], np.float64)
    parent = np.eye(4)
    parent[:3, :3] = Ry * 2
    parent[:3, 3] = [11, 2, 3]
    kid = parent @ np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 5], [0, 0, 0, 1]], np.float64)
    for k, want in [(0, kid), (1, parent)]:
        assert np.allclose(s.instances[k][1].reshape(4, 4), want, atol=1e-6), k
    lamp_m = s.instances[2][1].reshape(4, 4)
    assert np.array_equal(lamp_m, [[1, 0, 0, 10], [0, 1, 0, 4], [0, 0, 1, 0], [0, 0, 0, 1]])
    # lights: the lamp's two triangles in world space, node id 1 + 2 (root node first)
    assert len(s.tri_lights) == 2
    _, lt = s.meshes[lamp]
    assert list(lt["ltriIdx"]) == [0, 1] and np.allclose(lt["area"], [0.5, 0.5])
    L0 = s.tri_lights[0]
    assert np.array_equal(L0["vertex0"], quad[0] + [10, 4, 0]) and np.array_equal(L0["vertex1"], quad[2] + [10, 4, 0])
    assert np.allclose(L0["N"], [lt["Nx"][0], lt["Ny"][0], lt["Nz"][0]]) and abs(L0["area"] - 0.5) < 1e-6
    assert list(s.tri_lights["triIdx"]) == [0, 1] and np.all(s.tri_lights["instIdx"] == 3)
    assert np.allclose(s.tri_lights["radiance"], [5, 4, 3])


def test_rotated_instance_lights_are_world_space(tmp_path):
    """Lights of a rotated, scaled OBJ instance (add_instance_with_lights):
    vertices and normal transformed, area from the world-space vertices."""
    from test_loaders import write_scene
    s = scenes.Scene()
    mesh = L.load_obj(s, str(write_scene(tmp_path)))
    M = np.array([[0, 0, 2, 1], [0, 2, 0, 0], [-2, 0, 0, 0], [0, 0, 0, 1]], f32)   # 90 degrees about y, scale 2
    L.add_instance_with_lights(s, mesh, M.reshape(16))
    _, t = s.meshes[mesh]
    for k in range(2):
        tri = t[2 + k]
        Lt = s.tri_lights[k]
        for c in range(3):
            v = np.asarray(tri[f"vertex{c}"], np.float64)
            assert np.allclose(Lt[f"vertex{c}"], M[:3, :3] @ v + M[:3, 3], atol=1e-6)
        n = M[:3, :3] @ [tri["Nx"], tri["Ny"], tri["Nz"]]
        assert np.allclose(Lt["N"], n / np.linalg.norm(n), atol=1e-6)
        assert abs(Lt["area"] - 4 * tri["area"]) < 1e-5                        # scale 2: area x 4


def test_unsupported_inputs_raise(tmp_path):
    g = GltfBuilder()
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32)
    p = g.primitive(tri, [0, 1, 5])                                            # index past the vertices
    g.mesh([p])
    g.node(root=True, mesh=0)
    with pytest.raises(ValueError):
        L.load_gltf(scenes.Scene(), str(g.write_glb(tmp_path / "bad.glb")))
    g2 = GltfBuilder()
    p = g2.primitive(tri, [0, 1, 2])
    g2.doc["accessors"][p["attributes"]["POSITION"]]["type"] = "VEC2"
    g2.mesh([p])
    with pytest.raises(ValueError):
        L.load_gltf(scenes.Scene(), str(g2.write_glb(tmp_path / "bad2.glb")))


def animated_file(tmp_path, interp, times, keys, path="translation", emissive=False):
    g = GltfBuilder()
    m = g.material(pbrMetallicRoughness={"baseColorFactor": [5, 5, 5, 1] if emissive else [0.5, 0.5, 0.5, 1]})
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32)
    g.mesh([g.primitive(tri, [0, 1, 2], material=m)])
    g.node(root=True, mesh=0)
    tin = g.accessor(np.asarray(times, f32))
    tout = g.accessor(np.asarray(keys, f32))
    g.doc["animations"] = [{"samplers": [{"input": tin, "output": tout, "interpolation": interp}],
                            "channels": [{"sampler": 0, "target": {"node": 0, "path": path}}]}]
    return g.write_glb(tmp_path / f"anim_{interp}_{path}.glb")


def translation(s):
    return s.instances[0][1].reshape(4, 4)[:3, 3]


def test_linear_and_step_translation_keys_wrap_in_time(tmp_path):
    keys = [[0, 0, 0], [2, 0, 0], [2, 4, 0]]
    s = scenes.Scene()
    g = L.load_gltf(s, str(animated_file(tmp_path, "LINEAR", [0, 1, 2], keys)), animated=True)
    assert np.array_equal(translation(s), [0, 0, 0])
    for dt, want in [(0.5, [1, 0, 0]), (1.0, [2, 2, 0]), (1.0, [1, 0, 0])]:   # t = 0.5, 1.5, 2.5 -> 0.5
        g.update(dt)
        assert np.allclose(translation(s), want, atol=1e-6), dt
    s2 = scenes.Scene()
    g2 = L.load_gltf(s2, str(animated_file(tmp_path, "STEP", [0, 1, 2], keys)), animated=True)
    g2.update(1.2)
    assert np.array_equal(translation(s2), [2, 0, 0])


def test_cubic_spline_translation_is_hermite(tmp_path):
    # per key: in-tangent, value, out-tangent; tangents scaled by the key interval
    keys = [[0, 0, 0], [0, 0, 0], [1, 0, 0], [0, 0, 0], [4, 0, 0], [0, 0, 0]]
    s = scenes.Scene()
    g = L.load_gltf(s, str(animated_file(tmp_path, "CUBICSPLINE", [0, 2], keys)), animated=True)
    g.update(1.0)    # f = 0.5: h00 p0 + h01 p1 + h10 m0 = 0.5 * 0 + 0.5 * 4 + 0.125 * (out-tangent 1 * interval 2)
    assert np.allclose(translation(s), [2.25, 0, 0], atol=1e-6)


def test_rotation_slerp_stays_a_rotation_about_the_key_axis(tmp_path):
    h = np.sqrt(0.5)
    keys = [[0, 0, 0, 1], [0, h, 0, h]]                                     # identity -> 90 degrees about y
    s = scenes.Scene()
    g = L.load_gltf(s, str(animated_file(tmp_path, "LINEAR", [0, 1], keys, path="rotation")), animated=True)
    angles = []
    for _ in range(3):
        g.update(0.25)
        R = s.instances[0][1].reshape(4, 4)[:3, :3].astype(np.float64)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-5) and np.allclose(R[1], [0, 1, 0], atol=1e-6)
        angles.append(np.arctan2(R[0, 2], R[0, 0]))
    assert 0 < angles[0] < angles[1] < angles[2] < np.pi / 2
    # near-parallel keys take the linear branch (cos > 0.99): a small rotation stays small
    q = L.quat_slerp([0, 0, 0, 1], [0, 0.05, 0, np.sqrt(1 - 0.0025)], 0.5)
    assert abs(q[1] - 0.025) < 1e-6
    # the reference's first weight is sin(1 - t): at t = 0.5 between keys 90 degrees apart in quaternion angle
    q = L.quat_slerp([0, 0, 0, 1], [0, 1, 0, 0], 0.5)
    assert np.allclose(q, [0, np.sin(np.pi / 4), 0, np.sin(0.5)], atol=1e-6)


def test_animated_lamp_moves_its_lights(tmp_path):
    keys = [[0, 0, 0], [0, 3, 0]]
    s = scenes.Scene()
    g = L.load_gltf(s, str(animated_file(tmp_path, "LINEAR", [0, 1], keys, emissive=True)), animated=True)
    assert np.array_equal(s.tri_lights["vertex0"][0], [0, 0, 0])
    g.update(0.5)
    assert np.allclose(s.tri_lights["vertex0"][0], [0, 1.5, 0]) and np.allclose(s.tri_lights["vertex2"][0], [0, 2.5, 0])
    assert abs(s.tri_lights["area"][0] - 0.5) < 1e-6


def test_skinned_strip_follows_a_step_rotation(tmp_path):
    """Two joints (node 1 at the origin, node 2 one unit up), inverse binds
    identity and translation (0, -1, 0); the top row weighted to joint 2.
    A STEP rotation of 90 degrees about z on joint 2 takes a top vertex
    (x, 2, 0) to T(0, 1, 0) R T(0, -1, 0) (x, 2, 0) = (-1, 1 + x, 0); the
    lower rows stay put; face normals follow."""
    g = GltfBuilder()
    m = g.material(pbrMetallicRoughness={"baseColorFactor": [0.5, 0.5, 0.5, 1]})
    P = np.float32([[-0.5, y, 0] for y in (0, 1, 2)] + [[0.5, y, 0] for y in (0, 1, 2)])
    idx = [0, 3, 4, 0, 4, 1, 1, 4, 5, 1, 5, 2]
    prim = g.primitive(P, idx, material=m)
    prim["attributes"]["JOINTS_0"] = g.accessor(np.array([[0, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0]] * 2, "u2"))
    prim["attributes"]["WEIGHTS_0"] = g.accessor(np.float32([[1, 0, 0, 0]] * 6))
    g.mesh([prim])
    ibm = np.stack([np.eye(4, dtype=f32), np.eye(4, dtype=f32)])
    ibm[1, 1, 3] = -1
    view = g.view(ibm.transpose(0, 2, 1).astype(f32).tobytes())          # column-major in the file
    g.doc["accessors"].append({"bufferView": view, "componentType": 5126, "count": 2, "type": "MAT4"})
    g.doc["skins"] = [{"joints": [1, 2], "inverseBindMatrices": len(g.doc["accessors"]) - 1}]
    g.node(root=True, mesh=0, skin=0)
    g.node(root=True, children=[2])
    g.node(translation=[0, 1, 0])
    h = np.sqrt(0.5)
    tin = g.accessor(np.float32([0, 1, 2]))
    rot = g.accessor(np.float32([[0, 0, 0, 1], [0, 0, h, h], [0, 0, h, h]]))
    g.doc["animations"] = [{"samplers": [{"input": tin, "output": rot, "interpolation": "STEP"}],
                            "channels": [{"sampler": 0, "target": {"node": 2, "path": "rotation"}}]}]
    s = scenes.Scene()
    anim = L.load_gltf(s, str(g.write_glb(tmp_path / "skin.glb")), animated=True)
    v, t = s.meshes[0]
    assert np.allclose(v[:, :3], P[idx], atol=1e-6)                      # bind pose
    anim.update(1.5)
    v, t = s.meshes[0]
    want = P[idx].copy()
    top = want[:, 1] == 2
    want[top] = np.stack([np.full(top.sum(), -1.0), 1 + want[top, 0], np.zeros(top.sum())], 1)
    assert np.allclose(v[:, :3], want, atol=1e-5)
    assert np.allclose(t["vertex2"], v[2::3, :3])
    N = np.stack([t["Nx"], t["Ny"], t["Nz"]], -1)
    assert np.allclose(np.linalg.norm(N, axis=1), 1, atol=1e-5)


def test_morph_target_weight_animation(tmp_path):
    """One triangle with a morph target moving every corner by (0, 0, 1) and
    a LINEAR weights channel 0 -> 1 over one second: at t = 0.5 the corners
    sit at z = 0.5; vertex normals are the normalised sum of base and target
    normals (added unweighted, as the reference adds them)."""
    g = GltfBuilder()
    m = g.material(pbrMetallicRoughness={"baseColorFactor": [0.5, 0.5, 0.5, 1]})
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32)
    prim = g.primitive(tri, [0, 1, 2], np.tile([[0, 0, 1]], (3, 1)).astype(f32), None, m)
    prim["targets"] = [{"POSITION": g.accessor(np.tile([[0, 0, 1]], (3, 1)).astype(f32)),
                        "NORMAL": g.accessor(np.tile([[1, 0, 0]], (3, 1)).astype(f32))}]
    g.doc["meshes"].append({"primitives": [prim], "weights": [0.0]})
    g.node(root=True, mesh=0)
    tin = g.accessor(np.float32([0, 1]))
    wout = g.accessor(np.float32([0, 1]))
    g.doc["animations"] = [{"samplers": [{"input": tin, "output": wout}], "channels": [{"sampler": 0, "target": {"node": 0, "path": "weights"}}]}]
    s = scenes.Scene()
    anim = L.load_gltf(s, str(g.write_glb(tmp_path / "morph.glb")), animated=True)
    assert np.all(s.meshes[0][0][:, 2] == 0)
    anim.update(0.5)
    v, t = s.meshes[0]
    assert np.allclose(v[:, 2], 0.5) and np.allclose(t["vertex1"], [1, 0, 0.5])
    assert np.allclose(t["vN0"], [np.sqrt(0.5), 0, np.sqrt(0.5)], atol=1e-6)


def test_morph_without_normal_attribute_stays_finite(tmp_path):
    """A morphed triangle whose primitive has no NORMAL attribute (and whose
    target has none either) keeps its face normal through a weights update:
    the base normals are the ones the load produced, never a zero vector."""
    g = GltfBuilder()
    m = g.material(pbrMetallicRoughness={"baseColorFactor": [0.5, 0.5, 0.5, 1]})
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32)
    prim = g.primitive(tri, [0, 1, 2], None, None, m)
    prim["targets"] = [{"POSITION": g.accessor(np.tile([[0, 0, 1]], (3, 1)).astype(f32))}]
    g.doc["meshes"].append({"primitives": [prim], "weights": [0.0]})
    g.node(root=True, mesh=0)
    tin = g.accessor(np.float32([0, 1]))
    wout = g.accessor(np.float32([0, 1]))
    g.doc["animations"] = [{"samplers": [{"input": tin, "output": wout}], "channels": [{"sampler": 0, "target": {"node": 0, "path": "weights"}}]}]
    s = scenes.Scene()
    anim = L.load_gltf(s, str(g.write_glb(tmp_path / "morph_nonormal.glb")), animated=True)
    n_load = s.meshes[0][1]["vN0"].copy()
    anim.update(0.5)
    v, t = s.meshes[0]
    assert np.allclose(v[:, 2], 0.5)
    for c in ("vN0", "vN1", "vN2"):
        assert np.all(np.isfinite(t[c]))
    assert np.allclose(t["vN0"], n_load) and np.allclose(np.abs(t["vN0"][:, 2]), 1)
