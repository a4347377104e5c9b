"""The reference's own scene files (apps/_shareddata) through the producers,
read in place when the reference tree is present (skipped otherwise): glTF
files with strided, multi-buffer accessors, 1-bit palette PNGs, a TGA
texture named with different case than on disk, an OBJ whose every triangle
is emissive. Checked: the triangle counts the files hold, the strip sample's
soup against an independent decode of its buffers, and the CoreTri
invariants (finite unit normals and tangents, alphas finite and >= 0,
material and texture indices in range). CPU only."""
import base64
import json
import os
import struct

import numpy as np
import pytest

from lighthouse2_amd import loaders as L
from lighthouse2_amd import scenes

SHARED = "/root/reference/apps/_shareddata"
pytestmark = pytest.mark.skipif(not os.path.isdir(SHARED), reason="reference scene files not present")

CASES = [  # file, triangles, lights (None: not checked)
    ("strip.gltf", 8, 0),
    ("AnimatedMorphCube.glb", 12, 0),
    ("AnimatedMorphSphere.glb", 960, 0),
    ("animatedCube/AnimatedCube.gltf", 12, 0),
    ("bird/scene.gltf", None, 0),
    ("crystal/scene.gltf", 192, 0),
    ("earth/scene.gltf", 960, 0),
    ("statue/scene.gltf", 618, 0),
    ("lh2text.obj", 292, 292),
    ("legocar.obj", 10992, 0),
]


def load(name):
    s = scenes.Scene()
    path = os.path.join(SHARED, name)
    if name.endswith(".obj"):
        L.add_instance_with_lights(s, L.load_obj(s, path))
    else:
        L.load_gltf(s, path)
    return s


@pytest.mark.parametrize("name,ntris,nlights", CASES)
def test_reference_scene_invariants(name, ntris, nlights):
    s = load(name)
    if ntris is not None:
        assert s.triangle_count == ntris
    if nlights is not None:
        assert len(s.tri_lights) == nlights
    unit = lambda v: np.abs(np.linalg.norm(v, axis=-1) - 1) < 1e-4
    for verts, t in s.meshes:
        if len(t) == 0:
            continue
        assert np.isfinite(verts).all() and np.all(verts[:, 3] == 1)
        N = np.stack([t["Nx"], t["Ny"], t["Nz"]], -1)
        ok = np.isfinite(N).all(1)                  # degenerate triangles keep NaN normals, as in the reference
        assert ok.mean() > 0.98 and unit(N[ok]).all()
        fin = ok & np.isfinite(t["T"]).all(1) & np.isfinite(t["B"]).all(1)
        if not name.endswith(".obj"):               # the glTF path replaces NaN tangents; the OBJ path keeps them
            assert np.array_equal(fin, ok)
        assert unit(t["T"][fin]).all() and unit(t["B"][fin]).all()
        a = t["alpha"][ok]
        assert (np.isfinite(a) & (a >= 0)).all()
        assert (t["material"] < len(s.materials)).all()
    for key in ("color", "normals", "roughness"):
        ids = s.materials[key]["textureID"]
        assert ((ids == -1) | ((ids >= 0) & (ids < len(s.textures)))).all()
    for tex in s.textures:
        assert len(tex["pixels"]) >= tex["width"] * tex["height"]


def test_strip_sample_soup_matches_its_buffers():
    """SimpleSkin's strip: 24 u16 indices, 10 float3 positions in the first
    data: buffer, posed by its skin at load (HostNode::Update poses skinned
    meshes every update): both joints sit at translation (0, 1, 0) and both
    inverse bind matrices are translations by (-0.5, -1, 0), so every joint
    matrix is a shift by -0.5 in x and, the weights summing to one, the soup
    is positions[indices] shifted by (-0.5, 0, 0)."""
    doc = json.load(open(os.path.join(SHARED, "strip.gltf")))
    buf = base64.b64decode(doc["buffers"][0]["uri"].split(",", 1)[1])
    idx = struct.unpack("<24H", buf[0:48])
    pos = np.array(struct.unpack("<30f", buf[48:168]), np.float32).reshape(10, 3)
    s = load("strip.gltf")
    verts, t = s.meshes[0]
    assert np.allclose(verts[:, :3], pos[list(idx)] + np.float32([-0.5, 0, 0]), atol=1e-6)
    assert len(s.instances) == 1 and np.array_equal(s.instances[0][1], np.eye(4, dtype=np.float32).reshape(16))


def test_legocar_texture_is_read_from_tga():
    s = load("legocar.obj")
    assert len(s.textures) == 1
    tex = s.textures[0]
    assert tex["width"] > 0 and tex["height"] > 0
    textured = s.materials["color"]["textureID"] == 0
    assert textured.sum() == 1 and np.all(s.materials["color"]["value"][textured] == 1)


def test_reference_animations_play():
    """AnimatedCube's rotation channel and the strip's joint rotations advance
    without disturbing unit rotations; instance transforms move."""
    s = scenes.Scene()
    g = L.load_gltf(s, os.path.join(SHARED, "animatedCube/AnimatedCube.gltf"), animated=True)
    T0 = s.instances[0][1].copy()
    for _ in range(5):
        g.update(0.37)
        R = s.instances[0][1].reshape(4, 4)[:3, :3].astype(np.float64)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-5)
    assert not np.array_equal(s.instances[0][1], T0)
    s2 = scenes.Scene()
    g2 = L.load_gltf(s2, os.path.join(SHARED, "strip.gltf"), animated=True)
    g2.update(0.5)
    assert np.array_equal(s2.instances[0][1], np.eye(4, dtype=np.float32).reshape(16))   # only joint nodes animate


def test_strip_sample_skin_follows_its_joint_animation():
    """SimpleSkin's joint 2 rotates about z: vertices weighted to it move,
    those weighted only to joint 1 keep their bind pose; the strip stays finite."""
    s = scenes.Scene()
    g = L.load_gltf(s, os.path.join(SHARED, "strip.gltf"), animated=True)
    v0 = s.meshes[0][0].copy()
    g.update(0.25)
    v1 = s.meshes[0][0]
    moved = np.abs(v1 - v0).max(1) > 1e-4
    assert np.isfinite(v1).all() and moved.any() and not moved.all()
    assert np.all(v0[~moved][:, 1] <= 1.0 + 1e-6)     # the unmoved vertices are the strip's lower half


def test_morph_cube_weights_animate():
    """AnimatedMorphCube: its weights channel moves the cube's vertices."""
    s = scenes.Scene()
    g = L.load_gltf(s, os.path.join(SHARED, "AnimatedMorphCube.glb"), animated=True)
    v0 = s.meshes[0][0].copy()
    g.update(0.6)
    v1 = s.meshes[0][0]
    assert np.isfinite(v1).all() and np.abs(v1 - v0).max() > 1e-3
