"""Bump maps (HostTexture::BumpToNormalMap, lib/RenderSystem/host_texture.cpp:320-343,
reached from an MTL bump map with an optional `bump_height` line,
host_material.cpp:54-61). CPU only.

Pinned by closed forms: a flat height field gives the +z normal (128, 128, 255);
a linear ramp gives the normal of its slope, computed here in float64 from the
ramp's gradient (central difference inside, one-sided at the borders); the
sign conventions follow the reference's definitions (x: left minus right
neighbour, y: the next stored row minus the previous one); only MIP level 0
is converted; every material gets a texture of its own."""
import numpy as np
import pytest

from lighthouse2_amd import formats as F
from lighthouse2_amd import loaders as L
from lighthouse2_amd import scenes
from test_loaders import encode_png


def height_texture(hgt):
    """A texture whose red byte is `hgt` (stored rows as given)."""
    hgt = np.asarray(hgt, np.uint8)
    rgba = np.zeros(hgt.shape + (4,), np.uint8)
    rgba[..., 0] = hgt
    rgba[..., 1] = 77                     # green, blue and alpha play no part
    rgba[..., 2] = 200
    rgba[..., 3] = 9
    return L.texture_from_rgba(rgba)


def level0(t):
    w, h = t["width"], t["height"]
    p = t["pixels"][:w * h].reshape(h, w)
    return np.stack([(p >> (8 * k)) & 255 for k in range(4)], -1).astype(np.int64)


def expected_bytes(gx, gy, s):
    """round((normalize(gx s, gy s, 1) * 0.5 + 0.5) * 255) in float64 (gx, gy: height differences / 255)."""
    n = np.stack([gx * s, gy * s, np.ones_like(gx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    v = (n * 0.5 + 0.5) * 255.0
    return np.floor(v + 0.5).astype(np.int64), v


def test_flat_height_gives_plus_z():
    t = height_texture(np.full((6, 9), 140))
    b = level0(L.bump_to_normal_map(t, 3.0))
    assert np.all(b == [128, 128, 255, 255])           # 127.5 rounds away from zero
    assert t["flags"] & L.TEX_NORMALMAP and t["storage"] == F.NRM32


@pytest.mark.parametrize("s", [0.7, 1.3, 4.1])        # scales that keep the ramp values off .5 ties
def test_ramps_give_the_slope_normal(s):
    h, w = 7, 11
    y, x = np.mgrid[0:h, 0:w]
    hx = 5 * x + 2 * y                                   # rises to the right and along the stored rows
    t = height_texture(hx)
    got = level0(L.bump_to_normal_map(t, s))
    # reference convention: nx = left - right, ny = next row - previous row, one-sided at the borders
    gx = np.where((x > 0) & (x < w - 1), -10.0, -5.0) / 255.0
    gy = np.where((y > 0) & (y < h - 1), 4.0, 2.0) / 255.0
    exp, v = expected_bytes(gx, gy, s)
    safe = np.abs(v - np.floor(v) - 0.5) > 1e-3          # float32 and float64 agree away from .5 ties
    assert safe.mean() > 0.9
    assert np.array_equal(got[..., :3][safe], exp[safe])
    assert np.all(got[..., 3] == 255)
    # the normal leans against the slope: left in x, back along the stored rows (strictly inside, where
    # the central difference is twice the one-sided one at the borders)
    assert np.all(got[..., 0] < 128) and np.all(got[1:-1, :, 1] > 128) and np.all(got[..., 1] >= 128)


def test_only_level_zero_is_converted():
    rng = np.random.default_rng(3)
    t = height_texture(rng.integers(0, 256, (16, 16)))
    before = t["pixels"].copy()
    L.bump_to_normal_map(t, 2.0)
    n0 = 16 * 16
    assert not np.array_equal(t["pixels"][:n0], before[:n0])
    assert np.array_equal(t["pixels"][n0:], before[n0:])


def test_mtl_bump_map_with_height_scale(tmp_path):
    rng = np.random.default_rng(5)
    img = np.zeros((8, 12, 4), np.uint8)
    img[..., 0] = rng.integers(0, 256, (8, 12))
    img[..., 3] = 255
    (tmp_path / "h.png").write_bytes(encode_png(img, 6, [0, 1, 2, 3, 4]))
    (tmp_path / "q.mtl").write_text("newmtl a\nKd 0.5 0.5 0.5\nmap_bump -bm 9 h.png\nbump_height 2.5\n"
                                    "newmtl b\nKd 0.5 0.5 0.5\nbump h.png\n")
    (tmp_path / "q.obj").write_text("mtllib q.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                                    "usemtl a\nf 1/1 2/2 3/3\nusemtl b\nf 1/1 3/3 4/4\n")
    s = scenes.Scene()
    L.load_obj(s, str(tmp_path / "q.obj"))
    ids = [int(m["normals"]["textureID"]) for m in s.materials]
    assert len(ids) == 2 and ids[0] != ids[1]           # one texture per material
    for tid, scale in zip(ids, (2.5, 1.0)):             # -bm is not the height scale; bump_height is
        ref = L.bump_to_normal_map(L.load_texture(str(tmp_path / "h.png"), flipped=True), scale)
        assert np.array_equal(s.textures[tid]["pixels"], ref["pixels"])
        assert s.textures[tid]["storage"] == F.NRM32
