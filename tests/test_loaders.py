"""Scene-file producers (lighthouse2_amd/loaders.py): the PNG decoder against
an encoder written here with every filter type and colour type; OBJ + MTL
against closed forms of HostMesh::LoadGeometryFromOBJ (host_mesh.cpp:131-318):
polygon fans, face normals flipped towards the vertex normals, the
consistent-normal alpha acos(max(0.7, n.N)) (1 + 0.03632 (1 - n.N)^2), uv
tangents, the triangle LOD, material conversion (Kd, Kt, Ni, textures
FLIPPED + LINEARIZED) and emissive triangles becoming lights. CPU only."""
import struct
import zlib

import numpy as np
import pytest

from lighthouse2_amd import formats as F
from lighthouse2_amd import loaders as L
from lighthouse2_amd import scenes
import oracle as orc

f32 = np.float32


def encode_png(img, ctype, filters):
    """Test-side PNG encoder: img (h, w, c) uint8; filters: filter type per row."""
    h, w, c = img.shape
    rows = []
    prev = np.zeros(w * c, np.int32)
    for y in range(h):
        x = img[y].reshape(-1).astype(np.int32)
        a = np.concatenate([np.zeros(c, np.int32), x[:-c]])
        cc = np.concatenate([np.zeros(c, np.int32), prev[:-c]])
        b = prev
        f = filters[y % len(filters)]
        if f == 0:
            p = np.zeros_like(x)
        elif f == 1:
            p = a
        elif f == 2:
            p = b
        elif f == 3:
            p = (a + b) >> 1
        else:
            pa, pb, pc = np.abs(b - cc), np.abs(a - cc), np.abs(a + b - 2 * cc)
            p = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, cc))
        rows.append(bytes([f]) + ((x - p) & 255).astype(np.uint8).tobytes())
        prev = x
    chunk = lambda k, d: struct.pack(">I", len(d)) + k + d + struct.pack(">I", zlib.crc32(k + d) & 0xffffffff)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(b"".join(rows))) + chunk(b"IEND", b"")


@pytest.mark.parametrize("ctype,chans", [(0, 1), (2, 3), (4, 2), (6, 4)])
def test_png_decoder_all_filters(ctype, chans):
    rng = np.random.default_rng(ctype)
    img = rng.integers(0, 256, (13, 11, chans), dtype=np.uint8)
    rgba = L.decode_png(encode_png(img, ctype, [0, 1, 2, 3, 4]))
    want = np.full((13, 11, 4), 255, np.uint8)
    if chans in (1, 2):
        want[..., :3] = img[..., :1]
        if chans == 2:
            want[..., 3] = img[..., 1]
    else:
        want[..., :chans] = img
    assert np.array_equal(rgba, want)


def test_texture_flip_and_linearize():
    img = np.zeros((2, 4, 4), np.uint8)
    img[0, :, 0] = 200                                  # top row red 200
    img[1, :, 1] = 100                                  # bottom row green 100
    img[..., 3] = 255
    t = L.texture_from_rgba(img, flipped=True, linearized=True)
    px = t["pixels"][:8].reshape(2, 4)
    assert np.all((px[0] >> 8) & 255 == (100 * 100) >> 8)   # FLIPPED: the bottom row comes first
    assert np.all(px[1] & 255 == (200 * 200) >> 8)           # byte 0 is red (make_uchar4(r, g, b, a))
    assert np.all(px >> 24 == 255)
    t2 = L.texture_from_rgba(img)
    assert np.all(t2["pixels"][:4] & 255 == 200) and t2["storage"] == F.ARGB32 and t2["flags"] == L.TEX_LDR
    t3 = L.texture_from_rgba(img, normal_map=True)
    assert t3["storage"] == F.NRM32 and t3["flags"] == L.TEX_LDR | L.TEX_NORMALMAP


OBJ = """mtllib box.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
vn 0.6 0 0.8
vn 0 0 -1
usemtl tex
f 1/1/1 2/2/2 3/3/1 4/4/1
usemtl glow
f -5//3 -3//3 -4//3
f 1 2 5
"""
MTL = """newmtl tex
Kd 0.3 0.4 0.5
map_Kd checker.png
newmtl glow
Kd 4 3 2
Kt 0.1 0.2 0.3
Ni 1.4
"""


def write_scene(tmp_path):
    (tmp_path / "box.obj").write_text(OBJ)
    (tmp_path / "box.mtl").write_text(MTL)
    y, x = np.mgrid[0:8, 0:8]
    img = np.zeros((8, 8, 4), np.uint8)
    img[..., 0] = np.where((x + y) % 2, 220, 40)
    img[..., 1] = 120
    img[..., 2] = 20
    img[..., 3] = 255
    (tmp_path / "checker.png").write_bytes(encode_png(img, 6, [4, 1, 0, 2, 3]))
    return tmp_path / "box.obj"


def test_obj_closed_forms(tmp_path):
    s = scenes.Scene()
    mesh = L.load_obj(s, str(write_scene(tmp_path)))
    verts, tris = s.meshes[mesh]
    assert len(tris) == 4 and verts.shape == (12, 4)
    # fan of the quad: (1,2,3), (1,3,4)
    assert np.array_equal(verts[0:3, :3], [[0, 0, 0], [1, 0, 0], [1, 1, 0]])
    assert np.array_equal(verts[3:6, :3], [[0, 0, 0], [1, 1, 0], [0, 1, 0]])
    # face normal +z; vertex normals as given
    assert np.array_equal(np.stack([tris["Nx"], tris["Ny"], tris["Nz"]], 1)[:2], [[0, 0, 1], [0, 0, 1]])
    assert np.allclose(tris["vN1"][0], [0.6, 0, 0.8])
    # alpha of normal 2 (dot 0.8): acos(0.8) (1 + 0.03632 * 0.2^2); normal 1 (dot 1): 0
    a2 = np.float32(np.arccos(np.float32(0.8)) * (1 + 0.03632 * 0.2 * 0.2))
    assert abs(tris["alpha"][0][1] - a2) < 1e-6 and tris["alpha"][0][0] == 0
    # uv tangents: u along x, v along y
    assert np.allclose(tris["T"][:2], [[1, 0, 0], [1, 0, 0]]) and np.allclose(tris["B"][:2], [[0, 1, 0], [0, 1, 0]])
    # triangle 3: negative indices, vertex normals -z against the +z face: the face normal flips
    assert np.array_equal(verts[6:9, :3], [[0, 0, 0], [1, 1, 0], [1, 0, 0]])    # -5, -3, -4 -> vertices 1, 3, 2
    N3 = np.array([tris["Nx"][2], tris["Ny"][2], tris["Nz"][2]])
    assert np.allclose(N3, [0, 0, -1]) and np.allclose(tris["vN0"][2], [0, 0, -1])
    # triangle 4: no normals, no uv -> face normal everywhere, edge tangent, alpha 0
    assert np.allclose(tris["vN0"][3], [tris["Nx"][3], tris["Ny"][3], tris["Nz"][3]]) and np.all(tris["alpha"][3] == 0)
    assert np.allclose(tris["T"][3], [1, 0, 0])
    # materials: textured (colour 1, texture id, LOD), emissive with Kt and Ni
    m_tex, m_glow = tris["material"][0], tris["material"][2]
    assert s.materials[m_tex]["color"]["textureID"] == 0 and np.all(s.materials[m_tex]["color"]["value"] == 1)
    assert np.allclose(s.materials[m_glow]["color"]["value"], [4, 3, 2])
    assert np.allclose(s.materials[m_glow]["absorption"]["value"], [0.1, 0.2, 0.3]) and abs(s.materials[m_glow]["ior"]["value"] - 1.4) < 1e-6
    assert np.allclose(tris["LOD"][:2], scenes.triangle_lod(tris[:2], 8, 8)) and np.all(tris["LOD"][2:] == 0)
    # texture: FLIPPED + LINEARIZED
    tex = s.textures[0]
    px = tex["pixels"][:64].reshape(8, 8)
    # row 0 is the image's bottom row: pixel (x 0, y 7) is 220 red; row 7 the top row (40 red)
    assert (px[0, 0] & 255) == (220 * 220) >> 8 and (px[7, 0] & 255) == (40 * 40) >> 8
    assert ((px[0, 0] >> 8) & 255) == (120 * 120) >> 8
    # lights: the two glow triangles, areas by Heron, ltriIdx in list order
    inst = L.add_instance_with_lights(s, mesh)
    assert len(s.tri_lights) == 2
    _, tris = s.meshes[mesh]
    assert list(tris["ltriIdx"]) == [-1, -1, 0, 1]
    assert np.allclose(tris["area"][2:], [0.5, 0.5], atol=1e-6)
    assert list(s.tri_lights["triIdx"]) == [2, 3] and np.all(s.tri_lights["instIdx"] == inst)
    assert np.allclose(s.tri_lights["radiance"], [4, 3, 2])


def test_loaded_scene_renders_in_the_oracle(tmp_path):
    """The loaded triangles go through the boundary arrays into the oracle's
    renderer: an emissive triangle lights the textured quad."""
    s = scenes.Scene()
    mesh = L.load_obj(s, str(write_scene(tmp_path)))
    L.add_instance_with_lights(s, mesh)
    view = scenes.camera_view((0.5, 0.5, 2.0), (0.5, 0.5, 0.0), 60.0, 16, 16)
    r32, r64, n, _, _ = orc.render7(s, view, 16, 16, 1, 1, threads=4, bsdf=0)
    assert np.isfinite(r64).all() and r64[..., :3].sum() > 0


def png_raw(w, h, depth, ctype, rows, plte=None, trns=None):
    """PNG of already packed rows (filter type 0 each)."""
    chunk = lambda k, d: struct.pack(">I", len(d)) + k + d + struct.pack(">I", zlib.crc32(k + d) & 0xffffffff)
    body = b"".join(b"\x00" + r for r in rows)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if plte is not None:
        out += chunk(b"PLTE", bytes(plte))
    if trns is not None:
        out += chunk(b"tRNS", bytes(trns))
    return out + chunk(b"IDAT", zlib.compress(body)) + chunk(b"IEND", b"")


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_png_low_bit_depths(depth):
    rng = np.random.default_rng(depth)
    w, h = 13, 5
    v = rng.integers(0, 1 << depth, (h, w))
    per = 8 // depth

    def pack(row):
        padded = np.concatenate([row, np.zeros((-len(row)) % per, np.int64)]).reshape(-1, per)
        return bytes(int(sum(int(x) << (8 - depth * (k + 1)) for k, x in enumerate(g))) for g in padded)
    rows = [pack(v[y]) for y in range(h)]
    grey = L.decode_png(png_raw(w, h, depth, 0, rows))
    scale = {1: 255, 2: 85, 4: 17}[depth]
    assert np.array_equal(grey[..., 0], v * scale) and np.all(grey[..., 3] == 255)
    plte = rng.integers(0, 256, (1 << depth, 3)).astype(np.uint8)
    trns = [7, 200]
    pal = L.decode_png(png_raw(w, h, depth, 3, rows, plte.tobytes(), trns))
    assert np.array_equal(pal[..., :3], plte[v])
    assert np.array_equal(pal[..., 3], np.where(v == 0, 7, np.where(v == 1, 200, 255)))


def test_png_16_bit_keeps_high_byte_and_trns_key():
    rng = np.random.default_rng(16)
    w, h = 6, 4
    v = rng.integers(0, 65536, (h, w, 3)).astype(">u2")
    rgb = L.decode_png(png_raw(w, h, 16, 2, [v[y].tobytes() for y in range(h)]))
    assert np.array_equal(rgb[..., :3], (v.astype(np.int64) >> 8).astype(np.uint8))
    img = rng.integers(0, 256, (h, w, 3)).astype(np.uint8)
    img[1, 2] = (10, 20, 30)
    keyed = L.decode_png(png_raw(w, h, 8, 2, [img[y].tobytes() for y in range(h)], trns=[0, 10, 0, 20, 0, 30]))
    assert keyed[1, 2, 3] == 0 and (keyed[..., 3] == 0).sum() == (img == (10, 20, 30)).all(-1).sum()


def tga(img, rle=False, top_left=False):
    """Test-side TGA encoder: (h, w, c) uint8 RGB(A) or grey (c = 1)."""
    h, w, c = img.shape
    px = np.concatenate([img[..., 2::-1], img[..., 3:]], -1) if c >= 3 else img     # BGR(A)
    rows = px if top_left else px[::-1]
    flat = rows.reshape(-1, c)
    if rle:
        data, i = bytearray(), 0
        while i < len(flat):
            j = i
            while j + 1 < len(flat) and j - i < 127 and np.array_equal(flat[j + 1], flat[i]):
                j += 1
            if j > i:
                data += bytes([128 | (j - i)]) + flat[i].tobytes()
            else:
                data += bytes([0]) + flat[i].tobytes()
            i = j + 1
    else:
        data = flat.tobytes()
    itype = (3 if c == 1 else 2) + (8 if rle else 0)
    head = struct.pack("<BBBHHBHHHHBB", 0, 0, itype, 0, 0, 0, 0, 0, w, h, 8 * c, (0x20 if top_left else 0) | (8 if c == 4 else 0))
    return head + bytes(data)


@pytest.mark.parametrize("c,rle,top", [(3, False, False), (4, True, True), (1, True, False), (4, False, True)])
def test_tga_decoder(c, rle, top):
    rng = np.random.default_rng(c + 2 * rle)
    img = rng.integers(0, 4, (7, 9, c)).astype(np.uint8) * 60          # few values: RLE runs occur
    rgba = L.decode_tga(tga(img, rle, top))
    want = np.full((7, 9, 4), 255, np.uint8)
    if c == 1:
        want[..., :3] = img
    else:
        want[..., :c] = img
    assert np.array_equal(rgba, want)
