"""Scene files -> the boundary's arrays (CoreTri soup, CoreMaterial, textures,
CoreLightTri), the producers the reference's RenderSystem runs before it
calls a core:

* Wavefront OBJ + MTL: HostMesh::LoadGeometryFromOBJ (lib/RenderSystem/
  host_mesh.cpp:131-321) with HostMaterial::ConvertFrom (host_material.cpp:35-71);
* PNG textures: HostTexture::Load (host_texture.cpp:200-241) for 8-bit images,
  FLIPPED / LINEARIZED as the OBJ path requests them, MIP chain of
  ConstructMIPmaps;
* glTF 2.0 (.glb, .gltf with .bin files or data: URIs): HostScene::AddScene
  (host_scene.cpp:249-392) with HostMesh::ConvertFromGTLFMesh /
  BuildFromIndexedData (host_mesh.cpp:327-681), HostMaterial::ConvertFrom
  (host_material.cpp:76-170) and the HostNode hierarchy (host_node.cpp:77-199);
* lights of emissive triangles: HostNode::PrepareLights / UpdateLights
  (host_node.cpp:203-255) in world space, a material being emissive when a
  colour component exceeds 1 (HostMaterial::IsEmissive, host_material.h:82).

The file parsers are this project's (the reference uses tinyobjloader,
tinygltf and FreeImage): polygons are split as a fan, which is what tinyobjloader's ear
clipping gives for convex polygons; PNG is decoded with zlib (8-bit grey,
grey + alpha, RGB, RGBA and palette images, not interlaced); baseline and
extended-sequential JPEG by this project's decoder (jpeg.py), which follows
libjpeg's default arithmetic, the library behind FreeImage's JPEG plugin. All float
arithmetic of the producers is float32 in the reference's operation order.
"""
import os
import re
import struct
import zlib

import numpy as np

from . import formats as F
from . import scenes
from .jpeg import decode_jpeg

f32 = np.float32


# ---- PNG / TGA -------------------------------------------------------------------
def _unfilter(rows, bpp):
    """PNG row filters (None, Sub, Up, Average, Paeth) over bytes; bpp is the
    filter unit in bytes (at least 1)."""
    h, stride = rows.shape[0], rows.shape[1] - 1
    out = np.zeros((h, stride), np.uint8)
    prev = [0] * stride
    for y in range(h):
        f = int(rows[y, 0])
        line = rows[y, 1:]
        if f == 0:
            cur = line.tolist()
        elif f == 1:                                # Sub: a running sum per byte lane, mod 256
            pad = (-stride) % bpp
            lanes = np.concatenate([line, np.zeros(pad, np.uint8)]).reshape(-1, bpp).astype(np.int64)
            cur = (np.cumsum(lanes, 0) & 255).reshape(-1)[:stride].tolist()
        elif f == 2:
            cur = ((line.astype(np.int32) + np.asarray(prev, np.int32)) & 255).tolist()
        elif f in (3, 4):
            ln = line.tolist()
            cur = [0] * stride
            for x in range(stride):                 # depends on the reconstructed left neighbour
                a = cur[x - bpp] if x >= bpp else 0
                b = prev[x]
                if f == 3:
                    p = (a + b) >> 1
                else:
                    c = prev[x - bpp] if x >= bpp else 0
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[x] = (ln[x] + p) & 255
        else:
            raise ValueError(f"PNG filter type {f}")
        out[y] = cur
        prev = cur
    return out


def decode_png(data):
    """PNG bytes -> (h, w, 4) uint8 RGBA, rows top to bottom. Grey, grey +
    alpha, RGB, RGBA and palette images at every bit depth, not interlaced:
    16-bit samples keep their high byte, grey below 8 bits is scaled to
    0..255 and palette indices below 8 bits index the palette (what 8-bit
    loads of stb_image and FreeImage return); a tRNS chunk gives palette
    alphas or, for 8-bit grey and RGB, alpha 0 on the key colour."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, plte, trns = 8, [], None, None
    w = h = depth = ctype = interlace = None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif kind == b"tRNS":
            trns = np.frombuffer(body, np.uint8)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if w is None:
        raise ValueError("PNG without IHDR")
    if interlace != 0:
        raise ValueError("interlaced PNG images are not read")
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bits = depth * chans
    stride = (w * bits + 7) // 8
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    out = _unfilter(raw[:h * (stride + 1)].reshape(h, stride + 1), max(1, bits // 8))
    if depth == 8:
        px = out.reshape(h, w, chans)
    elif depth == 16:
        px = out.reshape(h, w, chans, 2)[..., 0]
    else:
        per = 8 // depth
        shifts = np.array([8 - depth * (k + 1) for k in range(per)], np.uint8)
        v = ((out[:, :, None] >> shifts) & ((1 << depth) - 1)).reshape(h, -1)[:, :w]
        if ctype == 0:
            v = v * {1: 255, 2: 85, 4: 17}[depth]
        px = v.astype(np.uint8)[..., None]
    rgba = np.full((h, w, 4), 255, np.uint8)
    if ctype == 0:
        rgba[..., :3] = px
        if trns is not None and depth == 8 and len(trns) >= 2:
            rgba[..., 3] = np.where(px[..., 0] == trns[1], 0, 255)
    elif ctype == 4:
        rgba[..., :3], rgba[..., 3] = px[..., :1], px[..., 1]
    elif ctype == 2:
        rgba[..., :3] = px
        if trns is not None and depth == 8 and len(trns) >= 6:
            key = trns[1::2][:3]
            rgba[..., 3] = np.where((px == key).all(-1), 0, 255)
    elif ctype == 6:
        rgba[:] = px
    else:
        if plte is None:
            raise ValueError("palette PNG without PLTE")
        idx = px[..., 0]
        if idx.max(initial=0) >= len(plte):
            raise ValueError("PNG palette index out of range")
        rgba[..., :3] = plte[idx]
        if trns is not None:
            alpha = np.full(256, 255, np.uint8)
            alpha[:len(trns)] = trns
            rgba[..., 3] = alpha[idx]
    return rgba


def decode_tga(data):
    """TGA bytes -> (h, w, 4) uint8 RGBA, rows top to bottom: true-colour
    (24/32-bit BGR(A)) and grey (8-bit) images, uncompressed or run-length
    encoded, either origin."""
    if len(data) < 18:
        raise ValueError("not a TGA file")
    idlen, cmtype, itype = data[0], data[1], data[2]
    cmlen, cmdepth = struct.unpack("<H", data[5:7])[0], data[7]
    w, h = struct.unpack("<HH", data[12:16])
    depth, desc = data[16], data[17]
    if itype not in (2, 3, 10, 11) or (itype in (2, 10) and depth not in (24, 32)) or (itype in (3, 11) and depth != 8):
        raise ValueError(f"TGA image type {itype} / {depth}-bit: only true-colour 24/32-bit and 8-bit grey are read")
    bpp = depth // 8
    pos = 18 + idlen + (cmlen * ((cmdepth + 7) // 8) if cmtype else 0)
    n = w * h * bpp
    if itype in (2, 3):
        pix = np.frombuffer(data[pos:pos + n], np.uint8)
    else:
        buf = bytearray()
        while len(buf) < n:
            head = data[pos]
            pos += 1
            count = (head & 127) + 1
            if head & 128:
                buf += data[pos:pos + bpp] * count
                pos += bpp
            else:
                buf += data[pos:pos + count * bpp]
                pos += count * bpp
        pix = np.frombuffer(bytes(buf[:n]), np.uint8)
    if len(pix) < n:
        raise ValueError("TGA pixel data is truncated")
    pix = pix.reshape(h, w, bpp)
    rgba = np.full((h, w, 4), 255, np.uint8)
    if bpp == 1:
        rgba[..., :3] = pix
    else:
        rgba[..., 0], rgba[..., 1], rgba[..., 2] = pix[..., 2], pix[..., 1], pix[..., 0]
        if bpp == 4:
            rgba[..., 3] = pix[..., 3]
    if not desc & 0x20:                             # bottom-left origin: first row stored is the bottom one
        rgba = rgba[::-1]
    if desc & 0x10:
        rgba = rgba[:, ::-1]
    return np.ascontiguousarray(rgba)


TEX_NORMALMAP, TEX_LDR = 2, 4                         # HostTexture flags (host_texture.h:39-44)


def texture_from_rgba(rgba, flipped=False, linearized=False, normal_map=False):
    """HostTexture::Load for an 8-bit image (host_texture.cpp:225-241): each
    texel is the bytes r, g, b, a in that order (make_uchar4 of the red,
    green, blue and alpha channels; the core reads byte k as component k,
    sampling_shared.h:29-33). The image library delivers rows bottom-up, so
    FLIPPED keeps the bottom row first and no flag stores the top row first;
    LINEARIZED squares r, g, b ((c * c) >> 8, sRGBtoLinear :73-81); 5 MIP
    levels (ConstructMIPmaps); flags LDR (| NORMALMAP, storage NRM32, for a
    normal map: ConvertToCoreTexDesc :43-67)."""
    a = np.asarray(rgba, np.uint8)
    h, w = a.shape[:2]
    rows = a[::-1] if flipped else a
    c = rows.astype(np.uint32)
    r, g, b, al = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    if linearized:
        r, g, b = (r * r) >> 8, (g * g) >> 8, (b * b) >> 8
    px = (al << 24) | (b << 16) | (g << 8) | r
    t = {"width": w, "height": h, "flags": TEX_LDR, "storage": F.ARGB32, "pixels": scenes.mip_chain(px)}
    if normal_map:
        mark_normal_map(t)
    return t


def bump_to_normal_map(tex, height_scale=1.0):
    """HostTexture::BumpToNormalMap (host_texture.cpp:320-343): the red byte
    of level 0 is a height h = r * (1/255); per texel n = normalize((h(x-1) -
    h(x+1)) * s, (h(y+1) - h(y-1)) * s, 1) with the texel itself standing in
    for a missing neighbour at the borders (rows in stored order), in float
    arithmetic (normalize multiplies by 1 / sqrtf of the dot product); the
    texel becomes round((n * 0.5 + 0.5) * 255) per component (in double,
    halves away from zero) with alpha 255. Only level 0 is rewritten: the
    MIP levels keep what the loader built from the height image, as in the
    reference, which converts after loading."""
    w, h = tex["width"], tex["height"]
    px = tex["pixels"]
    f32 = np.float32
    step = f32(1.0) / f32(255.0)
    z = (px[:w * h] & 255).reshape(h, w).astype(f32) * step
    x_prev = np.concatenate([z[:, :1], z[:, :-1]], axis=1)
    x_next = np.concatenate([z[:, 1:], z[:, -1:]], axis=1)
    y_prev = np.concatenate([z[1:], z[-1:]], axis=0)
    y_next = np.concatenate([z[:1], z[:-1]], axis=0)
    s = f32(height_scale)
    nx = (x_prev - x_next) * s
    ny = (y_prev - y_next) * s
    inv = f32(1.0) / np.sqrt((nx * nx + ny * ny) + f32(1.0))
    out = []
    for c in (nx * inv, ny * inv, inv):
        v = (c.astype(np.float64) * 0.5 + 0.5) * 255.0
        out.append(np.floor(v + 0.5).astype(np.uint32))        # round(): v >= 0 and v + 0.5 is exact
    px[:w * h] = (out[0] | (out[1] << 8) | (out[2] << 16) | np.uint32(255 << 24)).reshape(-1)
    mark_normal_map(tex)
    return tex


def mark_normal_map(tex):
    """HostMaterial sets NORMALMAP on a texture it uses as a normal map; the
    core then receives it as NRM32 texels (host_texture.cpp:61)."""
    tex["flags"] |= TEX_NORMALMAP
    tex["storage"] = F.NRM32


def decode_image(data, name=""):
    """PNG, JPEG or TGA bytes (by signature, then by file name) -> RGBA, top row first."""
    if data[:8] == b"\x89PNG\r\n\x1a\n":
        return decode_png(data)
    if data[:3] == b"\xff\xd8\xff":
        return decode_jpeg(data)
    if name.lower().endswith(".tga"):
        return decode_tga(data)
    raise ValueError(f"{name or 'image'}: only PNG, JPEG and TGA images are read")


def resolve_path(path):
    """The file at `path`, with backslashes read as separators and, when the
    exact name is absent, a case-insensitive match per path component (scene
    files written on Windows name their textures that way)."""
    path = path.replace("\\", "/")
    if os.path.exists(path):
        return path
    parts = os.path.abspath(path).split("/")
    cur = "/"
    for part in parts[1:]:
        nxt = os.path.join(cur, part)
        if not os.path.exists(nxt) and os.path.isdir(cur):
            match = [e for e in os.listdir(cur) if e.lower() == part.lower()]
            if match:
                nxt = os.path.join(cur, match[0])
        cur = nxt
    return cur


def load_texture(path, flipped=False, linearized=False, normal_map=False):
    path = resolve_path(path)
    with open(path, "rb") as fh:
        return texture_from_rgba(decode_image(fh.read(), path), flipped, linearized, normal_map)


# ---- skydome (Radiance HDR) ---------------------------------------------------


def decode_hdr(data):
    """Radiance RGBE image (the .hdr skydomes HostSkyDome::Load reads through
    FreeImage, host_skydome.cpp:103-133) -> (h, w, 3) float32, top row first.
    Scanlines are flat or new-style run-length (per channel: a count above
    128 repeats the next byte count - 128 times, otherwise count bytes
    follow); a pixel (r, g, b, e) is (r, g, b) * 2^(e - 136), zero for e = 0."""
    data = bytes(data)
    if not (data.startswith(b"#?RADIANCE") or data.startswith(b"#?RGBE")):
        raise ValueError("not a Radiance HDR file")
    pos = data.index(b"\n\n") + 2
    if b"FORMAT=" in data[:pos] and b"32-bit_rle_rgbe" not in data[:pos]:
        raise ValueError("Radiance HDR: only the RGBE pixel format is read")
    eol = data.index(b"\n", pos)
    res = data[pos:eol].split()
    if len(res) != 4 or res[0] != b"-Y" or res[2] != b"+X":
        raise ValueError("Radiance HDR: only the standard -Y H +X W orientation is read")
    h, w = int(res[1]), int(res[3])
    pos = eol + 1
    rgbe = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and data[pos] == 2 and data[pos + 1] == 2 and ((data[pos + 2] << 8) | data[pos + 3]) == w and not data[pos + 2] & 0x80:
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = data[pos]
                    if n > 128:
                        n -= 128
                        rgbe[y, x:x + n, c] = data[pos + 1]
                        pos += 2
                    else:
                        if n == 0:
                            raise ValueError("Radiance HDR: empty run")
                        rgbe[y, x:x + n, c] = np.frombuffer(data, np.uint8, n, pos + 1)
                        pos += 1 + n
                    x += n
                if x != w:
                    raise ValueError("Radiance HDR: run-length scanline overruns the width")
        else:
            rgbe[y] = np.frombuffer(data, np.uint8, 4 * w, pos).reshape(w, 4)
            pos += 4 * w
    e = rgbe[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136), 0).astype(f32)
    return (rgbe[..., :3].astype(f32) * f[..., None]).astype(f32)


def load_skydome(scene, path, scale=(1.0, 1.0, 1.0), transform=None):
    """HostSkyDome::Load (host_skydome.cpp:80-139): a .hdr file or the
    reference's cache of it (.bin: uint32 width, height, then float3 pixels,
    top row first), multiplied by scale; becomes the scene's sky
    (SetSkyData's pixels, width, height and world-to-sky transform,
    identity by default as HostSkyDome's worldToLight)."""
    path = resolve_path(path)
    with open(path, "rb") as fh:
        data = fh.read()
    if path.lower().endswith(".bin"):
        w, h = struct.unpack("<II", data[:8])
        px = np.frombuffer(data, f32, w * h * 3, 8).reshape(h, w, 3).copy()
    else:
        px = decode_hdr(data)
    sc = np.asarray(scale, f32)
    if np.any(sc != 1):
        px = (px * sc).astype(f32)
    scene.sky = (px, (np.eye(4, dtype=f32) if transform is None else np.asarray(transform, f32)).reshape(16))
    return px


# ---- OBJ / MTL ----------------------------------------------------------------
def _atof(text):
    """C atof: the longest leading decimal number, 0.0 if there is none."""
    m = re.match(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", text)
    return float(m.group(1)) if m else 0.0


def parse_mtl(text):
    """MTL text -> list of dicts (name, Kd, Kt, Ni, map_Kd, map_bump/norm, map_Ks, shading)."""
    mats, cur = [], None
    for line in text.splitlines():
        t = line.split("#", 1)[0].split()
        if not t:
            continue
        k, args = t[0], t[1:]
        if k == "newmtl":
            cur = {"name": " ".join(args), "Kd": (0.0, 0.0, 0.0), "Kt": (0.0, 0.0, 0.0), "Ni": 1.0}
            mats.append(cur)
        elif cur is None:
            continue
        elif k in ("Kd", "Kt", "Tf"):
            v = [float(x) for x in args[:3]]
            cur["Kt" if k == "Tf" else k] = tuple(v + [v[0]] * (3 - len(v)))
        elif k == "Ni":
            cur["Ni"] = float(args[0])
        elif k in ("map_Kd", "map_Ks", "norm", "map_bump", "bump"):
            name = args[-1]                           # options (-bm 1 ...) come before the file name
            cur["map_bump" if k in ("bump", "map_bump") else k] = name
        elif k == "bump_height":
            # an unknown MTL parameter, read with atof (host_material.cpp:57-59); a -bm option is not used
            cur["bump_height"] = _atof(" ".join(args))
        elif k == "shading":
            cur["shading"] = args[0] if args else ""
    return mats


def parse_obj(text):
    """OBJ text -> dict: v (n,3), vt (n,2), vn (n,3) float32; faces: int array
    (T, 3, 3) of (vertex, texcoord, normal) indices per triangle corner (-1:
    absent); mat: material name per triangle (None before any usemtl);
    mtllib: file names."""
    v, vt, vn, corners, mats, mtllib = [], [], [], [], [], []
    cur = None
    for line in text.splitlines():
        t = line.split("#", 1)[0].split()
        if not t:
            continue
        k = t[0]
        if k == "v":
            v.append([float(x) for x in t[1:4]])
        elif k == "vt":
            vt.append([float(x) for x in t[1:3]] + [0.0] * max(0, 3 - len(t)))
        elif k == "vn":
            vn.append([float(x) for x in t[1:4]])
        elif k == "f":
            poly = []
            for c in t[1:]:
                p = c.split("/")
                def idx(i, n):
                    if i >= len(p) or p[i] == "":
                        return -1
                    j = int(p[i])
                    return j - 1 if j > 0 else n + j
                poly.append((idx(0, len(v)), idx(1, len(vt)), idx(2, len(vn))))
            for i in range(1, len(poly) - 1):        # fan: tinyobjloader's ear clipping for convex polygons
                corners.append((poly[0], poly[i], poly[i + 1]))
                mats.append(cur)
        elif k == "usemtl":
            cur = " ".join(t[1:])
        elif k == "mtllib":
            mtllib += t[1:]
    return {"v": np.array(v, f32).reshape(-1, 3), "vt": np.array(vt, f32).reshape(-1, 2)[:, :2] if vt else np.zeros((0, 2), f32),
            "vn": np.array(vn, f32).reshape(-1, 3), "faces": np.array(corners, np.int64).reshape(-1, 3, 3), "mat": mats, "mtllib": mtllib}


def _dot(a, b):
    return ((a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]).astype(f32)


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1).astype(f32)


def _normalize(a):
    return (a * (f32(1.0) / np.sqrt(_dot(a, a)))[..., None]).astype(f32)


def obj_triangles(obj, material_ids, transform=None, flat_shaded=False, color_textures=None):
    """The vertex soup and CoreTri array of LoadGeometryFromOBJ
    (host_mesh.cpp:166-318). material_ids: scene material index per triangle;
    color_textures: material index -> (width, height) of its colour texture
    (for the triangle LOD)."""
    with np.errstate(invalid="ignore", divide="ignore"):   # degenerate triangles give NaN normals, as in the reference
        return _obj_triangles(obj, material_ids, transform, flat_shaded, color_textures)


def _obj_triangles(obj, material_ids, transform, flat_shaded, color_textures):
    faces = obj["faces"]
    T = len(faces)
    P = obj["v"]
    vi, ti, ni = faces[..., 0], faces[..., 1], faces[..., 2]
    M = np.asarray(F.IDENTITY if transform is None else transform, f32).reshape(4, 4)
    # per-normal alpha of consistent normal interpolation (:167-202)
    nN = len(obj["vn"])
    alphas = np.ones(nN, f32)
    has_n = ni[:, 0] > -1
    if flat_shaded:
        alphas[ni[ni > -1]] = 1
    elif nN:
        q = P[vi]                                     # untransformed positions
        N = _normalize(_cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0]))
        sel = np.nonzero(has_n)[0]
        vN = obj["vn"][ni[sel]]                       # (k, 3, 3)
        Ns = N[sel]
        d = np.stack([_dot(Ns, vN[:, c]) for c in range(3)], 1)
        flip = (d < 0).all(1)
        Ns = np.where(flip[:, None], -Ns, Ns).astype(f32)
        for c in range(3):                           # std::max(0.7f, d): a NaN dot gives 0.7
            d = _dot(vN[:, c], Ns)
            np.minimum.at(alphas, ni[sel, c], np.where(d > f32(0.7), d, f32(0.7)).astype(f32))
    w = f32(0.03632)
    alphas = (np.arccos(alphas).astype(f32) * (f32(1) + w * (f32(1) - alphas) * (f32(1) - alphas))).astype(f32)
    # soup: (v, 1) * transform (common_types.h operator*(float4, mat4))
    p = P[vi]
    tv = np.stack([((M[r, 0] * p[..., 0] + M[r, 1] * p[..., 1]) + M[r, 2] * p[..., 2]) + M[r, 3] for r in range(3)], -1).astype(f32)
    verts = np.ones((3 * T, 4), f32)
    verts[:, :3] = tv.reshape(-1, 3)
    tris = F.new_tris(T)
    v0, v1, v2 = tv[:, 0], tv[:, 1], tv[:, 2]
    tris["vertex0"], tris["vertex1"], tris["vertex2"] = v0, v1, v2
    e1, e2 = (v1 - v0).astype(f32), (v2 - v0).astype(f32)
    N = _normalize(_cross(e1, e2))
    if nN:
        vN = np.where(has_n[:, None, None], obj["vn"][np.maximum(ni, 0)], N[:, None, :]).astype(f32)
        N = np.where((has_n & (_dot(N, vN[:, 0]) < 0))[:, None], -N, N).astype(f32)
    else:
        vN = np.repeat(N[:, None, :], 3, 1)
    if flat_shaded:
        vN = np.repeat(N[:, None, :], 3, 1)
    tris["vN0"], tris["vN1"], tris["vN2"] = vN[:, 0], vN[:, 1], vN[:, 2]
    has_t = ti[:, 0] > -1
    Tn = _normalize(e1)
    Bn = _normalize(_cross(N, Tn))
    if len(obj["vt"]):
        uv = np.where(has_t[:, None, None], obj["vt"][np.maximum(ti, 0)], 0).astype(f32)
        for c in range(3):
            tris[f"u{c}"] = np.where(has_t, uv[:, c, 0], 0)
            tris[f"v{c}"] = np.where(has_t, uv[:, c, 1], 0)
        uv01, uv02 = (uv[:, 1] - uv[:, 0]).astype(f32), (uv[:, 2] - uv[:, 0]).astype(f32)
        d01 = (uv01[:, 0] * uv01[:, 0] + uv01[:, 1] * uv01[:, 1]).astype(f32)
        d02 = (uv02[:, 0] * uv02[:, 0] + uv02[:, 1] * uv02[:, 1]).astype(f32)
        use = has_t & (d01 != 0) & (d02 != 0)
        Tu = _normalize((e1 * uv02[:, 1:2] - e2 * uv01[:, 1:2]).astype(f32))
        Bu = _normalize((e2 * uv01[:, 0:1] - e1 * uv02[:, 0:1]).astype(f32))
        Tn = np.where(use[:, None], Tu, Tn)
        Bn = np.where(use[:, None], Bu, Bn)
    tris["T"], tris["B"] = Tn, Bn
    tris["Nx"], tris["Ny"], tris["Nz"] = N[:, 0], N[:, 1], N[:, 2]
    tris["material"] = material_ids
    tris["area"], tris["invArea"] = 0, 0              # set for lights only (:294-300)
    al = np.where(has_n[:, None], alphas[np.maximum(ni, 0)] if nN else 0, 0).astype(f32)
    tris["alpha"] = al
    for m, (tw, th) in (color_textures or {}).items():
        sel = tris["material"] == m
        if sel.any():
            tris["LOD"][sel] = scenes.triangle_lod(tris[sel], tw, th)
    return verts, tris


def _obj_material(m, scene, directory, tex_cache):
    """HostMaterial::ConvertFrom(tinyobj material) (host_material.cpp:35-71)."""
    mat = F.new_materials(1)
    mat["color"]["value"][0] = m["Kd"]
    mat["absorption"]["value"][0] = m["Kt"]
    mat["ior"]["value"] = m["Ni"]
    mat["roughness"]["value"] = 1.0
    mat["flags"] = 0 if m.get("shading") == "flat" else 1     # SMOOTH
    size = None

    def tex(name, **kw):
        key = (name, tuple(sorted(kw.items())))
        if key not in tex_cache:
            t = load_texture(os.path.join(directory, name), **kw)
            tex_cache[key] = (scene.add_texture(t), (t["width"], t["height"]))
        return tex_cache[key]
    if m.get("map_Kd"):
        tid, size = tex(m["map_Kd"], flipped=True, linearized=True)
        mat["color"]["textureID"] = tid
        mat["color"]["value"][0] = 1.0
    if m.get("norm"):
        mat["normals"]["textureID"] = tex(m["norm"], flipped=True, normal_map=True)[0]
    elif m.get("map_bump"):
        # a texture of its own per material (CreateTexture, not FindOrCreate: the height scale may differ)
        t = load_texture(os.path.join(directory, m["map_bump"]), flipped=True)
        bump_to_normal_map(t, m.get("bump_height", 1.0))
        mat["normals"]["textureID"] = scene.add_texture(t)
    if m.get("map_Ks"):
        mat["roughness"]["textureID"] = tex(m["map_Ks"], flipped=True)[0]
        mat["roughness"]["value"] = 1.0
    return mat, size


def load_obj(scene, path, transform=None, flat_shaded=False):
    """Adds the OBJ file's materials, textures and one mesh to `scene`
    (HostMesh::LoadGeometryFromOBJ); returns the mesh index. Triangles
    without a material get a default grey Lambert material."""
    directory = os.path.dirname(os.path.abspath(path))
    with open(path, "r") as fh:
        obj = parse_obj(fh.read())
    mtls = []
    for lib in obj["mtllib"]:
        with open(resolve_path(os.path.join(directory, lib)), "r") as fh:
            mtls += parse_mtl(fh.read())
    offset = len(scene.materials)
    by_name, sizes, cache = {}, {}, {}
    for k, m in enumerate(mtls):
        mat, size = _obj_material(m, scene, directory, cache)
        scene.add_material(mat)
        by_name.setdefault(m["name"], offset + k)
        if size is not None:
            sizes[offset + k] = size
    default = None
    ids = []
    for name in obj["mat"]:
        if name in by_name:
            ids.append(by_name[name])
        else:
            if default is None:
                default = scene.add_material(F.lambert_material((0.5, 0.5, 0.5)))
            ids.append(default)
    verts, tris = obj_triangles(obj, np.array(ids, np.int32), transform, flat_shaded, sizes)
    return scene.add_mesh(verts, tris)


# ---- glTF 2.0 -------------------------------------------------------------------
_GLTF_COMPONENT = {5120: "<i1", 5121: "<u1", 5122: "<i2", 5123: "<u2", 5124: "<i4", 5125: "<u4", 5126: "<f4"}
_GLTF_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
MODE_TRIANGLES, MODE_TRIANGLE_STRIP, MODE_TRIANGLE_FAN = 4, 5, 6
MAT_FROM_MTL, MAT_SINGLE_COLOR_COPY = 4, 16          # HostMaterial flags (host_material.h:68-74)


def read_gltf(path):
    """A .gltf (JSON; buffers and images from files beside it or data: URIs)
    or .glb (binary container, JSON chunk + BIN chunk) -> (document dict, list
    of buffer bytes, directory)."""
    import base64
    import json
    directory = os.path.dirname(os.path.abspath(path))
    with open(path, "rb") as fh:
        data = fh.read()
    glb_bin = None
    if data[:4] == b"glTF":
        version, length = struct.unpack("<II", data[4:12])
        if version != 2:
            raise ValueError(f"{path}: glb version {version}")
        pos, doc = 12, None
        while pos < min(length, len(data)):
            n, kind = struct.unpack("<II", data[pos:pos + 8])
            body = data[pos + 8:pos + 8 + n]
            if kind == 0x4E4F534A:
                doc = json.loads(body.decode("utf-8"))
            elif kind == 0x004E4942 and glb_bin is None:
                glb_bin = body
            pos += 8 + n
        if doc is None:
            raise ValueError(f"{path}: glb without a JSON chunk")
    else:
        doc = json.loads(data.decode("utf-8"))

    def fetch(uri):
        if uri.startswith("data:"):
            return base64.b64decode(uri.split(",", 1)[1])
        from urllib.parse import unquote
        with open(resolve_path(os.path.join(directory, unquote(uri))), "rb") as fh:
            return fh.read()
    bufs = []
    for b in doc.get("buffers", []):
        if "uri" in b:
            bufs.append(fetch(b["uri"]))
        elif glb_bin is not None:
            bufs.append(glb_bin)
        else:
            raise ValueError(f"{path}: buffer without data")
    doc["_fetch"] = fetch
    return doc, bufs, directory


def gltf_accessor(doc, bufs, index):
    """An accessor's elements as an (count, components) array, reading
    bufferView.byteOffset + accessor.byteOffset and stepping byteStride (the
    element size when the view gives none: Accessor::ByteStride)."""
    acc = doc["accessors"][index]
    if "bufferView" not in acc or "sparse" in acc:
        raise ValueError(f"accessor {index}: accessors without a buffer view, or sparse, are not read")
    view = doc["bufferViews"][acc["bufferView"]]
    dt = np.dtype(_GLTF_COMPONENT[acc["componentType"]])
    width = _GLTF_WIDTH[acc["type"]]
    stride = view.get("byteStride", 0) or dt.itemsize * width
    count = acc["count"]
    base = view.get("byteOffset", 0) + acc.get("byteOffset", 0)
    buf = bufs[view["buffer"]]
    if count and base + stride * (count - 1) + dt.itemsize * width > len(buf):
        raise ValueError(f"accessor {index} runs past the end of buffer {view['buffer']}")
    if count == 0:
        return np.zeros((0, width), dt)
    return np.ndarray((count, width), dt, buffer=buf, offset=base, strides=(stride, dt.itemsize)).copy()


def _gltf_float_attribute(doc, bufs, index, acc_type, what):
    acc = doc["accessors"][index]
    if acc["type"] != acc_type:
        raise ValueError(f"{what}: expected {acc_type} in the glTF file, got {acc['type']}")
    if acc["componentType"] != 5126:
        raise ValueError(f"{what}: only float components are read")
    return gltf_accessor(doc, bufs, index).astype(f32)


def gltf_primitive_indices(doc, bufs, prim):
    """Triangle corner indices of one primitive (host_mesh.cpp:333-374): the
    index accessor widened to int (signed for BYTE / SHORT), a fan expanded to
    (f0, f[i-1], f[i]) and a strip to (s[i-2], s[i-1], s[i]) (no winding
    alternation); None for a non-triangle mode. A primitive without indices
    uses 0 .. count-1 of its positions, as the glTF specification defines."""
    mode = prim.get("mode", MODE_TRIANGLES)
    if "indices" in prim:
        idx = gltf_accessor(doc, bufs, prim["indices"]).reshape(-1).astype(np.int64)
    else:
        idx = np.arange(doc["accessors"][prim["attributes"]["POSITION"]]["count"], dtype=np.int64)
    if mode == MODE_TRIANGLE_FAN:
        k = np.arange(2, len(idx))
        idx = np.stack([np.full(len(k), idx[0] if len(idx) else 0), idx[k - 1], idx[k]], 1).reshape(-1)
    elif mode == MODE_TRIANGLE_STRIP:
        k = np.arange(2, len(idx))
        idx = np.stack([idx[k - 2], idx[k - 1], idx[k]], 1).reshape(-1)
    elif mode != MODE_TRIANGLES:
        return None
    return idx[:len(idx) // 3 * 3].reshape(-1, 3)


def indexed_triangles(scene, idx, P, Nv=None, UV=None, UV2=None, material=0):
    """HostMesh::BuildFromIndexedData (host_mesh.cpp:520-681): the vertex soup
    and CoreTri array of indexed triangles. Per unique vertex the
    consistent-normal alpha is min over its triangles of dot(vN, N) (N flipped
    only when all three vertex normals face away; no 0.7 clamp), then
    acos(a) (1 + 0.03632 (1 - a)^2); the stored face normal is not flipped;
    tangents from the uvs, PBRT's frame for a degenerate uv mapping
    (pbrt-v3 triangle.cpp), the edge frame when they come out NaN or when
    there are no uvs; a triangle whose three uvs are one point of a textured
    material gets a single-colour copy of the material
    (HostScene::FindOrCreateMaterialCopy, host_scene.cpp:570-595). LOD and
    area stay 0 (the glTF path sets neither)."""
    with np.errstate(invalid="ignore", divide="ignore"):
        return _indexed_triangles(scene, np.asarray(idx, np.int64).reshape(-1, 3), np.asarray(P, f32), Nv, UV, UV2, material)


def _indexed_triangles(scene, idx, P, Nv, UV, UV2, material):
    T = len(idx)
    q = P[idx]                                        # (T, 3, 3)
    v0, v1, v2 = q[:, 0], q[:, 1], q[:, 2]
    e1, e2 = (v1 - v0).astype(f32), (v2 - v0).astype(f32)
    N = _normalize(_cross(e1, e2))
    alphas = np.ones(len(P), f32)
    if Nv is not None:
        vN = np.asarray(Nv, f32)[idx]
        flip = (_dot(N, vN[:, 0]) < 0) & (_dot(N, vN[:, 1]) < 0) & (_dot(N, vN[:, 2]) < 0)
        Na = np.where(flip[:, None], -N, N).astype(f32)
    else:
        vN = np.repeat(N[:, None, :], 3, 1)
        Na = N
    for c in range(3):                                # std::min(alpha, d) keeps alpha when d is NaN
        np.fmin.at(alphas, idx[:, c], _dot(vN[:, c], Na))
    w = f32(0.03632)
    alphas = (np.arccos(alphas).astype(f32) * (f32(1) + w * (f32(1) - alphas) * (f32(1) - alphas))).astype(f32)
    verts = np.ones((3 * T, 4), f32)
    verts[:, :3] = q.reshape(-1, 3)
    tris = F.new_tris(T)
    tris["material"] = material
    tris["Nx"], tris["Ny"], tris["Nz"] = N[:, 0], N[:, 1], N[:, 2]
    tris["vertex0"], tris["vertex1"], tris["vertex2"] = v0, v1, v2
    tris["alpha"] = alphas[idx]
    tris["vN0"], tris["vN1"], tris["vN2"] = vN[:, 0], vN[:, 1], vN[:, 2]
    Te = _normalize(e1)
    Be = _normalize(_cross(N, Te))
    if UV is not None:
        uv = np.asarray(UV, f32)[idx]                 # (T, 3, 2)
        for c in range(3):
            tris[f"u{c}"], tris[f"v{c}"] = uv[:, c, 0], uv[:, c, 1]
        _single_texel_materials(scene, tris, uv, material)
        uv01, uv02 = (uv[:, 1] - uv[:, 0]).astype(f32), (uv[:, 2] - uv[:, 0]).astype(f32)
        d01 = (uv01[:, 0] * uv01[:, 0] + uv01[:, 1] * uv01[:, 1]).astype(f32)
        d02 = (uv02[:, 0] * uv02[:, 0] + uv02[:, 1] * uv02[:, 1]).astype(f32)
        degenerate = (d01 == 0) | (d02 == 0)
        ax = np.abs(N[:, 0]) > np.abs(N[:, 1])
        sx = np.sqrt((N[:, 0] * N[:, 0] + N[:, 2] * N[:, 2]).astype(f32)).astype(f32)
        sy = np.sqrt((N[:, 1] * N[:, 1] + N[:, 2] * N[:, 2]).astype(f32)).astype(f32)
        zero = np.zeros(T, f32)
        Tp = np.where(ax[:, None], np.stack([-N[:, 2] / sx, zero / sx, N[:, 0] / sx], 1),
                      np.stack([zero / sy, N[:, 2] / sy, -N[:, 1] / sy], 1)).astype(f32)
        Bp = _normalize(_cross(N, Tp))
        Tu = _normalize((e1 * uv02[:, 1:2] - e2 * uv01[:, 1:2]).astype(f32))
        Bu = _normalize((e2 * uv01[:, 0:1] - e1 * uv02[:, 0:1]).astype(f32))
        Tn = np.where(degenerate[:, None], Tp, Tu)
        Bn = np.where(degenerate[:, None], Bp, Bu)
        ssum = Tn[:, 0] + Tn[:, 1]
        for v in (Tn[:, 2], Bn[:, 0], Bn[:, 1], Bn[:, 2]):
            ssum = (ssum + v).astype(f32)
        bad = np.isnan(ssum)
        Tn = np.where(bad[:, None], Te, Tn).astype(f32)
        Bn = np.where(bad[:, None], Be, Bn).astype(f32)
    else:
        Tn, Bn = Te, Be
    tris["T"], tris["B"] = Tn, Bn
    if UV2 is not None:
        uv2 = np.asarray(UV2, f32)[idx]
        for c in range(3):
            tris[f"u1_{c}"], tris[f"v1_{c}"] = uv2[:, c, 0], uv2[:, c, 1]
    return verts, tris


def _single_texel_materials(scene, tris, uv, material):
    """Triangles that read one texel of a textured material get an untextured
    copy of it whose colour is that texel (level 0, texel & 0xffffff, bytes
    2, 1, 0 as b, g, r scaled by 1/255 into x, y, z); an existing copy of
    the same colour is reused, whatever it was copied from."""
    tid = int(scene.materials["color"]["textureID"][material])
    if tid < 0:
        return
    one = (uv[:, 0, 0] == uv[:, 1, 0]) & (uv[:, 1, 0] == uv[:, 2, 0]) & (uv[:, 0, 1] == uv[:, 1, 1]) & (uv[:, 1, 1] == uv[:, 2, 1])
    if not one.any():
        return
    tex = scene.textures[tid]
    w, h = tex["width"], tex["height"]
    for t in np.nonzero(one)[0]:
        u = (int(np.trunc(f32(uv[t, 0, 0]) * f32(w))) & 0xffffffff) % w
        v = (int(np.trunc(f32(uv[t, 0, 1]) * f32(h))) & 0xffffffff) % h
        texel = int(tex["pixels"][u + v * w]) & 0xffffff
        r, g, b = (texel >> 16) & 255, (texel >> 8) & 255, texel & 255
        c = np.array([f32(b) * f32(1.0 / 255.0), f32(g) * f32(1.0 / 255.0), f32(r) * f32(1.0 / 255.0)], f32)
        mats = scene.materials
        hit = np.nonzero(((mats["flags"] & MAT_SINGLE_COLOR_COPY) != 0) & (mats["color"]["value"] == c).all(1))[0]
        if len(hit):
            tris["material"][t] = hit[0]
            continue
        m = mats[material:material + 1].copy()
        m["color"]["textureID"] = -1
        m["color"]["value"][0] = c
        m["flags"] |= MAT_SINGLE_COLOR_COPY
        tris["material"][t] = scene.add_material(m)


def _gltf_material(mat, tex_ids, scene):
    """HostMaterial::ConvertFrom(tinygltf material) (host_material.cpp:76-170):
    only the factors present in the file are set (absent ones stay "not
    set"): baseColorFactor -> colour, metallicFactor, roughnessFactor;
    baseColorTexture -> colour map; metallicRoughnessTexture -> roughness and
    metallic maps; normalTexture (index, scale) -> normal map, its texture
    marked as one. Emissive, occlusion and alpha-mode entries are ignored."""
    m = F.new_materials(1)
    m["flags"] = 1 | MAT_FROM_MTL                     # SMOOTH (the default) | FROM_MTL
    nt = mat.get("normalTexture")
    if nt is not None and nt.get("index", -1) > -1:
        m["normals"]["textureID"] = tex_ids[nt["index"]]
        m["normals"]["scale"] = nt.get("scale", 1.0)
        mark_normal_map(scene.textures[tex_ids[nt["index"]]])
    pbr = mat.get("pbrMetallicRoughness", {})
    if "baseColorFactor" in pbr:
        m["color"]["value"][0] = pbr["baseColorFactor"][:3]
    if "metallicFactor" in pbr:
        m["metallic"]["value"] = pbr["metallicFactor"]
    if "roughnessFactor" in pbr:
        m["roughness"]["value"] = pbr["roughnessFactor"]
    if "baseColorTexture" in pbr and "index" in pbr["baseColorTexture"]:
        m["color"]["textureID"] = tex_ids[int(pbr["baseColorTexture"]["index"])]
    if "metallicRoughnessTexture" in pbr and "index" in pbr["metallicRoughnessTexture"]:
        t = tex_ids[int(pbr["metallicRoughnessTexture"]["index"])]
        m["roughness"]["textureID"] = t
        m["metallic"]["textureID"] = t
    return m


def _gltf_image_rgba(doc, bufs, image):
    if "bufferView" in image:
        view = doc["bufferViews"][image["bufferView"]]
        off = view.get("byteOffset", 0)
        data = bytes(bufs[view["buffer"]][off:off + view["byteLength"]])
    elif "uri" in image:
        data = doc["_fetch"](image["uri"])
    else:
        raise ValueError("glTF image without data")
    return decode_image(data, image.get("uri", image.get("name", "")))


def trs_matrix(t, q, sc, M):
    """HostNode::UpdateTransformFromTRS (host_node.cpp:131-138): T * R * S *
    matrix in fp32; q = (x, y, z, w), rotated by quat::toMatrix
    (common_types.h:692-705)."""
    Tm = F.IDENTITY.copy()
    Tm[3], Tm[7], Tm[11] = np.asarray(t, f32)
    x, y, z, w = (f32(c) for c in q)
    two = f32(2)
    R = F.IDENTITY.copy()
    R[0] = (f32(1) - two * y * y) - two * z * z
    R[1] = two * x * y - two * w * z
    R[2] = two * x * z + two * w * y
    R[4] = two * x * y + two * w * z
    R[5] = (f32(1) - two * x * x) - two * z * z
    R[6] = two * y * z - two * w * x
    R[8] = two * x * z - two * w * y
    R[9] = two * y * z + two * w * x
    R[10] = (f32(1) - two * x * x) - two * y * y
    S = F.IDENTITY.copy()
    S[0], S[5], S[10] = np.asarray(sc, f32)
    return mat4_mul(mat4_mul(mat4_mul(Tm, R), S), M)


def _node_state(node):
    """A glTF node's matrix (column-major in the file, transposed into
    row-major cells), translation, rotation (x, y, z, w) and scale, and
    whether any was given (HostNode::ConvertFromGLTFNode, host_node.cpp:77-125)."""
    st = {"M": F.IDENTITY.copy(), "t": np.zeros(3, f32), "q": np.array([0, 0, 0, 1], f32), "s": np.ones(3, f32), "given": False}
    if len(node.get("matrix", [])) == 16:
        st["M"] = np.asarray(node["matrix"], np.float64).astype(f32).reshape(4, 4).T.reshape(16).copy()
        st["given"] = True
    for key, name, n in (("t", "translation", 3), ("q", "rotation", 4), ("s", "scale", 3)):
        if len(node.get(name, [])) == n:
            st[key] = np.asarray(node[name], np.float64).astype(f32)
            st["given"] = True
    return st


def gltf_node_transform(node):
    """A node's local transform: T * R * S * matrix when any of matrix,
    translation, rotation or scale is given, identity otherwise."""
    st = _node_state(node)
    return trs_matrix(st["t"], st["q"], st["s"], st["M"]) if st["given"] else st["M"]


# ---- glTF animation ------------------------------------------------------------------
def _quat_normalize(q):
    """quat::normalize: scale by 1 / sqrtf(((w w + x x) + y y) + z z); q = (x, y, z, w)."""
    x, y, z, w = (f32(c) for c in q)
    m = np.sqrt(f32(f32(f32(w * w + x * x) + y * y) + z * z)).astype(f32)
    r = f32(1) / m
    return np.array([x * r, y * r, z * r, w * r], f32)


def quat_slerp(a, b, t):
    """quat::slerp (common_types.h:738-762), q = (x, y, z, w): b flipped to
    a's hemisphere, a linear blend above cos 0.99, otherwise
    (sin(1 - t) a + sin(t angle) b) / sin(angle) (the reference's first
    weight is sin(1 - t), not sin((1 - t) angle))."""
    a = np.asarray(a, f32)
    r = np.asarray(b, f32).copy()
    t = f32(t)
    ct = f32(f32(f32(a[3] * r[3] + a[0] * r[0]) + a[1] * r[1]) + a[2] * r[2])
    if ct < 0:
        r, ct = (r * f32(-1)).astype(f32), -ct
    if ct > f32(0.99):
        return ((f32(1) - t) * a + t * r).astype(f32)
    ang = np.arccos(ct).astype(f32)
    s1, s2, s3 = np.sin(f32(1) - t).astype(f32), np.sin(t * ang).astype(f32), np.sin(ang).astype(f32)
    return ((s1 * a + s2 * r) / s3).astype(f32)


class AnimationSampler:
    """HostAnimation::Sampler (host_anim.cpp:31-181): key times, keys (vec3,
    quaternion (x, y, z, w) or scalar) and the interpolation (STEP,
    CUBICSPLINE with in-tangent, value, out-tangent per key, else LINEAR).
    Integer rotation / weight keys are scaled as the reference reads them
    (byte / 127 clamped at -1; unsigned byte read as a signed byte / 255;
    short and unsigned short read as their first byte, signed, / 32767
    clamped at -1 and / 65535)."""

    def __init__(self, doc, bufs, sampler):
        self.t = gltf_accessor(doc, bufs, sampler["input"]).reshape(-1).astype(f32)
        self.kind = {"STEP": "step", "CUBICSPLINE": "spline"}.get(sampler.get("interpolation", "LINEAR"), "linear")
        acc = doc["accessors"][sampler["output"]]
        raw = gltf_accessor(doc, bufs, sampler["output"])
        if acc["type"] == "VEC3":
            self.keys = raw.astype(f32)
        else:
            ct = acc["componentType"]
            if ct == 5126:
                v = raw.astype(f32)
            else:
                first = raw.astype(np.int64) & 255
                sb = np.where(first > 127, first - 256, first).astype(f32)       # the first byte, as a signed char
                v = {5120: np.maximum(sb / f32(127), f32(-1)), 5121: sb / f32(255),
                     5122: np.maximum(sb / f32(32767), f32(-1)), 5123: sb / f32(65535)}[ct].astype(f32)
            self.keys = v if acc["type"] == "VEC4" else v.reshape(-1)

    def sample_float(self, time, k, i, count):
        """SampleFloat (host_anim.cpp:95-118): weight i of count per key."""
        K, spline = self.keys, self.kind == "spline"
        if k == 0 and time < self.t[0]:
            return K[1] if spline else K[0]
        t0, t1 = self.t[k], self.t[k + 1]
        f = f32(f32(f32(time) - t0) / f32(t1 - t0))
        if f <= 0:
            return K[0]
        if spline:
            t2 = f * f
            t3 = t2 * f
            dt = f32(t1 - t0)
            p0, m0 = K[(k * count + i) * 3 + 1], f32(dt * K[(k * count + i) * 3 + 2])
            p1, m1 = K[((k + 1) * count + i) * 3 + 1], f32(dt * K[((k + 1) * count + i) * 3])
            return f32(f32(f32(m0 * f32((t3 - f32(2) * t2) + f) + p0 * f32((f32(2) * t3 - f32(3) * t2) + f32(1)))
                           + p1 * f32(f32(-2) * t3 + f32(3) * t2)) + m1 * f32(t3 - t2))
        if self.kind == "step":
            return K[k]
        return f32(f32((f32(1) - f) * K[k * count + i]) + f32(f * K[(k + 1) * count + i]))

    def sample(self, time, k, quat):
        """SampleVec3 / SampleQuat (:120-181)."""
        K, spline = self.keys, self.kind == "spline"
        if k == 0 and time < self.t[0]:
            return K[1] if spline else K[0]
        t0, t1 = self.t[k], self.t[k + 1]
        f = f32(f32(f32(time) - t0) / f32(t1 - t0))
        if f <= 0:
            return K[0]
        if spline:
            t2 = f * f
            t3 = t2 * f
            dt = f32(t1 - t0)
            p0, m0 = K[k * 3 + 1], (K[k * 3 + 2] * dt).astype(f32)
            p1, m1 = K[(k + 1) * 3 + 1], (K[(k + 1) * 3] * dt).astype(f32)
            out = (((m0 * f32((t3 - f32(2) * t2) + f) + p0 * f32((f32(2) * t3 - f32(3) * t2) + f32(1)))
                    + p1 * f32(f32(-2) * t3 + f32(3) * t2)) + m1 * f32(t3 - t2)).astype(f32)
            return _quat_normalize(_quat_normalize(out)) if quat else out
        if self.kind == "step":
            return _quat_normalize(K[k]) if quat else K[k]
        if quat:
            return _quat_normalize(_quat_normalize(quat_slerp(K[k], K[k + 1], f)))
        return ((f32(1) - f) * K[k] + f * K[k + 1]).astype(f32)


class AnimationChannel:
    """HostAnimation::Channel (host_anim.cpp:183-276): its own clock and key
    index; times wrap at the last key; a single key, or a last key at time
    0, holds the first key. Targets: translation, rotation, scale (morph
    weights are not applied: meshes are read in their base pose)."""

    def __init__(self, ch):
        self.sampler = ch["sampler"]
        self.node = ch["target"].get("node", -1)
        self.path = ch["target"]["path"]
        self.t, self.k = f32(0), 0

    def update(self, dt, smp, state):
        self.t = f32(self.t + f32(dt))
        if self.node < 0 or self.path not in ("translation", "rotation", "scale", "weights"):
            return False
        if self.path == "weights":
            w = state[self.node].get("w")
            if w is None:
                return False
            n = len(smp.t)
            dur = smp.t[n - 1]
            if dur == 0 or n == 1:
                w[:] = smp.keys[0]
            else:
                while self.t >= dur:
                    self.t, self.k = f32(self.t - dur), 0
                while self.t >= smp.t[(self.k + 1) % n]:
                    self.k += 1
                for i in range(len(w)):
                    w[i] = smp.sample_float(self.t, self.k, i, len(w))
            state[self.node]["morphed"] = True
            return True
        key = {"translation": "t", "rotation": "q", "scale": "s"}[self.path]
        n = len(smp.t)
        dur = smp.t[n - 1]
        if dur == 0 or n == 1:
            v = smp.keys[0]
        else:
            while self.t >= dur:
                self.t, self.k = f32(self.t - dur), 0
            while self.t >= smp.t[(self.k + 1) % n]:
                self.k += 1
            v = smp.sample(self.t, self.k, key == "q")
        state[self.node][key] = np.asarray(v, f32).copy()
        state[self.node]["given"] = True
        return True


class GltfScene:
    """The node hierarchy and animations of one glTF file added to a Scene:
    `instances` (scene instance indices, in HostNode::Update's children-first
    order), `update(dt)` advancing every animation channel by dt seconds
    (HostAnimation::Update) and re-placing the instances and the world-space
    lights of emissive meshes (HostNode::Update / UpdateLights)."""

    def __init__(self, scene, doc, bufs, mesh_ids, node_base, root, skin_data=None, morph_data=None):
        self.scene, self.mesh_ids, self.node_base = scene, mesh_ids, node_base
        self.morph = morph_data or {}                  # mesh -> (base positions, base normals, target offsets, target normals)
        # skins (HostSkin::ConvertFromGLTFSkin, host_mesh.cpp:46-66): joint nodes and
        # inverse bind matrices, column-major in the file, transposed into row-major cells
        self.skins = []
        for sk in doc.get("skins", []):
            ibm = None
            if sk.get("inverseBindMatrices", -1) > -1:
                ibm = gltf_accessor(doc, bufs, sk["inverseBindMatrices"]).astype(f32).reshape(-1, 4, 4).transpose(0, 2, 1).reshape(-1, 16)
            self.skins.append((list(sk.get("joints", [])), ibm))
        self.skinned = {}                              # mesh -> (original soup, original vertex normals, joints, weights)
        for m, (J, Wt) in (skin_data or {}).items():
            v, t = scene.meshes[m]
            vn = np.stack([t["vN0"], t["vN1"], t["vN2"]], 1).reshape(-1, 3)
            self.skinned[m] = (v.copy(), vn.astype(f32), J, Wt)
        self.docnodes = doc.get("nodes", [])
        self.state = [_node_state(n) for n in self.docnodes]
        for i, n in enumerate(self.docnodes):         # morph weights start at 0 (HostNode::ConvertFromGLTFNode)
            m = mesh_ids[n["mesh"]] if "mesh" in n else None
            self.state[i]["w"] = np.zeros(self.morph[m][2].shape[0], f32) if m in self.morph else None
            self.state[i]["morphed"] = False
        self.root = mat4_mul(F.IDENTITY, root)
        sc = doc.get("scenes", [])
        self.roots = sc[0].get("nodes", []) if sc else []
        self.samplers, self.channels = [], []
        for a in doc.get("animations", []):
            smp = [AnimationSampler(doc, bufs, x) for x in a.get("samplers", [])]
            self.channels.append([(AnimationChannel(c), smp) for c in a.get("channels", [])])
        combined, order = self._walk()
        self.order = order
        self.instances = [scene.add_instance(mesh_ids[self.docnodes[i]["mesh"]], combined[i]) for i in order]
        self.lights = {}
        for i in sorted(order):                      # lights in node order (PrepareLights at construction)
            self.lights[i] = instance_lights(scene, mesh_ids[self.docnodes[i]["mesh"]], combined[i], node_base + i)
        self._pose(combined)

    def _pose(self, combined):
        """HostNode::Update's skinning step (host_node.cpp:176-186) and
        HostMesh::SetPose(skin) (host_mesh.cpp:717-911, its scalar form):
        joint matrices inverse(mesh node) * joint node * inverse bind; per
        vertex the weighted sum of its four joint matrices moves the
        original position and vertex normal (normal renormalised), then the
        face normals are recomputed. The reference's vector path uses
        approximate reciprocal square roots, so skinned normals agree with it
        to about 12 bits, not bit for bit."""
        for i in self.order:
            node = self.docnodes[i]
            m = self.mesh_ids[node["mesh"]]
            if "skin" not in node or m not in self.skinned:
                continue
            joints, ibm = self.skins[node["skin"]]
            inv = np.linalg.inv(np.asarray(combined[i], np.float64).reshape(4, 4)).astype(f32).reshape(16)
            jm = np.stack([mat4_mul(mat4_mul(inv, combined.get(jn, F.IDENTITY)), ibm[k] if ibm is not None else F.IDENTITY)
                           for k, jn in enumerate(joints)]).reshape(-1, 4, 4)
            v0, vn0, J, Wt = self.skinned[m]
            S = (((Wt[:, 0, None, None] * jm[J[:, 0]] + Wt[:, 1, None, None] * jm[J[:, 1]]) + Wt[:, 2, None, None] * jm[J[:, 2]])
                 + Wt[:, 3, None, None] * jm[J[:, 3]]).astype(f32)
            p = np.einsum("nij,nj->ni", S.astype(np.float64), v0.astype(np.float64)).astype(f32)
            n4 = np.concatenate([vn0, np.zeros((len(vn0), 1), f32)], 1)
            nn = np.einsum("nij,nj->ni", S.astype(np.float64), n4.astype(np.float64))[:, :3].astype(f32)
            nn = _normalize(nn)
            _, t = self.scene.meshes[m]
            t = t.copy()
            q = p[:, :3].reshape(-1, 3, 3)
            t["vertex0"], t["vertex1"], t["vertex2"] = q[:, 0], q[:, 1], q[:, 2]
            with np.errstate(invalid="ignore", divide="ignore"):
                N = _normalize(_cross((q[:, 1] - q[:, 0]).astype(f32), (q[:, 2] - q[:, 0]).astype(f32)))
            nn = nn.reshape(-1, 3, 3)
            t["vN0"], t["vN1"], t["vN2"] = nn[:, 0], nn[:, 1], nn[:, 2]
            t["Nx"], t["Ny"], t["Nz"] = N[:, 0], N[:, 1], N[:, 2]
            verts = p.copy()
            verts[:, 3] = 1
            self.scene.meshes[m] = (verts, t)

    def _local(self, i):
        st = self.state[i]
        return trs_matrix(st["t"], st["q"], st["s"], st["M"]) if st["given"] else st["M"]

    def _walk(self):
        combined, order = {}, []

        def visit(i, parent):
            combined[i] = mat4_mul(parent, self._local(i))
            for c in self.docnodes[i].get("children", []):
                visit(c, combined[i])
            if "mesh" in self.docnodes[i]:
                order.append(i)
        for r in self.roots:
            visit(r, self.root)
        return combined, order

    def update(self, dt):
        moved = False
        for chans in self.channels:
            for ch, smp in chans:
                moved |= ch.update(dt, smp[ch.sampler], self.state)
        if not moved:
            return
        combined, _ = self._walk()
        for inst, i in zip(self.instances, self.order):
            self.scene.instances[inst] = (self.scene.instances[inst][0], combined[i])
        for i in self.order:
            st = self.state[i]
            if st["morphed"]:
                st["morphed"] = False
                self._morph(self.mesh_ids[self.docnodes[i]["mesh"]], st["w"])
        self._pose(combined)
        for i, slot in self.lights.items():
            if slot is not None:
                write_instance_lights(self.scene, self.mesh_ids[self.docnodes[i]["mesh"]], combined[i], self.node_base + i, *slot)

    def _morph(self, m, w):
        """HostMesh::SetPose(weights) (host_mesh.cpp:711-738): positions =
        base + sum of weight x target offset; vertex normals = normalise(base
        + every target's normal offset), the offsets added unweighted as the
        reference adds them; face normals are left as they were."""
        bp, bn, tp, tn = self.morph[m]
        pos = bp.copy()
        nrm = bn.copy()
        for j in range(len(w)):
            pos = (pos + f32(w[j]) * tp[j]).astype(f32)
            nrm = (nrm + tn[j]).astype(f32)
        with np.errstate(invalid="ignore", divide="ignore"):
            nrm = _normalize(nrm)
        verts, t = self.scene.meshes[m]
        verts, t = verts.copy(), t.copy()
        verts[:, :3] = pos
        q = pos.reshape(-1, 3, 3)
        n3 = nrm.reshape(-1, 3, 3)
        t["vertex0"], t["vertex1"], t["vertex2"] = q[:, 0], q[:, 1], q[:, 2]
        t["vN0"], t["vN1"], t["vN2"] = n3[:, 0], n3[:, 1], n3[:, 2]
        self.scene.meshes[m] = (verts, t)


def load_gltf(scene, path, transform=None, animated=False):
    """HostScene::AddScene for a glTF file (host_scene.cpp:249-392): adds the
    file's textures (8-bit PNG images, top row first, neither flipped nor
    linearised), materials, one mesh per glTF mesh (its triangle primitives
    concatenated, BuildFromIndexedData each), and the node hierarchy of scene
    0 under a root node holding `transform`: one instance per node with a
    mesh, in the order HostNode::Update visits them (children before their
    parent, host_node.cpp:141-199), at the combined transform, plus the
    lights of emissive triangles in world space in node order. Returns the
    list of instance indices. Without materials in the file every triangle
    uses scene material 0 (a grey Lambert material is added when the scene
    has none); a primitive without a material gets a grey Lambert one.
    Skinned meshes are posed by their skins; morph targets start at weight 0. With animated=True the GltfScene is returned,
    whose update(dt) plays the file's translation / rotation / scale
    animations (morph weights included) and re-poses skinned and morphed
    meshes."""
    doc, bufs, directory = read_gltf(path)
    tex_ids = []
    for t in doc.get("textures", []):
        if "source" not in t:
            raise ValueError("glTF texture without a source image")
        tex_ids.append(scene.add_texture(texture_from_rgba(_gltf_image_rgba(doc, bufs, doc["images"][t["source"]]))))
    mat_ids = [scene.add_material(_gltf_material(m, tex_ids, scene)) for m in doc.get("materials", [])]
    override = None
    if not mat_ids:
        if len(scene.materials) == 0:
            scene.add_material(F.lambert_material((0.5, 0.5, 0.5)))
        override = 0
    default = [None]

    def prim_material(prim):
        if override is not None:
            return override
        if prim.get("material", -1) < 0:
            if default[0] is None:
                default[0] = scene.add_material(F.lambert_material((0.5, 0.5, 0.5)))
            return default[0]
        return mat_ids[prim["material"]]
    mesh_ids, skin_data, morph_data = [], {}, {}
    for mesh in doc.get("meshes", []):
        parts, skin_parts, morph_parts = [], [], []
        for prim in mesh.get("primitives", []):
            idx = gltf_primitive_indices(doc, bufs, prim)
            if idx is None:
                continue
            at = prim["attributes"]
            P = _gltf_float_attribute(doc, bufs, at["POSITION"], "VEC3", "POSITION")
            Nv = _gltf_float_attribute(doc, bufs, at["NORMAL"], "VEC3", "NORMAL") if "NORMAL" in at else None
            UV = _gltf_float_attribute(doc, bufs, at["TEXCOORD_0"], "VEC2", "TEXCOORD_0") if "TEXCOORD_0" in at else None
            UV2 = _gltf_float_attribute(doc, bufs, at["TEXCOORD_1"], "VEC2", "TEXCOORD_1") if "TEXCOORD_1" in at else None
            if "TANGENT" in at:
                _gltf_float_attribute(doc, bufs, at["TANGENT"], "VEC4", "TANGENT")    # read and unused
            if len(idx) and (idx.min() < 0 or idx.max() >= len(P)):
                raise ValueError("glTF indices outside the vertex range")
            parts.append(indexed_triangles(scene, idx, P, Nv, UV, UV2, prim_material(prim)))
            nt = len(mesh.get("weights", []))
            if nt and len(prim.get("targets", [])) >= nt:
                # morph targets (host_mesh.cpp:462-508): per target, position and normal offsets per corner
                flat = idx.reshape(-1)
                tp = np.stack([_gltf_float_attribute(doc, bufs, prim["targets"][i]["POSITION"], "VEC3", "target POSITION")[flat]
                               if "POSITION" in prim["targets"][i] else np.zeros((len(flat), 3), f32) for i in range(nt)])
                tn = np.stack([_gltf_float_attribute(doc, bufs, prim["targets"][i]["NORMAL"], "VEC3", "target NORMAL")[flat]
                               if "NORMAL" in prim["targets"][i] else np.zeros((len(flat), 3), f32) for i in range(nt)])
                if Nv is not None:
                    bn = Nv[flat]
                else:
                    # no NORMAL attribute: the base normals are the corner normals the load produced
                    # (the face normal), so a morph update never normalises a zero vector
                    lt = parts[-1][1]
                    bn = np.stack([lt["vN0"], lt["vN1"], lt["vN2"]], axis=1).reshape(-1, 3).astype(f32)
                morph_parts.append((P[flat], bn, tp, tn))
            else:
                morph_parts.append(None)
            if "JOINTS_0" in at and "WEIGHTS_0" in at:
                J = gltf_accessor(doc, bufs, at["JOINTS_0"]).astype(np.int64)
                Wt = _gltf_float_attribute(doc, bufs, at["WEIGHTS_0"], "VEC4", "WEIGHTS_0")
                norm = f32(1) / (((Wt[:, 0] + Wt[:, 1]) + Wt[:, 2]) + Wt[:, 3])     # weights scaled to sum 1
                skin_parts.append((J[idx.reshape(-1)], (Wt * norm[:, None]).astype(f32)[idx.reshape(-1)]))
            else:
                skin_parts.append(None)
        if parts:
            verts = np.concatenate([p[0] for p in parts])
            tris = np.concatenate([p[1] for p in parts])
        else:
            verts, tris = np.zeros((0, 4), f32), F.new_tris(0)
        mesh_ids.append(scene.add_mesh(verts, tris))
        if morph_parts and all(x is not None for x in morph_parts):
            morph_data[mesh_ids[-1]] = tuple(np.concatenate([x[c] for x in morph_parts], axis=1 if c >= 2 else 0) for c in range(4))
        if skin_parts and all(x is not None for x in skin_parts):
            skin_data[mesh_ids[-1]] = (np.concatenate([x[0] for x in skin_parts]), np.concatenate([x[1] for x in skin_parts]))
    node_base = getattr(scene, "node_count", len(scene.instances)) + 1
    scene.node_count = node_base + len(doc.get("nodes", []))
    root = F.IDENTITY.copy() if transform is None else np.asarray(transform, f32).reshape(16)
    g = GltfScene(scene, doc, bufs, mesh_ids, node_base, root, skin_data, morph_data)
    return g.instances if not animated else g


def mat4_mul(a, b):
    """operator*(mat4, mat4) (lib/platform/system.cpp:57-69), fp32, cells
    row-major: r[i][j] = ((a[i][0] b[0][j] + a[i][1] b[1][j]) + a[i][2] b[2][j]) + a[i][3] b[3][j]."""
    a = np.asarray(a, f32).reshape(4, 4)
    b = np.asarray(b, f32).reshape(4, 4)
    r = ((a[:, 0:1] * b[0:1, :] + a[:, 1:2] * b[1:2, :]) + a[:, 2:3] * b[2:3, :]) + a[:, 3:4] * b[3:4, :]
    return r.astype(f32).reshape(16)


def _transform_points(M, p, w):
    """float4(p, w) * mat4 (system.cpp:97-103), rows 0..3."""
    M = np.asarray(M, f32).reshape(4, 4)
    return np.stack([((M[r, 0] * p[..., 0] + M[r, 1] * p[..., 1]) + M[r, 2] * p[..., 2]) + M[r, 3] * f32(w)
                     for r in range(4)], -1).astype(f32)


def _transformed_tris(tris, M):
    """TransformedHostTri (host_node.cpp:19-30): vertices (v, 1) * T, the
    face normal normalize((N, 0) * T) as a float4."""
    out = tris.copy()
    for k in ("vertex0", "vertex1", "vertex2"):
        out[k] = _transform_points(M, tris[k], 1.0)[:, :3]
    N4 = _transform_points(M, np.stack([tris["Nx"], tris["Ny"], tris["Nz"]], -1), 0.0)
    d = (((N4[:, 0] * N4[:, 0] + N4[:, 1] * N4[:, 1]) + N4[:, 2] * N4[:, 2]) + N4[:, 3] * N4[:, 3]).astype(f32)
    N4 = (N4 * (f32(1) / np.sqrt(d))[:, None]).astype(f32)
    out["Nx"], out["Ny"], out["Nz"] = N4[:, 0], N4[:, 1], N4[:, 2]
    return out


def instance_lights(scene, mesh, transform, node_id):
    """HostNode::PrepareLights / UpdateLights (host_node.cpp:203-255) for one
    instance: every triangle whose material is emissive (a colour component
    above 1, HostMaterial::IsEmissive host_material.h:82) gets its
    object-space area (CoreTri::UpdateArea) and its index in the light list
    (ltriIdx), and a CoreLightTri of the triangle transformed by the
    instance's transform is appended, radiance = material colour, instIdx =
    the node's id. Returns (first light index, triangle ids) or None."""
    verts, tris = scene.meshes[mesh]
    if len(tris) == 0:
        return None
    col = scene.materials["color"]["value"][tris["material"]]
    sel = np.nonzero((col > 1).any(1))[0]
    if len(sel) == 0:
        return None
    sub = tris[sel]
    scenes.add_emitters(scene, sub, np.zeros(3, f32), node_id, tri_ids=sel)   # areas, ltriIdx, light slots
    tris[sel] = sub
    first = len(scene.tri_lights) - len(sel)
    write_instance_lights(scene, mesh, transform, node_id, first, sel)
    return first, sel


def write_instance_lights(scene, mesh, transform, node_id, first, sel):
    """Rewrites an instance's light slots from its (new) transform."""
    _, tris = scene.meshes[mesh]
    world = _transformed_tris(tris[sel], transform)
    for k in range(len(sel)):
        rad = scene.materials["color"]["value"][tris["material"][sel[k]]]
        scene.tri_lights[first + k] = scenes.light_triangle(world[k], rad, int(sel[k]), node_id)[0]


def add_instance_with_lights(scene, mesh, transform=None):
    """SetInstance plus the lights of HostNode(meshIdx, transform)
    (host_node.cpp:40-47): the instance's emissive triangles become
    CoreLightTris in world space. Returns the instance index (which is also
    the node id when every node of the scene is an instance)."""
    inst = scene.add_instance(mesh, transform)
    instance_lights(scene, mesh, scene.instances[inst][1], inst)
    return inst
