"""Scene files -> the boundary's arrays (CoreTri soup, CoreMaterial, textures,
CoreLightTri), the producers the reference's RenderSystem runs before it
calls a core:

* Wavefront OBJ + MTL: HostMesh::LoadGeometryFromOBJ (lib/RenderSystem/
  host_mesh.cpp:131-321) with HostMaterial::ConvertFrom (host_material.cpp:35-71);
* PNG textures: HostTexture::Load (host_texture.cpp:200-241) for 8-bit images,
  FLIPPED / LINEARIZED as the OBJ path requests them, MIP chain of
  ConstructMIPmaps;
* lights of emissive triangles: HostNode::PrepareLights (host_node.cpp:203-233),
  a material being emissive when a colour component exceeds 1
  (HostMaterial::IsEmissive, host_material.h:82).

The file parsers are this project's (the reference uses tinyobjloader and
FreeImage): polygons are split as a fan, which is what tinyobjloader's ear
clipping gives for convex polygons; PNG is decoded with zlib (8-bit grey,
grey + alpha, RGB, RGBA and palette images, not interlaced). All float
arithmetic of the producers is float32 in the reference's operation order.
"""
import os
import struct
import zlib

import numpy as np

from . import formats as F
from . import scenes

f32 = np.float32


# ---- PNG -------------------------------------------------------------------
def decode_png(data):
    """PNG bytes -> (h, w, 4) uint8 RGBA, rows top to bottom."""
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
    if depth != 8 or interlace != 0:
        raise ValueError(f"PNG bit depth {depth} / interlace {interlace}: only 8-bit, non-interlaced images are read")
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    stride = w * chans
    rows = raw.reshape(h, stride + 1)
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f, line = rows[y, 0], rows[y, 1:].astype(np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 255
        elif f in (1, 3, 4):
            cur = np.zeros(stride, np.int32)
            for x in range(stride):                 # depends on the reconstructed left neighbour
                a = cur[x - chans] if x >= chans else 0
                b = prev[x]
                c = prev[x - chans] if x >= chans else 0
                if f == 1:
                    p = a
                elif f == 3:
                    p = (a + b) >> 1
                else:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[x] = (line[x] + p) & 255
        else:
            raise ValueError(f"PNG filter type {f}")
        out[y] = cur
        prev = cur
    px = out.reshape(h, w, chans)
    rgba = np.full((h, w, 4), 255, np.uint8)
    if ctype == 0:
        rgba[..., :3] = px
    elif ctype == 4:
        rgba[..., :3], rgba[..., 3] = px[..., :1], px[..., 1]
    elif ctype == 2:
        rgba[..., :3] = px
    elif ctype == 6:
        rgba[:] = px
    else:
        if plte is None:
            raise ValueError("palette PNG without PLTE")
        idx = px[..., 0]
        rgba[..., :3] = plte[idx]
        if trns is not None:
            alpha = np.full(256, 255, np.uint8)
            alpha[:len(trns)] = trns
            rgba[..., 3] = alpha[idx]
    return rgba


def texture_from_rgba(rgba, flipped=False, linearized=False, normal_map=False):
    """HostTexture::Load for an 8-bit image (host_texture.cpp:225-241): the
    image library delivers rows bottom-up, so FLIPPED keeps the bottom row
    first and no flag stores the top row first; LINEARIZED squares r, g, b
    ((c * c) >> 8, sRGBtoLinear :73-81); 5 MIP levels."""
    a = np.asarray(rgba, np.uint8)
    h, w = a.shape[:2]
    rows = a[::-1] if flipped else a
    c = rows.astype(np.uint32)
    r, g, b, al = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    if linearized:
        r, g, b = (r * r) >> 8, (g * g) >> 8, (b * b) >> 8
    px = scenes.pack_texels(r, g, b, al)
    return {"width": w, "height": h, "flags": 0, "storage": F.NRM32 if normal_map else F.ARGB32,
            "pixels": scenes.mip_chain(px)}


def load_texture(path, flipped=False, linearized=False, normal_map=False):
    if not path.lower().endswith(".png"):
        raise ValueError(f"{path}: only PNG textures are read")
    with open(path, "rb") as fh:
        return texture_from_rgba(decode_png(fh.read()), flipped, linearized, normal_map)


# ---- OBJ / MTL ----------------------------------------------------------------
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
            if k in ("map_bump", "bump") and "-bm" in args:
                cur["bump_height"] = float(args[args.index("-bm") + 1])
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
        raise ValueError(f"{m['name']}: bump maps (height to normal conversion) are not read; give a normal map (norm)")
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
        with open(os.path.join(directory, lib), "r") as fh:
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


def add_instance_with_lights(scene, mesh, transform=None):
    """SetInstance plus HostNode::PrepareLights for an identity-transformed
    instance: every triangle whose material is emissive becomes a
    CoreLightTri of radiance = material colour. Returns the instance index."""
    if transform is not None and not np.array_equal(np.asarray(transform, f32).reshape(16), F.IDENTITY):
        raise ValueError("lights of transformed instances are not produced here")
    inst = scene.add_instance(mesh, transform)
    verts, tris = scene.meshes[mesh]
    col = scene.materials["color"]["value"][tris["material"]]
    emissive = (col > 1).any(1)
    for m in np.unique(tris["material"][emissive]):
        sel = np.nonzero(tris["material"] == m)[0]
        sub = tris[sel]
        scenes.add_emitters(scene, sub, scene.materials["color"]["value"][m], inst, tri_ids=sel)
        tris[sel] = sub
    return inst
