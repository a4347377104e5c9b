"""numpy mirrors of the RenderCore_HIP boundary structs (include/lh2hip_types.h).

Each dtype reproduces the byte layout of the Lighthouse 2 type of the same name
(lib/RenderSystem/common_classes.h:57-97 CoreTri, :219-328 CoreMaterial,
:365-446 lights, :452-475 ViewPyramid; lib/RenderSystem/core_api_base.h:30-64
CoreStats), so arrays built here are passed to the C ABI as plain pointers.
"""
import numpy as np

F3 = ("<f4", (3,))
F2 = ("<f4", (2,))
U2 = ("<u4", (2,))

_CORETRI_FIELDS = [
    ("u0", "<f4"), ("u1", "<f4"), ("u2", "<f4"), ("ltriIdx", "<i4"),
    ("v0", "<f4"), ("v1", "<f4"), ("v2", "<f4"), ("material", "<u4"),
    ("vN0", F3), ("Nx", "<f4"), ("vN1", F3), ("Ny", "<f4"), ("vN2", F3), ("Nz", "<f4"),
    ("T", F3), ("area", "<f4"), ("B", F3), ("invArea", "<f4"),
    ("alpha", F3), ("LOD", "<f4"),
    ("vertex0", F3), ("dummy0", "<f4"), ("vertex1", F3), ("dummy1", "<f4"), ("vertex2", F3), ("dummy2", "<f4"),
    ("u1_0", "<f4"), ("u1_1", "<f4"), ("u1_2", "<f4"), ("dummy3", "<f4"),
    ("v1_0", "<f4"), ("v1_1", "<f4"), ("v1_2", "<f4"), ("dummy4", "<f4"),
]
CoreTri = np.dtype(_CORETRI_FIELDS)
assert CoreTri.itemsize == 208

Vec3Value = np.dtype({"names": ["value", "textureID", "scale", "uvscale", "uvoffset", "size"],
                      "formats": [F3, "<i4", "<f4", F2, F2, U2],
                      "offsets": [0, 12, 16, 24, 32, 40], "itemsize": 48})
ScalarValue = np.dtype({"names": ["value", "textureID", "component", "scale", "uvscale", "uvoffset", "size"],
                        "formats": ["<f4", "<i4", "<i4", "<f4", F2, F2, U2],
                        "offsets": [0, 4, 8, 12, 16, 24, 32], "itemsize": 40})

_MAT_MEMBERS = (
    [("color", Vec3Value), ("detailColor", Vec3Value), ("normals", Vec3Value), ("detailNormals", Vec3Value),
     ("flags", np.dtype("<u4")), ("absorption", Vec3Value)]
    + [(n, ScalarValue) for n in ("metallic", "subsurface", "specular", "roughness", "specularTint", "anisotropic",
                                  "sheen", "sheenTint", "clearcoat", "clearcoatGloss", "transmission", "eta",
                                  "reflection", "refraction", "ior")]
    + [("pbrtMaterialType", np.dtype("u1")), ("urough", ScalarValue), ("vrough", ScalarValue),
       ("Ks", Vec3Value), ("eta_rgb", Vec3Value), ("sigma", ScalarValue), ("thin", np.dtype("u1")),
       ("specTrans", ScalarValue), ("diffTrans", ScalarValue), ("scatterDistance", Vec3Value),
       ("flatness", ScalarValue), ("Kr", Vec3Value), ("opacity", Vec3Value)]
)


def _c_layout(members):
    """Offsets under C rules: each member aligned to its own alignment
    (8 for the value structs, which hold 8-aligned 2-vectors)."""
    names, formats, offsets, off, maxal = [], [], [], 0, 1
    for name, dt in members:
        al = 8 if dt.names and dt.itemsize in (40, 48) else dt.alignment
        off = (off + al - 1) // al * al
        names.append(name), formats.append(dt), offsets.append(off)
        off += dt.itemsize
        maxal = max(maxal, al)
    size = (off + maxal - 1) // maxal * maxal
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size})


CoreMaterial = _c_layout(_MAT_MEMBERS)
assert CoreMaterial.itemsize == 1344

CoreLightTri = np.dtype([("centre", F3), ("energy", "<f4"), ("N", F3), ("area", "<f4"),
                         ("radiance", F3), ("dummy2", "<i4"), ("vertex0", F3), ("triIdx", "<i4"),
                         ("vertex1", F3), ("instIdx", "<i4"), ("vertex2", F3), ("dummy1", "<i4")])
CorePointLight = np.dtype([("position", F3), ("energy", "<f4"), ("radiance", F3), ("dummy", "<i4")])
CoreSpotLight = np.dtype([("position", F3), ("cosInner", "<f4"), ("radiance", F3), ("cosOuter", "<f4"),
                          ("direction", F3), ("dummy", "<i4")])
CoreDirectionalLight = np.dtype([("direction", F3), ("energy", "<f4"), ("radiance", F3), ("dummy", "<i4")])
assert CoreLightTri.itemsize == 96 and CorePointLight.itemsize == 32
assert CoreSpotLight.itemsize == 48 and CoreDirectionalLight.itemsize == 32

ViewPyramid = np.dtype([("pos", F3), ("p1", F3), ("p2", F3), ("p3", F3), ("aperture", "<f4"),
                        ("spreadAngle", "<f4"), ("imagePlane", "<f4"), ("focalDistance", "<f4"), ("distortion", "<f4")])
assert ViewPyramid.itemsize == 68

CoreStats = np.dtype({"names": ["deviceName", "SMcount", "ccMajor", "ccMinor", "VRAM", "argb32TexelCount",
                                "argb128TexelCount", "nrm32TexelCount", "bvhBuildTime", "totalRays",
                                "totalExtensionRays", "totalShadowRays", "renderTime", "frameOverhead",
                                "primaryRayCount", "traceTime0", "bounce1RayCount", "traceTime1", "deepRayCount",
                                "traceTimeX", "shadowTraceTime", "shadeTime", "filterTime", "probedInstid",
                                "probedTriid", "probedDist", "probedWorldPos"],
                      "formats": ["<u8", "<u4", "<u4", "<u4", "<u4", "<u4", "<u4", "<u4", "<f4", "<u4", "<u4",
                                  "<u4", "<f4", "<f4", "<u4", "<f4", "<u4", "<f4", "<u4", "<f4", "<f4", "<f4",
                                  "<f4", "<i4", "<i4", "<f4", F3],
                      "offsets": [0, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80,
                                  84, 88, 92, 96, 100, 104, 108],
                      "itemsize": 120})

# Host texture descriptor (common_classes.h:336-359, host variant): texel
# pointer, width, height, flags (8 = HDR), pixelCount (all MIP levels),
# firstPixel, MIPlevels, storage (0 ARGB32, 1 ARGB128, 2 NRM32).
CoreTexDesc = np.dtype([("data", "<u8"), ("width", "<u4"), ("height", "<u4"), ("flags", "<u4"), ("pixelCount", "<u4"),
                        ("firstPixel", "<u4"), ("MIPlevels", "<u4"), ("storage", "<i4")], align=True)
ARGB32, ARGB128, NRM32 = 0, 1, 2
MIPLEVELCOUNT = 5   # common_settings.h:57


def texture_descs(textures):
    """CoreTexDesc array for a list of textures (dicts with width, height,
    flags, storage and `pixels`: uint32 words, or (n,4) float32 for ARGB128,
    all MIP levels included). The texel arrays must outlive the call that
    receives the descriptors; they are returned alongside."""
    d = np.zeros(len(textures), CoreTexDesc)
    keep = []
    for i, t in enumerate(textures):
        px = np.ascontiguousarray(t["pixels"], np.float32 if t["storage"] == ARGB128 else np.uint32)
        keep.append(px)
        d[i]["data"] = px.ctypes.data
        d[i]["width"], d[i]["height"], d[i]["flags"] = t["width"], t["height"], t.get("flags", 0)
        d[i]["pixelCount"] = len(px) if px.ndim == 1 else px.shape[0]
        d[i]["MIPlevels"] = t.get("mips", MIPLEVELCOUNT)
        d[i]["storage"] = t["storage"]
    return d, keep


# Build-tree node of the BVH download (lighthouse2_amd/csrc/bvh.h BuildNode).
BuildNode = np.dtype([("lo", F3), ("leftFirst", "<i4"), ("hi", F3), ("count", "<i4")])


def new_tris(n):
    """CoreTri array initialised like the reference's CoreTri() constructor
    (all zero, ltriIdx = -1)."""
    t = np.zeros(n, dtype=CoreTri)
    t["ltriIdx"] = -1
    return t


def new_materials(n):
    """CoreMaterial array with HostMaterial defaults (host_material.h:45-62):
    every value 1e-32 ("not set"), texture ids -1, scale 1, uv scale 1."""
    m = np.zeros(n, dtype=CoreMaterial)
    for name, dt in _MAT_MEMBERS:
        if dt in (Vec3Value, ScalarValue):
            m[name]["value"] = np.float32(1e-32)
            m[name]["textureID"] = -1
            m[name]["scale"] = 1.0
            m[name]["uvscale"] = 1.0
    return m


def lambert_material(color, roughness=1.0, transmission=0.0):
    """One CoreMaterial for the reference estimator's Lambert BSDF: an explicit
    roughness of 1 (ROUGHNESS = max(0.001, byte/255), tools_shared.h:43; an unset
    roughness would make the surface a pure specular)."""
    m = new_materials(1)
    m["color"]["value"][0] = np.asarray(color, np.float32)
    m["roughness"]["value"] = roughness
    m["transmission"]["value"] = transmission
    m["absorption"]["value"][0] = 0.0
    for name in ("metallic", "subsurface", "specular", "specularTint", "anisotropic", "sheen", "sheenTint",
                 "clearcoat", "clearcoatGloss"):
        m[name]["value"] = 0.0
    return m


def disney_material(color, metallic=0.0, subsurface=0.0, specular=0.5, roughness=0.5, specularTint=0.0, anisotropic=0.0,
                    sheen=0.0, sheenTint=0.0, clearcoat=0.0, clearcoatGloss=1.0, transmission=0.0, eta=1.5,
                    absorption=(0.0, 0.0, 0.0)):
    """One CoreMaterial with the Disney BSDF's parameters (HostMaterial fields
    of host_material.h, packed by SetMaterials into the 8-bit parameter words)."""
    m = new_materials(1)
    m["color"]["value"][0] = np.asarray(color, np.float32)
    m["absorption"]["value"][0] = np.asarray(absorption, np.float32)
    for name, v in (("metallic", metallic), ("subsurface", subsurface), ("specular", specular), ("roughness", roughness),
                    ("specularTint", specularTint), ("anisotropic", anisotropic), ("sheen", sheen), ("sheenTint", sheenTint),
                    ("clearcoat", clearcoat), ("clearcoatGloss", clearcoatGloss), ("transmission", transmission), ("eta", eta)):
        m[name]["value"] = v
    return m


IDENTITY = np.eye(4, dtype=np.float32).reshape(16)
