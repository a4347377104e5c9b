"""Host-side scene producers for RenderCore_HIP workloads.

These restate the parts of the Lighthouse 2 host scene model that create the
data crossing the core boundary, in float32 with the reference's operation
order:
  add_quad          HostScene::AddQuad        lib/RenderSystem/host_scene.cpp:401-451
  add_triangle      HostScene::AddTriToMesh   lib/RenderSystem/host_scene.cpp:225-244
  light_triangle    HostTriLight + ConvertToCoreLightTri  lib/RenderSystem/host_light.cpp:25-59
  camera_view       Camera::LookAt + GetView  lib/RenderSystem/camera.cpp:68-71,107-128,
                                              mat4::LookAt common_types.h:519-544
  mip_chain         HostTexture::ConstructMIPmaps lib/RenderSystem/host_texture.cpp:131-153
  triangle_lod      triangle LOD of HostMesh  lib/RenderSystem/host_mesh.cpp:308-317
and generate the seeded synthetic scenes the benchmarks and parity tests use.
"""
import numpy as np

from . import formats as F

f32 = np.float32


def _normalize(v):
    v = np.asarray(v, f32)
    d = f32(v[0] * v[0] + v[1] * v[1]) + f32(v[2] * v[2])
    inv = f32(1.0) / np.sqrt(f32(d))
    return (v * inv).astype(f32)


def _cross(a, b):
    a, b = np.asarray(a, f32), np.asarray(b, f32)
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], f32)


def concat(arrays, dtype):
    """Concatenate structured arrays keeping `dtype`'s padded C layout
    (np.concatenate would repack the fields)."""
    out = np.zeros(sum(len(a) for a in arrays), dtype)
    k = 0
    for a in arrays:
        out[k:k + len(a)] = a
        k += len(a)
    return out


class Scene:
    """Meshes (vertex soup + CoreTri), instances, materials and lights, in the
    order RenderSystem::SynchronizeSceneData hands them to a core."""

    def __init__(self):
        self.meshes = []          # (verts (3T,4) f32, CoreTri[T] or None)
        self.instances = []       # (mesh index, (16,) f32 row-major transform)
        self.materials = F.new_materials(0)
        self.tri_lights = np.zeros(0, F.CoreLightTri)
        self.point_lights = np.zeros(0, F.CorePointLight)
        self.spot_lights = np.zeros(0, F.CoreSpotLight)
        self.dir_lights = np.zeros(0, F.CoreDirectionalLight)
        self.textures = []        # dicts: width, height, flags, storage, pixels (all MIP levels)
        self.sky = None           # ((h, w, 3) float32 texels, (16,) float32 world-to-sky transform)

    def add_texture(self, tex):
        self.textures.append(tex)
        return len(self.textures) - 1

    def add_material(self, mat):
        self.materials = concat([self.materials, mat], F.CoreMaterial)
        return len(self.materials) - 1

    def add_mesh(self, verts, tris):
        self.meshes.append((np.ascontiguousarray(verts, f32), tris))
        return len(self.meshes) - 1

    def add_instance(self, mesh, transform=None):
        t = F.IDENTITY.copy() if transform is None else np.asarray(transform, f32).reshape(16)
        self.instances.append((mesh, t))
        return len(self.instances) - 1

    @property
    def triangle_count(self):
        return sum(len(v) // 3 for v, _ in self.meshes)


def add_quad(N, pos, width, height, mat):
    """Two triangles of a quad (HostScene::AddQuad): vertex soup (6,4) and CoreTri[2]."""
    N = _normalize(N)
    pos = np.asarray(pos, f32)
    tmp = np.array([0, 1, 0], f32) if abs(N[0]) > 0.9 else np.array([1, 0, 0], f32)
    T = (f32(0.5) * f32(width)) * _normalize(_cross(N, tmp))
    B = (f32(0.5) * f32(height)) * _normalize(_cross(_normalize(T), N))
    T, B = T.astype(f32), B.astype(f32)
    c = [pos - B - T, pos + B - T, pos - B + T, pos + B - T, pos + B + T, pos - B + T]
    verts = np.ones((6, 4), f32)
    for i, p in enumerate(c):
        verts[i, :3] = p
    tris = F.new_tris(2)
    tris["material"] = mat
    for k in ("vN0", "vN1", "vN2"):
        tris[k] = N
    tris["Nx"], tris["Ny"], tris["Nz"] = N[0], N[1], N[2]
    tris["vertex0"][0], tris["vertex1"][0], tris["vertex2"][0] = verts[0, :3], verts[1, :3], verts[2, :3]
    tris["vertex0"][1], tris["vertex1"][1], tris["vertex2"][1] = verts[3, :3], verts[4, :3], verts[5, :3]
    # tri.T = T / (0.5 * height); tri.B = B / (0.5 * width) in double (0.5 * height is a double)
    tris["T"] = (T.astype(np.float64) / (0.5 * float(height))).astype(f32)
    tris["B"] = (B.astype(np.float64) / (0.5 * float(width))).astype(f32)
    return verts, tris


def add_triangles(v0, v1, v2, mat):
    """Flat-shaded triangles (HostScene::AddTriToMesh), vectorised: v0/v1/v2 (T,3)."""
    v0, v1, v2 = (np.asarray(a, f32) for a in (v0, v1, v2))
    e1, e2 = v1 - v0, v2 - v0
    n = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1],
                  e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], 1).astype(f32)
    d = (n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2]
    inv = (f32(1.0) / np.sqrt(d)).astype(f32)
    N = (n * inv[:, None]).astype(f32)
    T = len(v0)
    verts = np.ones((3 * T, 4), f32)
    verts[0::3, :3], verts[1::3, :3], verts[2::3, :3] = v0, v1, v2
    tris = F.new_tris(T)
    tris["material"] = mat
    tris["vN0"], tris["vN1"], tris["vN2"] = N, N, N
    tris["Nx"], tris["Ny"], tris["Nz"] = N[:, 0], N[:, 1], N[:, 2]
    tris["vertex0"], tris["vertex1"], tris["vertex2"] = v0, v1, v2
    # tangent frame from the first edge, as HostMesh builds it for triangles without uv
    # (host_mesh.cpp:634-637)
    d = (e1[:, 0] * e1[:, 0] + e1[:, 1] * e1[:, 1]) + e1[:, 2] * e1[:, 2]
    Tn = (e1 * (f32(1.0) / np.sqrt(d))[:, None]).astype(f32)
    c = np.stack([N[:, 1] * Tn[:, 2] - N[:, 2] * Tn[:, 1], N[:, 2] * Tn[:, 0] - N[:, 0] * Tn[:, 2],
                  N[:, 0] * Tn[:, 1] - N[:, 1] * Tn[:, 0]], 1).astype(f32)
    dc = (c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1]) + c[:, 2] * c[:, 2]
    tris["T"] = Tn
    tris["B"] = (c * (f32(1.0) / np.sqrt(dc))[:, None]).astype(f32)
    return verts, tris


def light_triangle(tri, radiance, tri_idx, inst_idx):
    """CoreLightTri of an emissive triangle (host_light.cpp:25-59)."""
    L = np.zeros(1, F.CoreLightTri)
    v0, v1, v2 = (np.asarray(tri[k], f32) for k in ("vertex0", "vertex1", "vertex2"))
    L["vertex0"], L["vertex1"], L["vertex2"] = v0, v1, v2
    L["centre"] = f32(0.333333) * ((v0 + v1) + v2)
    L["N"] = np.array([tri["Nx"], tri["Ny"], tri["Nz"]], f32)
    ln = lambda a: np.sqrt(f32(a[0] * a[0] + a[1] * a[1]) + f32(a[2] * a[2])).astype(f32)
    a, b, c = ln(v1 - v0), ln(v2 - v1), ln(v0 - v2)
    s = f32(f32(a + b) + c) * f32(0.5)
    area = np.sqrt(f32(f32(f32(s * f32(s - a)) * f32(s - b)) * f32(s - c))).astype(f32)
    L["area"] = area
    rad = np.asarray(radiance, f32)
    L["radiance"] = rad
    E = rad * area
    L["energy"] = f32(f32(E[0] + E[1]) + E[2])
    L["triIdx"], L["instIdx"] = tri_idx, inst_idx
    return L


def add_emitters(scene, tris, radiance, inst_idx, tri_ids=None):
    """HostNode::PrepareLights (host_node.cpp:203-233) for one emissive mesh:
    each triangle gets its object-space area (CoreTri::UpdateArea, Heron's
    formula in fp32, common_classes.h:89-96) and its index in the light list
    (ltriIdx), and a CoreLightTri is appended. tri_ids: the triangles' indices
    in their mesh (default 0, 1, ...)."""
    ln = lambda a: np.sqrt(f32(a[0] * a[0] + a[1] * a[1]) + f32(a[2] * a[2])).astype(f32)
    lights = []
    for k in range(len(tris)):
        v0, v1, v2 = (np.asarray(tris[k][n], f32) for n in ("vertex0", "vertex1", "vertex2"))
        a, b, c = ln(v1 - v0), ln(v2 - v1), ln(v0 - v2)
        sp = f32(f32(a + b) + c) * f32(0.5)
        tris[k]["area"] = np.sqrt(f32(f32(f32(sp * f32(sp - a)) * f32(sp - b)) * f32(sp - c))).astype(f32)
        tris[k]["ltriIdx"] = len(scene.tri_lights) + k
        lights.append(light_triangle(tris[k], radiance, k if tri_ids is None else int(tri_ids[k]), inst_idx))
    scene.tri_lights = concat([scene.tri_lights] + lights, F.CoreLightTri)


def look_at(pos, target):
    """Camera transform of mat4::LookAt(pos, target, up = (0,1,0))."""
    z = _normalize(np.asarray(target, f32) - np.asarray(pos, f32))
    x = _normalize(_cross(z, _normalize([0, 1, 0])))
    y = _cross(x, z)
    return x, y, z


def camera_view(pos, target, fov=40.0, width=1920, height=1080, focal=5.0, aperture=0.0, distortion=0.0):
    """ViewPyramid of Camera::GetView (camera.cpp:107-128) with aspect = width/height."""
    right, up, forward = look_at(pos, target)
    pos = np.asarray(pos, f32)
    aspect = f32(width / height)
    PI = f32(3.14159265358979323846264)
    v = np.zeros(1, F.ViewPyramid)
    v["pos"] = pos
    v["spreadAngle"] = f32(f32(f32(fov) * PI) / f32(180)) / f32(height)
    screen = np.tan(f32(f32(fov) / f32(2)) / f32(f32(180) / PI)).astype(f32)
    fd = f32(focal)
    C = pos + fd * forward
    sr = ((screen * right) * fd) * aspect
    su = (screen * fd) * up
    v["p1"] = (C - sr) + su
    v["p2"] = (C + sr) + su
    v["p3"] = (C - sr) - su
    v["aperture"], v["focalDistance"], v["distortion"] = aperture, focal, distortion
    unit_r, unit_u = (screen * right) * aspect, screen * up
    p1u, p2u, p3u = (C - unit_r) + unit_u, (C + unit_r) + unit_u, (C - unit_r) - unit_u
    v["imagePlane"] = float(np.linalg.norm(p1u - p2u) * np.linalg.norm(p1u - p3u))
    return v


def uv_sphere(center, radius, slices, stacks):
    """Triangles of a latitude/longitude sphere: 2 * slices * (stacks - 1)
    triangles, wound outward. Returns v0, v1, v2 arrays (T,3) float32."""
    th = np.linspace(0.0, np.pi, stacks + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, slices + 1)
    P = np.stack([np.sin(th)[:, None] * np.cos(ph)[None, :],
                  np.cos(th)[:, None] * np.ones_like(ph)[None, :],
                  np.sin(th)[:, None] * np.sin(ph)[None, :]], -1) * radius + np.asarray(center)
    a, b, c = [], [], []
    for i in range(stacks):
        for j in range(slices):
            p00, p01, p10, p11 = P[i, j], P[i, j + 1], P[i + 1, j], P[i + 1, j + 1]
            if i > 0:
                a.append(p00), b.append(p01), c.append(p10)
            if i < stacks - 1:
                a.append(p01), b.append(p11), c.append(p10)
    return (np.asarray(a, f32), np.asarray(b, f32), np.asarray(c, f32))


def cornell_box(seed=1234, sphere_tris=20000):
    """Cornell-box parity scene: five axis-aligned walls of the unit box (open
    towards -z), a tessellated sphere of radius 0.3 at the centre (flat normals)
    and a 0.2 x 0.2 quad light under the ceiling facing down, radiance 15.
    Wall and sphere albedo U(0.2, 0.8) per channel from `seed`; roughness 1.
    Each wall, the sphere and the light are separate meshes with identity
    instances, as HostScene creates them."""
    rng = np.random.default_rng(seed)
    s = Scene()
    walls = [((0, 1, 0), (0.5, 0.0, 0.5)), ((0, -1, 0), (0.5, 1.0, 0.5)), ((0, 0, -1), (0.5, 0.5, 1.0)),
             ((1, 0, 0), (0.0, 0.5, 0.5)), ((-1, 0, 0), (1.0, 0.5, 0.5))]
    for N, pos in walls:
        m = s.add_material(F.lambert_material(rng.uniform(0.2, 0.8, 3).astype(f32)))
        v, t = add_quad(N, pos, 1.0, 1.0, m)
        s.add_instance(s.add_mesh(v, t))
    m = s.add_material(F.lambert_material(rng.uniform(0.2, 0.8, 3).astype(f32)))
    slices = 100
    stacks = sphere_tris // (2 * slices) + 1
    a, b, c = uv_sphere((0.5, 0.5, 0.5), 0.3, slices, stacks)
    v, t = add_triangles(a, b, c, m)
    s.add_instance(s.add_mesh(v, t))
    lm = s.add_material(F.lambert_material((15.0, 15.0, 15.0)))
    v, t = add_quad((0, -1, 0), (0.5, 0.999, 0.5), 0.2, 0.2, lm)
    li = s.add_instance(s.add_mesh(v, t))
    add_emitters(s, t, (15.0, 15.0, 15.0), li)
    return s


def disney_cornell_box(seed=1234, sphere_tris=4000):
    """The Cornell box with Disney materials: plastic floor, brushed-metal back
    wall (anisotropic), clear-coated left wall, sheen + subsurface right wall,
    diffuse ceiling, and a rough glass sphere with absorption."""
    s = cornell_box(seed=seed, sphere_tris=sphere_tris)
    rng = np.random.default_rng(seed + 1)
    c = lambda: rng.uniform(0.25, 0.85, 3).astype(f32)
    mats = [F.disney_material(c(), specular=0.5, roughness=0.35),
            F.disney_material(c(), specular=0.0, roughness=0.9),
            F.disney_material(c(), metallic=1.0, roughness=0.45, anisotropic=0.5),
            F.disney_material(c(), specular=0.3, roughness=0.6, clearcoat=1.0, clearcoatGloss=0.7),
            F.disney_material(c(), specular=0.2, roughness=0.8, sheen=0.8, sheenTint=0.5, subsurface=0.5),
            F.disney_material((0.9, 0.95, 1.0), roughness=0.15, transmission=1.0, eta=1.5, absorption=(0.3, 0.1, 0.05))]
    for k, m in enumerate(mats):
        s.materials[k] = m[0]
    return s


def displaced_sphere(n_tris, radius=1.0, seed=0, bumps=12, amplitude=0.15):
    """Noise-displaced sphere of about n_tris triangles: a latitude/longitude
    sphere whose radius is modulated by a seeded sum of smooth plane waves;
    spherical uv (4 x 2 repeats). Returns v0, v1, v2 (T,3) and uv (T,3,2), float32."""
    rng = np.random.default_rng(seed)
    stacks = max(4, int(round(np.sqrt(n_tris / 4.0))))
    slices = max(4, int(round(n_tris / (2.0 * (stacks - 1)))))
    th = np.linspace(0.0, np.pi, stacks + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, slices + 1)
    d = np.stack([np.sin(th)[:, None] * np.cos(ph)[None, :], np.cos(th)[:, None] * np.ones_like(ph)[None, :],
                  np.sin(th)[:, None] * np.sin(ph)[None, :]], -1)
    k = rng.normal(size=(bumps, 3))
    k *= rng.uniform(1.5, 6.0, (bumps, 1)) / np.linalg.norm(k, axis=1, keepdims=True)
    phase = rng.uniform(0, 2 * np.pi, bumps)
    wgt = rng.uniform(0.3, 1.0, bumps)
    disp = (np.sin(d @ k.T + phase) * wgt).sum(-1) / wgt.sum()
    P = d * (radius * (1.0 + amplitude * disp))[..., None]
    uvg = np.stack([np.broadcast_to(ph[None, :] / (2 * np.pi) * 4.0, th[:, None].shape[:1] + ph.shape),
                    np.broadcast_to(th[:, None] / np.pi * 2.0, th.shape + ph.shape[:1])], -1)
    i, j = np.meshgrid(np.arange(stacks), np.arange(slices), indexing="ij")
    i, j = i.reshape(-1), j.reshape(-1)
    top = i > 0                                   # triangle (00, 01, 10) exists below the north pole
    bot = i < stacks - 1                          # triangle (01, 11, 10) exists above the south pole
    a = np.concatenate([P[i[top], j[top]], P[i[bot], j[bot] + 1]])
    b = np.concatenate([P[i[top], j[top] + 1], P[i[bot] + 1, j[bot] + 1]])
    c = np.concatenate([P[i[top] + 1, j[top]], P[i[bot] + 1, j[bot]]])
    ua = np.concatenate([uvg[i[top], j[top]], uvg[i[bot], j[bot] + 1]])
    ub = np.concatenate([uvg[i[top], j[top] + 1], uvg[i[bot] + 1, j[bot] + 1]])
    uc = np.concatenate([uvg[i[top] + 1, j[top]], uvg[i[bot] + 1, j[bot]]])
    return a.astype(f32), b.astype(f32), c.astype(f32), np.stack([ua, ub, uc], 1).astype(f32)


def noise_texture(size=2048, seed=4):
    """Procedural ARGB32 texture (value noise at three octaves times seeded
    stripes), opaque, 5 MIP levels."""
    rng = np.random.default_rng(seed)
    n = min(size, 512)                            # noise resolution, replicated up to size
    img = np.zeros((n, n, 3), f32)
    t = (np.arange(n) / n).astype(f32)
    for octave, amp in ((8, 0.5), (32, 0.3), (128, 0.2)):
        g = rng.random((octave + 1, octave + 1, 3)).astype(f32)
        f = t * octave
        i = np.floor(f).astype(int)
        w = (f - i).astype(f32)
        gx = g[:, i] * (1 - w)[None, :, None] + g[:, i + 1] * w[None, :, None]          # (octave+1, n, 3)
        v = gx[i] * (1 - w)[:, None, None] + gx[i + 1] * w[:, None, None]              # (n, n, 3)
        img += f32(amp) * v
    r = size // n
    ang, freq = rng.uniform(0, np.pi), rng.uniform(6, 24)
    t = (np.arange(size) / size).astype(f32)
    stripes = f32(0.75) + f32(0.25) * np.sin((t[None, :] * f32(np.cos(ang)) + t[:, None] * f32(np.sin(ang))) * f32(freq * 2 * np.pi))
    ch = [np.clip(np.kron(img[..., k], np.ones((r, r), f32)) * stripes * f32(255), 0, 255) for k in range(3)]
    px = pack_texels(ch[0], ch[1], ch[2], np.full((size, size), 255))
    return {"width": size, "height": size, "flags": 0, "storage": F.ARGB32, "pixels": mip_chain(px)}


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced_scene(meshes=10, tris=(50000, 200000), instances=1000, tex_size=2048, lights=64, extent=50.0,
                    seeds=(2, 3, 4, 5)):
    """Instanced scene: `meshes` noise-displaced spheres (seeds 0..meshes-1;
    triangle counts uniform in `tris`, seed seeds[0]), `instances` instances
    with random affine transforms (uniform rotation, uniform scale U(0.5, 2),
    translation U[-extent, extent]^3; seed seeds[1]), one procedural ARGB32
    texture of tex_size^2 per mesh (seed seeds[2] + k), Disney materials
    (diffuse texture, GGX specular 0.5, roughness U(0.05, 1); seed seeds[3])
    and `lights` single-triangle area lights facing down above the cloud."""
    s = Scene()
    rng_t = np.random.default_rng(seeds[0])
    counts = rng_t.integers(tris[0], tris[1] + 1, meshes)
    rng_m = np.random.default_rng(seeds[3])
    for k in range(meshes):
        tex = s.add_texture(noise_texture(tex_size, seed=seeds[2] + k))
        mat = F.disney_material((1.0, 1.0, 1.0), specular=0.5, roughness=float(rng_m.uniform(0.05, 1.0)))
        mat["color"]["textureID"] = tex
        mat["color"]["uvscale"] = 1.0
        mat["color"]["uvoffset"] = 0.0
        m = s.add_material(mat)
        a, b, c, uv = displaced_sphere(int(counts[k]), 1.0, seed=k)
        v, t = add_triangles(a, b, c, m)
        for q in range(3):
            t[f"u{q}"], t[f"v{q}"] = uv[:, q, 0], uv[:, q, 1]
        t["LOD"] = triangle_lod(t, tex_size, tex_size)
        s.add_mesh(v, t)
    rng_i = np.random.default_rng(seeds[1])
    for _ in range(instances):
        M = np.eye(4)
        M[:3, :3] = _random_rotation(rng_i) * rng_i.uniform(0.5, 2.0)
        M[:3, 3] = rng_i.uniform(-extent, extent, 3)
        s.add_instance(int(rng_i.integers(0, meshes)), M.reshape(16).astype(f32))
    if lights:
        rng_l = np.random.default_rng(seeds[3] + 100)
        lm = s.add_material(F.lambert_material((20.0, 18.0, 15.0)))
        c = np.column_stack([rng_l.uniform(-extent, extent, lights), np.full(lights, 1.2 * extent), rng_l.uniform(-extent, extent, lights)])
        size = 0.08 * extent
        v0, v1, v2 = c + [-size, 0, -size], c + [size, 0, -size], c + [0, 0, size]   # wound to face -y
        v, t = add_triangles(v0, v1, v2, lm)
        li = s.add_instance(s.add_mesh(v, t))
        add_emitters(s, t, (20.0, 18.0, 15.0), li)
    return s


def instanced_view(width=1920, height=1080, extent=50.0):
    return camera_view((0.0, 0.25 * extent, -2.4 * extent), (0.0, 0.0, 0.0), 45.0, width, height)


def cornell_view(width=256, height=256):
    return camera_view((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), 40.0, width, height)


def random_triangles(n, seed=1, sigma=0.005):
    """Incoherent-traversal scene: n triangles, centroids U[0,1)^3, each vertex
    = centroid + N(0, sigma^2) per component. Vertex soup (3n, 4) float32."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 1, 3), dtype=np.float32)
    e = rng.standard_normal((n, 3, 3), dtype=np.float32) * f32(sigma)
    v = np.ones((n, 3, 4), f32)
    v[:, :, :3] = c + e
    return v.reshape(3 * n, 4)


def random_triangle_scene(n, seed=1, with_tris=False):
    s = Scene()
    v = random_triangles(n, seed)
    tris = None
    if with_tris:
        m = s.add_material(F.lambert_material((0.5, 0.5, 0.5)))
        _, tris = add_triangles(v[0::3, :3], v[1::3, :3], v[2::3, :3], m)
    s.add_instance(s.add_mesh(v, tris))
    return s


def microbench_view(width=1920, height=1080):
    """Camera of the incoherent-traversal workload: (0.5,0.5,-1.2) -> (0.5,0.5,0.5), FOV 40."""
    return camera_view((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), 40.0, width, height)


def mip_chain(base, levels=F.MIPLEVELCOUNT):
    """HostTexture::ConstructMIPmaps (host_texture.cpp:131-153): level i+1
    averages 2x2 texels of level i (r, g, b summed, shifted right by 2; alpha
    = min of the four). base: (h, w) uint32 words 0xAARRGGBB. Returns all
    levels concatenated (HostTexture::PixelsNeeded texels)."""
    src = np.asarray(base, np.uint32)
    out = [src.reshape(-1)]
    for _ in range(1, levels):
        h, w = src.shape[0] >> 1, src.shape[1] >> 1
        q = src[:2 * h, :2 * w].reshape(h, 2, w, 2)
        c = [q[:, 0, :, 0], q[:, 0, :, 1], q[:, 1, :, 0], q[:, 1, :, 1]]
        a = np.minimum(np.minimum(c[0] >> 24, c[1] >> 24), np.minimum(c[2] >> 24, c[3] >> 24))
        ch = lambda sh: sum((x >> sh) & 255 for x in c) >> 2
        src = ((a << 24) + (ch(16) << 16) + (ch(8) << 8) + ch(0)).astype(np.uint32)
        out.append(src.reshape(-1))
    return np.concatenate(out)


def pack_texels(r, g, b, a):
    """Bytes -> texel words: byte 0 (x) = b, 1 (y) = g, 2 (z) = r, 3 (w) = a."""
    c = lambda x: np.clip(np.asarray(x), 0, 255).astype(np.uint32)
    return (c(a) << 24) | (c(r) << 16) | (c(g) << 8) | c(b)


def checker_texture(size=64, cells=8, seed=4, alpha_hole=False):
    """ARGB32 checkerboard with seeded cell colours (and, optionally, a
    transparent square of alpha 0 in one corner), 5 MIP levels."""
    rng = np.random.default_rng(seed)
    col = rng.integers(40, 256, (cells, cells, 3))
    y, x = np.mgrid[0:size, 0:size]
    cy, cx = y * cells // size, x * cells // size
    c = col[cy, cx]
    a = np.full((size, size), 255)
    if alpha_hole:
        a[: size // 4, : size // 4] = 0
    px = pack_texels(c[..., 0], c[..., 1], c[..., 2], a)
    return {"width": size, "height": size, "flags": 0, "storage": F.ARGB32, "pixels": mip_chain(px)}


def normal_texture(size=64, seed=5):
    """NRM32 normal map: smooth bumps, byte k = (n_k * 0.5 + 0.5) * 255."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:size, 0:size] / size
    fx, fy = rng.uniform(1, 4, 2)
    nx, ny = 0.5 * np.sin(2 * np.pi * fx * x), 0.5 * np.cos(2 * np.pi * fy * y)
    nz = np.sqrt(np.maximum(0.0, 1 - nx * nx - ny * ny))
    enc = lambda v: np.round((v * 0.5 + 0.5) * 255)
    px = pack_texels(enc(nz), enc(ny), enc(nx), np.full((size, size), 255))
    return {"width": size, "height": size, "flags": 0, "storage": F.NRM32, "pixels": mip_chain(px)}


def roughness_texture(size=64, seed=6):
    """ARGB32 roughness/metalness map (glTF layout): byte 1 (y) roughness in
    [0.3, 1], byte 0 (x) metalness."""
    rng = np.random.default_rng(seed)
    rough = rng.integers(77, 256, (size, size))
    metal = rng.integers(0, 256, (size, size))
    px = pack_texels(0, rough, metal, np.full((size, size), 255))
    return {"width": size, "height": size, "flags": 0, "storage": F.ARGB32, "pixels": mip_chain(px)}


def triangle_lod(tris, tex_w, tex_h):
    """tri.LOD = 0.5 log2(Ta / Pa), Ta = texel area of the uv triangle, Pa =
    twice the world-space area (host_mesh.cpp:308-317), in float32."""
    ta = f32(tex_w * tex_h) * np.abs((tris["u1"] - tris["u0"]) * (tris["v2"] - tris["v0"]) -
                                     (tris["u2"] - tris["u0"]) * (tris["v1"] - tris["v0"])).astype(f32)
    e1 = (tris["vertex1"] - tris["vertex0"]).astype(f32)
    e2 = (tris["vertex2"] - tris["vertex0"]).astype(f32)
    cr = np.cross(e1, e2).astype(f32)
    pa = np.sqrt((cr * cr).sum(1)).astype(f32)
    return (f32(0.5) * np.log2(ta / pa)).astype(f32)


def sky_dome(width=64, height=32, seed=8, rotate_deg=30.0):
    """Procedural equirect sky: a vertical gradient plus seeded per-texel
    noise, and a world-to-sky rotation about the y axis."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height)[:, None] + 0.5) / height
    base = np.stack([0.2 + 0.8 * v, 0.3 + 0.5 * v, 1.0 - 0.6 * v], -1) * np.ones((1, width, 1))
    px = (base + rng.uniform(0, 0.3, (height, width, 3))).astype(f32)
    a = np.radians(rotate_deg)
    T = np.eye(4, dtype=f32)
    T[0, 0], T[0, 2], T[2, 0], T[2, 2] = np.cos(a), np.sin(a), -np.sin(a), np.cos(a)
    return px, T.reshape(16).astype(f32)


def _planar_uv(tris, axes, scale):
    for k in range(3):
        p = tris[f"vertex{k}"]
        tris[f"u{k}"] = (p[:, axes[0]] * scale).astype(f32)
        tris[f"v{k}"] = (p[:, axes[1]] * scale).astype(f32)


def textured_cornell_box(seed=1234, sphere_tris=4000, tex_size=64):
    """The Cornell box with texture maps and a skydome: floor with a diffuse
    checkerboard holding a transparent corner, back wall diffuse + detail
    diffuse, ceiling a roughness map, left wall a normal map, right wall two
    normal maps; planar uv (2 repeats per wall) and the reference's triangle
    LOD; the open front shows the sky."""
    s = cornell_box(seed=seed, sphere_tris=sphere_tris)
    checker = s.add_texture(checker_texture(tex_size, 8, seed=4, alpha_hole=True))
    detail = s.add_texture(checker_texture(tex_size, 4, seed=9))
    rough = s.add_texture(roughness_texture(tex_size, seed=6))
    nrm = s.add_texture(normal_texture(tex_size, seed=5))
    nrm2 = s.add_texture(normal_texture(tex_size, seed=7))
    axes = [(0, 2), (0, 2), (0, 1), (2, 1), (2, 1)]   # floor, ceiling, back, left, right
    for wall, ax in enumerate(axes):
        v, t = s.meshes[wall]
        _planar_uv(t, ax, 2.0)
        t["LOD"] = triangle_lod(t, tex_size, tex_size)
        m = t["material"][0]
        mat = s.materials[m:m + 1]
        use = {0: [("color", checker)], 1: [("roughness", rough)], 2: [("color", checker), ("detailColor", detail)],
               3: [("normals", nrm)], 4: [("normals", nrm), ("detailNormals", nrm2)]}[wall]
        for field, tid in use:
            mat[field]["textureID"] = tid
            mat[field]["uvscale"] = 1.0
            mat[field]["uvoffset"] = 0.0
        s.materials[m] = mat[0]
    s.sky = sky_dome()
    return s


def sky_scene(n=64, seed=3):
    """A few triangles in a mostly empty view, so most primary rays see the
    skydome."""
    s = Scene()
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.7, (n, 1, 3)).astype(f32)
    e = rng.normal(0, 0.02, (n, 3, 3)).astype(f32)
    v = (c + e).astype(f32)
    m = s.add_material(F.lambert_material((0.6, 0.5, 0.4)))
    verts, tris = add_triangles(v[:, 0], v[:, 1], v[:, 2], m)
    s.add_instance(s.add_mesh(verts, tris))
    s.sky = sky_dome(96, 48, seed=11, rotate_deg=75.0)
    return s


# ---- synthetic SVGF input (G-buffer of two planes) ----------------------------


def pack_normal2(n):
    """PackNormal2 (tools_shared.h:122-129), float32."""
    c = lambda v: np.clip(((v.astype(f32) + f32(1)) * f32(511)).astype(np.int64), 0, 1023).astype(np.uint32)
    return (c(n[..., 0]) << 2) + (c(n[..., 1]) << 12) + (c(n[..., 2]) << 22)


def hdr_to_rgb32(c):
    """HDRtoRGB32 (tools_shared.h:162-168), float32."""
    q = lambda v, s: (f32(s) * np.minimum(f32(1), v.astype(f32))).astype(np.uint32)
    return (q(c[..., 0], 1023) << 22) + (q(c[..., 1], 2047) << 11) + q(c[..., 2], 2047)


def gbuffer(view, w, h, seed=0, noise=0.5, spec_strip=True, samples=1, const=None):
    """The G-buffer of two planes (floor y = 0, back wall z = 1) seen through
    a ViewPyramid, packed as RenderCore_Optix7Filter's shade kernel packs it
    (PackFeatures, kernels/pathtracer.h:44-58), with noisy direct and
    indirect radiance: (accumulator (2, h, w, 4), features (h, w, 4) uint32,
    world positions (h, w, 4), depth derivatives (h, w, 4))."""
    v = view[0] if view.shape else view
    pos, p1, p2, p3 = (np.asarray(v[k], np.float64) for k in ("pos", "p1", "p2", "p3"))
    x = (np.arange(w) + 0.5) / w
    y = (np.arange(h) + 0.5) / h
    P = p1[None, None] + x[None, :, None] * (p2 - p1)[None, None] + y[:, None, None] * (p3 - p1)[None, None]
    D = P - pos
    D /= np.linalg.norm(D, axis=2, keepdims=True)
    t_floor = np.where(D[..., 1] < -1e-6, (0.0 - pos[1]) / np.where(D[..., 1] == 0, 1, D[..., 1]), np.inf)
    t_wall = np.where(D[..., 2] > 1e-6, (1.0 - pos[2]) / np.where(D[..., 2] == 0, 1, D[..., 2]), np.inf)
    floor = t_floor < t_wall
    t = np.where(floor, t_floor, t_wall)
    t = np.where(np.isfinite(t), t, 50.0)
    I = pos + D * t[..., None]
    N = np.where(floor[..., None], np.array([0, 1, 0.0]), np.array([0, 0, -1.0]))
    rng = np.random.default_rng(seed)
    if const is None:
        checker = ((np.floor(I[..., 0] * 4) + np.floor(I[..., 2] * 4) + np.floor(I[..., 1] * 4)) % 2)
        albedo = np.stack([0.3 + 0.5 * checker, 0.5 + 0.3 * floor, 0.6 - 0.3 * checker], -1)
        light = 0.5 + 0.5 * np.cos(I[..., 0] * 3) ** 2
        direct = albedo * light[..., None] * rng.gamma(1 / max(noise, 1e-3), max(noise, 1e-3), (h, w, 1))
        indirect = albedo * 0.3 * rng.gamma(1 / max(noise, 1e-3), max(noise, 1e-3), (h, w, 1))
    else:
        albedo = np.full((h, w, 3), const[0])
        direct = np.full((h, w, 3), const[1])
        indirect = np.full((h, w, 3), const[2])
    spec = np.zeros((h, w), np.uint32)
    if spec_strip:
        spec[:, w // 3: w // 3 + 3] = 1
    matid = floor.astype(np.uint32) + 1
    pn = pack_normal2(N * -1.0) + spec
    feat = np.zeros((h, w, 4), np.uint32)
    feat[..., 0] = hdr_to_rgb32(albedo)
    feat[..., 1] = pn
    feat[..., 2] = t.astype(f32).view(np.uint32)
    feat[..., 3] = (spec << 4) + (matid << 6)
    wp = np.zeros((h, w, 4), f32)
    wp[..., :3] = I
    wp[..., 3] = pn.view(f32)
    dd = np.zeros((h, w, 4), f32)
    dd[..., 2] = np.gradient(t, axis=1)
    dd[..., 3] = np.gradient(t, axis=0)
    acc = np.zeros((2, h, w, 4), f32)
    acc[0, ..., :3] = direct * samples
    acc[1, ..., :3] = indirect * samples
    return acc, feat, wp, dd
