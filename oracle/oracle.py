"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker of RenderCore_HIP. Build it with
`make -C oracle`.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None
P, I, U = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        sigs = {
            "orc_create": (P, []), "orc_destroy": (None, [P]),
            "orc_add_mesh": (I, [P, P, I, P]), "orc_add_instance": (I, [P, I, P]), "orc_finalize": (None, [P]),
            "orc_set_materials": (None, [P, P, I]), "orc_set_lights": (None, [P, P, I, P, I, P, I, P, I]),
            "orc_trace": (None, [P, P, I, P, I, I, P, I]),
            "orc_bvh_nodes": (I, [P, I, P, P, P]), "orc_tlas_nodes": (I, [P, P, P]),
            "orc_bvh_layout": (None, [P, I, P, I, P, P, P, P, P, P]),
            "orc_render_frame": (None, [P, P, I, I, I, I, ctypes.c_float, U, I, P, I]),
            "orc_primary_rays": (None, [P, I, I, U, ctypes.c_float, P]),
            "orc_bounce": (I, [P, P, P, P, I, U, ctypes.c_float, P, P]),
            "orc_wang_hash": (U, [U]), "orc_xor32_seq": (None, [U, I, P]),
            "orc_sincos": (None, [ctypes.c_float, P, P]),
            "orc_float_to_half": (ctypes.c_uint16, [ctypes.c_float]), "orc_half_to_float": (ctypes.c_float, [ctypes.c_uint16]),
            "orc_mat4_inverse": (None, [P, P]), "orc_convert_material": (None, [P, P]), "orc_sizeof_omat": (I, []),
            "orc_tri_intersect": (I, [P, P, P, P]),
            "orc_set_textures": (None, [P, P, I]), "orc_set_sky": (None, [P, P, I, I, P]),
            "orc_sample_sky": (None, [P, P, I, P, I]), "orc_fetch_texels": (None, [P, I, P, I, I, ctypes.c_float, P]),
            "orc_math": (None, [I, P, P, I, P]),
            "orc_svgf_create": (P, [I, I]), "orc_svgf_destroy": (None, [P]),
            "orc_svgf_frame": (None, [P, P, P, P, P, P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, I,
                                      ctypes.c_float, ctypes.c_float, I, P, I]),
            "orc_svgf_read": (None, [P, I, P]),
            "orc_svgf_pass": (None, [P, I, P, P, P, P, P, P, I]),
            "orc_set_blue_noise": (None, [P, P, P, P, U]),
            "orc_blue_noise4": (None, [P, P, I, P]), "orc_normals": (None, [I, P, P, I]),
            "orc_disney_eval": (None, [P, P, P, P, P, P, P, I, P]), "orc_disney_sample": (None, [P, P, P, P, P, P, P, I, P]),
            "orc_ggx": (None, [I, P, P, P, I, ctypes.c_float, ctypes.c_float, P]),
            "orc_set_brute": (None, [P, I]),
            "orc_render_frameF": (I, [P, P, I, I, I, I, I, ctypes.c_float, ctypes.c_float, U, P, P, P, P, I, I, ctypes.c_float,
                                      ctypes.c_float]),
            "orc_render_frame7": (I, [P, P, I, I, I, I, ctypes.c_float, ctypes.c_float, U, U, P, I, I]),
        }
        for n, (r, a) in sigs.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def bvh_layout(nodes, prims):
    """The traversal layout (oracle_layout.cpp) of a build tree as the device
    downloads it: (f32 nodes as uint32 rows of 16, q8 nodes likewise, leaf
    triangle order, last-of-leaf flags, counts [f32, q8, f32 root, q8 root,
    binary depth, 4-wide depth], q8 map: per q8 node its build node and its
    children's build nodes, -1 past the last)."""
    nodes = np.ascontiguousarray(nodes)
    prims = np.ascontiguousarray(prims, np.uint32)
    cap = max(1, (len(nodes) - 1) // 2)
    f32 = np.zeros((cap, 16), np.uint32)
    q8 = np.zeros((cap, 16), np.uint32)
    order = np.zeros(len(prims), np.uint32)
    last = np.zeros(len(prims), np.uint8)
    q8map = np.zeros((cap, 5), np.int32)
    counts = np.zeros(6, np.int32)
    lib().orc_bvh_layout(_p(nodes), len(nodes), _p(prims), len(prims), _p(f32), _p(q8), _p(q8map), _p(order), _p(last), _p(counts))
    return f32[:counts[0]], q8[:counts[1]], order, last, counts, q8map[:counts[1]]


class OracleScene:
    """CPU restatement of a RenderCore_HIP scene (BVH per mesh, TLAS, materials, lights)."""

    def __init__(self, scene=None):
        self.h = lib().orc_create()
        self._keep = []
        if scene is not None:
            self.load(scene)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def add_mesh(self, verts, tris=None):
        v = np.ascontiguousarray(verts, np.float32)
        return lib().orc_add_mesh(self.h, _p(v), len(v) // 3, _p(tris))

    def add_instance(self, mesh, T=None):
        t = np.eye(4, dtype=np.float32).reshape(16) if T is None else np.ascontiguousarray(T, np.float32).reshape(16)
        return lib().orc_add_instance(self.h, mesh, _p(t))

    def finalize(self):
        lib().orc_finalize(self.h)

    def load(self, scene):
        for v, t in scene.meshes:
            self.add_mesh(v, None if t is None else np.ascontiguousarray(t))
        for m, T in scene.instances:
            self.add_instance(m, T)
        self.finalize()
        if getattr(scene, "sky", None) is not None:
            self.set_sky(*scene.sky)
        if getattr(scene, "textures", None):
            self.set_textures(scene.textures)
        mats = np.ascontiguousarray(scene.materials)
        lib().orc_set_materials(self.h, _p(mats), len(mats))
        a = [np.ascontiguousarray(x) for x in (scene.tri_lights, scene.point_lights, scene.spot_lights, scene.dir_lights)]
        self._keep += [mats] + a
        lib().orc_set_lights(self.h, _p(a[0]), len(a[0]), _p(a[1]), len(a[1]), _p(a[2]), len(a[2]), _p(a[3]), len(a[3]))

    def set_textures(self, textures):
        from lighthouse2_amd import formats as F
        d, keep = F.texture_descs(textures)
        lib().orc_set_textures(self.h, _p(d), len(d))
        del keep

    def set_sky(self, pixels, T=None):
        px = np.ascontiguousarray(pixels, np.float32)
        t = None if T is None else np.ascontiguousarray(T, np.float32).reshape(16)
        lib().orc_set_sky(self.h, _p(px), px.shape[1], px.shape[0], _p(t))

    def sample_sky(self, dirs, small=False):
        """SampleSkydome, or SampleSmallSkydome (the 64x64-averaged dome) with small=True."""
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros_like(d)
        lib().orc_sample_sky(self.h, _p(d), len(d), _p(out), 1 if small else 0)
        return out

    def fetch_texels(self, tex_id, uv, lam=None):
        u = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((len(u), 4), np.float32)
        lib().orc_fetch_texels(self.h, tex_id, _p(u), len(u), 0 if lam is None else 1, 0.0 if lam is None else lam, _p(out))
        return out

    def trace(self, rays, any_hit=False, brute=False, threads=8, counters=False):
        """rays (n, 8) float32 {O, tmin, D, tmax}. Closest: (n, 4) records with
        uint32 views (uv, inst, prim) and t; any: (n,) bool."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        h = np.zeros((len(r), 4), np.float32)
        c = np.zeros(3, np.uint64) if counters else None
        lib().orc_trace(self.h, _p(r), len(r), _p(h), 1 if any_hit else 0, 1 if brute else 0, _p(c), threads)
        out = h[:, 0] != 0 if any_hit else h
        return (out, c) if counters else out

    def bvh_nodes(self, mesh=0):
        from lighthouse2_amd.formats import BuildNode
        n = lib().orc_bvh_nodes(self.h, mesh, None, None, None)
        nodes = np.zeros(n, BuildNode)
        depth = ctypes.c_int(0)
        lib().orc_bvh_nodes(self.h, mesh, _p(nodes), None, ctypes.byref(depth))
        prims = np.zeros(int(nodes["count"].sum()), np.uint32)
        lib().orc_bvh_nodes(self.h, mesh, None, _p(prims), None)
        return nodes, prims, depth.value

    def render_frame(self, view, w, h, spp, max_depth, eps, R0gen, pass_, threads=8):
        acc = np.zeros((spp * w * h, 4), np.float32)
        v = np.ascontiguousarray(view)
        lib().orc_render_frame(self.h, _p(v), w, h, spp, max_depth, eps, R0gen, pass_, _p(acc), threads)
        return acc

    def set_brute_force(self, on=True):
        """Render with brute-force traversal (every triangle) instead of the BVH."""
        lib().orc_set_brute(self.h, 1 if on else 0)

    def set_blue_noise(self, tables=None, seed=0x2545F491):
        """tables: (sequence[65536], scrambling[131072], ranking[131072]) uint8,
        or None for the core's seeded default tables."""
        if tables is None:
            lib().orc_set_blue_noise(self.h, None, None, None, seed)
            return
        t = [np.ascontiguousarray(x, np.uint8) for x in tables]
        assert t[0].size == 65536 and t[1].size == 131072 and t[2].size == 131072
        lib().orc_set_blue_noise(self.h, _p(t[0]), _p(t[1]), _p(t[2]), 0)

    def blue_noise4(self, xysd):
        """blueNoiseSampler4 at rows (x, y, sampleIndex, dimension)."""
        q = np.ascontiguousarray(xysd, np.int32).reshape(-1, 4)
        out = np.zeros((len(q), 4), np.float32)
        lib().orc_blue_noise4(self.h, _p(q), len(q), _p(out))
        return out

    def render_frame7(self, view, w, h, spp, pass_, eps, clamp_value, shift, cam_seed, threads=8, bsdf=1):
        """One optix7-estimator frame: per-path (xyz radiance, w connections) and
        the number of shade waves it ran."""
        acc = np.zeros((spp * w * h, 4), np.float32)
        v = np.ascontiguousarray(view)
        waves = lib().orc_render_frame7(self.h, _p(v), w, h, spp, pass_, eps, clamp_value, shift, cam_seed, _p(acc), threads, bsdf)
        if waves < 0:
            raise RuntimeError("set_blue_noise first")
        return acc, waves

    def render_frameF(self, view, w, h, spp, pass_, blue_slot, eps, clamp_value, cam_seed, gbuffer=None, threads=8, bsdf=1, jitter=(0.0, 0.0)):
        """One filter-estimator frame: per-path direct and indirect sums (2, spp*w*h, 4)
        and the waves it ran; gbuffer = (features (h,w,4) uint32, world_pos, delta_depth)
        arrays updated in place (None: filter off); jitter: the TAA view jitter in pixels."""
        acc = np.zeros((2, spp * w * h, 4), np.float32)
        v = np.ascontiguousarray(view)
        f, wp, dd = gbuffer if gbuffer is not None else (None, None, None)
        waves = lib().orc_render_frameF(self.h, _p(v), w, h, spp, pass_, blue_slot, eps, clamp_value, cam_seed, _p(acc), _p(f), _p(wp),
                                        _p(dd), threads, bsdf, jitter[0], jitter[1])
        if waves < 0:
            raise RuntimeError("set_blue_noise first")
        return acc, waves

    def bounce(self, rays, hits, path_ids, R0, tmin):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        h = np.ascontiguousarray(hits, np.float32).reshape(-1, 4)
        p = np.ascontiguousarray(path_ids, np.uint32)
        out = np.zeros_like(r)
        po = np.zeros_like(p)
        k = lib().orc_bounce(self.h, _p(r), _p(h), _p(p), len(r), R0, tmin, _p(out), _p(po))
        return out[:k], po[:k]


def xor32_host(seed):
    """RandomUInt(seed) of lib/platform/system.cpp:51: returns (value, new seed)."""
    s = seed & 0xffffffff
    s ^= (s << 13) & 0xffffffff
    s ^= s >> 17
    s ^= (s << 5) & 0xffffffff
    return s, s


RESTART, CONVERGE = 1, 0


def render_frames(scene_or_oracle, view, w, h, spp=1, frames=1, max_depth=3, eps=1e-3, seed=0x12345678, threads=8, modes=None):
    """Oracle of a sequence of Render calls, one result per call. `modes` lists
    Restart/Converge per call (default: one Restart, then `frames` Converge
    calls). Restates the convergence state machine of RenderCore::Render
    (lib/RenderCore_PrimeRef/rendercore.cpp:543-550; rendercore.h:129-134):
    a Restart, or the first Converge call after it, clears the accumulator and
    reseeds the camera RNG. Yields (acc32 (h,w,4) float32, acc64 float64,
    samplesTaken) after each call (the arrays are updated in place)."""
    o = scene_or_oracle if isinstance(scene_or_oracle, OracleScene) else OracleScene(scene_or_oracle)
    if modes is None:
        modes = [RESTART] + [CONVERGE] * frames
    acc32 = np.zeros((w * h, 4), np.float32)
    acc64 = np.zeros((w * h, 4), np.float64)
    cam, taken, first = seed, 0, False
    for m in modes:
        if m == RESTART or first:
            acc32[:] = 0
            acc64[:] = 0
            taken, first, cam = 0, True, seed
        if m == CONVERGE:
            first = False
        R0, cam = xor32_host(cam)
        pa = o.render_frame(view, w, h, spp, max_depth, eps, R0, taken, threads)
        for s in range(spp):
            blk = pa[s * w * h:(s + 1) * w * h]
            acc32 += blk
            acc64 += blk.astype(np.float64)
        taken += spp
        yield acc32.reshape(h, w, 4), acc64.reshape(h, w, 4), taken


def render(scene_or_oracle, view, w, h, spp=1, frames=1, max_depth=3, eps=1e-3, seed=0x12345678, threads=8, modes=None):
    """The state after the last call of render_frames: (acc32 (h,w,4) float32,
    acc64 float64, samplesTaken)."""
    out = None
    for out in render_frames(scene_or_oracle, view, w, h, spp, frames, max_depth, eps, seed, threads, modes):
        pass
    return out[0].copy(), out[1].copy(), out[2]


class OracleSvgf:
    """CPU restatement of SVGF (oracle_svgf.cpp) with history across frames."""

    def __init__(self, w, h):
        self.w, self.h = w, h
        self.hd = lib().orc_svgf_create(w, h)

    def __del__(self):
        if getattr(self, "hd", None):
            lib().orc_svgf_destroy(self.hd)
            self.hd = None

    def frame(self, acc, features, world_pos, delta_depth, prev_view, j=(0.0, 0.0), prev_j=(0.0, 0.0), samples=1,
              clamp_direct=2.5, clamp_indirect=15.0, stationary=0, out=None, taa=False):
        a = np.ascontiguousarray(acc, np.float32)
        f = np.ascontiguousarray(features, np.uint32)
        wp = np.ascontiguousarray(world_pos, np.float32)
        dd = np.ascontiguousarray(delta_depth, np.float32)
        v = np.ascontiguousarray(prev_view)
        o = np.zeros((self.h, self.w, 4), np.float32) if out is None else out
        lib().orc_svgf_frame(self.hd, _p(a), _p(f), _p(wp), _p(dd), _p(v), j[0], j[1], prev_j[0], prev_j[1], samples,
                             clamp_direct, clamp_indirect, stationary, _p(o), 1 if taa else 0)
        return o

    def apply_pass(self, phase, combined, features, world_pos, delta_depth, moments, last=False):
        """One a-trous pass (applyFilterKernel) on the given inputs, without
        history; returns the (h, w, 4) float32 output (combined 5:11 format,
        or the sqrt image with last=True)."""
        a = np.ascontiguousarray(combined, np.float32)
        f = np.ascontiguousarray(features, np.uint32)
        wp = np.ascontiguousarray(world_pos, np.float32)
        dd = np.ascontiguousarray(delta_depth, np.float32)
        m = np.ascontiguousarray(moments, np.float32)
        o = np.zeros((self.h, self.w, 4), np.float32)
        lib().orc_svgf_pass(self.hd, phase, _p(a), _p(f), _p(wp), _p(dd), _p(m), _p(o), 1 if last else 0)
        return o

    def read(self, which):
        """which: 0 features, 1 moments, 2 motion, 3 phase-1 history, 4 last pass, 5 world positions, 6 TAA history."""
        shape = {0: ((self.h, self.w, 4), np.uint32), 2: ((self.h, self.w, 2), np.float32)}.get(which, ((self.h, self.w, 4), np.float32))
        out = np.zeros(shape[0], shape[1])
        lib().orc_svgf_read(self.hd, which, _p(out))
        return out


def math(which, a, b=None):
    """The oracle's own transcendental functions: 'atan2' (a = y, b = x), 'acos', 'log2', 'exp'."""
    k = {"atan2": 0, "acos": 1, "log2": 2, "exp": 3}[which]
    a = np.ascontiguousarray(a, np.float32)
    bb = np.ascontiguousarray(a if b is None else b, np.float32)
    out = np.zeros_like(a)
    lib().orc_math(k, _p(a), _p(bb), len(a), _p(out))
    return out


def disney_params(metallic=0.0, subsurface=0.0, specular=0.5, roughness=0.5, spectint=0.0, anisotropic=0.0, sheen=0.0,
                  sheentint=0.0, clearcoat=0.0, clearcoatgloss=0.0, transmission=0.0, eta=1.5):
    """Packed ShadingData.parameters words (tools_shared.h:33-36) from values in [0, 1]."""
    b = lambda v: int(round(min(max(v, 0.0), 1.0) * 255))
    x = b(metallic) | b(subsurface) << 8 | b(specular) << 16 | b(roughness) << 24
    y = b(spectint) | b(anisotropic) << 8 | b(sheen) << 16 | b(sheentint) << 24
    z = b(clearcoat) | b(clearcoatgloss) << 8 | b(transmission) << 16
    w = int(np.float32(eta).view(np.uint32))
    return np.array([x, y, z, w], np.uint32)


def disney_eval(color, params, wo, wi, N=(0, 0, 1), T=(1, 0, 0), transmittance=(0, 0, 0)):
    """EvaluateBSDF rows: (n, 4) value rgb and pdf."""
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    out = np.zeros((len(wo), 4), np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (color, transmittance, N, T)]
    lib().orc_disney_eval(_p(a[0]), _p(np.ascontiguousarray(params, np.uint32)), _p(a[1]), _p(a[2]), _p(a[3]), _p(wo), _p(wi), len(wo), _p(out))
    return out


def disney_sample(color, params, wo, r3, N=(0, 0, 1), T=(1, 0, 0), transmittance=(0, 0, 0)):
    """SampleBSDF rows: (n, 8) value rgb, pdf, wi xyz, specular flag."""
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(r3, np.float32).reshape(-1, 3)
    out = np.zeros((len(wo), 8), np.float32)
    a = [np.ascontiguousarray(x, np.float32) for x in (color, transmittance, N, T)]
    lib().orc_disney_sample(_p(a[0]), _p(np.ascontiguousarray(params, np.uint32)), _p(a[1]), _p(a[2]), _p(a[3]), _p(wo), _p(r), len(wo), _p(out))
    return out


def ggx(which, v, m, r2, ax, ay):
    """0 D(m), 1 visible-normal pdf(v, m), 2 sample(v, r), 3 GTR1 D(m), 4 GTR1 sample(r), 5 G1(v)."""
    k = {"D": 0, "pdf": 1, "sample": 2, "gtr1_D": 3, "gtr1_sample": 4, "G1": 5}[which]
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    m = np.ascontiguousarray(m, np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(r2, np.float32).reshape(-1, 2)
    n = max(len(v), len(m), len(r))
    out = np.zeros((n, 3), np.float32)
    lib().orc_ggx(k, _p(v), _p(m), _p(r), n, ax, ay, _p(out))
    return out


def pack_normals(n3):
    """PackNormal of (n, 3) unit vectors -> uint32."""
    n = np.ascontiguousarray(n3, np.float32).reshape(-1, 3)
    out = np.zeros(len(n), np.uint32)
    lib().orc_normals(1, _p(n), _p(out), len(n))
    return out


def unpack_normals(packed):
    p = np.ascontiguousarray(packed, np.uint32)
    out = np.zeros((len(p), 3), np.float32)
    lib().orc_normals(0, _p(out), _p(p), len(p))
    return out


def render7(scene_or_oracle, view, w, h, spp=1, frames=1, eps=1e-3, seed=0x12345678, clamp_value=10.0, threads=8,
            modes=None, blue_noise=None, shift_seed=0x11331445, bsdf=1):
    """Oracle of a sequence of Render calls with the optix7 estimator
    (rendercore_optix7/rendercore.cpp:822-905): the blue-noise shift seed
    advances once per frame and is never reset; a Restart (or the first
    Converge call after one) clears the accumulator and resets the camera seed;
    the camera seed advances once per shade wave the frame ran. Returns
    (acc32 (h,w,4), acc64, samplesTaken, camSeed, shiftSeed)."""
    o = scene_or_oracle if isinstance(scene_or_oracle, OracleScene) else OracleScene(scene_or_oracle)
    o.set_blue_noise(blue_noise)
    if modes is None:
        modes = [RESTART] + [CONVERGE] * frames
    acc32 = np.zeros((w * h, 4), np.float32)
    acc64 = np.zeros((w * h, 4), np.float64)
    cam, shift, taken, first = seed, shift_seed, 0, False
    for m in modes:
        if m == RESTART or first:
            acc32[:] = 0
            acc64[:] = 0
            taken, first, cam = 0, True, seed
        if m == CONVERGE:
            first = False
        shift, _ = xor32_host(shift)
        pa, waves = o.render_frame7(view, w, h, spp, taken, eps, clamp_value, shift, cam, threads, bsdf)
        for _ in range(waves):
            cam, _ = xor32_host(cam)
        for s in range(spp):
            blk = pa[s * w * h:(s + 1) * w * h]
            acc32 += blk
            acc64 += blk.astype(np.float64)
        taken += spp
    return acc32.reshape(h, w, 4), acc64.reshape(h, w, 4), taken, cam, shift


HALTON_X, HALTON_Y = (0.3, 0.7, 0.2, 0.8), (0.2, 0.8, 0.7, 0.3)   # RenderCore_Optix7Filter/rendercore.cpp:755


class RenderF:
    """Oracle of a sequence of Render calls with the filter estimator
    (RenderCore_Optix7Filter/rendercore.cpp:740-935): the camera seed is never
    reset and advances once per shade wave; the blue-noise slot advances once
    per frame; with the filter on, the G-buffer of sample 0 and SVGF (prepare,
    three a-trous passes, then TAA + unsharpen or the plain display) give the
    image, the history carried from frame to frame. With TAA and the filter on
    (the reference's default) the view is jittered by a 4-frame Halton cycle."""

    def __init__(self, scene_or_oracle, w, h, spp=1, eps=1e-3, seed=0x12345678, clamp_value=10.0, filter_on=True,
                 clamp_direct=2.5, clamp_indirect=15.0, bsdf=1, blue_noise=None, threads=8, taa=True):
        from lighthouse2_amd import formats as F
        self.o = scene_or_oracle if isinstance(scene_or_oracle, OracleScene) else OracleScene(scene_or_oracle)
        self.o.set_blue_noise(blue_noise)
        self.w, self.h, self.spp, self.eps, self.clamp, self.filter = w, h, spp, eps, clamp_value, filter_on
        self.cd, self.ci, self.bsdf, self.threads = clamp_direct, clamp_indirect, bsdf, threads
        self.cam, self.slot, self.taken, self.first = seed, 0, 0, False
        self.acc32 = np.zeros((2, h * w, 4), np.float32)
        self.acc64 = np.zeros((2, h * w, 4), np.float64)
        self.features = np.zeros((h, w, 4), np.uint32)
        self.world_pos = np.zeros((h, w, 4), np.float32)
        self.delta_depth = np.zeros((h, w, 4), np.float32)
        self.prev_view = np.zeros(1, F.ViewPyramid)
        self.svgf = OracleSvgf(w, h) if filter_on else None
        self.image = np.zeros((h, w, 4), np.float32)
        self.taa, self.cycle, self.j, self.prev_j = taa, 0, (0.0, 0.0), (0.0, 0.0)

    def frame(self, view, mode=RESTART):
        w, h, spp = self.w, self.h, self.spp
        if mode == RESTART or self.first:
            self.acc32[:] = 0
            self.acc64[:] = 0
            self.taken, self.first = 0, True
        if mode == CONVERGE:
            self.first = False
        on = self.taa and self.filter
        self.j = (HALTON_X[self.cycle], HALTON_Y[self.cycle]) if on else (0.0, 0.0)
        self.cycle = (self.cycle + 1) & 3
        gb = (self.features, self.world_pos, self.delta_depth) if self.filter else None
        pa, waves = self.o.render_frameF(view, w, h, spp, self.taken, self.slot, self.eps, self.clamp, self.cam, gb, self.threads, self.bsdf,
                                         jitter=self.j)
        for _ in range(waves):
            self.cam, _ = xor32_host(self.cam)
        for half in range(2):
            for s in range(spp):
                blk = pa[half, s * w * h:(s + 1) * w * h]
                self.acc32[half] += blk
                self.acc64[half] += blk.astype(np.float64)
        self.taken += spp
        self.slot = (self.slot + 1) & 255
        if self.filter:
            self.image = self.svgf.frame(self.acc32.reshape(2, h, w, 4), self.features, self.world_pos, self.delta_depth, self.prev_view,
                                         j=self.j, prev_j=self.prev_j, samples=self.taken, clamp_direct=self.cd,
                                         clamp_indirect=self.ci, stationary=0 if self.taken == spp else 1, taa=self.taa)
            self.prev_j = self.j
        else:
            self.image = (self.acc32[0] * np.float32(1.0 / self.taken)).reshape(h, w, 4)
        self.prev_view = np.ascontiguousarray(view).copy()
        return self.image


def sincos(x):
    s, c = ctypes.c_float(), ctypes.c_float()
    lib().orc_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def wang_hash(s):
    return lib().orc_wang_hash(s)


def xor32_seq(seed, n):
    out = np.zeros(n, np.uint32)
    lib().orc_xor32_seq(seed, n, _p(out))
    return out


def primary_rays(view, w, h, R0, tmin):
    out = np.zeros((w * h, 8), np.float32)
    v = np.ascontiguousarray(view)
    lib().orc_primary_rays(_p(v), w, h, R0, tmin, _p(out))
    return out
