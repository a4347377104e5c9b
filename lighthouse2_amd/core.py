"""Python binding of RenderCore_HIP (librendercore_hip.so) over its C ABI.

`RenderCore` mirrors the CoreAPI_Base methods (lib/RenderSystem/core_api_base.h:81-119)
with the same names, argument meaning and call order; errors raise
`LH2HIPError` instead of aborting the process. `load_scene` performs the call
sequence of RenderSystem::SynchronizeSceneData (lib/RenderSystem/rendersystem.cpp:203-211).
There is no fallback: importing fails loudly when the HIP library is missing.
"""
import ctypes
import os

import numpy as np

from . import formats as F

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librendercore_hip.so")

LH2_CONVERGE, LH2_RESTART = 0, 1


class LH2HIPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lh2hip error {code}: {msg}")
        self.code = code


_lib = None
P = ctypes.c_void_p
I = ctypes.c_int

_SIGS = {
    "lh2hip_last_error": (ctypes.c_char_p, []),
    "lh2hip_version": (I, []),
    "lh2hip_create": (I, [ctypes.POINTER(P)]),
    "lh2hip_destroy": (I, [P]),
    "lh2hip_set_target": (I, [P, I, I, I]),
    "lh2hip_setting": (I, [P, ctypes.c_char_p, ctypes.c_float]),
    "lh2hip_set_geometry": (I, [P, I, P, I, I, P]),
    "lh2hip_set_instance": (I, [P, I, I, P]),
    "lh2hip_finalize_instances": (I, [P]),
    "lh2hip_set_materials": (I, [P, P, I]),
    "lh2hip_set_lights": (I, [P, P, I, P, I, P, I, P, I]),
    "lh2hip_set_textures": (I, [P, P, I]),
    "lh2hip_set_sky": (I, [P, P, I, I, P]),
    "lh2hip_set_blue_noise": (I, [P, P, P, P]),
    "lh2hip_read_buffer": (I, [P, I, P, ctypes.c_size_t]),
    "lh2hip_set_probe_pos": (I, [P, I, I]),
    "lh2hip_render": (I, [P, P, I, I]),
    "lh2hip_wait": (I, [P]),
    "lh2hip_read_image": (I, [P, P]),
    "lh2hip_read_accumulator_f64": (I, [P, P]),
    "lh2hip_get_stats": (I, [P, P]),
    "lh2hip_trace_closest": (I, [P, P, P, I, P]),
    "lh2hip_trace_any": (I, [P, P, P, I, P]),
    "lh2hip_trace_closest_counted": (I, [P, P, P, I, P, P]),
    "lh2hip_generate_primary": (I, [P, I, I, ctypes.c_uint32, ctypes.c_float, P, P, P]),
    "lh2hip_bounce_diffuse": (I, [P, P, P, P, I, ctypes.c_uint32, ctypes.c_float, P, P, P, P]),
    "lh2hip_bvh_info": (I, [P, I, P]),
    "lh2hip_bvh_download": (I, [P, I, P, I, P, I, P]),
    "lh2hip_bvh_layout_download": (I, [P, I, P, I, P, I, P, I]),
    "lh2hip_dn_conv3x3": (I, [P, I, P, P, P, I, I, I, I, P]),
    "lh2hip_dn_dw_workspace_bytes": (ctypes.c_uint64, [I, I]),
    "lh2hip_dn_conv3x3_dw": (I, [P, I, P, I, I, I, P, P, P, P]),
    "lh2hip_dn_dw_scratch_bytes": (ctypes.c_uint64, []),
    "lh2hip_dn_conv3x3_dw_noalloc": (I, [P, I, P, I, I, I, P, P, P]),
    "lh2hip_dn_relu_bwd": (I, [P, P, P, ctypes.c_uint64, P]),
    "lh2hip_dn_pool2": (I, [P, I, I, I, P, P]),
    "lh2hip_dn_pool2_bwd": (I, [P, I, I, I, P, I, P]),
    "lh2hip_dn_upcat": (I, [P, I, P, I, I, I, P, P]),
    "lh2hip_dn_conv3x3_cat": (I, [P, P, P, P, P, I, I, I, I, P]),
    "lh2hip_dn_conv3x3_pool": (I, [P, I, P, P, P, I, I, I, P, P]),
    "lh2hip_dn_upcat_bwd": (I, [P, I, I, I, I, P, P, P]),
    "lh2hip_dn_pack_input": (I, [P, P, P, P, I, I, P, P]),
    "lh2hip_dn_relmse": (I, [P, I, P, I, I, P, P, P]),
    "lh2hip_dn_adam": (I, [P, P, P, P, ctypes.c_uint64, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, I,
                           P, P, I, I, P]),
    "lh2hip_svgf_create": (I, [I, I, ctypes.POINTER(P)]),
    "lh2hip_svgf_destroy": (I, [P]),
    "lh2hip_svgf_frame": (I, [P, P, P, P, P, P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, I,
                              ctypes.c_float, ctypes.c_float, I, P, P]),
    "lh2hip_svgf_read": (I, [P, I, P]),
    "lh2hip_svgf_set_taa": (I, [P, I]),
    "lh2hip_server_create": (I, [P, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(P)]),
    "lh2hip_server_destroy": (I, [P]),
    "lh2hip_server_submit": (I, [P, P, I, I, I, I, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "lh2hip_server_submit_denoised": (I, [P, P, I, I, I, I, ctypes.c_uint32, I, ctypes.POINTER(ctypes.c_uint64)]),
    "lh2hip_server_done": (I, [P, ctypes.c_uint64, ctypes.POINTER(I)]),
    "lh2hip_server_wait": (I, [P, ctypes.c_uint64, ctypes.c_double]),
    "lh2hip_server_read": (I, [P, ctypes.c_uint64, P, P]),
    "lh2hip_server_times": (I, [P, ctypes.c_uint64, P]),
    "lh2hip_server_set_denoiser": (I, [P, P, ctypes.c_uint64]),
    "lh2hip_plugin_read_image": (I, [P, P]),
    "lh2hip_plugin_read_accumulator_f64": (I, [P, P]),
    "lh2hip_server_stats": (I, [P, P]),
    "lh2hip_server_pause": (I, [P, I]),
    "lh2hip_dn_weights_bytes": (ctypes.c_uint64, []),
    "lh2hip_dn_frame_input": (I, [P, P, ctypes.c_float, I, I, I, I, P, P]),
    "lh2hip_dn_conv3x3_display": (I, [P, P, P, P, I, I, P, I, I, P]),
    "lh2hip_dn_frame_output": (I, [P, I, I, P, I, I, P, P]),
}
EXPORTED = sorted(_SIGS) + ["CreateCore"]


def library():
    """Loads librendercore_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        # PyTorch ships its own libamdhip64 (same soname): load it first so the
        # core binds to that runtime and device pointers are shared.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise LH2HIPError(rc, library().lh2hip_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data


_p = _ptr


class RenderCore:
    """One RenderCore_HIP instance (CoreAPI_Base::CreateCoreAPI + Init)."""

    def __init__(self):
        h = P()
        _check(library().lh2hip_create(ctypes.byref(h)))
        self._h = h
        self._keep = []

    def Shutdown(self):
        if self._h:
            _check(library().lh2hip_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.Shutdown()
        except Exception:
            pass

    # ---- CoreAPI_Base -------------------------------------------------------
    def SetTarget(self, width, height, spp=1):
        self.width, self.height, self.spp = width, height, spp
        _check(library().lh2hip_set_target(self._h, width, height, spp))

    def Setting(self, name, value):
        _check(library().lh2hip_setting(self._h, name.encode(), float(value)))

    def SetGeometry(self, meshIdx, vertexData, triangles=None):
        v = np.ascontiguousarray(vertexData, np.float32)
        if v.ndim != 2 or v.shape[1] != 4 or len(v) % 3 != 0:
            raise ValueError("SetGeometry: vertexData must be (3n, 4) float32 (three float4 per triangle)")
        t = None if triangles is None else np.ascontiguousarray(triangles, F.CoreTri)
        if t is not None and (t.ndim != 1 or len(t) != len(v) // 3):
            raise ValueError(f"SetGeometry: {len(v) // 3} triangles in vertexData but {t.size} CoreTri records")
        _check(library().lh2hip_set_geometry(self._h, meshIdx, _ptr(v), len(v), len(v) // 3, _ptr(t)))

    def SetInstance(self, instanceIdx, meshIdx, transform=None):
        m = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(16)
        _check(library().lh2hip_set_instance(self._h, instanceIdx, meshIdx, _ptr(m)))

    def FinalizeInstances(self):
        _check(library().lh2hip_finalize_instances(self._h))

    def SetMaterials(self, materials):
        m = np.ascontiguousarray(materials, F.CoreMaterial)
        _check(library().lh2hip_set_materials(self._h, _ptr(m), len(m)))

    def SetLights(self, tri=None, point=None, spot=None, directional=None):
        a = [np.ascontiguousarray(x if x is not None else np.zeros(0, d), d)
             for x, d in ((tri, F.CoreLightTri), (point, F.CorePointLight), (spot, F.CoreSpotLight),
                          (directional, F.CoreDirectionalLight))]
        _check(library().lh2hip_set_lights(self._h, _ptr(a[0]), len(a[0]), _ptr(a[1]), len(a[1]),
                                           _ptr(a[2]), len(a[2]), _ptr(a[3]), len(a[3])))

    def SetTextures(self, textures):
        """CoreAPI_Base::SetTextures: list of texture dicts (formats.texture_descs)."""
        d, keep = F.texture_descs(textures)
        _check(library().lh2hip_set_textures(self._h, _ptr(d) if len(d) else None, len(d)))
        del keep

    def SetSkyData(self, pixels, width, height, worldToLight=None):
        p = np.ascontiguousarray(pixels, np.float32)
        m = None if worldToLight is None else np.ascontiguousarray(worldToLight, np.float32).reshape(16)
        _check(library().lh2hip_set_sky(self._h, _ptr(p), width, height, _ptr(m)))

    def SetBlueNoise(self, sequence, scrambling, ranking):
        """Blue-noise tables of the optix7 estimator (8-bit entries: 65536
        sequence, 131072 scrambling, 131072 ranking)."""
        t = [np.ascontiguousarray(x, np.uint8).reshape(-1) for x in (sequence, scrambling, ranking)]
        if t[0].size != 65536 or t[1].size != 131072 or t[2].size != 131072:
            raise ValueError("SetBlueNoise: sizes 65536, 131072, 131072")
        _check(library().lh2hip_set_blue_noise(self._h, _ptr(t[0]), _ptr(t[1]), _ptr(t[2])))

    def SetProbePos(self, x, y):
        _check(library().lh2hip_set_probe_pos(self._h, x, y))

    def Render(self, view, converge=LH2_RESTART, async_=False):
        v = np.ascontiguousarray(view, F.ViewPyramid)
        _check(library().lh2hip_render(self._h, _ptr(v), converge, 1 if async_ else 0))

    def WaitForRender(self):
        _check(library().lh2hip_wait(self._h))

    def GetCoreStats(self):
        s = np.zeros(1, F.CoreStats)
        _check(library().lh2hip_get_stats(self._h, _ptr(s)))
        return s[0]

    # ---- extensions -----------------------------------------------------------
    def read_image(self):
        out = np.zeros((self.height, self.width, 4), np.float32)
        _check(library().lh2hip_read_image(self._h, _ptr(out)))
        return out

    def read_accumulator_f64(self):
        out = np.zeros((self.height, self.width, 4), np.float64)
        _check(library().lh2hip_read_accumulator_f64(self._h, _ptr(out)))
        return out

    def read_buffer(self, which):
        """Internal frame buffers: 'acc64' (2, h, w, 4) float64 (direct, indirect halves),
        'features' (h, w, 4) uint32, 'world_pos' / 'delta_depth' (h, w, 4) float32,
        'acc32' (2, h, w, 4) float32."""
        k, shape, dt = {"acc64": (0, (2, self.height, self.width, 4), np.float64),
                        "features": (1, (self.height, self.width, 4), np.uint32),
                        "world_pos": (2, (self.height, self.width, 4), np.float32),
                        "delta_depth": (3, (self.height, self.width, 4), np.float32),
                        "acc32": (4, (2, self.height, self.width, 4), np.float32)}[which]
        out = np.zeros(shape, dt)
        _check(library().lh2hip_read_buffer(self._h, k, _ptr(out), out.nbytes))
        return out

    def trace_closest(self, rays_ptr, hits_ptr, count, stream=None):
        """rays/hits: device pointers (e.g. torch tensor .data_ptr())."""
        _check(library().lh2hip_trace_closest(self._h, rays_ptr, hits_ptr, count, stream))

    def trace_closest_counted(self, rays_ptr, hits_ptr, count, stream=None):
        """As trace_closest; returns (nodes visited, triangles tested, node
        steps of all waves, lanes visiting a node in them, leaf steps, lanes
        testing a triangle in them, loop iterations and lanes holding a ray
        while rays remain, the same after the last ray was handed out, ns
        from the first wave's start to the feed running dry, ns to the end)."""
        out = np.zeros(12, np.uint64)
        _check(library().lh2hip_trace_closest_counted(self._h, rays_ptr, hits_ptr, count, stream, _p(out)))
        return tuple(int(x) for x in out)

    def trace_any(self, rays_ptr, mask_ptr, count, stream=None):
        _check(library().lh2hip_trace_any(self._h, rays_ptr, mask_ptr, count, stream))

    def bounce_diffuse(self, rays_in, hits, path_in, count, R0, tmin, rays_out, path_out, count_out, stream=None):
        _check(library().lh2hip_bounce_diffuse(self._h, rays_in, hits, path_in, count, R0, float(tmin),
                                               rays_out, path_out, count_out, stream))

    def bvh_info(self, mesh=0):
        out = np.zeros(8, np.int64)
        _check(library().lh2hip_bvh_info(self._h, mesh, _ptr(out)))
        keys = ("build_nodes", "leaves", "depth", "nodes_f32", "nodes_q8", "depth_q8", "build_us", "triangles")
        return dict(zip(keys, (int(x) for x in out)))

    def bvh_download(self, mesh=0):
        info = self.bvh_info(mesh)
        nodes = np.zeros(info["build_nodes"], F.BuildNode)
        prims = np.zeros(info["triangles"], np.uint32)
        cnt = ctypes.c_int(0)
        _check(library().lh2hip_bvh_download(self._h, mesh, _ptr(nodes), len(nodes), _ptr(prims), len(prims),
                                             ctypes.byref(cnt)))
        return nodes, prims

    def bvh_layout_download(self, mesh=0):
        """The device-built traversal layout: f32 nodes and q8 nodes (64 bytes
        each, as uint32 rows of 16) and the leaf triangles (float32 rows of 12)."""
        info = self.bvh_info(mesh)
        f32 = np.zeros((info["nodes_f32"], 16), np.uint32)
        q8 = np.zeros((info["nodes_q8"], 16), np.uint32)
        tris = np.zeros((info["triangles"], 12), np.float32)
        _check(library().lh2hip_bvh_layout_download(self._h, mesh, _ptr(f32), len(f32), _ptr(q8), len(q8), _ptr(tris), len(tris)))
        return f32, q8, tris


class RenderServer:
    """Render job server over a core's current scene (lh2hip_server_*): submit
    jobs of any size/spp/depth/seed, wait, read images. A job's image equals
    RenderCore.Render(view, LH2_RESTART) with the same settings."""

    def __init__(self, core, budget_bytes=8 << 30, pass_paths=1 << 21):
        h = P()
        _check(library().lh2hip_server_create(core._h, budget_bytes, pass_paths, ctypes.byref(h)))
        self._h, self._sizes = h, {}
        self._core = core          # the server reads the core's scene: keep the core alive as long as the server

    def submit(self, view, width, height, spp=1, max_depth=3, seed=0x12345678, denoise=0):
        """denoise 1: filter estimator + SVGF for this job; 2: the same frame
        through the learned denoiser (set_denoiser first)."""
        v = np.ascontiguousarray(view, F.ViewPyramid)
        jid = ctypes.c_uint64()
        if denoise:
            _check(library().lh2hip_server_submit_denoised(self._h, _ptr(v), width, height, spp, max_depth, seed, denoise, ctypes.byref(jid)))
        else:
            _check(library().lh2hip_server_submit(self._h, _ptr(v), width, height, spp, max_depth, seed, ctypes.byref(jid)))
        self._sizes[jid.value] = (height, width)
        return jid.value

    def set_denoiser(self, blob):
        """Weights of the learned denoiser (UNetDenoiser.weights_blob())."""
        b = np.ascontiguousarray(np.frombuffer(blob, np.uint8) if isinstance(blob, (bytes, bytearray)) else blob)
        _check(library().lh2hip_server_set_denoiser(self._h, _ptr(b), b.nbytes))

    def pause(self, paused=True):
        """Paused: no new pass starts (the running one finishes) until pause(False)."""
        _check(library().lh2hip_server_pause(self._h, 1 if paused else 0))

    def stats(self):
        """Batching counters: passes, job parts, passes shared by several jobs,
        most jobs in one pass, paths (eye rays), extension rays traced, and
        the device ms of the extension-traversal and shade launches."""
        o = np.zeros(8, np.uint64)
        _check(library().lh2hip_server_stats(self._h, _ptr(o)))
        d = dict(zip(("passes", "parts", "shared_passes", "max_parts", "paths", "rays"), (int(x) for x in o[:6])))
        d["trace_ms"], d["shade_ms"] = float(o[6]) * 1e-3, float(o[7]) * 1e-3
        return d

    def done(self, jid):
        d = ctypes.c_int()
        _check(library().lh2hip_server_done(self._h, jid, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, jid, timeout=-1.0):
        _check(library().lh2hip_server_wait(self._h, jid, timeout))

    def times(self, jid):
        t = np.zeros(3, np.float64)
        _check(library().lh2hip_server_times(self._h, jid, _ptr(t)))
        return t

    def read(self, jid, with_acc=False):
        if jid not in self._sizes:
            raise LH2HIPError(-1, "server: unknown job (never submitted or already read)")
        h, w = self._sizes[jid]
        img = np.zeros((h, w, 4), np.float32)
        acc = np.zeros((h, w, 4), np.float64) if with_acc else None
        _check(library().lh2hip_server_read(self._h, jid, _ptr(img), _ptr(acc) if with_acc else None))
        del self._sizes[jid]
        return (img, acc) if with_acc else img

    def close(self):
        if getattr(self, "_h", None):
            library().lh2hip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Svgf:
    """SVGF filter state on the GPU (lh2hip_svgf_*); frame() takes device
    pointers, read() returns internal buffers as numpy arrays."""

    def __init__(self, width, height, taa=False):
        h = P()
        _check(library().lh2hip_svgf_create(width, height, ctypes.byref(h)))
        self._h, self.w, self.h = h, width, height
        if taa:
            _check(library().lh2hip_svgf_set_taa(self._h, 1))

    def frame(self, acc_ptr, features_ptr, world_pos_ptr, delta_depth_ptr, prev_view, out_ptr, j=(0.0, 0.0), prev_j=(0.0, 0.0),
              samples=1, clamp_direct=2.5, clamp_indirect=15.0, stationary=0, stream=None):
        v = np.ascontiguousarray(prev_view, F.ViewPyramid)
        _check(library().lh2hip_svgf_frame(self._h, acc_ptr, features_ptr, world_pos_ptr, delta_depth_ptr, _ptr(v), j[0], j[1],
                                           prev_j[0], prev_j[1], samples, clamp_direct, clamp_indirect, stationary, out_ptr, stream))

    def read(self, which):
        shape = {0: ((self.h, self.w, 4), np.uint32), 2: ((self.h, self.w, 2), np.float32)}.get(which, ((self.h, self.w, 4), np.float32))
        out = np.zeros(shape[0], shape[1])
        _check(library().lh2hip_svgf_read(self._h, which, _ptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            library().lh2hip_svgf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def generate_primary(view, width, height, R0, tmin, rays_ptr, path_ptr=None, stream=None):
    v = np.ascontiguousarray(view, F.ViewPyramid)
    _check(library().lh2hip_generate_primary(_ptr(v), width, height, R0, float(tmin), rays_ptr, path_ptr, stream))


def load_scene(core, scene):
    """RenderSystem::SynchronizeSceneData call order (rendersystem.cpp:46-211):
    sky, textures, materials, geometry per mesh, lights, instances (+
    truncation), FinalizeInstances."""
    if getattr(scene, "sky", None) is not None:
        px, T = scene.sky
        core.SetSkyData(px.reshape(-1, 3), px.shape[1], px.shape[0], T)
    if getattr(scene, "textures", None):
        core.SetTextures(scene.textures)
    core.SetMaterials(scene.materials if len(scene.materials) else F.new_materials(1))
    for i, (v, t) in enumerate(scene.meshes):
        core.SetGeometry(i, v, t)
    core.SetLights(scene.tri_lights, scene.point_lights, scene.spot_lights, scene.dir_lights)
    for i, (m, T) in enumerate(scene.instances):
        core.SetInstance(i, m, T)
    core.SetInstance(len(scene.instances), -1)
    core.FinalizeInstances()
