"""Headless RenderSystem over a core: the per-frame synchronisation and render
forwarding of lib/RenderSystem/rendersystem.cpp, for hosts that drive the
core from Python (scene arrays from `scenes` / `loaders`).

* `synchronize()` is RenderSystem::SynchronizeSceneData (rendersystem.cpp:
  46-211): sky, textures, materials, meshes, lights, then the instance list
  and FinalizeInstances, sending only what changed since the previous call.
  The reference detects changes with a crc64 of each object's bytes
  (TRACKCHANGES, lib/platform/system.h:161-167); here a crc32 of each
  array's bytes does the same job. As in the reference, all textures are
  sent when any changed (SynchronizeTextures :59-72), all materials when any
  changed (:79-97), each changed mesh on its own (SynchronizeMeshes :106-118;
  with the core's "refit" setting a mesh of unchanged size is refitted), all
  lights when any changed (:162-184), and the whole instance list when a
  transform or mesh index changed, a mesh was re-sent, or the count changed
  (UpdateSceneGraph :127-155); FinalizeInstances runs on every call.
* `render()` is RenderSystem::Render (:217-227): the RenderSettings
  (rendersystem.h:65-72) and the camera's clamp value (camera.h:43) are
  forwarded as core settings before each frame.
"""
import zlib
from dataclasses import dataclass

import numpy as np

from . import formats as F
from .core import LH2_CONVERGE


@dataclass
class RenderSettings:
    """RenderSettings (rendersystem.h:65-72) plus Camera::clampValue."""
    geometryEpsilon: float = 1.0e-3
    filterDirectClamp: float = 15.0
    filterIndirectClamp: float = 2.5
    filterEnabled: int = 1
    TAAEnabled: int = 1
    clampValue: float = 10.0


def _crc(*arrays):
    c = 0
    for a in arrays:
        if a is None:
            c = zlib.crc32(b"\0none", c)
            continue
        a = np.ascontiguousarray(a)
        c = zlib.crc32(str((a.dtype.str, a.shape)).encode(), c)
        c = zlib.crc32(a.view(np.uint8).reshape(-1).data, c)
    return c


class RenderSystem:
    def __init__(self, core, scene, settings=None):
        self.core, self.scene = core, scene
        self.settings = settings or RenderSettings()
        self._sky = self._tex = self._mat = self._lights = None
        self._meshes = []
        self._instances = None
        self.sent = []          # what the last synchronize() sent, in call order

    def _send(self, what, *args):
        getattr(self.core, what)(*args)
        self.sent.append((what,) + tuple(a for a in args if isinstance(a, (int, np.integer))))

    def synchronize(self):
        s, self.sent = self.scene, []
        sky = getattr(s, "sky", None)
        if sky is not None:
            c = _crc(sky[0], np.asarray(sky[1], np.float32))
            if c != self._sky:
                px, T = sky
                self._send("SetSkyData", px.reshape(-1, 3), px.shape[1], px.shape[0], T)
                self._sky = c
        textures = getattr(s, "textures", None) or []
        c = _crc(*[np.array([t["width"], t["height"], t["flags"], t["storage"]], np.int64) for t in textures],
                 *[t["pixels"] for t in textures])
        if textures and c != self._tex:
            self._send("SetTextures", textures)
            self._tex = c
        mats = s.materials if len(s.materials) else F.new_materials(1)
        c = _crc(mats)
        if c != self._mat:
            self._send("SetMaterials", mats)
            self._mat = c
        meshes_changed = False
        for i, (v, t) in enumerate(s.meshes):
            c = _crc(v, t)
            if i >= len(self._meshes) or self._meshes[i] != c:
                self._send("SetGeometry", i, v, t)
                if i >= len(self._meshes):
                    self._meshes.append(c)
                else:
                    self._meshes[i] = c
                meshes_changed = True
        c = _crc(s.tri_lights, s.point_lights, s.spot_lights, s.dir_lights)
        if c != self._lights:
            self._send("SetLights", s.tri_lights, s.point_lights, s.spot_lights, s.dir_lights)
            self._lights = c
        inst = [(int(m), np.asarray(T, np.float32).reshape(16).copy()) for m, T in s.instances]
        key = _crc(np.array([m for m, _ in inst], np.int64), *[T for _, T in inst])
        if meshes_changed or self._instances != key:
            for i, (m, T) in enumerate(inst):
                self._send("SetInstance", i, m, T)
            self._send("SetInstance", len(inst), -1)
            self._instances = key
        self._send("FinalizeInstances")
        return self.sent

    def render(self, view, converge=LH2_CONVERGE, async_=False):
        st = self.settings
        for key, value in (("epsilon", st.geometryEpsilon), ("clampValue", st.clampValue),
                           ("clampDirect", st.filterDirectClamp), ("clampIndirect", st.filterIndirectClamp),
                           ("filter", st.filterEnabled), ("TAA", st.TAAEnabled)):
            self.core.Setting(key, value)
        self.core.Render(view, converge, async_)

    def wait(self):
        self.core.WaitForRender()
