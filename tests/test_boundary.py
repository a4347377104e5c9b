"""The drop-in boundary: librendercore_hip.so loads, exports every entry point
include/lh2hip.h declares plus CreateCore, and the numpy mirrors of the
boundary structs match the C layouts. CPU only (no kernel launches)."""
import os
import re
import subprocess

import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import formats as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "lh2hip.h")).read()
    return sorted(set(re.findall(r"LH2HIP_API\s+[\w\s\*]+?\b(lh2hip_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = lh.library()
    names = declared_entry_points()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert hasattr(lib, "CreateCore")
    assert set(names) <= set(lh.core.EXPORTED)
    assert lib.lh2hip_version() == 1
    out = subprocess.run(["nm", "-D", "--defined-only", lh.core.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    assert set(names) | {"CreateCore"} <= exported


def test_struct_layouts_match_c():
    prog = r'''
#include "lh2hip_types.h"
#include <cstdio>
#include <cstddef>
using namespace lh2hip;
#define P(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main() {
  printf("CoreTri size %zu\nCoreMaterial size %zu\nCoreLightTri size %zu\nViewPyramid size %zu\nCoreStats size %zu\n",
    sizeof(CoreTri), sizeof(CoreMaterial), sizeof(CoreLightTri), sizeof(ViewPyramid), sizeof(CoreStats));
  P(CoreTri, material) P(CoreTri, vN0) P(CoreTri, alpha) P(CoreTri, vertex2) P(CoreTri, v1_2)
  P(CoreMaterial, absorption) P(CoreMaterial, roughness) P(CoreMaterial, eta) P(CoreMaterial, urough)
  P(CoreMaterial, thin) P(CoreMaterial, specTrans) P(CoreMaterial, opacity)
  P(CoreLightTri, radiance) P(CoreLightTri, vertex2) P(ViewPyramid, distortion)
  P(CoreStats, bvhBuildTime) P(CoreStats, probedWorldPos) P(CoreSpotLight, direction)
}
'''
    exe = os.path.join(ROOT, "_build", "layout_probe")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    srcf = exe + ".cpp"
    open(srcf, "w").write(prog)
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", exe, srcf], check=True)
    lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    dt = {"CoreTri": F.CoreTri, "CoreMaterial": F.CoreMaterial, "CoreLightTri": F.CoreLightTri,
          "ViewPyramid": F.ViewPyramid, "CoreStats": F.CoreStats, "CoreSpotLight": F.CoreSpotLight}
    for ln in filter(None, lines):
        t, f, v = ln.split()
        if f == "size":
            assert dt[t].itemsize == int(v), ln
        else:
            assert dt[t].fields[f][1] == int(v), ln


def test_create_without_device_reports_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(lh.LH2HIPError) as e:
        lh.RenderCore()
    assert e.value.code in (-2, -5)


def test_scene_producers():
    v, t = lh.scenes.add_quad((0, -1, 0), (0.5, 1.0, 0.5), 1.0, 1.0, 3)
    assert v.shape == (6, 4) and (v[:, 1] == 1.0).all() and (t["material"] == 3).all()
    assert np.allclose(t["Ny"], -1) and np.allclose(np.abs(t["T"]).sum(1), 1) and np.allclose(np.abs(t["B"]).sum(1), 1)
    s = lh.scenes.cornell_box()
    assert s.triangle_count == 5 * 2 + 20000 + 2 and len(s.instances) == 7
    L = s.tri_lights
    assert np.allclose(L["area"].sum(), 0.04, rtol=1e-5) and np.allclose(L["energy"], 15 * 3 * L["area"], rtol=1e-5)
    view = lh.scenes.cornell_view(256, 256)
    assert np.allclose(view["pos"], (0.5, 0.5, -1.2))
    c = (view["p1"] + (view["p2"] - view["p1"]) * 0.5 + (view["p3"] - view["p1"]) * 0.5)[0]
    d = c - view["pos"][0]
    assert np.allclose(d / np.linalg.norm(d), (0, 0, 1), atol=1e-6)  # image centre looks at the target


def test_set_geometry_rejects_mismatched_arrays():
    """SetGeometry checks its arrays on the host before the C call reads
    them: three float4 per triangle and exactly one CoreTri per triangle
    (a shorter CoreTri array would be over-read by the upload)."""
    class Dummy:
        _h = None
    v = np.zeros((6, 4), np.float32)
    with pytest.raises(ValueError):
        lh.RenderCore.SetGeometry(Dummy(), 0, np.zeros((5, 4), np.float32))
    with pytest.raises(ValueError):
        lh.RenderCore.SetGeometry(Dummy(), 0, np.zeros((6, 3), np.float32))
    with pytest.raises(ValueError):
        lh.RenderCore.SetGeometry(Dummy(), 0, v, F.new_tris(1))
    with pytest.raises(ValueError):
        lh.RenderCore.SetGeometry(Dummy(), 0, v, F.new_tris(3))
