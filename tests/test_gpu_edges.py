"""Edge sizes through the boundary on the GPU, each against the oracle: a
1 x 1 and a 3 x 2 target for the reference and optix7 estimators, a ragged
7 x 5 filter-estimator sequence (SVGF's border rows and columns, TAA), a
target too small for SVGF (an error, not a fault), and a scene whose only
mesh has no triangles (a black frame)."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import formats as F
from lighthouse2_amd import scenes
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h", [(1, 1), (3, 2)])
@pytest.mark.parametrize("estimator", [0, 1])
def test_tiny_targets(torch_cuda, w, h, estimator):
    s = scenes.cornell_box(sphere_tris=500)
    view = scenes.cornell_view(w, h)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(w, h, 3)
    core.Setting("estimator", estimator)
    core.Setting("maxDepth", 3)
    for m in [lh.LH2_RESTART, lh.LH2_CONVERGE]:
        core.Render(view, m)
    acc = core.read_accumulator_f64()
    core.Shutdown()
    if estimator == 0:
        r64 = orc.render(s, view, w, h, 3, 1, max_depth=3)[1]
    else:
        r64 = orc.render7(s, view, w, h, 3, 1)[1]
    assert acc.shape == (h, w, 4) and np.array_equal(acc, r64)


def test_ragged_filter_sequence(torch_cuda):
    s = scenes.cornell_box(sphere_tris=500)
    v0 = scenes.cornell_view(7, 5)
    v1 = scenes.camera_view((0.51, 0.5, -1.2), (0.5, 0.5, 0.5), 40.0, 7, 5)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(7, 5, 1)
    core.Setting("estimator", 2)
    core.Setting("bsdf", 0)
    ref = orc.RenderF(s, 7, 5, spp=1, bsdf=0)
    for v, m in [(v0, orc.RESTART), (v0, orc.CONVERGE), (v1, orc.RESTART)]:
        core.Render(v, lh.LH2_RESTART if m == orc.RESTART else lh.LH2_CONVERGE)
        want = ref.frame(v, m)
        img = core.read_image()
        assert np.array_equal(img[1:-1, 1:-1].view(np.uint32), want[1:-1, 1:-1].view(np.uint32))
        assert np.array_equal(core.read_buffer("acc64")[0], ref.acc64.reshape(2, 5, 7, 4)[0])
    core.Shutdown()


def test_target_too_small_for_svgf_is_an_error(torch_cuda):
    s = scenes.cornell_box(sphere_tris=200)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(2, 2, 1)
    core.Setting("estimator", 2)
    with pytest.raises(lh.LH2HIPError):
        core.Render(scenes.cornell_view(2, 2), lh.LH2_RESTART)
    core.Setting("filter", 0)                        # without the filter the size is fine
    core.Render(scenes.cornell_view(2, 2), lh.LH2_RESTART)
    assert np.isfinite(core.read_image()).all()
    core.Shutdown()


def test_empty_mesh_renders_black(torch_cuda):
    s = scenes.Scene()
    s.add_material(F.lambert_material((0.5, 0.5, 0.5)))
    s.add_instance(s.add_mesh(np.zeros((0, 4), np.float32), None))
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(16, 8, 1)
    core.Render(scenes.cornell_view(16, 8), lh.LH2_RESTART)
    img = core.read_image()
    core.Shutdown()
    assert img.shape == (8, 16, 4) and np.all(img[..., :3] == 0)
