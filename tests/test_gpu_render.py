"""GPU render parity: the wavefront kernels (generate, extend, shade, connect,
resolve, finalize) against the oracle's restatement of the reference
estimator, through the C ABI. Per-path arithmetic is reproduced bit for bit,
so the fp64 accumulators must be identical and the fp32 image within 1e-4
relative of the fp64 oracle (tolerance of the parity contract)."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import scenes
import oracle as orc

pytestmark = pytest.mark.gpu
REL_TOL = 1e-4


def render_gpu(scene, view, w, h, spp, frames, depth, eps=1e-3, seed=0x12345678, fmt=1):
    core = lh.RenderCore()
    core.Setting("nodeFormat", fmt)
    lh.load_scene(core, scene)
    core.SetTarget(w, h, spp)
    core.Setting("maxDepth", depth)
    core.Setting("epsilon", eps)
    if seed != 0x12345678:
        core.Setting("seed", seed)         # a float setting: seeds up to 2^24 are exact
    core.Render(view, lh.LH2_RESTART)          # camera moved: preview frame
    for f in range(frames):
        core.Render(view, lh.LH2_CONVERGE)     # the first converging frame restarts accumulation
    img, acc = core.read_image(), core.read_accumulator_f64()
    stats = core.GetCoreStats()
    core.Shutdown()
    return img, acc, stats


def errors(img, ref):
    rel = np.abs(img[..., :3].astype(np.float64) - ref[..., :3]) / np.maximum(np.abs(ref[..., :3]), 1e-6)
    return float(rel.mean()), float(rel.max())


@pytest.mark.parametrize("fmt", [1, 0])
def test_cornell_box_matches_oracle(torch_cuda, fmt):
    """Cornell box, 256 x 256, 1 spp x 16 frames, max depth 3, seed 1234."""
    s = scenes.cornell_box(seed=1234)
    view = scenes.cornell_view(256, 256)
    img, acc, st = render_gpu(s, view, 256, 256, 1, 16, 3, seed=1234, fmt=fmt)
    r32, r64, n = orc.render(s, view, 256, 256, 1, 16, 3, 1e-3, seed=1234, threads=16)
    mean_e, max_e = errors(img, r64 / n)
    print(f"fmt {fmt}: image vs fp64 oracle: mean rel {mean_e:.3e} max rel {max_e:.3e}; "
          f"radiance mean {(r64[..., :3] / n).mean():.5f}")
    assert max_e <= REL_TOL
    assert np.array_equal(acc, r64)                      # fp64 accumulators identical
    assert np.array_equal(img, (r32 * np.float32(1.0 / n)).astype(np.float32))  # fp32 image identical
    assert n == 16 and st["primaryRayCount"] == 256 * 256


def test_multi_sample_resolve_is_ordered(torch_cuda):
    s = scenes.cornell_box(seed=7, sphere_tris=4000)
    view = scenes.cornell_view(48, 40)
    img, acc, _ = render_gpu(s, view, 48, 40, 4, 3, 4, seed=99)
    r32, r64, n = orc.render(s, view, 48, 40, 4, 3, 4, 1e-3, seed=99)
    assert n == 12
    assert np.array_equal(acc, r64)
    assert errors(img, r64 / n)[1] <= REL_TOL
    img2, acc2, _ = render_gpu(s, view, 48, 40, 4, 3, 4, seed=99)
    assert np.array_equal(acc, acc2)  # deterministic across runs


def test_probe_and_call_order(torch_cuda):
    s = scenes.cornell_box(sphere_tris=2000)
    view = scenes.cornell_view(32, 32)
    core = lh.RenderCore()
    with pytest.raises(lh.LH2HIPError):
        core.Render(view)                      # no target yet
    core.SetTarget(32, 32, 1)
    core.Render(view)                          # no scene yet: no-op
    lh.load_scene(core, s)
    core.SetProbePos(16, 16)
    core.Setting("maxDepth", 2)
    core.Render(view)
    st = core.GetCoreStats()
    rays = orc.primary_rays(view, 32, 32, orc.xor32_host(0x12345678)[0], 1e-3)
    h = orc.OracleScene(s).trace(rays[16 * 32 + 16:16 * 32 + 17]).view(np.int32)
    assert (st["probedInstid"], st["probedTriid"]) == (h[0, 1], h[0, 2])
    core.Shutdown()


def test_no_lights_renders_only_emitters(torch_cuda):
    s = scenes.cornell_box(sphere_tris=2000)
    s.tri_lights = s.tri_lights[:0]
    view = scenes.cornell_view(32, 32)
    img, acc, _ = render_gpu(s, view, 32, 32, 1, 2, 3)
    r32, r64, n = orc.render(s, view, 32, 32, 1, 2, 3, 1e-3)
    assert np.array_equal(acc, r64)
    vals = set(np.unique(img[..., :3]).tolist())
    assert vals <= {0.0, 7.5, 15.0} and 15.0 in vals  # only directly seen emitters (1 or 2 of 2 frames)


def test_textured_cornell_box_with_sky(torch_cuda):
    """Diffuse (trilinear, alpha-masked corner), detail diffuse, normal and
    roughness maps and the skydome through the open front: accumulators
    identical to the oracle's."""
    s = scenes.textured_cornell_box(sphere_tris=4000)
    view = scenes.cornell_view(96, 96)
    img, acc, _ = render_gpu(s, view, 96, 96, 2, 3, 4, seed=77)
    r32, r64, n = orc.render(s, view, 96, 96, 2, 3, 4, 1e-3, seed=77)
    assert np.array_equal(acc, r64)
    assert errors(img, r64 / n)[1] <= REL_TOL
    plain = scenes.cornell_box(seed=1234, sphere_tris=4000)       # same box without maps and sky
    _, p64, _ = orc.render(plain, view, 96, 96, 2, 3, 4, 1e-3, seed=77)
    assert not np.array_equal(p64, r64)


def test_sky_seen_by_primary_rays(torch_cuda):
    s = scenes.sky_scene()
    view = scenes.camera_view((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), 70.0, 128, 96)
    img, acc, _ = render_gpu(s, view, 128, 96, 1, 2, 2, seed=5)
    r32, r64, n = orc.render(s, view, 128, 96, 1, 2, 2, 1e-3, seed=5)
    assert np.array_equal(acc, r64)
    sky_px = s.sky[0].reshape(-1, 3)
    seen = np.isin(img[..., 0], sky_px[:, 0])
    assert seen.mean() > 0.5                 # most pixels show a sky texel exactly
