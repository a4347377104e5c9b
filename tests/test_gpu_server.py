"""Render job server: jobs of mixed resolution, spp, depth and seed,
scheduled in quanta under an HBM budget, must each produce exactly the image
RenderCore.Render(view, Restart) produces for them alone; a job denoised by
SVGF exactly the filtered first frame of RenderCore.Render with the filter
estimator and TAA off."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import scenes

pytestmark = pytest.mark.gpu


def solo(scene, view, w, h, spp, depth, seed):
    core = lh.RenderCore()
    lh.load_scene(core, scene)
    core.SetTarget(w, h, spp)
    core.Setting("maxDepth", depth)
    core.Setting("seed", seed)
    core.Render(view, lh.LH2_RESTART)
    img, acc = core.read_image(), core.read_accumulator_f64()
    core.Shutdown()
    return img, acc


def test_job_alone_equals_render(torch_cuda):
    s = scenes.cornell_box(sphere_tris=2000)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 30, pass_paths=1 << 20)
    view = scenes.cornell_view(64, 48)
    j = srv.submit(view, 64, 48, spp=3, max_depth=4, seed=1234)
    srv.wait(j, timeout=60)
    t = srv.times(j)
    img, acc = srv.read(j, with_acc=True)
    ref_img, ref_acc = solo(s, view, 64, 48, 3, 4, 1234)
    assert np.array_equal(acc, ref_acc) and np.array_equal(img, ref_img)
    assert t[0] <= t[1] <= t[2]
    srv.close()
    core.Shutdown()


def test_mixed_jobs_match_their_solo_images(torch_cuda):
    s = scenes.cornell_box(sphere_tris=2000)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    jobs = [(96, 64, 4, 3, 11), (32, 32, 1, 5, 12), (128, 96, 2, 4, 13), (48, 40, 8, 2, 14), (64, 64, 3, 6, 15)]
    # budget for about two jobs at a time, quanta of 4096 paths: jobs interleave and wait for admission
    srv = lh.RenderServer(core, budget_bytes=2 * 128 * 96 * 64, pass_paths=4096)
    ids = [srv.submit(scenes.cornell_view(w, h), w, h, spp, d, seed) for (w, h, spp, d, seed) in jobs]
    for j in ids:
        srv.wait(j, timeout=120)
    times = [srv.times(j) for j in ids]
    out = [srv.read(j, with_acc=True) for j in ids]
    for (w, h, spp, d, seed), (img, acc) in zip(jobs, out):
        ri, ra = solo(s, scenes.cornell_view(w, h), w, h, spp, d, seed)
        assert np.array_equal(acc, ra), (w, h, spp, d, seed)
        assert np.array_equal(img, ri)
    assert max(t[1] for t in times) > min(t[0] for t in times)   # someone waited for admission
    srv.close()
    core.Shutdown()


def test_server_errors(torch_cuda):
    s = scenes.cornell_box(sphere_tris=500)
    core = lh.RenderCore()
    with pytest.raises(lh.LH2HIPError):
        lh.RenderServer(core)                      # no scene yet
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=32 * 32 * 64, pass_paths=1024)
    with pytest.raises(lh.LH2HIPError):
        srv.submit(scenes.cornell_view(64, 64), 64, 64)   # larger than the budget
    with pytest.raises(lh.LH2HIPError):
        srv.submit(scenes.cornell_view(8, 8), 8, 8, spp=0)
    j = srv.submit(scenes.cornell_view(32, 32), 32, 32, spp=2, max_depth=2)
    srv.wait(j, timeout=60)
    srv.read(j)
    with pytest.raises(lh.LH2HIPError):
        srv.read(j)                                # already released
    with pytest.raises(lh.LH2HIPError):
        srv.done(12345)
    srv.close()
    core.Shutdown()


def solo_filtered(scene, view, w, h, spp, seed):
    core = lh.RenderCore()
    lh.load_scene(core, scene)
    core.SetTarget(w, h, spp)
    core.Setting("estimator", 2)
    core.Setting("TAA", 0)
    core.Setting("seed", seed)
    core.Render(view, lh.LH2_RESTART)
    img, acc = core.read_image(), core.read_accumulator_f64()
    core.Shutdown()
    return img, acc


def test_svgf_jobs_match_filtered_render(torch_cuda):
    s = scenes.cornell_box(sphere_tris=2000)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 30, pass_paths=3000)
    jobs = [(64, 48, 2, 21, 1), (40, 40, 1, 22, 0), (64, 48, 3, 23, 1)]     # two SVGF jobs of one size share a filter
    ids = [srv.submit(scenes.cornell_view(w, h), w, h, spp, 4, seed, denoise=dn) for (w, h, spp, seed, dn) in jobs]
    for j in ids:
        srv.wait(j, timeout=120)
    out = [srv.read(j, with_acc=True) for j in ids]
    for (w, h, spp, seed, dn), (img, acc) in zip(jobs, out):
        ri, ra = solo_filtered(s, scenes.cornell_view(w, h), w, h, spp, seed) if dn else solo(s, scenes.cornell_view(w, h), w, h, spp, 4, seed)
        assert np.array_equal(acc, ra), (w, h, spp, seed, dn)
        assert np.array_equal(img, ri), (w, h, spp, seed, dn)
        assert np.isfinite(img).all() and img[..., :3].sum() > 0
    with pytest.raises(lh.LH2HIPError):
        srv.submit(scenes.cornell_view(2, 2), 2, 2, 1, 4, 1, denoise=1)
    srv.close()
    core.Shutdown()


def test_more_persistent_blocks_after_server_start(torch_cuda):
    """traceBlocksPerCU raised after the server was created: the server grows
    its traversal overflow stacks before the next quantum; the job still
    equals its solo render."""
    s = scenes.cornell_box(sphere_tris=2000)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 30, pass_paths=1 << 16)
    core.Setting("traceBlocksPerCU", 12)
    j = srv.submit(scenes.cornell_view(48, 40), 48, 40, spp=2, max_depth=4, seed=77)
    srv.wait(j, timeout=60)
    img, acc = srv.read(j, with_acc=True)
    ri, ra = solo(s, scenes.cornell_view(48, 40), 48, 40, 2, 4, 77)
    assert np.array_equal(acc, ra) and np.array_equal(img, ri)
    srv.close()
    core.Shutdown()
