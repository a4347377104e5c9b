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


def test_jobs_on_reused_buffers_match_their_solo_images(torch_cuda):
    """Finished jobs' buffers go back to a pool keyed by size and serve later
    jobs, also in another role: at 64 x 48 a plain job's fp64 accumulator has
    the size of a denoised job's fp32 accumulator, and its image and fp32
    accumulator the size of each G-buffer plane. Rounds of jobs that wait for
    the previous round must still equal their solo renders exactly."""
    s = scenes.cornell_box(sphere_tris=2000)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 30, pass_paths=1 << 20)
    view = scenes.cornell_view(64, 48)
    rounds = [[(2, 31, 1), (3, 32, 0)], [(2, 33, 0), (3, 34, 1)], [(1, 35, 1), (4, 36, 0), (2, 37, 0)]]
    for jobs in rounds:
        ids = [srv.submit(view, 64, 48, spp, 4, seed, denoise=dn) for (spp, seed, dn) in jobs]
        for j in ids:
            srv.wait(j, timeout=120)
        for (spp, seed, dn), j in zip(jobs, ids):
            img, acc = srv.read(j, with_acc=True)
            ri, ra = solo_filtered(s, view, 64, 48, spp, seed) if dn else solo(s, view, 64, 48, spp, 4, seed)
            assert np.array_equal(acc, ra), (spp, seed, dn)
            assert np.array_equal(img, ri), (spp, seed, dn)
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


def test_jobs_share_passes_and_match_solo(torch_cuda):
    """Small jobs of one estimator ride in the same passes (the batching
    counters say so) and each still equals its solo render bit for bit."""
    s = scenes.cornell_box(sphere_tris=2000)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 30, pass_paths=1 << 16)
    jobs = [(32, 32, 2, 3, 31), (40, 24, 1, 5, 32), (24, 24, 4, 4, 33), (48, 32, 3, 2, 34), (16, 16, 8, 6, 35), (64, 40, 2, 4, 36)]
    srv.pause()                                   # a burst: every job is queued before the first pass
    ids = [srv.submit(scenes.cornell_view(w, h), w, h, spp, d, seed) for (w, h, spp, d, seed) in jobs]
    srv.pause(False)
    for j in ids:
        srv.wait(j, timeout=120)
    st = srv.stats()
    out = [srv.read(j, with_acc=True) for j in ids]
    for (w, h, spp, d, seed), (img, acc) in zip(jobs, out):
        ri, ra = solo(s, scenes.cornell_view(w, h), w, h, spp, d, seed)
        assert np.array_equal(acc, ra), (w, h, spp, d, seed)
        assert np.array_equal(img, ri)
    assert st["shared_passes"] >= 1 and st["max_parts"] >= 2, st
    assert st["paths"] == sum(w * h * spp for (w, h, spp, _, _) in jobs)
    srv.close()
    core.Shutdown()


def test_scene_is_locked_while_jobs_are_pending(torch_cuda):
    """Changing the scene while jobs are queued fails (the jobs hold its device
    buffers); so does destroying a core whose server still exists; once the
    jobs are done the scene may change again."""
    from lighthouse2_amd import formats as F
    s = scenes.cornell_box(sphere_tris=2000)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 30, pass_paths=1 << 12)
    srv.pause()                                   # the jobs stay pending while the scene calls are tried
    ids = [srv.submit(scenes.cornell_view(96, 96), 96, 96, spp=16, max_depth=6, seed=40 + k) for k in range(4)]
    with pytest.raises(lh.LH2HIPError):
        core.SetMaterials(s.materials)
    with pytest.raises(lh.LH2HIPError):
        core.FinalizeInstances()
    srv.pause(False)
    for j in ids:
        srv.wait(j, timeout=120)
        srv.read(j)
    core.SetMaterials(s.materials)          # no job pending: allowed again
    assert library_status_destroy_fails(core)
    srv.close()
    core.Shutdown()


def library_status_destroy_fails(core):
    from lighthouse2_amd.core import library
    return library().lh2hip_destroy(core._h) == -3   # LH2HIP_ERR_STATE: a server still uses the core


def test_target_size_limit_of_32_bit_path_indices(torch_cuda):
    core = lh.RenderCore()
    with pytest.raises(lh.LH2HIPError):
        core.SetTarget(16384, 16384, 17)       # 2^28 * 17 paths > 2^32 (rejected before any allocation)
    with pytest.raises(lh.LH2HIPError):
        core.SetTarget(4096, 4096, 257)        # 2^24 * 257 > 2^32
    core.SetTarget(64, 64, 1)
    s = scenes.cornell_box(sphere_tris=500)
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 40, pass_paths=1 << 16)
    with pytest.raises(lh.LH2HIPError):
        srv.submit(scenes.cornell_view(4096, 4096), 4096, 4096, spp=257)
    srv.close()
    core.Shutdown()


def _solo_frame_buffers(scene, view, w, h, spp, seed):
    core = lh.RenderCore()
    lh.load_scene(core, scene)
    core.SetTarget(w, h, spp)
    core.Setting("estimator", 2)
    core.Setting("TAA", 0)
    core.Setting("seed", seed)
    core.Render(view, lh.LH2_RESTART)
    img, acc = core.read_image(), core.read_accumulator_f64()
    acc32, feat = core.read_buffer("acc32"), core.read_buffer("features")
    core.Shutdown()
    return img, acc, acc32, feat


def _learned_solo(torch, net_params, acc32, feat, spp, w, h):
    from lighthouse2_amd.denoiser import UNetDenoiser
    Hp, Wp = (h + 15) // 16 * 16, (w + 15) // 16 * 16
    net = UNetDenoiser(Hp, Wp, params=net_params)
    a = torch.from_numpy(np.ascontiguousarray(acc32)).cuda()
    f = torch.from_numpy(np.ascontiguousarray(feat).view(np.int32)).cuda()
    img = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    net.set_input_frame(a, f, spp, w, h)
    net.forward()
    net.output_frame(f, w, h, img)
    torch.cuda.synchronize()
    return img.cpu().numpy(), net


def test_instanced_mix_with_both_denoisers(torch_cuda):
    """A reduced C5 mix on the instanced scene (displaced-sphere meshes, random
    affine instances, textures, area lights): a 1080p job, 512^2 at 16 spp,
    256^2 at 64 spp, SVGF and learned-denoiser jobs and plain jobs of mixed
    depth. Every job equals its solo render (learned: the solo frame through
    the same network from Python), jobs share passes."""
    torch = torch_cuda
    from lighthouse2_amd.denoiser import init_params
    s = scenes.instanced_scene(meshes=3, tris=(4000, 12000), instances=60, tex_size=256, lights=8, extent=20.0)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    params = init_params(seed=3)
    srv = lh.RenderServer(core, budget_bytes=4 << 30, pass_paths=1 << 22)
    from lighthouse2_amd.denoiser import UNetDenoiser
    blob_net = UNetDenoiser(16, 16, params=params)
    srv.set_denoiser(blob_net.weights_blob())
    view = lambda w, h: scenes.instanced_view(w, h, extent=20.0)
    # (w, h, spp, depth, seed, denoise)
    jobs = [(1920, 1080, 1, 4, 51, 1), (256, 256, 64, 5, 52, 0), (512, 512, 16, 8, 53, 2), (256, 256, 4, 6, 54, 1),
            (512, 512, 1, 4, 55, 2), (256, 256, 16, 7, 56, 0)]
    srv.pause()
    ids = [srv.submit(view(w, h), w, h, spp, d, seed, denoise=dn) for (w, h, spp, d, seed, dn) in jobs]
    srv.pause(False)
    for j in ids:
        srv.wait(j, timeout=300)
    st = srv.stats()
    out = [srv.read(j, with_acc=True) for j in ids]
    for (w, h, spp, d, seed, dn), (img, acc) in zip(jobs, out):
        print("job", w, h, spp, d, seed, dn, flush=True)
        if dn == 0:
            ri, ra = solo(s, view(w, h), w, h, spp, d, seed)
            assert np.array_equal(acc, ra) and np.array_equal(img, ri), (w, h, spp, d, seed)
            continue
        ri, ra, acc32, feat = _solo_frame_buffers(s, view(w, h), w, h, spp, seed)
        assert np.array_equal(acc, ra), (w, h, spp, seed, dn)
        if dn == 1:
            assert np.array_equal(img, ri), (w, h, spp, seed)
        else:
            li, _ = _learned_solo(torch, params, acc32, feat, spp, w, h)
            assert np.array_equal(img, li), (w, h, spp, seed, float(np.abs(img - li).max()))
            assert np.isfinite(img).all() and img[..., :3].max() > 0
    assert st["shared_passes"] >= 1, st
    srv.close()
    core.Shutdown()


def test_learned_job_needs_weights(torch_cuda):
    s = scenes.cornell_box(sphere_tris=500)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    srv = lh.RenderServer(core, budget_bytes=1 << 28, pass_paths=1 << 16)
    with pytest.raises(lh.LH2HIPError):
        srv.submit(scenes.cornell_view(32, 32), 32, 32, 1, 4, 1, denoise=2)
    with pytest.raises(lh.LH2HIPError):
        srv.set_denoiser(b"\0" * 16)              # wrong size
    srv.close()
    core.Shutdown()
