#!/usr/bin/env python3
"""Benchmarks of RenderCore_HIP on one MI355X; the default run prints one JSON
line whose headline is the incoherent-extend rate, with the other two metrics
of the workload set and the parity run as sub-objects.

Headline (incoherent traversal microbenchmark): 4,000,000 random triangles
(centroids U[0,1)^3, vertex offsets N(0, 0.005^2), seed 1), binned-SAH BVH
built on the GPU; 1920x1080 primary rays from (0.5,0.5,-1.2) towards
(0.5,0.5,0.5), FOV 40, then 3 cosine-weighted diffuse bounces (misses die).
One step = primary generation + 4 extend waves + 3 bounce kernels.
metric: "Mrays/s incoherent extend" = rays of waves 2..4 / time of their extend
launches (HIP events on the launch stream), the reference's rays/traceTime
definition (lib/RenderCore_PrimeRef/rendercore.cpp:568-576,
apps/benchmarkapp/main.cpp:479-482); q8 nodes, fp32 nodes reported beside.

Sub-objects: "frame" (ms/frame, 1080p 1 spp + SVGF on the instanced scene),
"serve" (p50/p99 job latency of a Poisson job stream through the job
server), "c1_parity" (the Cornell-box parity run against the FP64 oracle)."""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NODE_BYTES = {"q8": 64, "f32": 64}
TRI_BYTES = 48          # leaf triangle: 3 float4
RAY_IO_BYTES = 48       # 32 B ray in + 16 B hit record out
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=4_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--format", choices=["q8", "f32"], default="q8")
    ap.add_argument("--loop", choices=["merged", "ww", "greedy"], default="merged", help="traversal loop kind")
    ap.add_argument("--tri-bias", type=int, default=16, help="greedy loop weight of leaf lanes, in 1/16")
    ap.add_argument("--refill-lanes", type=int, default=16, help="idle lanes that make a wave fetch rays")
    ap.add_argument("--chunk", type=int, default=32, help="rays a wave reserves per atomic")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--shadow-lane", type=int, default=1, choices=[0, 1],
                    help="frame: trace shadow rays on a second stream beside the next extension wave")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="persistent traversal blocks per CU (0: core default)")
    ap.add_argument("--min-rays-per-lane", type=float, default=0.0, help="shrink the grid of small waves (0: off)")
    ap.add_argument("--workload", choices=["extend", "denoise", "svgf", "serve", "frame", "learned"], default="extend",
                    help="extend: the headline traversal metric; denoise: U-Net forward/train at 1088x1920; "
                         "svgf: filter ms/frame at 1920x1080; serve: job-server latency under Poisson arrivals; "
                         "frame: instanced scene, 1080p 1 spp + SVGF; learned: denoiser trained on rendered pairs vs SVGF")
    ap.add_argument("--serve-jobs", type=int, default=0, help="jobs of the serving stream (0: 200 in the default run, 2000 alone)")
    ap.add_argument("--headline-only", action="store_true", help="skip the frame, serve and parity sub-objects")
    ap.add_argument("--serve-weights", default="", help="learned-denoiser weight blob for the serving stream (default: seeded init)")
    ap.add_argument("--poses", type=int, default=2000, help="learned: training poses (held-out poses: 16)")
    ap.add_argument("--train-steps", type=int, default=8000, help="learned: Adam steps (batch 1)")
    ap.add_argument("--oracle-crops", type=int, default=4, help="learned: held-out poses checked against FP64-oracle crops")
    ap.add_argument("--weights-out", default="", help="learned: write the trained weight blob here")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (FETCH_SIZE/WRITE_SIZE calibration)")
    return ap.parse_args()


class Waves:
    """Device buffers and launches of one step."""

    def __init__(self, torch, core, view, W, H):
        self.torch, self.core, self.view, self.W, self.H = torch, core, view, W, H
        n = W * H
        self.n = n
        f = dict(device="cuda")
        self.rays = [torch.zeros((n, 8), dtype=torch.float32, **f) for _ in range(2)]
        self.pid = [torch.zeros(n, dtype=torch.int32, **f) for _ in range(2)]
        self.hits = torch.zeros((n, 4), dtype=torch.float32, **f)
        self.cnt = torch.zeros(1, dtype=torch.int32, **f)
        self.stream = torch.cuda.current_stream()
        self.sh = self.stream.cuda_stream

    def step(self, events=None, keep=None):
        """Runs one step; returns [n1, n2, n3, n4] ray counts per wave."""
        from lighthouse2_amd import core as C
        torch = self.torch
        C.generate_primary(self.view, self.W, self.H, 1234, 1e-5, self.rays[0].data_ptr(), self.pid[0].data_ptr(), self.sh)
        counts, cur, n = [], 0, self.n
        for wave in range(4):
            counts.append(n)
            if events is not None:
                events[wave][0].record(self.stream)
            self.core.trace_closest(self.rays[cur].data_ptr(), self.hits.data_ptr(), n, self.sh)
            if events is not None:
                events[wave][1].record(self.stream)
            if keep is not None and wave >= 1:
                keep.append((self.rays[cur][:n].clone(), self.hits[:n].clone()))
            if wave == 3:
                break
            self.core.bounce_diffuse(self.rays[cur].data_ptr(), self.hits.data_ptr(), self.pid[cur].data_ptr(), n,
                                     (wave + 1) * 91771, 1e-5, self.rays[cur ^ 1].data_ptr(), self.pid[cur ^ 1].data_ptr(),
                                     self.cnt.data_ptr(), self.sh)
            n = int(self.cnt.item())
            cur ^= 1
        return counts


def _events(torch, stream, fn, reps):
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def bench_denoise(a, np, torch, lh, scenes):
    """U-Net denoiser at 1088x1920 (1080p padded to the 16-pixel grid of 4
    pooling steps): forward ms and TFLOP/s (multiply-add FLOPs of the 19
    convolutions), and one training step (forward + backward + Adam)."""
    from lighthouse2_amd import denoiser as D
    H, W = 1088, 1920
    net = D.UNetDenoiser(H, W, seed=0)
    rng = np.random.default_rng(0)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    rad, alb, nrm = (dev(rng.uniform(0, 1, (H, W, 4))) for _ in range(3))
    dep = dev(rng.uniform(0, 10, (H, W)))
    tgt = dev(rng.uniform(0.1, 1, (H, W, 4)))
    net.set_input(rad, alb, nrm, dep)
    st = net.stream
    for _ in range(a.warmup):
        net.forward()
    fwd_ms = _events(torch, st, net.forward, a.steps)

    def train():
        net.forward()
        net.loss_backward(tgt)
        net.step(lr=1e-4)
    for _ in range(2):
        train()
    train_ms = _events(torch, st, train, max(2, a.steps // 2))
    flops = net.flops_forward()
    tf = flops / (fwd_ms * 1e-3) / 1e12
    # a whole denoised 1080p frame from the filter core's buffers (input packing + forward + display image),
    # against SVGF (prepare + 3 a-trous passes + display, no TAA) on the same synthetic frame
    g0, v1 = svgf_gbuffer(np, torch, scenes)
    img = torch.zeros((1080, 1920, 4), dtype=torch.float32, device="cuda")

    def unfused_frame():
        net.set_input_frame(g0[0], g0[1], 1, 1920, 1080)
        net.forward()
        net.output_frame(g0[1], 1920, 1080, img)

    def learned_frame():            # the output convolution writes the display image
        net.denoise_frame(g0[0], g0[1], 1, 1920, 1080, img)
    for _ in range(2):
        unfused_frame()
        learned_frame()
    unfused_ms = _events(torch, st, unfused_frame, a.steps)
    frame_ms = _events(torch, st, learned_frame, a.steps)
    sv_ms = svgf_ms(a, np, torch, lh, scenes, g0, v1)
    print(json.dumps({"metric": "ms/frame learned denoiser forward 1080p", "value": round(fwd_ms, 3), "unit": "ms",
                      "higher_is_better": False, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "dtype": "bf16 (fp32 accumulate)",
                      "data": "synthetic", "vs_baseline": None,
                      "config": {"workload": "5-level 3x3 conv U-Net, 32 channels, 1088x1920", "layers": len(D.layer_specs())},
                      "learned_frame_ms": round(frame_ms, 3), "learned_frame_unfused_ms": round(unfused_ms, 3), "svgf_frame_ms": round(sv_ms, 3),
                      "train_step_ms": round(train_ms, 3), "forward_gflop": round(flops / 1e9, 1),
                      "roofline": {"bound": "mfma", "achieved": round(tf, 1), "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tf / 2500.0, 4),
                                   "traffic": None}}))


def svgf_gbuffer(np, torch, scenes, W=1920, H=1080):
    """A synthetic 1080p filter-core frame on the device: accumulators
    (direct, indirect), packed features, world position, depth derivatives,
    and the view of the next (moved) camera."""
    v0 = scenes.camera_view((0.5, 0.6, -1.0), (0.5, 0.5, 1.0), 50.0, W, H)
    v1 = scenes.camera_view((0.505, 0.6, -0.999), (0.5, 0.5, 1.0), 50.0, W, H)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    gb = scenes.gbuffer(v0, W, H, seed=1)
    g0 = [dev(x) for x in gb]
    g0[1] = dev(gb[1].view(np.int32))
    return g0, v1


def svgf_ms(a, np, torch, lh, scenes, g0, v1, W=1920, H=1080):
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    f = lh.Svgf(W, H)
    st = torch.cuda.current_stream()

    def frame():
        f.frame(g0[0].data_ptr(), g0[1].data_ptr(), g0[2].data_ptr(), g0[3].data_ptr(), v1, out.data_ptr(), j=(0.1, -0.1),
                samples=1, stationary=0, stream=st.cuda_stream)
    for _ in range(a.warmup):
        frame()
    return _events(torch, st, frame, a.steps)


def bench_svgf(a, np, torch, lh, scenes):
    """SVGF filter (prepare + 3 a-trous passes + display) on a 1920x1080
    synthetic G-buffer, moving camera: ms per frame."""
    W, H = 1920, 1080
    g0, v1 = svgf_gbuffer(np, torch, scenes, W, H)
    ms = svgf_ms(a, np, torch, lh, scenes, g0, v1, W, H)
    # compulsory bytes per pixel over all passes: prepare reads 2x16 acc + 16 feat + 16 pos + 16 prevpos + 16 moments (history),
    # writes 16 shading + 8 motion + 16 moments + 16 feat; each a-trous pass reads ~16 A + 16 feat + 16 moments + 16 depth;
    # phase 1 adds 16 motion-history; plus 16 out: about 240 B/px
    bpp = 240
    gbs = W * H * bpp / (ms * 1e-3) / 1e9
    print(json.dumps({"metric": "ms/frame SVGF 1080p", "value": round(ms, 3), "unit": "ms", "higher_is_better": False, "n_gpus": 1,
                      "steps": a.steps, "warmup": a.warmup, "dtype": "f32", "data": "synthetic G-buffer", "vs_baseline": None,
                      "config": {"workload": "SVGF prepare + 3 a-trous passes + display, 1920x1080, moving camera"},
                      "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s", "frac": round(gbs / 8000.0, 4),
                                   "traffic": None, "bytes_per_pixel": bpp}}))


# algorithmic bytes of one filter-estimator shade launch per path: in 76 (ray 32, state 16, aux 8,
# path id 4, hit 16), triangle 128 (8 of 13 float4), material 128, texels 32 (a trilinear fetch),
# out 140 (extension ray + state + aux + id 60, shadow ray + potential 48, path-slot read-modify-write
# 32); the first wave also writes the 48-byte G-buffer (features, world position, depth derivatives)
SHADE_BYTES_WAVE1 = 552
SHADE_BYTES = 504


def frame_section(a, np, torch, lh, scenes):
    """One real-time frame of the instanced workload at 1920x1080, 1 spp:
    10 noise-displaced spheres (50k-200k triangles), 1000 instances with
    random affine transforms (two-level BVH), ten 2048^2 textures, Disney
    materials (Lambert diffuse + GGX, roughness U(0.05, 1)), 64 area lights;
    the filter estimator (eye rays, shade with G-buffer, NEE, its path cap)
    followed by SVGF and TAA, the reference's defaults. Camera fixed; each
    frame restarts accumulation, as a moving camera would. ms/frame is wall
    time around Render; the roofline is the shade kernel's (HIP events)."""
    t0 = time.time()
    s = scenes.instanced_scene()
    t_scene = time.time() - t0
    core = lh.RenderCore()
    t0 = time.time()
    lh.load_scene(core, s)
    W, H = 1920, 1080
    core.SetTarget(W, H, 1)
    core.Setting("estimator", 2)
    core.Setting("shadowLane", a.shadow_lane)
    view = scenes.instanced_view(W, H)
    core.Render(view, lh.LH2_RESTART)
    core.WaitForRender()
    t_load = time.time() - t0
    for _ in range(a.warmup):
        core.Render(view, lh.LH2_RESTART)
    keys = ("traceTime0", "traceTime1", "traceTimeX", "shadeTime", "shadowTraceTime", "filterTime")
    parts = {k: 0.0 for k in keys}
    shaded = [0, 0]          # first-wave paths, later-wave paths
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(a.steps):
        core.Render(view, lh.LH2_RESTART)
        st = core.GetCoreStats()
        for k in keys:
            parts[k] += float(st[k])
        shaded[0] += int(st["primaryRayCount"])
        shaded[1] += int(st["bounce1RayCount"]) + int(st["deepRayCount"])
    core.WaitForRender()
    ms = (time.time() - t0) * 1e3 / a.steps
    shade_s = parts["shadeTime"]
    parts = {k: round(v * 1e3 / a.steps, 3) for k, v in parts.items()}
    img = core.read_image()
    tris = sum(len(v) // 3 for v, _ in s.meshes)
    gbs = (shaded[0] * SHADE_BYTES_WAVE1 + shaded[1] * SHADE_BYTES) / max(shade_s, 1e-9) / 1e9
    line = {"metric": "ms/frame 1080p 1spp+denoise", "value": round(ms, 3), "unit": "ms", "higher_is_better": False, "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "dtype": "f32", "data": "synthetic scene", "vs_baseline": None,
            "config": {"workload": "instanced scene: 10 displaced spheres 50k-200k tris, 1000 instances, 2048^2 textures, "
                                   "Disney (diffuse + GGX), 64 area lights; 1920x1080, 1 spp, filter estimator + SVGF + TAA",
                       "mesh_triangles": int(tris), "instances": len(s.instances), "path_cap": "the filter core's MAXPATHLENGTH (7)"},
            "breakdown_ms": parts, "paths_shaded_per_frame": int(sum(shaded) / a.steps),
            "roofline": {"bound": "hbm", "kernel": "k_shadeF", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "bytes": "algorithmic",
                         "bytes_per_path": {"first_wave": SHADE_BYTES_WAVE1, "later_waves": SHADE_BYTES}},
            "image_mean": float(img[1:-1, 1:-1, :3].mean()),
            "scene_build_s": round(t_scene, 1), "load_and_first_frame_s": round(t_load, 1)}
    core.Shutdown()
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        threads = min(16, os.cpu_count() or 1)
        w, h = W // 8, H // 8                        # 1/64 of the pixels, same view frustum
        ref = orc.RenderF(s, w, h, spp=1, threads=threads)
        t0 = time.time()
        ref.frame(scenes.instanced_view(w, h), orc.RESTART)
        cpu_s = time.time() - t0
        line["cpu_baseline"] = {"value": round(cpu_s * 64 * 1e3, 1), "unit": "ms", "cores": threads, "cpu": cpu_model(), "kind": "port",
                                "sample": f"oracle filter estimator + SVGF + TAA at {w}x{h} (1/64 of the pixels), scaled by 64"}
    return line


def bench_frame(a, np, torch, lh, scenes):
    print(json.dumps(frame_section(a, np, torch, lh, scenes)))


def crop_view(np, view, W, H, x0, y0, cw, ch):
    """ViewPyramid of the pixels [x0, x0 + cw) x [y0, y0 + ch) of a W x H view
    (same pixel size and spread angle, screen corners moved)."""
    v = view.copy()
    p1, p2, p3 = (np.asarray(view[k][0], np.float64) for k in ("p1", "p2", "p3"))
    at = lambda x, y: p1 + (x / W) * (p2 - p1) + (y / H) * (p3 - p1)
    v["p1"], v["p2"], v["p3"] = at(x0, y0), at(x0 + cw, y0), at(x0, y0 + ch)
    return v


def bench_learned(a, np, torch, lh, scenes):
    """The learned denoiser against SVGF on the instanced workload: training
    pairs rendered by the core (1 spp filter-estimator frames with their
    G-buffer vs 1024-spp renders of the same poses, 256x256; 2,000 training
    poses by default), trained with the library's own kernels; relMSE on
    held-out poses for the 1-spp frame, SVGF (one frame, and after 4 still
    frames of history) and the network against the 1024-spp renders, and, on
    the CPU leg, against FP64-oracle renders of 64x64 crops of held-out poses;
    ms/frame of the network's forward pass at 1080p."""
    from lighthouse2_amd import denoiser as D
    from lighthouse2_amd import training as T
    t0 = time.time()
    s = scenes.instanced_scene()
    w = h = 256
    n_train, n_eval = a.poses, 16
    pr = T.PairRenderer(s, w, h, ref_spp=1024, spp_per_frame=128)
    views = T.views_near_instances(s, n_train + n_eval, w, h, seed=6)
    samples = []
    for k, v in enumerate(views[:n_train]):
        d = pr.render(v)
        samples.append((T.network_inputs(torch, d), T.network_target(torch, d)))
        if (k + 1) % 100 == 0:
            print(f"rendered {k + 1} / {n_train} training pairs ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    data = [pr.render(v) for v in views[n_train:]]
    hist = [pr.render(v, history=4)["svgf"] for v in views[n_train:]]
    pr.close()
    t_data = time.time() - t0
    net = D.UNetDenoiser(h, w, seed=0)
    t0 = time.time()
    steps = a.train_steps
    T.train(net, torch, samples, steps, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    print(f"trained {steps} steps in {t_train:.0f} s", file=sys.stderr, flush=True)
    err = {"1spp": [], "svgf": [], "svgf_4_frames": [], "learned": []}
    outs = []
    for d, sv4 in zip(data, hist):
        net.set_input(*T.network_inputs(torch, d))
        out = T.network_radiance(net.forward().cpu().numpy(), d)
        outs.append(out)
        err["1spp"].append(T.relmse(d["radiance"], d["ref"]))
        err["svgf"].append(T.relmse(d["svgf"][1:-1, 1:-1], d["ref"][1:-1, 1:-1]))
        err["svgf_4_frames"].append(T.relmse(sv4[1:-1, 1:-1], d["ref"][1:-1, 1:-1]))
        err["learned"].append(T.relmse(out, d["ref"]))
    err = {k: round(float(np.mean(v)), 5) for k, v in err.items()}
    if a.weights_out:
        open(a.weights_out, "wb").write(net.weights_blob())
    # ms/frame at 1080p: network forward (1088x1920 padded)
    net2 = D.UNetDenoiser(1088, 1920, seed=0, params=net.params_numpy())
    z = torch.zeros((1088, 1920, 4), dtype=torch.float32, device="cuda")
    net2.set_input(z, z, z, z[..., 0].contiguous())
    for _ in range(3):
        net2.forward()
    fwd_ms = _events(torch, net2.stream, net2.forward, 10)
    line = {"metric": "relMSE vs 1024 spp, held-out poses", "value": err["learned"], "unit": "relMSE", "higher_is_better": False,
            "n_gpus": 1, "steps": steps, "warmup": 0, "dtype": "bf16 (fp32 accumulate)", "data": "rendered by the core",
            "vs_baseline": None, "relmse": err, "learned_forward_ms_1080p": round(fwd_ms, 3),
            "config": {"workload": f"instanced scene, {w}x{h}, {n_train} training / {n_eval} held-out poses, "
                                   f"1 spp inputs vs 1024 spp targets, {steps} Adam steps (batch 1)"},
            "data_s": round(t_data, 1), "train_s": round(t_train, 1)}
    if not a.no_cpu:
        # FP64 oracle: 64x64 centre crops of held-out poses at 1024 spp (filter estimator, no SVGF)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        threads = min(16, os.cpu_count() or 1)
        cw = ch = 64
        x0 = y0 = (w - cw) // 2
        o = orc.OracleScene(s)
        crop = {"1spp": [], "svgf": [], "learned": [], "gpu_1024spp_target": []}
        t0 = time.time()
        for d, out, v in list(zip(data, outs, views[n_train:]))[:a.oracle_crops]:
            ref = orc.RenderF(o, cw, ch, spp=1024, filter_on=False, taa=False, threads=threads)
            ref.frame(crop_view(np, v, w, h, x0, y0, cw, ch), orc.RESTART)
            r64 = ((ref.acc64[0] + ref.acc64[1]) / 1024.0).reshape(ch, cw, 4)[..., :3]
            sl = (slice(y0, y0 + ch), slice(x0, x0 + cw))
            crop["1spp"].append(T.relmse(d["radiance"][sl], r64))
            crop["svgf"].append(T.relmse(d["svgf"][sl], r64))
            crop["learned"].append(T.relmse(out[sl], r64))
            crop["gpu_1024spp_target"].append(T.relmse(d["ref"][sl], r64))
            print(f"oracle crop {len(crop['learned'])} done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
        line["relmse_vs_fp64_oracle_crops"] = {k: round(float(np.mean(v)), 5) for k, v in crop.items()}
        line["oracle_crops"] = f"{len(crop['learned'])} held-out poses, {cw}x{ch} centre crop, 1024 spp, FP64 accumulation, {threads} threads"
        line["cpu_baseline"] = {"value": round((time.time() - t0) / max(1, len(crop["learned"])), 1), "unit": "s per 64x64 crop at 1024 spp",
                                "cores": threads, "cpu": cpu_model(), "kind": "port", "sample": line["oracle_crops"]}
    print(json.dumps(line))


def serve_section(a, np, torch, lh, scenes, n_jobs):
    """Job server under Poisson arrivals (lambda 20 jobs/s, seed 7), the
    first n_jobs of the stream: resolutions {256^2, 512^2, 1920x1080,
    3840x2160}, spp {1, 4, 16, 64}, depth U{4..8}, half the jobs denoised by
    SVGF and half by the learned U-Net, on the instanced scene (10 displaced
    spheres, 1,000 instances, 2048^2 textures, 64 area lights); admission by a
    64 GB budget, passes of up to 2^22 paths shared by up to 16 jobs.
    latency = completion - arrival. Every job is rendered by the filter
    estimator (its path cap replaces the drawn depth)."""
    import threading
    from lighthouse2_amd.denoiser import UNetDenoiser
    s = scenes.instanced_scene()
    core = lh.RenderCore()
    lh.load_scene(core, s)
    pass_paths = 1 << 22
    srv = lh.RenderServer(core, budget_bytes=64 << 30, pass_paths=pass_paths)
    if a.serve_weights:
        blob, wsrc = open(a.serve_weights, "rb").read(), os.path.basename(a.serve_weights)
    else:
        blob, wsrc = UNetDenoiser(16, 16, seed=0).weights_blob(), "seeded initial weights (forward cost does not depend on them)"
    srv.set_denoiser(blob)
    rng = np.random.default_rng(7)
    res = [(256, 256), (512, 512), (1920, 1080), (3840, 2160)]
    jobs = [(res[rng.integers(4)], int(rng.choice([1, 4, 16, 64])), int(rng.integers(4, 9)), int(1 + rng.integers(2)))
            for _ in range(n_jobs)]
    gaps = rng.exponential(1 / 20.0, n_jobs)
    ids = []
    t_start = time.time()

    def submitter():
        t = time.time()
        for k, (((w, h), spp, d, dn), gap) in enumerate(zip(jobs, gaps)):
            t += gap
            time.sleep(max(0.0, t - time.time()))
            ids.append(srv.submit(scenes.instanced_view(w, h), w, h, spp, d, 1000 + k, denoise=dn))
            if (k + 1) % 200 == 0:
                print(f"serve: {k + 1} / {n_jobs} jobs submitted ({time.time() - t_start:.0f} s)", file=sys.stderr, flush=True)
    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    lat, paths = [], 0
    for k, (j, ((w, h), spp, d, dn)) in enumerate(zip(ids, jobs)):
        srv.wait(j)
        if (k + 1) % 100 == 0:
            print(f"serve: {k + 1} / {n_jobs} jobs done ({time.time() - t_start:.0f} s)", file=sys.stderr, flush=True)
        t = srv.times(j)
        lat.append(t[2] - t[0])
        paths += w * h * spp
        srv.read(j)
    makespan = time.time() - t_start
    st = srv.stats()
    srv.close()
    core.Shutdown()
    lat = np.array(lat)
    later = st["rays"] - st["paths"]
    gbs = (st["paths"] * SHADE_BYTES_WAVE1 + later * SHADE_BYTES) / max(st["shade_ms"] * 1e-3, 1e-9) / 1e9
    line = {"metric": "p50/p99 job latency", "value": round(float(np.percentile(lat, 50)), 3), "unit": "s",
            "p99_s": round(float(np.percentile(lat, 99)), 3), "higher_is_better": False, "n_gpus": 1,
            "jobs": len(lat), "lambda_jobs_per_s": 20, "makespan_s": round(makespan, 2),
            "mrays_s": round(st["rays"] / makespan / 1e6, 1), "mpaths_s": round(paths / makespan / 1e6, 1),
            "dtype": "f32", "data": "synthetic", "vs_baseline": None,
            "batching": {k: st[k] for k in ("passes", "parts", "shared_passes", "max_parts")},
            "device_ms": {"extend": round(st["trace_ms"], 1), "shade": round(st["shade_ms"], 1)},
            "roofline": {"bound": "hbm", "kernel": "k_shadeF (passes shared by several jobs)", "achieved": round(gbs, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "bytes": "algorithmic"},
            "config": {"workload": f"Poisson job stream (first {n_jobs} jobs of seed 7) on the instanced scene (10 displaced spheres "
                                   "50k-200k tris, 1000 instances, 2048^2 textures, 64 area lights), mixed resolution/spp, "
                                   "filter estimator per job, SVGF or learned denoiser 50/50",
                       "budget_gb": 64, "pass_paths": pass_paths, "denoiser_weights": wsrc}}
    if not a.no_cpu:
        # CPU: 1% of the jobs (at least 2), each rendered by the oracle filter estimator + SVGF at 1/64 of its pixels, scaled by 64
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        threads = min(16, os.cpu_count() or 1)
        pick = np.random.default_rng(8).choice(len(jobs), size=max(2, len(jobs) // 100), replace=False)
        o = orc.OracleScene(s)
        times = []
        for k in pick:
            (w, h), spp, d, dn = jobs[k]
            ref = orc.RenderF(o, w // 8, h // 8, spp=spp, threads=threads, taa=False, seed=1000 + int(k))
            t0 = time.time()
            ref.frame(scenes.instanced_view(w // 8, h // 8), orc.RESTART)
            times.append((time.time() - t0) * 64)
        line["cpu_baseline"] = {"value": round(float(np.median(times)), 2), "unit": "s per job (median)", "cores": threads, "cpu": cpu_model(),
                                "kind": "port", "sample": f"{len(pick)} jobs of the stream ({', '.join(f'{jobs[k][0][0]}x{jobs[k][0][1]} {jobs[k][1]} spp' for k in pick)}): "
                                                         "oracle filter estimator + SVGF at 1/64 of the pixels, scaled by 64"}
    return line


def bench_serve(a, np, torch, lh, scenes):
    print(json.dumps(serve_section(a, np, torch, lh, scenes, a.serve_jobs or 2000)))


def c1_section(a, np, torch, lh, scenes):
    """The parity run: Cornell box (5 wall quads + a 20,000-triangle sphere,
    0.2 x 0.2 light of radiance 15), 256 x 256, 1 spp x 16 frames, max depth 3,
    seed 1234, the reference estimator. Every image (the accumulated image
    after each frame) against the FP64 oracle: mean and max relative error;
    the fp64 accumulators compared bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    s = scenes.cornell_box(seed=1234, sphere_tris=20000)
    W = H = 256
    view = scenes.cornell_view(W, H)
    core = lh.RenderCore()
    lh.load_scene(core, s)
    core.SetTarget(W, H, 1)
    core.Setting("maxDepth", 3)
    core.Setting("epsilon", 1e-3)
    core.Setting("seed", 1234)
    threads = min(16, os.cpu_count() or 1)
    o = orc.OracleScene(s)
    modes = [orc.RESTART] + [orc.CONVERGE] * 16
    frames, gpu_ms, cpu_s, identical = [], [], 0.0, True
    gen = orc.render_frames(o, view, W, H, 1, 16, 3, 1e-3, seed=1234, threads=threads, modes=modes)
    for k, m in enumerate(modes):
        torch.cuda.synchronize()
        t0 = time.time()
        core.Render(view, lh.LH2_RESTART if m == orc.RESTART else lh.LH2_CONVERGE)
        core.WaitForRender()
        gpu_ms.append((time.time() - t0) * 1e3)
        img, acc = core.read_image(), core.read_accumulator_f64()
        t0 = time.time()
        _, r64, n = next(gen)
        cpu_s += time.time() - t0
        if k == 0:
            continue                      # the preview frame before the converging sequence
        ref = r64[..., :3] / n
        rel = np.abs(img[..., :3].astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-6)
        frames.append({"frame": k, "samples": int(n), "mean_rel": float(f"{rel.mean():.3e}"), "max_rel": float(f"{rel.max():.3e}")})
        identical = identical and bool(np.array_equal(acc, r64))
    core.Shutdown()
    return {"metric": "per-image relative error vs FP64 oracle", "value": max(f["max_rel"] for f in frames), "unit": "max rel",
            "tolerance": 1e-4, "fp64_accumulators_identical": identical, "images": frames,
            "config": {"workload": "Cornell box 256x256, 1 spp x 16 frames, max depth 3, seed 1234, 20,000-triangle sphere, reference estimator"},
            "gpu_ms_per_frame": round(float(np.median(gpu_ms[1:])), 3),
            "cpu_baseline": {"value": round(cpu_s / len(modes) * 1e3, 1), "unit": "ms/frame", "cores": threads, "cpu": cpu_model(), "kind": "port",
                             "sample": "the same 17 frames (preview + 16) on the oracle: fp32 estimator, fp64 accumulation"}}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measure_extend(a, np, torch, core, wv, steps, warmup):
    """Times `steps` steps after `warmup`; returns per-step ray counts and the
    per-wave extend times (ms) from HIP events on the launch stream."""
    for _ in range(warmup):
        wv.step()
    torch.cuda.synchronize()
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(4)] for _ in range(steps)]
    per_step = []
    torch.cuda.synchronize()
    tw0 = time.time()
    for s in range(steps):
        per_step.append(wv.step(ev[s]))
    torch.cuda.synchronize()
    wall = time.time() - tw0
    ms = [[ev[s][w][0].elapsed_time(ev[s][w][1]) for w in range(4)] for s in range(steps)]
    return per_step, ms, wall


def bench_extend(a, np, torch, lh, scenes):
    t0 = time.time()
    verts = scenes.random_triangles(a.tris, seed=1)
    t_gen = time.time() - t0
    core = lh.RenderCore()
    core.Setting("nodeFormat", 1 if a.format == "q8" else 0)
    core.Setting("traceLoop", {"merged": 0, "ww": 1, "greedy": 2}[a.loop])
    core.Setting("traceTriBias", a.tri_bias)
    core.Setting("traceRefillLanes", a.refill_lanes)
    core.Setting("traceChunk", a.chunk)
    if a.blocks_per_cu:
        core.Setting("traceBlocksPerCU", a.blocks_per_cu)
    if a.min_rays_per_lane:
        core.Setting("traceMinRaysPerLane", a.min_rays_per_lane)
    t0 = time.time()
    core.SetGeometry(0, verts, None)
    core.SetInstance(0, 0)
    core.FinalizeInstances()
    t_build = time.time() - t0
    info = core.bvh_info(0)
    view = scenes.microbench_view(a.width, a.height)
    wv = Waves(torch, core, view, a.width, a.height)

    per_step, ms, wall = measure_extend(a, np, torch, core, wv, a.steps, a.warmup)
    inco_rays = sum(sum(c[1:]) for c in per_step)
    inco_ms = sum(sum(m[1:]) for m in ms)
    prim_rays = sum(c[0] for c in per_step)
    prim_ms = sum(m[0] for m in ms)
    value = inco_rays / (inco_ms * 1e-3) / 1e6

    # algorithmic work per incoherent ray, counted by the kernel on the same rays (untimed)
    keep = []
    wv.step(keep=keep)
    torch.cuda.synchronize()
    nodes = tris = nrays = 0
    it = np.zeros(10)
    scratch = torch.zeros_like(wv.hits)
    for r, _h in keep:
        c = core.trace_closest_counted(r.data_ptr(), scratch.data_ptr(), len(r), wv.sh)
        nodes, tris, nrays = nodes + c[0], tris + c[1], nrays + len(r)
        it += np.array(c[2:], np.float64)
    nodes_per_ray, tris_per_ray = nodes / nrays, tris / nrays
    bytes_per_ray = RAY_IO_BYTES + nodes_per_ray * NODE_BYTES[a.format] + tris_per_ray * TRI_BYTES
    launch_ms = [m for st in ms for m in st[1:]]
    launch_rays = [c for st in per_step for c in st[1:]]
    achieved = float(np.mean([r * bytes_per_ray / (m * 1e-3) / 1e9 for r, m in zip(launch_rays, launch_ms)]))

    # the other node format on the same triangles and rays: fp32 BVH2 nodes (q8 BVH4 is the default)
    other = "f32" if a.format == "q8" else "q8"
    core.Setting("nodeFormat", 0 if other == "f32" else 1)
    core.FinalizeInstances()
    ps2, ms2, _ = measure_extend(a, np, torch, core, wv, max(3, a.steps // 2), 2)
    other_value = sum(sum(c[1:]) for c in ps2) / (sum(sum(m[1:]) for m in ms2) * 1e-3) / 1e6
    # the same incoherent rays (the compaction order of a wave varies from run to run) traced with
    # the other node format: hit records must be identical bit for bit
    other_equal, other_rays = True, 0
    for r, h in keep:
        h2 = torch.empty_like(h)
        core.trace_closest(r.data_ptr(), h2.data_ptr(), len(r), wv.sh)
        torch.cuda.synchronize()
        other_equal = other_equal and torch.equal(h.view(torch.int32), h2.view(torch.int32))
        other_rays += len(r)

    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", "k_trace_merged_incoherent_4m_traffic.json")
    if os.path.exists(tf) and a.format == "q8":
        tj = json.load(open(tf))
        traffic, traffic_src = tj.get("bytes_per_launch"), tj.get("source")

    cpu = None
    parity = None
    if not a.no_cpu:
        cpu, parity = cpu_baseline(a, verts, keep, np)
    out = {
        "metric": "Mrays/s incoherent extend",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random triangles, generated on the host)",
        "config": {"workload": "incoherent traversal: 4M random triangles, 1920x1080 primary + 3 diffuse bounces",
                   "triangles": a.tris, "width": a.width, "height": a.height, "node_format": a.format,
                   "bvh": "binned SAH 16 bins, leaf <= 4, GPU build", "seed": 1,
                   "blocks_per_cu": f"{a.blocks_per_cu or 8} requested, capped at the kernel occupancy", "min_rays_per_lane": a.min_rays_per_lane,
                   "loop": a.loop, "ray_chunk": a.chunk, "refill_lanes": a.refill_lanes},
        # achieved = algorithmic bytes per ray (ray + hit, nodes and triangles the kernel visits on these
        # rays) x rays / launch time; the kernel is co-limited by VALU issue and the vector-memory address
        # and data units (SQ and TA/TD counters in profiles/), and
        # the HBM traffic that reaches memory is far below the algorithmic bytes (MALL and L2 hits)
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "bytes": "algorithmic", "limiter": "VALU issue + vector-memory units",
                     "bytes_per_ray": round(bytes_per_ray, 1), "nodes_per_ray": round(nodes_per_ray, 2),
                     "tris_per_ray": round(tris_per_ray, 2)},
        "cpu_baseline": cpu,
        "parity": parity,
        "node_formats": {a.format: round(value, 2), other: round(other_value, 2), "hits_equal": bool(other_equal), "rays_compared": int(other_rays)},
        "primary_mrays_s": round(prim_rays / (prim_ms * 1e-3) / 1e6, 2),
        "incoherent_rays_per_step": int(inco_rays / a.steps),
        "extend_ms_per_launch": round(float(np.mean(launch_ms)), 4),
        "per_wave": [{"wave": w + 1, "rays": int(np.mean([c[w] for c in per_step])),
                      "ms": round(float(np.mean([m[w] for m in ms])), 4),
                      "mrays_s": round(float(np.mean([c[w] for c in per_step]) / np.mean([m[w] for m in ms]) / 1e3), 1)}
                     for w in range(4)],
        "simd_utilisation": {"node_loop": round(it[1] / max(64 * it[0], 1), 3), "leaf_loop": round(it[3] / max(64 * it[2], 1), 3),
                             "node_iters_per_ray": round(it[0] * 64 / nrays, 2), "leaf_iters_per_ray": round(it[2] * 64 / nrays, 2),
                             "lanes_holding_ray": round(it[5] / max(64 * it[4], 1), 3), "loop_iters_per_ray": round(it[4] * 64 / nrays, 2),
                             "tail_lanes_holding_ray": round(it[7] / max(64 * it[6], 1), 3), "tail_iters_per_ray": round(it[6] * 64 / nrays, 2),
                             "tail_time_share": round(1.0 - it[8] / max(it[9], 1), 3)},
        "wall_mrays_s_all_waves": round((inco_rays + prim_rays) / wall / 1e6, 2),
        "bvh": {"build_s": round(t_build, 3), "scene_gen_s": round(t_gen, 3), **info},
    }
    core.Shutdown()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (requested --gpus {a.gpus}); "
              "the traversal would shard by image tile with the scene replicated", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import lighthouse2_amd as lh
    from lighthouse2_amd import scenes

    torch.cuda.set_device(0)
    if a.workload != "extend":
        {"denoise": bench_denoise, "svgf": bench_svgf, "serve": bench_serve, "frame": bench_frame,
         "learned": bench_learned}[a.workload](a, np, torch, lh, scenes)
        return
    if a.copy_bytes:
        src = torch.ones(a.copy_bytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.copy_bytes, "dispatches": 3}))
        return
    out = bench_extend(a, np, torch, lh, scenes)
    print("headline done: %.1f Mrays/s" % out["value"], file=sys.stderr, flush=True)
    if not a.headline_only:
        out["frame"] = frame_section(a, np, torch, lh, scenes)
        print("frame done: %.3f ms" % out["frame"]["value"], file=sys.stderr, flush=True)
        out["serve"] = serve_section(a, np, torch, lh, scenes, a.serve_jobs or 200)
        print("serve done: p50 %.3f s" % out["serve"]["value"], file=sys.stderr, flush=True)
        out["c1_parity"] = c1_section(a, np, torch, lh, scenes)
    print(json.dumps(out), flush=True)


def cpu_baseline(a, verts, keep, np):
    """The oracle's BVH traversal (oracle/, a C++ restatement of the same
    closest-hit contract and BVH rules) on a bounded sample of the same
    incoherent rays, on this host's cores; also checks the GPU hits of that
    sample bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    s = __import__("lighthouse2_amd").scenes.Scene()
    s.add_instance(s.add_mesh(verts, None))
    t0 = time.time()
    o = orc.OracleScene(s)
    t_build = time.time() - t0
    rays = np.concatenate([r.cpu().numpy() for r, _ in keep])
    hits = np.concatenate([h.cpu().numpy() for _, h in keep])
    rng = np.random.default_rng(0)
    idx = rng.choice(len(rays), size=min(len(rays), 200_000), replace=False)
    t0 = time.time()
    ref = o.trace(rays[idx], threads=threads)
    dt = time.time() - t0
    mism = int((ref.view(np.uint32) != hits[idx].view(np.uint32)).any(1).sum())
    # scale the sample so the timed CPU work is about cpu_seconds
    reps = max(1, int(a.cpu_seconds / max(dt, 1e-3)))
    n_total, t_total = len(idx), dt
    sample = rays[idx]
    for _ in range(reps - 1):
        t0 = time.time()
        o.trace(sample, threads=threads)
        t_total += time.time() - t0
        n_total += len(idx)
        if t_total > a.cpu_seconds:
            break
    cpu = {"value": round(n_total / t_total / 1e6, 3), "unit": "Mrays/s", "cores": threads, "cpu": cpu_model(), "kind": "port",
           "sample": f"{n_total} incoherent rays (random 200k of waves 2-4, repeated) traced by the oracle BVH on "
                     f"{threads} threads in {t_total:.1f} s; oracle BVH build {t_build:.1f} s excluded"}
    return cpu, {"checked": int(len(idx)), "mismatches": mism}


if __name__ == "__main__":
    main()
