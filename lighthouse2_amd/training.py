"""Training data and evaluation for the learned denoiser, rendered by the core
itself: for each camera pose, a 1-spp frame of the filter estimator gives the
network's inputs (radiance = direct + indirect, and the G-buffer of the
first diffuse vertex: albedo, normal, depth) and SVGF's output on the same
frame; a long converged render of the same pose gives the target. Error is
relMSE against the converged render, the metric the network trains on.
"""
import numpy as np

from . import core as C
from . import scenes

f32 = np.float32


def decode_features(feat):
    """SVGF feature words (h, w, 4) uint32 -> albedo (h, w, 3), normal (h, w, 3), depth (h, w)
    (HDRtoRGB32, PackNormal2 and the hit distance of tools_shared.h:122-183)."""
    f = np.asarray(feat, np.uint32)
    a = f[..., 0]
    albedo = np.stack([(a >> 22) / 1023.0, ((a >> 11) & 2047) / 2047.0, (a & 2047) / 2047.0], -1).astype(f32)
    n = f[..., 1]
    normal = np.stack([((n >> 2) & 1023) / 511.0 - 1, ((n >> 12) & 1023) / 511.0 - 1, (n >> 22) / 511.0 - 1], -1).astype(f32)
    depth = f[..., 2].copy().view(f32)
    return albedo, normal, depth


def random_views(n, w, h, seed=6, extent=50.0, fov=50.0):
    """Camera poses inside the instanced scene's cloud looking at seeded targets."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        pos = rng.uniform(-0.9 * extent, 0.9 * extent, 3)
        tgt = pos + rng.normal(size=3) * extent * 0.5
        out.append(scenes.camera_view(tuple(pos), tuple(tgt), fov, w, h))
    return out


def views_near_instances(scene, n, w, h, seed=6, dist=(4.0, 12.0), fov=50.0):
    """Camera poses looking at seeded instances of the scene from a few units
    away (frames filled with geometry rather than the empty background)."""
    rng = np.random.default_rng(seed)
    centres = np.array([T[[3, 7, 11]] for m, T in scene.instances if len(scene.meshes[m][0]) > 30], np.float64)
    out = []
    for _ in range(n):
        c = centres[int(rng.integers(len(centres)))]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pos = c + d * rng.uniform(*dist)
        out.append(scenes.camera_view(tuple(pos), tuple(c + rng.normal(size=3) * 0.5), fov, w, h))
    return out


def relmse(x, ref):
    x, ref = np.asarray(x, np.float64)[..., :3], np.asarray(ref, np.float64)[..., :3]
    return float(((x - ref) ** 2 / (ref * ref + 0.01)).mean())


class PairRenderer:
    """Two cores on one scene: a 1-spp filter-estimator core (its Render runs
    SVGF, without TAA) and a converging core for the targets."""

    def __init__(self, scene, w, h, ref_spp=1024, spp_per_frame=64, bsdf=1):
        self.w, self.h, self.ref_spp, self.per = w, h, ref_spp, spp_per_frame
        self.fast = C.RenderCore()
        C.load_scene(self.fast, scene)
        self.fast.SetTarget(w, h, 1)
        self.fast.Setting("estimator", 2)
        self.fast.Setting("bsdf", bsdf)
        self.fast.Setting("TAA", 0)                  # SVGF alone; no view jitter
        self.slow = C.RenderCore()
        C.load_scene(self.slow, scene)
        self.slow.SetTarget(w, h, spp_per_frame)
        self.slow.Setting("estimator", 2)
        self.slow.Setting("bsdf", bsdf)                # same eye rays as the fast core (pixel corners)
        self.slow.Setting("TAA", 0)                    # no jitter between the converging frames

    def render(self, view, history=1):
        """Returns dict: radiance (h,w,3) 1 spp, albedo, normal, depth, svgf (h,w,3)
        linear radiance after SVGF (`history` 1-spp frames of a still camera),
        ref (h,w,3) converged radiance."""
        for _ in range(history):
            self.fast.Render(view, C.LH2_RESTART)
        acc = self.fast.read_buffer("acc64")
        radiance = (acc[0, ..., :3] + acc[1, ..., :3]).astype(f32)
        albedo, normal, depth = decode_features(self.fast.read_buffer("features"))
        disp = self.fast.read_image()[..., :3].astype(np.float64)
        svgf = (disp ** 4).astype(f32)              # display = lastPass^(1/2), lastPass = radiance^(1/2)
        self.slow.Render(view, C.LH2_RESTART)
        for _ in range(self.ref_spp // self.per):
            self.slow.Render(view, C.LH2_CONVERGE)   # the first converging frame restarts: ref_spp samples in all
        acc = self.slow.read_buffer("acc64")
        ref = ((acc[0, ..., :3] + acc[1, ..., :3]) / self.ref_spp).astype(f32)
        return dict(radiance=radiance, albedo=albedo, normal=normal, depth=depth, svgf=svgf, ref=ref)

    def close(self):
        self.fast.Shutdown()
        self.slow.Shutdown()


ALBEDO_FLOOR = f32(0.02)


def _demod(x, albedo):
    """Radiance -> log(1 + radiance / albedo): the network works on the
    untextured, range-compressed signal, as SVGF filters albedo-demodulated
    illumination."""
    return np.log1p(np.maximum(x, 0) / np.maximum(albedo, ALBEDO_FLOOR)).astype(f32)


def network_inputs(torch, sample):
    """Device tensors for UNetDenoiser.set_input: demodulated log radiance, albedo,
    normal (h, w, 4) and scaled log depth (h, w)."""
    h, w = sample["radiance"].shape[:2]
    pad = lambda a: np.concatenate([a, np.zeros((h, w, 1), f32)], -1)
    depth = np.log1p(np.minimum(np.where(np.isfinite(sample["depth"]), sample["depth"], 1e4), 1e4)) * f32(0.1)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, f32)).cuda()
    rad = _demod(sample["radiance"], sample["albedo"])
    return dev(pad(rad)), dev(pad(sample["albedo"])), dev(pad(sample["normal"])), dev(depth.astype(f32))


def network_target(torch, sample):
    """Training target in the network's output space (h, w, 4) on the device."""
    t = _demod(sample["ref"], sample["albedo"])
    t = np.concatenate([t, np.ones(t.shape[:2] + (1,), f32)], -1)
    return torch.from_numpy(np.ascontiguousarray(t)).cuda()


def network_radiance(out, sample):
    """Network output (h, w, >=3) -> radiance: (exp(out) - 1) * albedo."""
    o = np.asarray(out, f32)[..., :3]
    return (np.expm1(np.maximum(o, 0)) * np.maximum(sample["albedo"], ALBEDO_FLOOR)).astype(f32)


def train(net, torch, samples, steps, lr=2e-3, seed=0, log_every=0):
    """Adam on single-image batches drawn from `samples` (prepared with
    network_inputs and a device target). Returns the loss history."""
    rng = np.random.default_rng(seed)
    losses = []
    for i in range(steps):
        inp, tgt = samples[int(rng.integers(len(samples)))]
        net.set_input(*inp)
        net.forward()
        loss = net.loss_backward(tgt)
        net.step(lr=lr * (0.5 if i > steps * 0.7 else 1.0))
        if log_every and (i % log_every == 0 or i == steps - 1):
            losses.append((i, float(loss.item())))
    return losses
