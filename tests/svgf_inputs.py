"""Synthetic SVGF inputs for the tests: the G-buffer of two planes (floor y =
0, back wall z = 1) seen through a ViewPyramid, packed as the reference's
shade kernel packs it (PackFeatures / HDRtoRGB32 / PackNormal2,
lib/RenderCore_Optix7Filter/kernels/pathtracer.h:44-58; tools_shared.h:122-183),
with noisy direct and indirect radiance."""
import numpy as np

f32 = np.float32


def pack_normal2(n):
    c = lambda v: np.clip(((v.astype(f32) + f32(1)) * f32(511)).astype(np.int64), 0, 1023).astype(np.uint32)
    return (c(n[..., 0]) << 2) + (c(n[..., 1]) << 12) + (c(n[..., 2]) << 22)


def hdr_to_rgb32(c):
    q = lambda v, s: (f32(s) * np.minimum(f32(1), v.astype(f32))).astype(np.uint32)
    return (q(c[..., 0], 1023) << 22) + (q(c[..., 1], 2047) << 11) + q(c[..., 2], 2047)


def gbuffer(view, w, h, seed=0, noise=0.5, spec_strip=True, samples=1, const=None):
    v = view[0] if view.shape else view
    pos, p1, p2, p3 = (np.asarray(v[k], np.float64) for k in ("pos", "p1", "p2", "p3"))
    x = (np.arange(w) + 0.5) / w
    y = (np.arange(h) + 0.5) / h
    P = p1[None, None] + x[None, :, None] * (p2 - p1)[None, None] + y[:, None, None] * (p3 - p1)[None, None]
    D = P - pos
    D /= np.linalg.norm(D, axis=2, keepdims=True)
    t_floor = np.where(D[..., 1] < -1e-6, (0.0 - pos[1]) / np.where(D[..., 1] == 0, 1, D[..., 1]), np.inf)
    t_wall = np.where(D[..., 2] > 1e-6, (1.0 - pos[2]) / np.where(D[..., 2] == 0, 1, D[..., 2]), np.inf)
    floor = t_floor < t_wall
    t = np.where(floor, t_floor, t_wall)
    t = np.where(np.isfinite(t), t, 50.0)
    I = pos + D * t[..., None]
    N = np.where(floor[..., None], np.array([0, 1, 0.0]), np.array([0, 0, -1.0]))
    rng = np.random.default_rng(seed)
    if const is None:
        checker = ((np.floor(I[..., 0] * 4) + np.floor(I[..., 2] * 4) + np.floor(I[..., 1] * 4)) % 2)
        albedo = np.stack([0.3 + 0.5 * checker, 0.5 + 0.3 * floor, 0.6 - 0.3 * checker], -1)
        light = 0.5 + 0.5 * np.cos(I[..., 0] * 3) ** 2
        direct = albedo * light[..., None] * rng.gamma(1 / max(noise, 1e-3), max(noise, 1e-3), (h, w, 1))
        indirect = albedo * 0.3 * rng.gamma(1 / max(noise, 1e-3), max(noise, 1e-3), (h, w, 1))
    else:
        albedo = np.full((h, w, 3), const[0])
        direct = np.full((h, w, 3), const[1])
        indirect = np.full((h, w, 3), const[2])
    spec = np.zeros((h, w), np.uint32)
    if spec_strip:
        spec[:, w // 3: w // 3 + 3] = 1
    matid = floor.astype(np.uint32) + 1
    pn = pack_normal2(N * -1.0) + spec
    feat = np.zeros((h, w, 4), np.uint32)
    feat[..., 0] = hdr_to_rgb32(albedo)
    feat[..., 1] = pn
    feat[..., 2] = t.astype(f32).view(np.uint32)
    feat[..., 3] = (spec << 4) + (matid << 6)
    wp = np.zeros((h, w, 4), f32)
    wp[..., :3] = I
    wp[..., 3] = pn.view(f32)
    dd = np.zeros((h, w, 4), f32)
    dd[..., 2] = np.gradient(t, axis=1)
    dd[..., 3] = np.gradient(t, axis=0)
    acc = np.zeros((2, h, w, 4), f32)
    acc[0, ..., :3] = direct * samples
    acc[1, ..., :3] = indirect * samples
    return acc, feat, wp, dd
