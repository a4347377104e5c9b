"""GPU parity of the optix7 estimator ("estimator" setting 1: rendercore_optix7's
shadeKernel with the shared Lambert BSDF, blue noise, MIS, firefly clamp) against
the oracle through the C ABI: fp64 accumulators identical, fp32 image within
1e-4 relative of the fp64 oracle. Covers the frame state machine (blue-noise
shift per frame, camera seed per executed wave), the random-number fallback
past 64 samples, alpha cut-outs, the small skydome after a bounce, a distorted
thin-lens camera, a scene without lights, and both BSDFs of the core (Lambert,
and Disney: diffuse, sheen, GGX specular, GTR1 clear coat, rough dielectric)."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import formats as F
from lighthouse2_amd import scenes
import oracle as orc

pytestmark = pytest.mark.gpu
REL_TOL = 1e-4


def render_gpu7(scene, view, w, h, spp, frames, eps=1e-3, seed=0x12345678, tables=None, clamp=None, modes=None, bsdf=1):
    core = lh.RenderCore()
    lh.load_scene(core, scene)
    core.SetTarget(w, h, spp)
    core.Setting("estimator", 1)
    core.Setting("bsdf", bsdf)
    core.Setting("epsilon", eps)
    if clamp is not None:
        core.Setting("clampValue", clamp)
    if seed != 0x12345678:
        core.Setting("seed", seed)
    if tables is not None:
        core.SetBlueNoise(*tables)
    for m in modes or [lh.LH2_RESTART] + [lh.LH2_CONVERGE] * frames:
        core.Render(view, m)
    img, acc = core.read_image(), core.read_accumulator_f64()
    core.Shutdown()
    return img, acc


def check(img, acc, r32, r64, n):
    assert np.array_equal(acc, r64)
    assert np.array_equal(img, (r32 * np.float32(1.0 / n)).astype(np.float32))
    ref = r64[..., :3] / n
    rel = np.abs(img[..., :3].astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-6)
    assert rel.max() <= REL_TOL


def _tables(seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 256, 65536, dtype=np.uint8), rng.integers(0, 256, 131072, dtype=np.uint8),
            rng.integers(0, 256, 131072, dtype=np.uint8))


@pytest.mark.parametrize("bsdf", [0, 1])
def test_cornell_box_default_tables(torch_cuda, bsdf):
    """Lambert (0) and Disney (1) BSDFs on the Lambert-parameterised box."""
    s = scenes.cornell_box(seed=1234, sphere_tris=8000)
    view = scenes.cornell_view(96, 96)
    img, acc = render_gpu7(s, view, 96, 96, 1, 6, bsdf=bsdf)
    r32, r64, n, _, _ = orc.render7(s, view, 96, 96, 1, 6, threads=16, bsdf=bsdf)
    print(f"radiance mean {(r64[..., :3] / n).mean():.5f}; connections per pixel {(r64[..., 3] / n).mean():.3f}")
    check(img, acc, r32, r64, n)
    assert (r64[..., 3] > 0).mean() > 0.5


def test_textured_box_past_64_samples(torch_cuda):
    """8 spp x 9 frames: samples 0..71 cross the blue-noise range (64);
    textured walls with an alpha hole and a 256 x 128 sky (4 x 2 small dome)."""
    s = scenes.textured_cornell_box(sphere_tris=3000)
    s.sky = scenes.sky_dome(256, 128, seed=12, rotate_deg=40.0)
    view = scenes.cornell_view(40, 32)
    tables = _tables(5)
    img, acc = render_gpu7(s, view, 40, 32, 8, 9, seed=4321, tables=tables)
    r32, r64, n, _, _ = orc.render7(s, view, 40, 32, 8, 9, seed=4321, blue_noise=tables, threads=16)
    assert n == 72
    check(img, acc, r32, r64, n)


def test_disney_materials(torch_cuda):
    """Plastic, anisotropic metal, clear coat, sheen + subsurface and a rough
    absorbing glass sphere through the Disney BSDF."""
    s = scenes.disney_cornell_box(sphere_tris=6000)
    view = scenes.cornell_view(80, 64)
    img, acc = render_gpu7(s, view, 80, 64, 2, 4, seed=777)
    r32, r64, n, _, _ = orc.render7(s, view, 80, 64, 2, 4, seed=777, threads=16)
    print(f"radiance mean {(r64[..., :3] / n).mean():.5f}")
    check(img, acc, r32, r64, n)
    assert np.isfinite(r64).all() and (r64[..., :3] > 0).mean() > 0.5


def test_state_machine_and_clamp(torch_cuda):
    """Restart / Converge sequence: the shift seed keeps advancing across
    restarts while the camera seed resets; a low firefly clamp."""
    s = scenes.cornell_box(seed=3, sphere_tris=2000)
    view = scenes.cornell_view(32, 32)
    R, C = lh.LH2_RESTART, lh.LH2_CONVERGE
    modes = [R, C, C, R, C, C, C]
    img, acc = render_gpu7(s, view, 32, 32, 2, 0, clamp=0.5, modes=modes)
    r32, r64, n, _, _ = orc.render7(s, view, 32, 32, 2, modes=[orc.RESTART if m == R else orc.CONVERGE for m in modes],
                                     clamp_value=0.5)
    check(img, acc, r32, r64, n)
    assert (r64[..., :3] / n).max() <= 0.5 * 3 + 1e-6     # sky (none), emitters and NEE are all clamped


def test_distorted_thin_lens_camera(torch_cuda):
    s = scenes.cornell_box(seed=8, sphere_tris=2000)
    view = scenes.camera_view((0.5, 0.5, -1.2), (0.5, 0.5, 0.5), 40.0, 48, 40, aperture=0.02, distortion=0.15)
    img, acc = render_gpu7(s, view, 48, 40, 2, 2)
    r32, r64, n, _, _ = orc.render7(s, view, 48, 40, 2, 2)
    check(img, acc, r32, r64, n)


def test_no_lights_accepts_emitters_directly(torch_cuda):
    """Without lights there is no NEE and an emitter hit returns its colour
    unweighted (pathtracer.h:133-137)."""
    s = scenes.cornell_box(sphere_tris=2000)
    s.tri_lights = s.tri_lights[:0]
    view = scenes.cornell_view(32, 32)
    img, acc = render_gpu7(s, view, 32, 32, 1, 2)
    r32, r64, n, _, _ = orc.render7(s, view, 32, 32, 1, 2)
    check(img, acc, r32, r64, n)
    assert np.all(r64[..., 3] == 0)


def test_uniform_sky_over_plane(torch_cuda):
    """Closed form on the GPU: albedo x sky radiance at every pixel."""
    a = np.array([0.25, 0.5, 0.75], np.float32)
    s = scenes.Scene()
    m = s.add_material(F.lambert_material(a, roughness=0.4))
    v, t = scenes.add_quad((0, 1, 0), (0, 0, 0), 400.0, 400.0, m)
    s.add_instance(s.add_mesh(v, t))
    s.sky = (np.full((128, 256, 3), 0.75, np.float32), F.IDENTITY.copy())
    view = scenes.camera_view((0.0, 1.0, -2.0), (0.0, 0.0, 0.0), 40.0, 32, 32)
    img, acc = render_gpu7(s, view, 32, 32, 2, 2, bsdf=0)
    assert np.allclose(acc[..., :3] / 4, a * 0.75, rtol=2e-6, atol=0)
