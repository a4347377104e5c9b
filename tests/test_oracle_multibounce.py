"""Closed forms for the reference estimator's multi-bounce path: Russian
roulette and the throughput update of RenderCore_PrimeRef's shadeKernel
(lib/RenderCore_PrimeRef/kernels/pathtracer.h:150-161), with
SurvivalProbability (lib/CUDA/shared_kernel_code/tools_shared.h:213-216) and
the Lambert branch of SampleBSDF / EvaluateBSDF (kernels/bsdf.h:40-101),
against the oracle (oracle/oracle_render.cpp). CPU only.

A diffuse bounce of colour c survives with p = min(1, max(c)) where the
"albedo" handed to SurvivalProbability is the BSDF value c / pi, and the
survivor's throughput is multiplied by (c / pi) |cos| / (p * cos / pi)
= pi c / max(c): RR is unbiased, each bounce multiplies the expected
throughput by c. At pathLength == maxDepth the survival probability is 0,
and the connection (NEE) of the last vertex is not traced.

1. A Lambert floor under a uniform sky, no lights: a path hits the floor,
   survives with p = max(c)/pi and then sees the sky, so every path slot
   is either 0 or exactly pi c / max(c) * L, and the pixel mean is c L
   (maxDepth >= 2; 0 at maxDepth 1).
2. Inside a closed diffuse sphere (radius R, albedo c) with a point light of
   intensity I at its centre: every wall point receives I / R^2 directly
   and, from the rest of the wall, pi times the wall's radiance (a sphere's
   interior form factors are uniform), so with NEE at path vertices
   1 .. maxDepth - 1 a pixel sees (c / pi)(I / R^2)(1 + c + ... + c^(d-2)).
"""
import numpy as np
import pytest

from lighthouse2_amd import formats as F, scenes
import oracle as orc

f32 = np.float32
ALBEDO = np.array([0.625, 0.5, 0.375])   # exact in fp16 (the material's storage) and fp32


def _sky_floor(L):
    s = scenes.Scene()
    m = s.add_material(F.lambert_material(ALBEDO.astype(f32)))
    v, t = scenes.add_quad((0, 1, 0), (0, 0, 0), 1.0e3, 1.0e3, m)
    s.add_instance(s.add_mesh(v, t))
    s.sky = (np.broadcast_to(np.asarray(L, f32), (8, 16, 3)).copy(), None)
    return s


@pytest.mark.parametrize("max_depth", [1, 2, 3])
def test_uniform_sky_over_a_floor_path_by_path(max_depth):
    L = np.array([1.5, 2.0, 0.75])
    s = _sky_floor(L)
    o = orc.OracleScene(s)
    w = h = 64
    spp = 32
    view = scenes.camera_view((0.0, 1.0, -1.0), (0.0, 0.0, 0.0), 40.0, w, h)
    pa = o.render_frame(view, w, h, spp, max_depth, 1e-3, 0x9e3779b9, 0).astype(np.float64)[:, :3]
    if max_depth == 1:
        assert np.all(pa == 0)
        return
    lit = np.any(pa > 0, axis=1)
    # each survivor carries exactly pi c / max(c) * L (fp32 arithmetic: 1e-5)
    expect = np.pi * ALBEDO / ALBEDO.max() * L
    assert np.allclose(pa[lit], expect, rtol=1e-5), (pa[lit][:4], expect)
    assert np.all(pa[~lit] == 0)
    # survival probability max(c) / pi, binomial within 4 sigma
    n, p = len(pa), ALBEDO.max() / np.pi
    assert abs(lit.mean() - p) < 4 * np.sqrt(p * (1 - p) / n), (lit.mean(), p)
    # the estimator's mean: c L
    m = pa.mean(0)
    se = pa.std(0) / np.sqrt(n)
    assert np.all(np.abs(m - ALBEDO * L) < 4 * se + 1e-9), (m, ALBEDO * L, se)


def _sphere_with_point_light(intensity, radius=1.0, slices=256, stacks=128):
    s = scenes.Scene()
    m = s.add_material(F.lambert_material(ALBEDO.astype(f32)))
    a, b, c = scenes.uv_sphere((0.0, 0.0, 0.0), radius, slices, stacks)
    v, t = scenes.add_triangles(a, b, c, m)
    s.add_instance(s.add_mesh(v, t))
    pl = np.zeros(1, F.CorePointLight)
    pl["position"] = (0.0, 0.0, 0.0)
    pl["radiance"] = intensity
    pl["energy"] = float(np.sum(intensity))
    s.point_lights = pl
    return s


@pytest.mark.parametrize("max_depth", [2, 3, 5])
def test_integrating_sphere_series(max_depth):
    I = np.array([2.0, 3.0, 1.0])
    s = _sphere_with_point_light(I)
    o = orc.OracleScene(s)
    w = h = 64
    spp = 256
    # the camera at the centre (beside the light: its shadow rays end short of the centre)
    view = scenes.camera_view((0.0, 0.0, 0.0), (0.3, 0.2, 1.0), 60.0, w, h)
    pa = o.render_frame(view, w, h, spp, max_depth, 1e-4, 0x2545f491, 0).astype(np.float64)[:, :3]
    n = len(pa)
    m = pa.mean(0)
    se = pa.std(0) / np.sqrt(n)
    L1 = ALBEDO / np.pi * I          # radius 1
    series = sum(ALBEDO ** k for k in range(max_depth - 1))
    expect = L1 * series
    # flat facets of the 65k-triangle sphere bias the closed form by < 2e-3 relative
    tol = 4 * se + 2e-3 * expect
    assert np.all(np.abs(m - expect) < tol), (max_depth, m, expect, se)
    # the next term of the series (a connection at the last vertex, or one more
    # bounce) would move the mean by far more than the tolerance
    assert np.all(ALBEDO[:2] ** (max_depth - 1) * L1[:2] > 3 * tol[:2])
