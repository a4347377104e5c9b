"""The oracle's Disney BSDF (sharedBSDFs/disney.h, ggxmdf.h, frosted.h) pinned
by what the mathematics fixes: the GGX and GTR1 distributions are normalised
(projected area 1), the visible-normal pdf integrates to 1 and the sampler
draws from it, the sampled pdf and value equal the evaluated ones, and
sampling is unbiased: E[f cos / pdf] equals the quadrature of f cos."""
import numpy as np
import pytest

import oracle as orc

f32 = np.float32


def _sphere_grid(n_theta=400, n_phi=800, upper=True):
    """Midpoint quadrature on the (upper) hemisphere or the sphere: directions and solid angles."""
    t_hi = np.pi / 2 if upper else np.pi
    th = (np.arange(n_theta) + 0.5) * (t_hi / n_theta)
    ph = (np.arange(n_phi) + 0.5) * (2 * np.pi / n_phi)
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (t_hi / n_theta) * (2 * np.pi / n_phi)).reshape(-1)
    return d.astype(f32), dw


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.2, 0.6), (0.8, 0.8)])
def test_ggx_projected_area_is_one(ax, ay):
    m, dw = _sphere_grid()
    D = orc.ggx("D", m, m, np.zeros((len(m), 2)), ax, ay)[:, 0].astype(np.float64)
    assert abs((D * m[:, 2] * dw).sum() - 1) < 2e-3


@pytest.mark.parametrize("alpha", [0.05, 0.3, 0.9])
def test_gtr1_projected_area_is_one(alpha):
    m, dw = _sphere_grid()
    D = orc.ggx("gtr1_D", m, m, np.zeros((len(m), 2)), alpha, alpha)[:, 0].astype(np.float64)
    assert abs((D * m[:, 2] * dw).sum() - 1) < 3e-3


@pytest.mark.parametrize("ax,ay,theta", [(0.3, 0.3, 0.3), (0.5, 0.2, 1.1), (0.7, 0.7, 1.45)])
def test_visible_normal_pdf_and_sampler(ax, ay, theta):
    v = np.array([np.sin(theta) * 0.6, np.sin(theta) * 0.8, np.cos(theta)], f32)
    m, dw = _sphere_grid()
    V = np.repeat(v[None], len(m), 0)
    pdf = orc.ggx("pdf", V, m, np.zeros((len(m), 2)), ax, ay)[:, 0].astype(np.float64)
    pdf = np.where(m @ v > 0, pdf, 0)                  # the visible normals: v . m > 0
    assert abs((pdf * dw).sum() - 1) < 3e-3
    # the sampler draws from that pdf: compare moments
    rng = np.random.default_rng(1)
    r = rng.random((200000, 2)).astype(f32)
    s = orc.ggx("sample", np.repeat(v[None], len(r), 0), np.zeros((len(r), 3)), r, ax, ay).astype(np.float64)
    assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-5) and (s[:, 2] >= 0).all()
    for g in (lambda x: x[:, 2], lambda x: x[:, 0], lambda x: x[:, 0] * x[:, 1], lambda x: x[:, 2] ** 4):
        want = (g(m.astype(np.float64)) * pdf * dw).sum()
        got = g(s).mean()
        assert abs(got - want) < 4e-3, (got, want)


def test_gtr1_sampler_draws_D_cos():
    alpha = 0.2
    m, dw = _sphere_grid()
    D = orc.ggx("gtr1_D", m, m, np.zeros((len(m), 2)), alpha, alpha)[:, 0].astype(np.float64)
    rng = np.random.default_rng(2)
    r = rng.random((200000, 2)).astype(f32)
    s = orc.ggx("gtr1_sample", np.zeros((len(r), 3)), np.zeros((len(r), 3)), r, alpha, alpha).astype(np.float64)
    for g in (lambda x: x[:, 2], lambda x: x[:, 2] ** 8, lambda x: x[:, 0] ** 2):
        want = (g(m.astype(np.float64)) * D * m[:, 2] * dw).sum()
        assert abs(g(s).mean() - want) < 4e-3


MATERIALS = {
    "diffuse": dict(specular=0.0, roughness=0.8),
    "plastic": dict(specular=0.5, roughness=0.55),
    "metal": dict(metallic=1.0, roughness=0.6, anisotropic=0.4),
    "coated": dict(specular=0.3, roughness=0.7, clearcoat=1.0, clearcoatgloss=0.3),
    "subsurface": dict(subsurface=0.6, specular=0.4, roughness=0.9, spectint=0.5),
}


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_sample_matches_evaluate_and_is_unbiased(name):
    color = np.array([0.7, 0.45, 0.25], f32)
    params = orc.disney_params(**MATERIALS[name])
    wo = np.array([0.35, -0.2, 0.915], f32)
    wo /= np.linalg.norm(wo)
    rng = np.random.default_rng(3)
    n = 200000
    r = rng.random((n, 3)).astype(f32)
    s = orc.disney_sample(color, params, np.repeat(wo[None], n, 0), r)
    ok = s[:, 3] > 1e-4
    wi = s[ok, 4:7]
    e = orc.disney_eval(color, params, np.repeat(wo[None], ok.sum(), 0), wi)
    # the pdf and value returned by the sampler are the evaluated ones (no sheen: no quirk);
    # below the horizon the half vector of a reflection is ill-conditioned (|wo + wi| small)
    up = wi[:, 2] > 0.02
    assert up.mean() > 0.8
    assert np.allclose(s[ok, 3][up], e[up, 3], rtol=2e-4, atol=1e-6)
    assert np.allclose(s[ok, :3][up], e[up, :3], rtol=2e-4, atol=1e-6)
    # unbiased: E[f cos / pdf ; wi above the surface] equals the quadrature of f cos over the
    # hemisphere (microfacet reflections below the horizon are kept by the reference and left out here)
    w = s[ok, :3].astype(np.float64) * np.abs(wi[:, 2:3]) / s[ok, 3:4] * (wi[:, 2:3] > 0)
    est = w.sum(0) / n
    d, dw = _sphere_grid(300, 600)
    f = orc.disney_eval(color, params, np.repeat(wo[None], len(d), 0), d)[:, :3].astype(np.float64)
    quad = (f * d[:, 2:3] * dw[:, None]).sum(0)
    stderr = w.std(0) / np.sqrt(n)
    print(name, est, quad, stderr)
    assert np.all(np.abs(est - quad) < 5 * stderr + 3e-3 * quad)


def test_evaluate_rejects_pure_specular_and_counts_quirk():
    color = np.array([0.5, 0.5, 0.5], f32)
    wo = np.array([[0, 0.6, 0.8]], f32)
    wi = np.array([[0.3, -0.2, 0.93]], f32)
    e = orc.disney_eval(color, orc.disney_params(roughness=0.0), wo, wi)      # ROUGHNESS <= 0.001
    assert e[0, 3] == 0 and np.all(e[0, :3] == 0)
    # sheen replaces the diffuse value in EvaluateBSDF (disney.h:327-330) but not in SampleBSDF
    p0 = orc.disney_params(specular=0.0, roughness=0.8)
    p1 = orc.disney_params(specular=0.0, roughness=0.8, sheen=1.0)
    d0 = orc.disney_eval(color, p0, wo, wi)
    d1 = orc.disney_eval(color, p1, wo, wi)
    assert d1[0, 0] < d0[0, 0]


def test_dielectric_transmits_and_reflects():
    color = np.array([1.0, 1.0, 1.0], f32)
    params = orc.disney_params(transmission=1.0, roughness=0.3, eta=1.5)
    wo = np.repeat(np.array([[0.2, 0.1, 0.97]], f32) / np.linalg.norm([0.2, 0.1, 0.97]), 4000, 0)
    r = np.random.default_rng(4).random((4000, 3)).astype(f32)
    s = orc.disney_sample(color, params, wo, r)
    assert np.all(s[:, 7] == 1)                        # the dielectric branch reports a specular bounce
    good = s[:, 3] > 1e-4
    below = (s[good, 6] < 0).mean()
    F0 = ((1.5 - 1) / (1.5 + 1)) ** 2
    assert 0.85 < below < 1.0 and abs((1 - below) - F0) < 0.05   # mostly refraction; reflection near F0
