"""The learned denoiser trained on frames the core renders (1 spp of the
filter estimator with its G-buffer, against converged renders of the same
poses) by the library's own forward, backward and Adam kernels: on held-out
poses it must cut the relMSE of the 1-spp frame (the method without it) by
more than half; SVGF's relMSE on the same frames is reported beside it."""
import numpy as np
import pytest

from lighthouse2_amd import scenes
from lighthouse2_amd import denoiser as D
from lighthouse2_amd import training as T

pytestmark = pytest.mark.gpu


def test_trained_on_rendered_pairs(torch_cuda):
    torch = torch_cuda
    s = scenes.instanced_scene(meshes=4, tris=(4000, 12000), instances=300, tex_size=128, lights=16, extent=20.0)
    w, h = 128, 128
    pr = T.PairRenderer(s, w, h, ref_spp=256, spp_per_frame=64)
    views = T.views_near_instances(s, 30, w, h, seed=11)
    data = [pr.render(v) for v in views]
    pr.close()
    train, held = data[:24], data[24:]
    samples = [(T.network_inputs(torch, d), T.network_target(torch, d)) for d in train]
    net = D.UNetDenoiser(h, w, seed=3)
    T.train(net, torch, samples, 600, lr=1e-3)
    noisy, svgf, learned = [], [], []
    for d in held:
        net.set_input(*T.network_inputs(torch, d))
        out = T.network_radiance(net.forward().cpu().numpy(), d)
        noisy.append(T.relmse(d["radiance"], d["ref"]))
        svgf.append(T.relmse(d["svgf"][1:-1, 1:-1], d["ref"][1:-1, 1:-1]))
        learned.append(T.relmse(out, d["ref"]))
    print(f"held-out relMSE: 1 spp {np.mean(noisy):.4f}, SVGF {np.mean(svgf):.4f}, learned {np.mean(learned):.4f}")
    assert np.isfinite(learned).all()
    assert np.mean(learned) < 0.5 * np.mean(noisy)
