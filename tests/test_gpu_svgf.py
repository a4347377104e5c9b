"""SVGF on the GPU against the CPU restatement
(oracle/oracle_svgf.cpp), bit for bit: a three-frame sequence (fresh
history, static camera, moving jittered camera with a specular strip that
takes the diamond search), comparing the display image, the last-pass
image, luminance moments, motion vectors, history counts and the phase-1
history after every frame."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import scenes
import oracle as orc
from lighthouse2_amd.scenes import gbuffer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("taa", [False, True])
def test_svgf_sequence_matches_oracle(torch_cuda, taa):
    """taa: TAApass + unsharpenTAA after the last pass, the TAA history
    (read 6) compared too."""
    torch = torch_cuda
    w, h = 72, 40
    views = [scenes.camera_view((0.5, 0.6, -1.0), (0.5, 0.5, 1.0), 50.0, w, h)] * 2 + \
        [scenes.camera_view((0.52, 0.61, -0.99), (0.5, 0.5, 1.0), 50.0, w, h)]
    frames = [dict(stationary=0, j=(0.0, 0.0), pj=(0.0, 0.0)), dict(stationary=1, j=(0.0, 0.0), pj=(0.0, 0.0)),
              dict(stationary=0, j=(0.25, -0.125), pj=(0.0, 0.0))]
    g = lh.Svgf(w, h, taa=taa)
    o = orc.OracleSvgf(w, h)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out_d = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    prev = views[0]
    for k, (v, f) in enumerate(zip(views, frames)):
        acc, feat, wp, dd = gbuffer(v, w, h, seed=10 + k, samples=2)
        a, fe, p, d = dev(acc), dev(feat.view(np.int32)), dev(wp), dev(dd)
        g.frame(a.data_ptr(), fe.data_ptr(), p.data_ptr(), d.data_ptr(), prev, out_d.data_ptr(), j=f["j"], prev_j=f["pj"],
                samples=2, stationary=f["stationary"])
        torch.cuda.synchronize()
        ref = o.frame(acc, feat, wp, dd, prev, j=f["j"], prev_j=f["pj"], samples=2, stationary=f["stationary"], taa=taa)
        got = out_d.cpu().numpy()
        assert np.array_equal(got[1:-1, 1:-1].view(np.uint32), ref[1:-1, 1:-1].view(np.uint32)), k
        for which in range(7 if taa else 6):
            assert np.array_equal(g.read(which).view(np.uint32), o.read(which).view(np.uint32)), (k, which)
        prev = v
    g.close()
