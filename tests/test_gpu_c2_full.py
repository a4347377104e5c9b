"""The incoherent-traversal workload at its full size: 4,000,000 random
triangles (centroids U[0,1)^3, vertex offsets N(0, 0.005^2), seed 1), a
binned-SAH BVH built on the GPU, 1920x1080 primary rays from (0.5, 0.5, -1.2)
towards (0.5, 0.5, 0.5) and 3 cosine-weighted diffuse bounces generated on
the device. Every hit record of the four waves is traced with both node
formats (8-bit quantised BVH4 and fp32 BVH2) and must be identical bit for
bit; a random sample of the incoherent waves must equal the oracle's BVH
traversal; every hit lies in its ray's open interval."""
import numpy as np
import pytest

import lighthouse2_amd as lh
from lighthouse2_amd import core as C, scenes
import oracle as orc

pytestmark = pytest.mark.gpu

W, H, N_TRIS = 1920, 1080, 4_000_000


def test_full_size_waves_equal_across_node_formats_and_oracle(torch_cuda):
    torch = torch_cuda
    verts = scenes.random_triangles(N_TRIS, seed=1)
    core = lh.RenderCore()
    core.Setting("nodeFormat", 1)
    core.SetGeometry(0, verts, None)
    core.SetInstance(0, 0)
    core.FinalizeInstances()
    n = W * H
    rays = [torch.zeros((n, 8), dtype=torch.float32, device="cuda") for _ in range(2)]
    pid = [torch.zeros(n, dtype=torch.int32, device="cuda") for _ in range(2)]
    hits = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    sh = torch.cuda.current_stream().cuda_stream
    C.generate_primary(scenes.microbench_view(W, H), W, H, 1234, 1e-5, rays[0].data_ptr(), pid[0].data_ptr(), sh)
    waves, cur, m = [], 0, n
    for wave in range(4):
        core.trace_closest(rays[cur].data_ptr(), hits.data_ptr(), m, sh)
        waves.append((rays[cur][:m].clone(), hits[:m].clone()))
        if wave == 3:
            break
        core.bounce_diffuse(rays[cur].data_ptr(), hits.data_ptr(), pid[cur].data_ptr(), m, (wave + 1) * 91771, 1e-5,
                            rays[cur ^ 1].data_ptr(), pid[cur ^ 1].data_ptr(), cnt.data_ptr(), sh)
        m = int(cnt.item())
        cur ^= 1
    print("rays per wave:", [len(r) for r, _ in waves], flush=True)
    assert len(waves[1][0]) > 500_000 and len(waves[3][0]) > 100_000

    # the same rays with fp32 BVH2 nodes: identical hit records
    core.Setting("nodeFormat", 0)
    core.FinalizeInstances()
    for r, h in waves:
        h2 = torch.empty_like(h)
        core.trace_closest(r.data_ptr(), h2.data_ptr(), len(r), sh)
        torch.cuda.synchronize()
        assert torch.equal(h.view(torch.int32), h2.view(torch.int32))

    # every hit inside its ray's open interval (tmin, tmax), primitive ids in range
    for r, h in waves:
        rr, hh = r.cpu().numpy(), h.cpu().numpy()
        prim = hh[:, 2].view(np.int32)
        hit = prim >= 0
        assert np.all(prim[hit] < N_TRIS)
        assert np.all(hh[hit, 1].view(np.int32) == 0)
        assert np.all((hh[hit, 3] > rr[hit, 3]) & (hh[hit, 3] < rr[hit, 7]))

    # a random sample of the incoherent waves against the oracle's BVH traversal
    s = scenes.Scene()
    s.add_instance(s.add_mesh(verts, None))
    o = orc.OracleScene(s)
    inco_r = np.concatenate([r.cpu().numpy() for r, _ in waves[1:]])
    inco_h = np.concatenate([h.cpu().numpy() for _, h in waves[1:]])
    idx = np.random.default_rng(0).choice(len(inco_r), size=100_000, replace=False)
    ref = o.trace(inco_r[idx], threads=16)
    assert np.array_equal(inco_h[idx].view(np.uint32), ref.view(np.uint32))
    core.Shutdown()
