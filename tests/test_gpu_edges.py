"""Edge cases on the GPU against the FP64 oracle: the largest supported extent along each
axis (nx = 2048, N.y = 2048, N.z = 128), the smallest meshes (2x1x1, a single row, two
planes), and ragged radix-3 sizes through the integrator."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api
from test_gpu_ode import rel_l2, tilted_x, bench_like_params

pytestmark = pytest.mark.gpu


def nz_fft(nz):
    N = 1
    while nz > 1 and N < 2 * nz:
        N *= 2
    return N if nz > 1 else 0


@pytest.mark.parametrize("n", [(2048, 8, 1), (8, 1024, 1), (16, 8, 64), (2, 1, 1), (2, 1, 3), (6, 1, 2)])
def test_demag_extreme_sizes(gpu, n):
    h = (4e-9, 3e-9, 5e-9)
    ncell = n[0] * n[1] * n[2]
    M = api.random_magnetisation(ncell, 8e5, ncell + 3).astype(np.float32)
    H = api.Demag(n, h).convolute(torch.from_numpy(M).cuda()).cpu().numpy()
    ref = O.OracleDemag(n, h, Nz_override=nz_fft(n[2]))(M.astype(np.float64))
    assert rel_l2(H, ref) <= 1e-6


@pytest.mark.parametrize("path", ["row", "unfused"])
def test_rkdp54_ragged_radix3_mesh(gpu, path, monkeypatch):
    monkeypatch.setenv("BM_ROW_KERNEL", "1" if path == "row" else "0")
    n, h, p = (96, 24, 3), (4e-9, 4e-9, 4e-9), bench_like_params()
    M0 = tilted_x(n[0] * n[1] * n[2], p.Ms, 17, tilt=0.3)
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p, Nz_override=8)
    o.set_M(M0.astype(np.float64))
    s.run_to_iteration(20)
    o.run_to_iteration(20)
    assert s.state()["rejections"] == o.state()["rejections"]
    assert abs(s.state()["time"] - o.state()["time"]) <= 1e-3 * o.state()["time"]
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-4


def test_smallest_mesh_rk4(gpu):
    n, h = (2, 2, 1), (5e-9, 5e-9, 5e-9)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, Ha=(1e4, 2e4, 0.0),
                     modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)
    M0 = tilted_x(4, p.Ms, 3, tilt=0.5)
    s = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    o = O.OracleSim(api.BM_EVAL_RK4, n, h, p)
    o.set_M(M0.astype(np.float64))
    for x in (s, o):
        x.set_dt(1e-13)
    s.run_to_iteration(50)
    o.run_to_iteration(50)
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-5
