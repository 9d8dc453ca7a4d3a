"""Launch variants that must not change results, against the default path:

* the integrator's row kernel at nx = 512 (one wavefront per 512-point x transform): the
  wave-independent variant (BM_ROW_WAVE=1 at creation) against the barrier-synchronised
  default, and against the FP64 oracle.  Both run the same arithmetic in the same order, so
  their trajectories agree bit for bit; meshes with ny not a multiple of the 8-row tile (a
  wave then holds one or no row of the last tile) and odd ny are included;
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu


def params():
    return api.FMParams(Ms=1e6, A=1.5e-11, alpha=0.1, K1=5e5, ea=(0.0, 0.0, 1.0), Ha=(0.0, 0.0, 39788.736),
                        modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)


def run(monkeypatch, wave, method, n, M0, steps):
    monkeypatch.setenv("BM_ROW_KERNEL", "1")
    monkeypatch.setenv("BM_ROW_WAVE", "1" if wave else "0")
    s = api.FMSolver(method, n, (4e-9, 4e-9, 4e-9), params(), M0)
    s.run_to_iteration(steps)
    torch.cuda.synchronize()
    out = (s.M.cpu().clone(), s.state())
    s.close()
    return out


@pytest.mark.parametrize("n", [(512, 16, 2), (512, 13, 3), (512, 6, 1), (512, 19, 1)])
@pytest.mark.parametrize("method", [api.BM_EVAL_RKDP54, api.BM_EVAL_RK4, api.BM_EVAL_TEULER])
def test_wave_row_kernel_equals_block_row_kernel(gpu, monkeypatch, n, method):
    M0 = api.random_magnetisation(n[0] * n[1] * n[2], 1e6, 3).astype(np.float32)
    Mw, sw = run(monkeypatch, True, method, n, M0, 8)
    Mb, sb = run(monkeypatch, False, method, n, M0, 8)
    assert sw["iteration"] == sb["iteration"] == 8
    assert sw["rejections"] == sb["rejections"]
    assert sw["time"] == sb["time"]
    assert torch.equal(Mw, Mb)


def test_wave_row_kernel_matches_oracle(gpu, monkeypatch):
    n = (512, 13, 3)
    M0 = api.random_magnetisation(n[0] * n[1] * n[2], 1e6, 4).astype(np.float32)
    Mw, sw = run(monkeypatch, True, api.BM_EVAL_RKDP54, n, M0, 6)
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, (4e-9, 4e-9, 4e-9), params(), Nz_override=8)
    o.set_M(M0.astype(np.float64))
    o.run_to_iteration(6)
    so = o.state()
    assert sw["rejections"] == so["rejections"]
    assert abs(sw["time"] - so["time"]) <= 1e-3 * so["time"]
    Mo = o.get_M()
    e = float(np.linalg.norm(Mw.numpy().astype(np.float64) - Mo) / np.linalg.norm(Mo))
    print(f"wave row kernel vs FP64 oracle after 6 RKDP54 steps: rel L2 {e:.2e}")
    assert e <= 5e-4

