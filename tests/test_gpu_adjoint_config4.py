"""The differentiation workload as stated: a 128x128x4 film of 4 nm cells with a Gaussian-blob
Ms map in [5e5, 1.1e6] A/m, A 1.3e-11 J/m, alpha 0.05, a 1 ns sinc pulse of 20 mT in a seeded
direction, smooth seeded m0 and m_target; 4000 RK4 steps of 250 fs with checkpoints every
100; loss 1 - <m_final . m_target>; gradients with respect to the Ms map and the pulse
amplitude by the GPU's discrete adjoint.

Checked against
* the FP64 oracle's discrete adjoint of the same rollout on the same (FP32-rounded) inputs
  (tests/golden/adjoint_config4.npz, written by tests/golden/make_adjoint_config4.py; the
  oracle adjoint is itself pinned by central differences to 1e-6 with quadratic convergence,
  tests/test_oracle_adjoint.py): over all cells, at 32 seeded cells, the pulse amplitude and
  the directional derivative along a seeded sign pattern on the 32 cells;
* central finite differences of the GPU's own FP32 forward rollout along that direction, at
  h = 1e-2 and 1e-3 relative.  A one-cell perturbation moves the loss by about 1e-7 here,
  which is the FP32 rollout's own rounding noise, so one-cell central differences in FP32 do
  not resolve 1e-4; the 32-cell direction moves it by 4e-6 (1e-3) to 4e-5 (1e-2).
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import bench
from boris2_amd import api

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "adjoint_config4.npz"
STEPS, DT, CKPT = 4000, 250e-15, 100


def rollout(n, h, p, Ms, m0t, mtt, backward=False):
    r = api.RK4Adjoint(n, h, p, torch.tensor(Ms, dtype=torch.float32), DT, STEPS, CKPT)
    L, _ = r.forward(m0t, mtt)
    g = ga = None
    if backward:
        g, ga = r.backward()
        g = g.cpu().numpy().astype(np.float64)
    r.close()
    return L, g, ga


def test_config4_adjoint_against_fp64_oracle_adjoint_and_fd(gpu):
    f = np.load(FIX)
    assert int(f["steps"]) == STEPS and float(f["dt"]) == DT
    n, h, p, Ms, m0, mt = bench.adjoint_problem()
    ncell = n[0] * n[1] * n[2]
    Ms = np.asarray(Ms, np.float32)
    m0t, mtt = torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32)
    L, g, ga = rollout(n, h, p, Ms, m0t, mtt, backward=True)
    g0, ga0, L0 = f["gMs"], float(f["gamp"]), float(f["loss"])
    eL = abs(L - L0)
    eg = float(np.linalg.norm(g - g0) / np.linalg.norm(g0))
    ea = abs(ga - ga0) / abs(ga0)
    rng = np.random.default_rng(32)
    cells = rng.choice(ncell, 32, replace=False)
    ec = np.abs(g[cells] - g0[cells]) / np.abs(g0[cells])
    sgn = rng.choice([-1.0, 1.0], 32)
    v = np.zeros(ncell)
    v[cells] = sgn * Ms[cells]
    dir_adj, dir_ora = float(g @ v), float(g0 @ v)
    print(f"loss {L:.9f} vs {L0:.9f} ({eL:.1e}); dL/dMs rel L2 {eg:.2e}; dL/damp {ga:.6e} vs {ga0:.6e} ({ea:.1e}); "
          f"32 cells: median {np.median(ec):.1e} max {ec.max():.1e}; direction {dir_adj:.6e} vs {dir_ora:.6e} "
          f"({abs(dir_adj - dir_ora) / abs(dir_ora):.1e})", flush=True)
    # central differences of the GPU forward along v (Ms * (1 +- hr * sign) on the 32 cells)
    fd = {}
    for hr in (1e-2, 1e-3):
        Mp = (Ms.astype(np.float64) + hr * v).astype(np.float32)
        Mm = (Ms.astype(np.float64) - hr * v).astype(np.float32)
        dv = (Mp.astype(np.float64) - Mm.astype(np.float64)) / 2
        Lp, _, _ = rollout(n, h, p, Mp, m0t, mtt)
        Lm, _, _ = rollout(n, h, p, Mm, m0t, mtt)
        fd[hr] = ((Lp - Lm) / 2, float(g @ dv), float(g0 @ dv))
        print(f"FD h {hr:g}: {fd[hr][0]:.6e}, adjoint {fd[hr][1]:.6e} ({abs(fd[hr][0] - fd[hr][1]) / abs(fd[hr][1]):.1e}), "
              f"oracle adjoint {fd[hr][2]:.6e}", flush=True)
    # measured: loss 1.4e-7, Ms map 1.4e-4 (L2 over all cells), amplitude 8.7e-5, the
    # 32-cell direction 4.6e-5, per cell median 4.3e-5 (cells with small gradients up to
    # 2.2e-3): the FP32 rollout's departure from the FP64 one over 4000 steps
    assert eL <= 1e-6
    assert eg <= 5e-4
    assert ea <= 2e-4
    assert abs(dir_adj - dir_ora) <= 1e-4 * abs(dir_ora)
    assert np.median(ec) <= 2e-4
    # FD of the FP32 rollout: at h = 1e-2 it differs from the exact (oracle) gradient by its
    # own truncation (1.3e-4) and from the adjoint by 1.7e-4; at h = 1e-3 the FP32 noise of
    # the loss (a few 1e-9 against a 7.7e-7 change) leaves 8e-3
    d_fd, d_adj, d_ora = fd[1e-2]
    assert abs(d_fd - d_adj) <= 5e-4 * abs(d_adj)
    assert abs(d_fd - d_ora) <= 5e-4 * abs(d_ora)
    d_fd, d_adj, _ = fd[1e-3]
    assert abs(d_fd - d_adj) <= 2e-2 * abs(d_adj)
