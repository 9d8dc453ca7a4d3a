"""GPU discrete RK4 adjoint against the FP64 oracle adjoint (itself pinned by central
finite differences), through the C-ABI and through the torch.autograd.Function."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu


def problem(n=(16, 8, 2), seed=2):
    rng = np.random.default_rng(seed)
    ncell = n[0] * n[1] * n[2]
    Ms = rng.uniform(5e5, 1.1e6, ncell)
    d = rng.normal(size=3); d /= np.linalg.norm(d)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN,
                     wave=1, amp=20e-3 / api.MU0, w=2 * np.pi / 1e-9, t0=2e-12, dir=tuple(d))
    m0 = rng.normal(size=(ncell, 3)); m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    mt = rng.normal(size=(ncell, 3)); mt /= np.linalg.norm(mt, axis=1, keepdims=True)
    return n, p, Ms, m0, mt


def oracle_grads(n, p, Ms, m0, mt, dt, steps):
    a = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps, Nz_override=4 if n[2] > 1 else 0)
    L, _ = a.forward(m0, mt)
    g, ga = a.backward(m0, mt)
    return L, g, ga


@pytest.mark.parametrize("ckpt", [7, 40])
def test_gpu_adjoint_matches_oracle(gpu, ckpt):
    n, p, Ms, m0, mt = problem()
    dt, steps = 250e-15, 40
    L0, g0, ga0 = oracle_grads(n, p, Ms, m0, mt, dt, steps)
    r = api.RK4Adjoint(n, (4e-9,) * 3, p, torch.tensor(Ms, dtype=torch.float32), dt, steps, ckpt)
    L, _ = r.forward(torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32))
    g, ga = r.backward()
    g = g.cpu().numpy().astype(np.float64)
    assert abs(L - L0) <= 1e-5
    assert np.linalg.norm(g - g0) / np.linalg.norm(g0) <= 2e-3
    assert abs(ga - ga0) <= 2e-3 * abs(ga0)
    r.close()


def test_autograd_function_matches_direct(gpu):
    n, p, Ms, m0, mt = problem(n=(8, 8, 1), seed=4)
    dt, steps = 250e-15, 20
    Mst = torch.tensor(Ms, dtype=torch.float32, device="cuda", requires_grad=True)
    amp = torch.tensor(p.amp, dtype=torch.float64, requires_grad=True)
    m0t, mtt = torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32)
    loss = api.RK4RolloutLoss.apply(Mst, amp, n, (4e-9,) * 3, p, dt, steps, m0t, mtt, 5)
    (3.0 * loss).backward()
    r = api.RK4Adjoint(n, (4e-9,) * 3, p, Mst.detach(), dt, steps, 5)
    L, _ = r.forward(m0t, mtt)
    g, ga = r.backward()
    assert abs(loss.item() - L) <= 1e-12
    assert torch.allclose(Mst.grad, 3.0 * g, rtol=1e-6, atol=0)
    assert abs(amp.grad.item() - 3.0 * ga) <= 1e-9 * abs(ga)
