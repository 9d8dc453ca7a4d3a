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
    eg, ea = np.linalg.norm(g - g0) / np.linalg.norm(g0), abs(ga - ga0) / abs(ga0)
    print(f"ckpt {ckpt}: loss {abs(L - L0):.2e}, dL/dMs rel L2 {eg:.2e}, dL/damp rel {ea:.2e}")
    # measured 5.5e-6 (dL/dMs) and 4.8e-6 (dL/d amp) after 40 steps: the FP32 forward's own
    # departure from the FP64 trajectory (1.6e-6 relative) carried into the gradient
    assert abs(L - L0) <= 1e-6
    assert eg <= 5e-5
    assert ea <= 5e-5
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


HZ, Z0 = [8e-9, 4e-9, 16e-9], [0.0, 12e-9, 20e-9]
LAY = dict(hz=HZ, z0=Z0, A=[1.3e-11, 1.0e-11, 1.6e-11], alpha=[0.02, 0.08, 0.03], K1=[0.0, 3e5, 1e5])


def stack_problem(nx=16, ny=12, seed=6):
    rng = np.random.default_rng(seed)
    ncell = 3 * nx * ny
    Ms = np.repeat(rng.uniform(6e5, 1.2e6, 3), nx * ny)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, K2=2e4, ea=(0.3, 0.0, 0.95),
                     modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN,
                     Ha=(-1e4, 5e3, 2e3), wave=2, slope=(4e14, -2e14, 1e14))
    m0 = rng.normal(size=(ncell, 3)); m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    mt = rng.normal(size=(ncell, 3)); mt /= np.linalg.norm(mt, axis=1, keepdims=True)
    return (nx, ny, 3), p, Ms, m0, mt


def test_gpu_layer_stack_adjoint_matches_oracle(gpu):
    """Layer stack with per-layer A, alpha, K1, anisotropy with K2 on a tilted axis and a field
    ramp: per-cell Ms gradients and the ramp-slope gradients against the FP64 oracle adjoint."""
    n, p, Ms, m0, mt = stack_problem()
    dt, steps = 200e-15, 40
    q = api.FMParams(**{**p.__dict__, "modules": p.modules & ~api.BM_MOD_DEMAG})
    o = O.OracleAdjoint(n, (4e-9,) * 3, q, Ms, dt, steps)
    o.set_layers(4e-9, 4e-9, HZ, Z0, with_demag=True, A=LAY["A"], alpha=LAY["alpha"], K1=LAY["K1"])
    L0, _ = o.forward(m0, mt)
    g0, _ = o.backward(m0, mt)
    gz0 = o.gz.copy()
    r = api.RK4Adjoint(n, (4e-9, 4e-9, 16e-9), p, torch.tensor(Ms, dtype=torch.float32), dt, steps, 9, layers=LAY)
    L, _ = r.forward(torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32))
    g, _ = r.backward()
    gz = r.zeeman_grads()
    g = g.cpu().numpy().astype(np.float64)
    eg = np.linalg.norm(g - g0) / np.linalg.norm(g0)
    print(f"stack: loss {abs(L - L0):.2e}, dL/dMs rel L2 {eg:.2e}, slopes {gz[1:]} vs {gz0[1:]}")
    assert abs(L - L0) <= 1e-6
    assert eg <= 5e-5
    np.testing.assert_allclose(gz[1:], gz0[1:], rtol=5e-5, atol=5e-5 * np.abs(gz0[1:]).max())
    r.close()


def test_layer_stack_autograd_matches_direct(gpu):
    n, p, Ms, m0, mt = stack_problem(nx=8, ny=8, seed=7)
    dt, steps = 200e-15, 20
    Ms_l = torch.tensor(Ms.reshape(3, -1)[:, 0], dtype=torch.float64, requires_grad=True)
    slope = torch.tensor(p.slope, dtype=torch.float64, requires_grad=True)
    m0t, mtt = torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32)
    loss = api.LayerStackRolloutLoss.apply(Ms_l, slope, n[0], n[1], 4e-9, 4e-9, LAY, p, dt, steps, m0t, mtt, 5)
    (2.0 * loss).backward()
    r = api.RK4Adjoint(n, (4e-9, 4e-9, 16e-9), p, torch.tensor(Ms, dtype=torch.float32), dt, steps, 5, layers=LAY)
    L, _ = r.forward(m0t, mtt)
    g, _ = r.backward()
    gz = r.zeeman_grads()
    assert abs(loss.item() - L) <= 1e-12
    gl = g.double().cpu().reshape(3, -1).sum(1)
    assert torch.allclose(Ms_l.grad, 2.0 * gl, rtol=1e-6, atol=0)
    assert torch.allclose(slope.grad, 2.0 * torch.tensor(gz[1:]), rtol=1e-9, atol=0)


def test_batched_adjoint_equals_one_by_one(gpu):
    """Four layer stacks with different per-cell Ms, starting states, targets and ramps: the
    lockstep multi-stream adjoint gives each the same loss, Ms gradient and ramp-slope
    gradient, bit for bit, as its own adjoint run alone."""
    dt, steps, ck = 200e-15, 30, 7
    probs = []
    for b in range(4):
        n, p, Ms, m0, mt = stack_problem(nx=24, ny=16, seed=30 + b)
        p.slope = (1e13 * (b + 1), -5e12, 2e12 * b)
        probs.append((n, p, Ms, m0, mt))
    mk = lambda n, p, Ms: api.RK4Adjoint(n, (4e-9, 4e-9, 16e-9), p, torch.tensor(Ms, dtype=torch.float32), dt, steps, ck, layers=LAY)
    alone = []
    for n, p, Ms, m0, mt in probs:
        r = mk(n, p, Ms)
        L, Mf = r.forward(torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32))
        g, _ = r.backward()
        alone.append((L, Mf.cpu(), g.cpu(), r.zeeman_grads()))
        r.close()
    rs = [mk(n, p, Ms) for n, p, Ms, _, _ in probs]
    bt = api.RK4AdjointBatch(rs)
    losses, Mfs = bt.forward([torch.tensor(m0, dtype=torch.float32) for *_, m0, _ in probs],
                             [torch.tensor(mt, dtype=torch.float32) for *_, mt in probs])
    grads = bt.backward()
    for b, (L, Mf, g, gz) in enumerate(alone):
        assert losses[b] == L
        assert torch.equal(Mfs[b].cpu(), Mf)
        assert torch.equal(grads[b][0].cpu(), g)
        assert np.array_equal(rs[b].zeeman_grads(), gz)
    for r in rs:
        r.close()


@pytest.mark.parametrize("n,pbc", [((16, 8, 1), (3, 2, 0)), ((8, 8, 4), (2, 0, 2))], ids=["2d_xy", "3d_xz"])
def test_gpu_periodic_adjoint_matches_oracle(gpu, n, pbc):
    """Periodic meshes (bm_adj_set_pbc with a demag handle of the same images) against the
    oracle's periodic adjoint, itself pinned by central differences."""
    n, p, Ms, m0, mt = problem(n=n, seed=7)
    dt, steps = 250e-15, 30
    a = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps, pbc=pbc)
    L0, _ = a.forward(m0, mt)
    g0, ga0 = a.backward(m0, mt)
    r = api.RK4Adjoint(n, (4e-9,) * 3, p, torch.tensor(Ms, dtype=torch.float32), dt, steps, 8, pbc=pbc)
    L, _ = r.forward(torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32))
    g, ga = r.backward()
    g = g.cpu().numpy().astype(np.float64)
    eg, ea = np.linalg.norm(g - g0) / np.linalg.norm(g0), abs(ga - ga0) / abs(ga0)
    print(f"pbc {pbc}: loss {abs(L - L0):.2e}, dL/dMs rel L2 {eg:.2e}, dL/damp rel {ea:.2e}")
    assert abs(L - L0) <= 1e-6
    assert eg <= 5e-5 and ea <= 5e-5
    r.close()


def test_gpu_adjoint_pbc_needs_matching_demag(gpu):
    n, p, Ms, m0, mt = problem(n=(8, 8, 1), seed=3)
    from boris2_amd import _lib
    r = api.RK4Adjoint(n, (4e-9,) * 3, p, torch.tensor(Ms, dtype=torch.float32), 250e-15, 4, 2)
    with pytest.raises(RuntimeError, match="error 5"):
        _lib.check(r.L.bm_adj_set_pbc(r.handle, 1, 0, 0))   # its demag handle has no images
    r.close()
