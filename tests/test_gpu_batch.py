"""GPU tests of the batched fixed-step integrator: every mesh of a batch of uneven sizes
must follow the same trajectory as the plain single-mesh path (itself pinned against the
oracle), the graph replay must equal direct launches, and one mesh is checked against the
oracle directly."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu

H = (4e-9, 4e-9, 4e-9)


def batch_meshes(seed=7):
    rng = np.random.default_rng(seed)
    sizes = [(64, 32, 1), (48, 64, 2), (64, 32, 1), (32, 32, 4), (96, 48, 1), (32, 24, 2)]
    meshes = []
    for n in sizes:
        p = api.FMParams(Ms=float(rng.uniform(4e5, 1.2e6)), A=1.3e-11, alpha=float(rng.uniform(0.01, 0.5)),
                         modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE)
        meshes.append((n, H, p))
    M0 = [api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 11 + b).astype(np.float32) for b, (n, _, p) in enumerate(meshes)]
    return meshes, M0


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("method", [api.BM_EVAL_RK4, api.BM_EVAL_TEULER])
def test_batch_equals_plain_single_mesh_runs(gpu, method):
    meshes, M0 = batch_meshes()
    dt, steps = 50e-15, 20
    bs = api.BatchSolver(method, meshes, M0)
    bs.set_dt(dt)
    bs.run(steps)
    torch.cuda.synchronize()
    assert bs.state()["iteration"] == steps
    for b, (n, h, p) in enumerate(meshes):
        s = api.FMSolver(method, n, h, p, M0[b])
        s.set_dt(dt)
        s.run_to_iteration(steps)
        e = rel_l2(bs.mesh(b).cpu().numpy(), s.M.cpu().numpy().astype(np.float64))
        assert e <= 2e-6, (b, n, e)
        s.close()
    bs.close()


def test_graph_replay_equals_direct_launches(gpu):
    meshes, M0 = batch_meshes(3)
    a = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, use_graph=True)
    b = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, use_graph=False)
    for x in (a, b):
        x.set_dt(50e-15)
        x.run(5)
    torch.cuda.synchronize()
    assert torch.equal(a.M, b.M)
    # a second run continues from the same state
    a.run(3); b.run(3)
    torch.cuda.synchronize()
    assert torch.equal(a.M, b.M)


def test_batch_mesh_matches_oracle(gpu):
    meshes, M0 = batch_meshes(5)
    dt, steps = 50e-15, 10
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    bs.set_dt(dt)
    bs.run(steps)
    torch.cuda.synchronize()
    for b in (1, 3):
        n, h, p = meshes[b]
        Nz = 1
        while n[2] > 1 and Nz < 2 * n[2]:
            Nz *= 2
        o = O.OracleSim(api.BM_EVAL_RK4, n, h, p, Nz_override=Nz if n[2] > 1 else 0)
        o.set_M(M0[b].astype(np.float64))
        o.set_dt(dt)
        o.run_to_iteration(steps)
        assert rel_l2(bs.mesh(b).cpu().numpy(), o.get_M()) <= 1e-5


def stack_batch(seed=21):
    """Layer stacks (the multilayer benchmark's 2/1/4-cell layers with 1-cell spacers) of
    uneven lateral sizes, with per-layer Ms, A, alpha, K1 and a per-stack field ramp, batched
    with a plain 2D and a plain 3D mesh (one with a sinc pulse)."""
    rng = np.random.default_rng(seed)
    hz, z0 = [8e-9, 4e-9, 16e-9], [0.0, 12e-9, 20e-9]
    items = []
    for nx, ny in ((32, 32), (48, 32), (32, 32)):
        lay = dict(hz=hz, z0=z0, Ms=list(rng.uniform(6e5, 1.2e6, 3)), A=[1.3e-11] * 3, alpha=[0.02, 0.05, 0.02],
                   K1=[0.0, 3e5, 0.0])
        slope = tuple(float(v) for v in rng.normal(0, 4e13, 3))
        p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, ea=(0.0, 0.0, 1.0), Ha=tuple(float(v) for v in rng.normal(0, 1e4, 3)),
                         wave=2, slope=slope,
                         modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)
        items.append(((nx, ny, 3), (4e-9, 4e-9, 16e-9), p, lay))
    d = rng.normal(size=3); d /= np.linalg.norm(d)
    items.append(((64, 32, 1), H, api.FMParams(Ms=8e5, alpha=0.05, wave=1, amp=1.5e4, w=2 * np.pi / 2e-11, t0=1e-12, dir=tuple(d),
                                               modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN), None))
    items.append(((32, 16, 4), H, api.FMParams(Ms=1e6, alpha=0.1, Ha=(1e4, 0.0, 0.0), wave=2, slope=(0.0, 3e13, 0.0),
                                               modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN), None))
    M0 = []
    for b, (n, h, p, lay) in enumerate(items):
        ms = np.repeat(lay["Ms"], n[0] * n[1]) if lay else np.full(n[0] * n[1] * n[2], p.Ms)
        rng2 = np.random.default_rng(100 + b)
        th, ph = 0.5 * rng2.random(ms.size), 2 * np.pi * rng2.random(ms.size)
        M0.append((np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], 1) * ms[:, None]).astype(np.float32))
    return items, M0


def test_batched_layer_stacks_equal_single_runs(gpu):
    items, M0 = stack_batch()
    dt, steps = 100e-15, 40
    bs = api.BatchSolver(api.BM_EVAL_RK4, [(n, h, p) for n, h, p, _ in items], M0, layers=[lay for *_, lay in items])
    bs.set_dt(dt)
    bs.run(steps)
    torch.cuda.synchronize()
    for b, (n, h, p, lay) in enumerate(items):
        if lay:
            s = api.LayerStackSolver(api.BM_EVAL_RK4, n[0], n[1], h[0], h[1], lay["hz"], lay["z0"], p,
                                     Ms=lay["Ms"], A=lay["A"], alpha=lay["alpha"], K1=lay["K1"], M0=M0[b])
        else:
            s = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0[b])
        s.set_dt(dt)
        s.run_to_iteration(steps)
        torch.cuda.synchronize()
        e = rel_l2(bs.mesh(b).cpu().numpy(), s.M.cpu().numpy().astype(np.float64))
        # the device evaluates the time-dependent field in FP32 from the step time, the single
        # path in FP64 on the host: identical up to the field's rounding
        assert e <= 2e-5, (b, n, e)
        s.close()
    assert abs(bs.state()["time"] - steps * dt) <= 1e-6 * steps * dt
    bs.close()


def test_batched_layer_stack_graph_equals_direct(gpu):
    items, M0 = stack_batch(5)
    args = ([(n, h, p) for n, h, p, _ in items], M0)
    a = api.BatchSolver(api.BM_EVAL_RK4, *args, use_graph=True, layers=[lay for *_, lay in items])
    b = api.BatchSolver(api.BM_EVAL_RK4, *args, use_graph=False, layers=[lay for *_, lay in items])
    for x in (a, b):
        x.set_dt(100e-15)
        x.run(6)
    torch.cuda.synchronize()
    assert torch.equal(a.M, b.M)


def test_autotuned_batch_equals_default(gpu):
    """Launch shapes change only how work items are spread over workgroups, not their
    arithmetic: the autotuned batch reproduces the default batch bit for bit."""
    meshes, M0 = batch_meshes(9)
    items, S0 = stack_batch(9)
    meshes = meshes + [(n, h, p) for n, h, p, _ in items]
    M0 = M0 + S0
    layers = [None] * (len(meshes) - len(items)) + [lay for *_, lay in items]
    a = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, layers=layers, autotune=True)
    b = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, layers=layers)
    for x in (a, b):
        x.set_dt(50e-15)
        x.run(4)
    torch.cuda.synchronize()
    assert torch.equal(a.M, b.M)
    rep = a.autotune_report()
    kinds = {l.split()[0] for l in rep}
    assert {"x_fwd", "x_inv", "y_fwd", "y_inv", "z", "y2d", "ml_mult"} <= kinds, rep
    assert all(float(l.split()[-2]) > 0 for l in rep), rep
    assert b.autotune_report() == []
