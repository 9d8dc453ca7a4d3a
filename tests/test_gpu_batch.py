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
