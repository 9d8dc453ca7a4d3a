"""GPU tests of batched RKDP54 with per-mesh adaptive steps controlled on the device: every
mesh must follow the plain single-mesh RKDP54 path (same accepted and rejected counts, same
time, same state), the captured graph must equal direct launches bit for bit, a run split in
two must equal one run, and meshes are checked against the FP64 oracle directly."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api
from test_gpu_batch import H, rel_l2
from test_gpu_ode import tilted_x

pytestmark = pytest.mark.gpu


def rkdp_meshes(seed=7, wave=0, tilt=None):
    """Six meshes of uneven sizes (2D and 3D) from random states (or cones of `tilt` rad about
    x): the default first step is too long for them, so the controller rejects and shrinks dT
    per mesh."""
    rng = np.random.default_rng(seed)
    sizes = [(64, 32, 1), (48, 64, 2), (64, 32, 1), (32, 32, 4), (96, 48, 1), (32, 24, 2)]
    meshes = []
    for n in sizes:
        p = api.FMParams(Ms=float(rng.uniform(4e5, 1.2e6)), A=1.3e-11, alpha=float(rng.uniform(0.01, 0.5)),
                         K1=float(rng.uniform(0, 3e5)), ea=(0.0, 0.0, 1.0), Ha=(0.0, 0.0, 2e4),
                         modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)
        if wave:
            p.wave, p.amp, p.w, p.t0, p.dir = 1, 2e4, np.pi / 5e-12, 3e-12, (0.6, 0.8, 0.0)
        meshes.append((n, H, p))
    if tilt is None:
        M0 = [api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 21 + b).astype(np.float32) for b, (n, _, p) in enumerate(meshes)]
    else:
        M0 = [tilted_x(n[0] * n[1] * n[2], p.Ms, 21 + b, tilt=tilt) for b, (n, _, p) in enumerate(meshes)]
    return meshes, M0


@pytest.mark.parametrize("wave", [0, 1])
def test_rkdp_batch_equals_plain_single_mesh_runs(gpu, wave):
    meshes, M0 = rkdp_meshes(wave=wave)
    steps = 15
    bs = api.BatchSolver(api.BM_EVAL_RKDP54, meshes, M0)
    bs.run(steps)
    torch.cuda.synchronize()
    assert bs.state()["iteration"] == steps
    total_rej = 0
    for b, (n, h, p) in enumerate(meshes):
        s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0[b])
        s.run_to_iteration(steps)
        ms, st = bs.mesh_state(b), s.state()
        assert ms["iteration"] == st["iteration"] == steps
        assert ms["rejections"] == st["rejections"], (b, ms, st)
        assert abs(ms["time"] - st["time"]) <= 1e-5 * st["time"], (b, ms, st)
        e = rel_l2(bs.mesh(b).cpu().numpy(), s.M.cpu().numpy().astype(np.float64))
        assert e <= 1e-4, (b, n, e)
        total_rej += ms["rejections"]
        s.close()
    assert total_rej >= 3          # the controller had work to do
    bs.close()


def test_rkdp_batch_graph_equals_direct_and_split_run(gpu):
    meshes, M0 = rkdp_meshes(3)
    a = api.BatchSolver(api.BM_EVAL_RKDP54, meshes, M0, use_graph=True)
    b = api.BatchSolver(api.BM_EVAL_RKDP54, meshes, M0, use_graph=False)
    c = api.BatchSolver(api.BM_EVAL_RKDP54, meshes, M0, use_graph=True)
    a.run(12)
    b.run(12)
    c.run(5)
    c.run(7)
    torch.cuda.synchronize()
    assert torch.equal(a.M, b.M)
    assert torch.equal(a.M, c.M)
    for m in range(len(meshes)):
        assert a.mesh_state(m) == b.mesh_state(m) == c.mesh_state(m)


def test_rkdp_batch_matches_oracle(gpu):
    """From 0.3 rad cones (random states make the first rejections depend on lte values that
    FP32 rounding perturbs, see test_gpu_ode)."""
    meshes, M0 = rkdp_meshes(5, tilt=0.3)
    steps = 12
    bs = api.BatchSolver(api.BM_EVAL_RKDP54, meshes, M0)
    bs.set_astep(err_fail=5e-5)
    bs.run(steps)
    torch.cuda.synchronize()
    for b in (0, 3):
        n, h, p = meshes[b]
        Nz = 1
        while n[2] > 1 and Nz < 2 * n[2]:
            Nz *= 2
        o = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p, Nz_override=Nz if n[2] > 1 else 0)
        o.set_M(M0[b].astype(np.float64))
        o.set_astep(err=5e-5)
        o.run_to_iteration(steps)
        ms, so = bs.mesh_state(b), o.state()
        assert ms["rejections"] == so["rejections"]
        assert abs(ms["time"] - so["time"]) <= 1e-3 * so["time"]
        assert rel_l2(bs.mesh(b).cpu().numpy(), o.get_M()) <= 1e-4
