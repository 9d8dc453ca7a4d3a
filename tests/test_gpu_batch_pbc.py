"""GPU tests of periodic meshes in a batch (bm_batch_set_pbc): every periodic mesh must follow
its plain single-mesh periodic run (bm_ode_set_pbc, itself checked against the oracle) with
RK4 and RKDP54, next to open meshes of the same sizes, whose passes then form separate
groups; one periodic mesh is checked against the FP64 oracle directly, and the setter's
state rules are enforced."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api
from test_gpu_batch import H, rel_l2

pytestmark = pytest.mark.gpu

# (n, pbc): 2D xy-periodic, 2D with z images, 3D x- and z-periodic, 3D y-periodic, and open
# meshes of the same sizes
CASES = [((32, 32, 1), (2, 2, 0)), ((32, 32, 1), None), ((48, 32, 1), (3, 0, 1)), ((16, 24, 4), (1, 0, 3)),
         ((16, 24, 4), None), ((24, 16, 2), (0, 2, 0)), ((24, 16, 2), None)]


def meshes_and_M0(seed=3):
    rng = np.random.default_rng(seed)
    meshes = []
    for n, _ in CASES:
        p = api.FMParams(Ms=float(rng.uniform(5e5, 1.1e6)), A=1.3e-11, alpha=float(rng.uniform(0.02, 0.3)),
                         K1=float(rng.uniform(0, 2e5)), ea=(0.0, 0.0, 1.0), Ha=(0.0, 1e4, 2e4),
                         modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)
        meshes.append((n, H, p))
    M0 = [api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 60 + b).astype(np.float32) for b, (n, _, p) in enumerate(meshes)]
    return meshes, M0


def pbcs():
    return [pb for _, pb in CASES]


def test_periodic_meshes_in_rk4_batch_equal_single_runs(gpu):
    meshes, M0 = meshes_and_M0()
    dt, steps = 40e-15, 20
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, pbc=pbcs())
    bs.set_dt(dt)
    bs.run(steps)
    torch.cuda.synchronize()
    for b, (n, h, p) in enumerate(meshes):
        s = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0[b], pbc=CASES[b][1] or (0, 0, 0))
        s.set_dt(dt)
        s.run_to_iteration(steps)
        e = rel_l2(bs.mesh(b).cpu().numpy(), s.M.cpu().numpy().astype(np.float64))
        print(f"{n} pbc {CASES[b][1]}: rel L2 {e:.2e}")
        assert e <= 2e-6, (b, n, e)
        s.close()
    bs.close()


def test_periodic_meshes_in_rkdp54_batch_equal_single_runs(gpu):
    meshes, M0 = meshes_and_M0(5)
    steps = 15
    bs = api.BatchSolver(api.BM_EVAL_RKDP54, meshes, M0, pbc=pbcs())
    bs.run(steps)
    torch.cuda.synchronize()
    for b, (n, h, p) in enumerate(meshes):
        s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0[b], pbc=CASES[b][1] or (0, 0, 0))
        s.run_to_iteration(steps)
        ms, st = bs.mesh_state(b), s.state()
        assert ms["iteration"] == st["iteration"] == steps
        assert ms["rejections"] == st["rejections"], (b, ms, st)
        assert abs(ms["time"] - st["time"]) <= 1e-5 * st["time"], (b, ms, st)
        e = rel_l2(bs.mesh(b).cpu().numpy(), s.M.cpu().numpy().astype(np.float64))
        assert e <= 1e-4, (b, n, e)
        s.close()
    bs.close()


def test_periodic_batch_mesh_matches_oracle(gpu):
    meshes, M0 = meshes_and_M0(7)
    dt, steps = 40e-15, 10
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, pbc=pbcs())
    bs.set_dt(dt)
    bs.run(steps)
    torch.cuda.synchronize()
    for b in (0, 3):
        n, h, p = meshes[b]
        o = O.OracleSim(api.BM_EVAL_RK4, n, h, p)
        o.set_pbc(CASES[b][1])
        o.set_M(M0[b].astype(np.float64))
        o.set_dt(dt)
        o.run_to_iteration(steps)
        e = rel_l2(bs.mesh(b).cpu().numpy(), o.get_M())
        print(f"{n} pbc {CASES[b][1]} vs oracle: rel L2 {e:.2e}")
        assert e <= 2e-6, (b, e)
    bs.close()


def test_periodic_groups_are_tuned_and_pinned(gpu):
    """Periodic passes appear as "_pbc" groups in the autotuner's report, and a report applied
    to a new batch reproduces the tuned run bit for bit."""
    meshes, M0 = meshes_and_M0(9)
    a = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, pbc=pbcs(), autotune="radix")
    a.set_dt(40e-15)
    a.run(4)
    rep = a.autotune_report()
    names = {l.split()[0] for l in rep}
    assert {"x_fwd_pbc", "x_inv_pbc", "y_fwd_pbc", "y_inv_pbc", "z_pbc", "y2d_pbc", "y2d", "z"} <= names, rep
    assert all(" radix=0 " in l for l in rep if l.split()[0].endswith("_pbc"))
    b = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0, pbc=pbcs())
    b.apply_report(rep)
    b.set_dt(40e-15)
    b.run(4)
    torch.cuda.synchronize()
    assert torch.equal(a.M, b.M)
    a.close(); b.close()


def test_set_pbc_state_rules(gpu):
    meshes, M0 = meshes_and_M0()
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes[:2], M0[:2])
    with pytest.raises(RuntimeError, match="error 1"):
        api._lib.check(bs.L.bm_batch_set_pbc(bs.handle, 2, 1, 0, 0))   # no mesh 2
    bs.set_schedule("y2d", 64, per_cu=4)
    api._lib.check(bs.L.bm_batch_set_pbc(bs.handle, 0, 2, 2, 0))
    with pytest.raises(RuntimeError, match="error 1"):
        bs.set_schedule("y2d_pbc", 32, radix=1)                         # periodic: default schedule
    bs.set_schedule("y2d_pbc", 32, per_cu=4)
    bs.set_dt(40e-15)
    bs.run(1)
    with pytest.raises(RuntimeError, match="error 5"):
        api._lib.check(bs.L.bm_batch_set_pbc(bs.handle, 1, 1, 0, 0))   # after the first run
    bs.close()
