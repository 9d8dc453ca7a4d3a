"""GPU tests of the fused thin-mesh pass (the y, z, y passes of a mesh with N.z = 4 or 8 and a
power-of-two N.y as one kernel that never writes the full spectra; by default from 512 x 512
cells per plane, here forced at every size with BM_YZY=2): the demag field against the FP64
oracle and against the separate passes (BM_YZY=0), the add and energy output modes, an
RKDP54 thin-film trajectory against the oracle, and batches whose thin meshes take it."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu

H = (4e-9, 4e-9, 4e-9)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def demag(n, monkeypatch, fused):
    # 2: the fused pass at any size (by default it starts at 512 x 512 cells per plane)
    monkeypatch.setenv("BM_YZY", "2" if fused else "0")
    return api.Demag(n, H)


# every N.y the pass serves (16 .. 2048) with N.z 4 and 8; odd and ragged in-plane sizes
SIZES = [(24, 8, 2), (16, 16, 3), (33, 32, 2), (64, 64, 4), (20, 100, 2), (48, 128, 3), (10, 256, 4), (12, 300, 2),
         (6, 1000, 2), (8, 512, 4)]


@pytest.mark.parametrize("n", SIZES, ids=lambda n: "x".join(map(str, n)))
def test_fused_thin_pass_matches_oracle_and_separate_passes(gpu, monkeypatch, n):
    M = api.random_magnetisation(n[0] * n[1] * n[2], 8e5, 3 + n[1]).astype(np.float32)
    Mt = torch.from_numpy(M).cuda()
    d1 = demag(n, monkeypatch, True)
    N, _ = d1.info()
    assert N[2] in (4, 8)
    Hf = d1.convolute(Mt).cpu().numpy()
    d0 = demag(n, monkeypatch, False)
    Hs = d0.convolute(Mt).cpu().numpy()
    ref = O.OracleDemag(n, H, Nz_override=N[2])(M.astype(np.float64))
    ef, es = rel_l2(Hf, ref), rel_l2(Hf, Hs.astype(np.float64))
    print(f"{n} N {N}: fused vs oracle {ef:.2e}, vs separate passes {es:.2e}")
    assert ef <= 1e-6 and es <= 1e-6
    d1.close(); d0.close()


def test_fused_thin_pass_add_and_energy_modes(gpu, monkeypatch):
    n = (64, 128, 2)
    M = api.random_magnetisation(n[0] * n[1] * n[2], 8e5, 5).astype(np.float32)
    Mt = torch.from_numpy(M).cuda()
    base = torch.from_numpy(np.random.default_rng(1).normal(size=M.shape).astype(np.float32) * 1e4).cuda()
    out = {}
    for fused in (True, False):
        d = demag(n, monkeypatch, fused)
        H = base.clone()
        e = torch.zeros(1, dtype=torch.float32, device="cuda")
        d.convolute(Mt, H, clear_out=False, energy=e)
        out[fused] = (H.cpu().numpy().astype(np.float64), float(e.item()))
        d.close()
    assert rel_l2(out[True][0] - base.cpu().numpy(), out[False][0] - base.cpu().numpy()) <= 1e-6
    assert abs(out[True][1] - out[False][1]) <= 1e-6 * abs(out[False][1])
    assert out[True][1] != 0.0


def test_thin_film_rkdp54_matches_oracle(gpu, monkeypatch):
    monkeypatch.setenv("BM_YZY", "2")
    n = (64, 32, 2)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, K1=3e5, ea=(0.0, 0.0, 1.0), Ha=(2e4, 0.0, 0.0),
                     modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)
    M0 = api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 9).astype(np.float32)
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, H, p, M0)
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, H, p, Nz_override=4)
    o.set_M(M0.astype(np.float64))
    s.run_to_iteration(30)
    o.run_to_iteration(30)
    st, so = s.state(), o.state()
    assert st["iteration"] == so["iteration"] == 30
    assert st["rejections"] == so["rejections"]
    dm = np.abs(s.average_m() - o.average_m()).max()
    print(f"thin film RKDP54: <m> {dm:.2e}, rejections {st['rejections']}")
    assert dm <= 1e-4
    s.close()


def test_batched_thin_meshes_take_the_fused_pass(gpu, monkeypatch):
    monkeypatch.setenv("BM_YZY", "2")
    # N.y 64 / 128 / 256 with N.z 4 and 8 take the fused groups; N.y 48 (radix 3) the separate passes
    sizes = [(64, 32, 2), (48, 64, 4), (32, 128, 3), (48, 24, 2), (64, 32, 2)]
    rng = np.random.default_rng(4)
    meshes = [(n, H, api.FMParams(Ms=float(rng.uniform(5e5, 1e6)), A=1.3e-11, alpha=0.1,
                                  modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE)) for n in sizes]
    M0 = [api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 20 + b).astype(np.float32) for b, (n, _, p) in enumerate(meshes)]
    for method, steps in ((api.BM_EVAL_RK4, 10), (api.BM_EVAL_RKDP54, 12)):
        bs = api.BatchSolver(method, meshes, M0, autotune=True)
        if method == api.BM_EVAL_RK4:
            bs.set_dt(50e-15)
        bs.run(steps)
        torch.cuda.synchronize()
        if method == api.BM_EVAL_RK4:
            rep = bs.autotune_report()
            kinds = {(l.split()[0], int(l.split()[1])) for l in rep}
            assert {("yzy4", 64), ("yzy8", 128), ("yzy8", 256)} <= kinds, rep
            assert ("y_fwd", 48) in kinds and ("z", 4) in kinds, rep
        for b, (n, h, p) in enumerate(meshes):
            s = api.FMSolver(method, n, h, p, M0[b])
            if method == api.BM_EVAL_RK4:
                s.set_dt(50e-15)
            s.run_to_iteration(steps)
            torch.cuda.synchronize()
            if method == api.BM_EVAL_RKDP54:
                ms, st = bs.mesh_state(b), s.state()
                assert ms["rejections"] == st["rejections"] and abs(ms["time"] - st["time"]) <= 1e-5 * st["time"], (b, ms, st)
            e = rel_l2(bs.mesh(b).cpu().numpy(), s.M.cpu().numpy().astype(np.float64))
            assert e <= 2e-6, (method, b, n, e)
            s.close()
        bs.close()
    # the fused groups can be pinned like the others
    bs = api.BatchSolver(api.BM_EVAL_RK4, meshes, M0)
    bs.set_schedule("yzy4", 64, per_cu=4)
    with pytest.raises(RuntimeError, match="error 1"):
        bs.set_schedule("yzy4", 64, radix=1)
    bs.close()
