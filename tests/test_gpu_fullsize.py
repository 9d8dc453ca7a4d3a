"""GPU checks at the benchmark's full sizes, through properties that do not depend on size
and one direct FP64 oracle comparison the host finishes in seconds.

* 512x512x32 cells of 4 nm (the plain-method benchmark mesh):
  - demag of a uniformly magnetised prism: the cell-averaged field equals Aharoni's
    closed-form demag factors (exact for the Newell tensor, whatever the mesh size);
  - the demag operator is linear and symmetric (<M1, N M2> = <M2, N M1> for equal cells);
  - the demag field of the BorisRand random state against the FP64 oracle (1e-6);
  - one RKDP54 step from the BorisRand random state against the FP64 oracle;
  - the fused row-kernel path and the unfused path follow the same 10-step trajectory.
* the 64 films of uneven sizes (the batched workload): batched RK4 against each film run
  on its own.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api
from test_oracle_demag import aharoni_Dz

pytestmark = pytest.mark.gpu

N_BIG, H_BIG = (512, 512, 32), (4e-9, 4e-9, 4e-9)


def big_params():
    return api.FMParams(Ms=1e6, A=1.5e-11, alpha=0.1, K1=5e5, ea=(0.0, 0.0, 1.0), Ha=(0.0, 0.0, 50e-3 / api.MU0),
                        modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)


def big_m0(seed=1):
    ncell = N_BIG[0] * N_BIG[1] * N_BIG[2]
    return api.random_magnetisation(ncell, 1e6, seed).astype(np.float32)


def test_fullsize_demag_uniform_prism_and_symmetry(gpu):
    n, h = N_BIG, H_BIG
    ncell = n[0] * n[1] * n[2]
    d = api.Demag(n, h)
    a, b, c = n[0] * h[0] / 2, n[1] * h[1] / 2, n[2] * h[2] / 2
    Nf = {0: aharoni_Dz(b, c, a), 1: aharoni_Dz(c, a, b), 2: aharoni_Dz(a, b, c)}
    for comp in range(3):
        M = torch.zeros((ncell, 3), dtype=torch.float32, device="cuda")
        M[:, comp] = 1.0
        H = d.convolute(M)
        mean = H[:, comp].double().mean().item()
        assert abs(mean + Nf[comp]) <= 2e-6, (comp, mean, Nf[comp])
        # the other two components average to zero
        others = [H[:, q].double().mean().item() for q in range(3) if q != comp]
        assert max(abs(v) for v in others) <= 2e-6
    g = torch.Generator(device="cuda").manual_seed(5)
    M1 = torch.randn((ncell, 3), generator=g, device="cuda")
    M2 = torch.randn((ncell, 3), generator=g, device="cuda")
    H1, H2 = d.convolute(M1), d.convolute(M2)
    s12 = (M1.double() * H2.double()).sum().item()
    s21 = (M2.double() * H1.double()).sum().item()
    s11 = (M1.double() * H1.double()).sum().item()
    # white-noise states: the cross terms are ~sqrt(n) |N| while FP32 rounding moves them by
    # ~1e-4 of that; a wrong sign or transposed off-diagonal kernel term shows at percent level
    assert abs(s12 - s21) <= 1e-4 * abs(s12) + 1e-2, (s12, s21)
    assert s11 < 0                              # demag energy of any state is positive
    H3 = d.convolute(2.5 * M1 - 0.75 * M2)
    ref = 2.5 * H1.double() - 0.75 * H2.double()
    assert ((H3.double() - ref).norm() / ref.norm()).item() <= 1e-6


def test_fullsize_demag_matches_fp64_oracle(gpu):
    """The demag field of the 512x512x32 mesh in the BorisRand random state against the FP64
    oracle's convolution of the same M: relative L2 <= 1e-6."""
    n, h = N_BIG, H_BIG
    M0 = big_m0()
    d = api.Demag(n, h)
    Hg = d.convolute(torch.from_numpy(M0).cuda()).cpu().numpy().astype(np.float64)
    d.close()
    torch.cuda.empty_cache()
    ref = O.OracleDemag(n, h, Nz_override=64)(M0.astype(np.float64))
    e = float(np.linalg.norm(Hg - ref) / np.linalg.norm(ref))
    print(f"512x512x32 demag rel L2 vs FP64 oracle: {e:.2e}")
    assert e <= 1e-6, e


def test_fullsize_rkdp54_first_step_matches_oracle(gpu):
    """One accepted RKDP54 step (six field evaluations) of the benchmark workload against the
    FP64 oracle run on the host's cores."""
    n, h, p = N_BIG, H_BIG, big_params()
    M0 = big_m0()
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
    s.run_to_iteration(1)
    Mg = s.M.cpu().numpy().astype(np.float64)
    s.close()
    torch.cuda.empty_cache()
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p, Nz_override=64)
    o.set_M(M0.astype(np.float64))
    o.run_to_iteration(1)
    Mo = o.get_M()
    assert o.state()["iteration"] == 1
    e = float(np.linalg.norm(Mg - Mo) / np.linalg.norm(Mo))
    assert e <= 1e-5, e
    m = Mg.reshape(-1, 3)
    assert np.abs(np.linalg.norm(m, axis=1) - 1e6).max() <= 1.0     # renormalised to Ms


def test_fullsize_row_kernel_path_equals_unfused(gpu, monkeypatch):
    n, h, p = N_BIG, H_BIG, big_params()
    M0 = big_m0(2)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("BM_ROW_KERNEL", flag)
        s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
        s.run_to_iteration(10)
        out[flag] = (s.M.cpu().numpy().astype(np.float64), s.state())
        s.close()
        torch.cuda.empty_cache()
    (Ma, sa), (Mb, sb) = out["1"], out["0"]
    assert sa["rejections"] == sb["rejections"]
    assert abs(sa["time"] - sb["time"]) <= 1e-5 * sb["time"]
    assert float(np.linalg.norm(Ma - Mb) / np.linalg.norm(Mb)) <= 2e-5


def test_fullsize_batched_films_equal_single_runs(gpu):
    """The 64 films (nx, ny from {64..256}, nz from {1, 2, 4, 8}, seeded Ms and alpha) of the
    batched workload: 5 RK4 steps at 50 fs, batched vs each film on its own."""
    from bench import batched_meshes        # the repository root is on sys.path (conftest)
    meshes = batched_meshes()
    bs = api.BatchSolver(api.BM_EVAL_RK4, [(n, hh, pp) for n, hh, pp, _ in meshes], [M0 for *_, M0 in meshes])
    bs.set_dt(50e-15)
    bs.run(5)
    torch.cuda.synchronize()
    worst = 0.0
    for b, (n, hh, pp, M0) in enumerate(meshes):
        s = api.FMSolver(api.BM_EVAL_RK4, n, hh, pp, M0)
        s.set_dt(50e-15)
        s.run_to_iteration(5)
        ref = s.M.cpu().numpy().astype(np.float64)
        worst = max(worst, float(np.linalg.norm(bs.mesh(b).cpu().numpy() - ref) / np.linalg.norm(ref)))
        s.close()
    bs.close()
    assert worst <= 2e-6, worst
