"""GPU parity of the field + LLG stage paths against the oracle.

Both evaluation paths are covered: the fused row kernel (x-C2R of H_demag + field + stage +
x-R2C of the stage's output M; the default) and the unfused one (planar H_demag written by
the convolution, then the stage kernel; BM_ROW_KERNEL=0 at solver creation).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu

SP4_N = (128, 32, 1)
SP4_H = (3.90625e-9, 3.90625e-9, 3e-9)


def tilted_x(ncell, Ms, seed, tilt=0.05):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, tilt, ncell)
    ph = rng.uniform(0, 2 * np.pi, ncell)
    m = np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], axis=1)
    return (m * Ms).astype(np.float32)


def sp4_params():
    return api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, Ha=(-24.6e-3 / api.MU0, 4.3e-3 / api.MU0, 0.0),
                        modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)


@pytest.fixture(params=["row", "unfused"])
def path(request, monkeypatch):
    monkeypatch.setenv("BM_ROW_KERNEL", "1" if request.param == "row" else "0")
    return request.param


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("modules", [
    api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN,
    api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN | api.BM_MOD_ANIS_UNI])
def test_effective_field_matches_oracle(gpu, path, modules):
    n, h = (32, 24, 4), (4e-9, 4e-9, 4e-9)
    p = api.FMParams(Ms=1e6, A=1.5e-11, alpha=0.1, K1=5e5, ea=(0.0, 0.0, 1.0), Ha=(1e3, -2e3, 39788.736), modules=modules)
    ncell = n[0] * n[1] * n[2]
    M0 = api.random_magnetisation(ncell, p.Ms, 7).astype(np.float32)
    s = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    s.keep_heff(True)
    s.iterate_stage()
    H = s.heff().cpu().numpy()
    o = O.OracleSim(api.BM_EVAL_RK4, n, h, p)
    o.set_M(M0.astype(np.float64))
    o.update_fields()
    assert rel_l2(H, o.heff()) <= 2e-6


@pytest.mark.parametrize("method", [api.BM_EVAL_RKDP54, api.BM_EVAL_RK4, api.BM_EVAL_TEULER, api.BM_EVAL_AHEUN])
def test_short_run_matches_oracle(gpu, path, method):
    n, h, p = SP4_N, SP4_H, sp4_params()
    ncell = n[0] * n[1] * n[2]
    M0 = tilted_x(ncell, p.Ms, 11)
    s = api.FMSolver(method, n, h, p, M0)
    o = O.OracleSim(method, n, h, p)
    o.set_M(M0.astype(np.float64))
    steps = 40
    s.run_to_iteration(steps)
    o.run_to_iteration(steps)
    st, so = s.state(), o.state()
    assert st["iteration"] == so["iteration"] == steps
    assert abs(st["time"] - so["time"]) <= 1e-3 * so["time"]
    mg, mo = s.average_m(), o.average_m()
    assert np.abs(mg - mo).max() <= 1e-4, (mg, mo)
    Mg = s.M.cpu().numpy()
    assert rel_l2(Mg, o.get_M()) <= 1e-4


def bench_like_params():
    return api.FMParams(Ms=1e6, A=1.5e-11, alpha=0.1, K1=5e5, ea=(0.0, 0.0, 1.0), Ha=(0.0, 0.0, 50e-3 / api.MU0),
                        modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN)


@pytest.mark.parametrize("method", [api.BM_EVAL_RKDP54, api.BM_EVAL_RK4, api.BM_EVAL_TEULER, api.BM_EVAL_AHEUN])
def test_3d_run_matches_oracle(gpu, path, method):
    """3D mesh (z pass + kernel multiply), 12 steps of all four methods from the BorisRand
    random state.  <m> within 1e-4 of the FP64 oracle, equal rejection counts.  Per cell:
    fixed-step methods within 1e-4; the adaptive ones within 5e-4, because their step sizes
    come from an lte that FP32 rounding moves (the per-cell difference with the step sizes
    pinned is checked by test_adaptive_dt_replayed_on_oracle)."""
    n, h, p = (64, 48, 6), (4e-9, 4e-9, 4e-9), bench_like_params()
    ncell = n[0] * n[1] * n[2]
    M0 = api.random_magnetisation(ncell, p.Ms, 5).astype(np.float32)
    s = api.FMSolver(method, n, h, p, M0)
    # the oracle with the same N.z (next power of two of 2 nz) as the GPU path
    o = O.OracleSim(method, n, h, p, Nz_override=16)
    o.set_M(M0.astype(np.float64))
    steps = 12
    s.run_to_iteration(steps)
    o.run_to_iteration(steps)
    st, so = s.state(), o.state()
    assert st["iteration"] == so["iteration"] == steps
    assert st["rejections"] == so["rejections"]
    assert abs(st["time"] - so["time"]) <= 1e-3 * so["time"]
    dm = float(np.abs(s.average_m() - o.average_m()).max())
    e = rel_l2(s.M.cpu().numpy(), o.get_M())
    print(f"method {method}: <m> {dm:.2e}, per-cell rel L2 {e:.2e}, dT {st['dT']:.6e} vs {so['dT']:.6e}")
    assert dm <= 1e-4
    assert e <= (5e-4 if method in (api.BM_EVAL_RKDP54, api.BM_EVAL_AHEUN) else 1e-4)


@pytest.mark.parametrize("method", [api.BM_EVAL_RKDP54, api.BM_EVAL_AHEUN])
def test_adaptive_dt_replayed_on_oracle(gpu, path, method):
    """The adaptive run of test_3d_run_matches_oracle, with the GPU's accepted step sizes
    replayed on the FP64 oracle as fixed steps: with the same step sizes the per-cell
    difference is the stages' FP32 rounding alone (rejected attempts restore M exactly, so
    the accepted path does not depend on them).  The accepted step sizes are the differences
    of the time reached at each iteration; a rejection can undo the previous step (RKDP54
    judges a step at the next step's first stage), so a later call overwrites that entry."""
    n, h, p = (64, 48, 6), (4e-9, 4e-9, 4e-9), bench_like_params()
    ncell = n[0] * n[1] * n[2]
    M0 = api.random_magnetisation(ncell, p.Ms, 5).astype(np.float32)
    s = api.FMSolver(method, n, h, p, M0)
    o = O.OracleSim(method, n, h, p, Nz_override=16)
    o.set_M(M0.astype(np.float64))
    t_at = {0: 0.0}
    calls = 0
    while s.state()["iteration"] < 12:
        calls += 1
        assert calls <= 100
        it0 = s.state()["iteration"]
        s.advance_time()
        st = s.state()
        assert st["iteration"] <= it0 + 1
        t_at = {k: v for k, v in t_at.items() if k < st["iteration"]}
        t_at[st["iteration"]] = st["time"]
    for k in range(1, 13):
        dT = t_at[k] - t_at[k - 1]
        o.set_dt(dT)
        o.set_astep(dtmin=dT, dtmax=dT)
        o.advance_time()
    assert s.state()["iteration"] == o.state()["iteration"] == 12
    assert abs(s.state()["time"] - o.state()["time"]) <= 1e-12 * o.state()["time"]
    e = rel_l2(s.M.cpu().numpy(), o.get_M())
    print(f"method {method}: per-cell rel L2 with replayed step sizes {e:.2e}")
    assert e <= 2e-5


def config2_small_oracle_run():
    """FP64 oracle reference for test_rkdp54_random_start_config2_params, computed once."""
    global _CFG2_CACHE
    try:
        return _CFG2_CACHE
    except NameError:
        pass
    n, h, p = (128, 128, 16), (4e-9, 4e-9, 4e-9), bench_like_params()
    M0 = api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 1).astype(np.float32)
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p, Nz_override=32)
    o.set_M(M0.astype(np.float64))
    marks = []
    for it in (25, 50, 75, 100):
        o.run_to_iteration(it)
        print(f"oracle at {it} steps", flush=True)
        marks.append((it, o.average_m(), o.state()))
    _CFG2_CACHE = (n, h, p, M0, marks)
    return _CFG2_CACHE


def test_rkdp54_random_start_config2_params(gpu, path):
    """The large-3D workload's parameters (Ms 1e6, A 1.5e-11, K1 5e5 along z, alpha 0.1,
    Hz 50 mT, 4 nm cubes) and BorisRand random start on a reduced 128x128x16 mesh: 100
    adaptive RKDP54 steps; <m> within 1e-4 of the FP64 oracle at every 25th step and equal
    accepted and rejected counts."""
    n, h, p, M0, marks = config2_small_oracle_run()
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
    worst = 0.0
    for it, m_o, st_o in marks:
        s.run_to_iteration(it)
        st = s.state()
        assert st["iteration"] == st_o["iteration"] == it
        assert st["rejections"] == st_o["rejections"]
        assert abs(st["time"] - st_o["time"]) <= 1e-3 * st_o["time"]
        worst = max(worst, float(np.abs(s.average_m() - m_o).max()))
    print(f"<m> worst {worst:.2e}, rejections {st['rejections']}, time {st['time']:.4e}")
    assert worst <= 1e-4
    s.close()


@pytest.mark.parametrize("wave", ["sinc", "ramp"])
@pytest.mark.parametrize("method", [api.BM_EVAL_RKDP54, api.BM_EVAL_RK4])
def test_time_dependent_field_matches_oracle(gpu, path, method, wave):
    """Zeeman waveforms evaluated at each step's start time (Boris/Zeeman.cpp:164-300): a
    20 mT sinc pulse of 50 ps centred at 0.1 ns, or a field ramp, on the small film."""
    n, h = SP4_N, SP4_H
    p = sp4_params()
    if wave == "sinc":
        p.wave, p.amp, p.w, p.t0, p.dir = 1, 20e-3 / api.MU0, np.pi / 50e-12, 0.1e-9, (0.6, 0.0, 0.8)
    else:
        p.wave, p.slope = 2, (2e14, -1e14, 5e13)
    ncell = n[0] * n[1] * n[2]
    M0 = tilted_x(ncell, p.Ms, 13)
    s = api.FMSolver(method, n, h, p, M0)
    o = O.OracleSim(method, n, h, p)
    o.set_M(M0.astype(np.float64))
    steps = 60
    s.run_to_iteration(steps)
    o.run_to_iteration(steps)
    assert abs(s.state()["time"] - o.state()["time"]) <= 1e-3 * o.state()["time"]
    assert np.abs(s.average_m() - o.average_m()).max() <= 1e-4
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-4


def test_rkdp54_rejections_match_oracle(gpu, path):
    """A 0.2 rad random cone about x under a 5e5 J/m^3 z anisotropy makes the default first
    steps too long: rejected steps (restore of M, re-primed FSAL) in both."""
    n, h, p = (64, 32, 4), (4e-9, 4e-9, 4e-9), bench_like_params()
    ncell = n[0] * n[1] * n[2]
    M0 = tilted_x(ncell, p.Ms, 9, tilt=0.2)
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p, Nz_override=8)
    o.set_M(M0.astype(np.float64))
    s.run_to_iteration(8)
    o.run_to_iteration(8)
    st, so = s.state(), o.state()
    assert so["rejections"] >= 2
    assert st["rejections"] == so["rejections"]
    assert abs(st["time"] - so["time"]) <= 1e-4 * so["time"]
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 2e-5


def test_caller_writes_m_between_steps(gpu, path):
    """The solver must pick up a magnetisation the caller wrote between two calls."""
    n, h, p = (64, 32, 4), (4e-9, 4e-9, 4e-9), bench_like_params()
    ncell = n[0] * n[1] * n[2]
    M0 = api.random_magnetisation(ncell, p.Ms, 3).astype(np.float32)
    M1 = api.random_magnetisation(ncell, p.Ms, 4).astype(np.float32)
    s = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    o = O.OracleSim(api.BM_EVAL_RK4, n, h, p, Nz_override=8)
    o.set_M(M0.astype(np.float64))
    s.advance_time()
    o.advance_time()
    s.M.copy_(torch.from_numpy(M1).to(s.M.device))
    o.set_M(M1.astype(np.float64))
    for _ in range(3):
        s.advance_time()
        o.advance_time()
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-5


@pytest.mark.slow
def test_sp4_2000_step_trajectory(gpu):
    """The small-film workload: 128x32x1 cells of 3.90625 x 3.90625 x 3 nm, Ms 8e5, A 1.3e-11,
    alpha 0.02, demag + exchange + Zeeman (-24.6, 4.3, 0) mT, m along +x with a seeded tilt
    <= 0.05 rad, 2000 RKDP54 steps at tolerance 1e-5: <m> within 1e-4 of the FP64 oracle at
    every 100th step."""
    n, h, p = SP4_N, SP4_H, sp4_params()
    ncell = n[0] * n[1] * n[2]
    M0 = tilted_x(ncell, p.Ms, 2026)
    s = api.FMSolver(api.BM_EVAL_RKDP54, n, h, p, M0)
    o = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p)
    o.set_M(M0.astype(np.float64))
    o32 = O.OracleSim(api.BM_EVAL_RKDP54, n, h, p, fp32=True)     # the same algorithm in FP32 on the CPU
    o32.set_M(M0.astype(np.float64))
    for x in (s, o, o32):
        x.set_astep(1e-5)
    worst = worst32 = 0.0
    for it in range(100, 2001, 100):
        for x in (s, o, o32):
            x.run_to_iteration(it)
        worst = max(worst, float(np.abs(s.average_m() - o.average_m()).max()))
        worst32 = max(worst32, float(np.abs(s.average_m() - o32.average_m()).max()))
    assert worst <= 1e-4, worst
    assert worst32 <= 1e-4, worst32
    assert abs(s.state()["time"] - o.state()["time"]) <= 1e-3 * o.state()["time"]
