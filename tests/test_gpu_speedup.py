"""GPU parity of the demag evaluation speedup (saved demag fields, Lagrange extrapolation in
time, self term subtracted on save and added back; DemagCUDA::UpdateField,
Boris/DemagCUDA.cpp:106-468) against the FP64 oracle with the same speedup mode.

Bars: <m> within 1e-4 and M within 1e-4 relative L2; the same number of convolutions, saved
fields and (to the step-time tolerance) save times as the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from boris2_amd import api
from test_gpu_ode import SP4_H, SP4_N, bench_like_params, rel_l2, sp4_params, tilted_x

pytestmark = pytest.mark.gpu


def _pair(method, n, h, p, M0, mode, dTs=0.0, Nz=0, dt=None, fixed=False):
    s = api.FMSolver(method, n, h, p, M0)
    o = O.OracleSim(method, n, h, p, Nz_override=Nz)
    o.set_M(M0.astype(np.float64))
    if dt is not None:
        s.set_dt(dt)
        o.set_dt(dt)
    if fixed:
        s.set_astep(dT_min=dt, dT_max=dt)
        o.set_astep(dtmin=dt, dtmax=dt)
    s.set_eval_speedup(mode, dTs)
    o.set_speedup(mode, dTs)
    return s, o


def _compare(s, o, steps, time_tol=1e-3):
    s.run_to_iteration(steps)
    o.run_to_iteration(steps)
    st, so = s.state(), o.state()
    assert st["iteration"] == so["iteration"] == steps
    assert st["rejections"] == so["rejections"]
    assert abs(st["time"] - so["time"]) <= time_tol * so["time"]
    sg, sc = s.speedup_state(), o.speedup_info()
    assert sg["num_saved"] == sc["num_saved"]
    assert sg["convolutions"] == sc["convolutions"], (sg, sc)
    np.testing.assert_allclose(sg["tdm"], sc["tdm"], rtol=time_tol, atol=1e-18)
    mg, mo = s.average_m(), o.average_m()
    assert np.abs(mg - mo).max() <= 1e-4, (mg, mo)
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-4


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("method", [api.BM_EVAL_RKDP54, api.BM_EVAL_RK4])
def test_speedup_2d_film_matches_oracle(gpu, method, mode):
    n, h, p = SP4_N, SP4_H, sp4_params()
    M0 = tilted_x(n[0] * n[1] * n[2], p.Ms, 11)
    s, o = _pair(method, n, h, p, M0, mode)
    _compare(s, o, 40)


@pytest.mark.parametrize("mode", [1, 3, 6])
@pytest.mark.parametrize("method", [api.BM_EVAL_RKDP54, api.BM_EVAL_RK4, api.BM_EVAL_TEULER, api.BM_EVAL_AHEUN])
def test_speedup_3d_matches_oracle(gpu, method, mode):
    """3D mesh with anisotropy; the adaptive methods from a 0.3 rad cone (see test_gpu_ode)."""
    n, h, p = (64, 48, 6), (4e-9, 4e-9, 4e-9), bench_like_params()
    ncell = n[0] * n[1] * n[2]
    if method in (api.BM_EVAL_RKDP54, api.BM_EVAL_AHEUN):
        M0 = tilted_x(ncell, p.Ms, 5, tilt=0.3)
    else:
        M0 = api.random_magnetisation(ncell, p.Ms, 5).astype(np.float32)
    s, o = _pair(method, n, h, p, M0, mode, Nz=16)
    _compare(s, o, 12)


def test_speedup_unlinked_time_step_matches_oracle(gpu):
    """dTspeedup independent of dT: a convolution whenever 2.5 dT of simulated time passed."""
    n, h, p = (64, 32, 4), (4e-9, 4e-9, 4e-9), bench_like_params()
    M0 = api.random_magnetisation(n[0] * n[1] * n[2], p.Ms, 8).astype(np.float32)
    dt = 1e-13
    s, o = _pair(api.BM_EVAL_RK4, n, h, p, M0, 3, dTs=2.5 * dt, Nz=8, dt=dt)
    _compare(s, o, 30, time_tol=1e-9)
    # every evaluation convolves until 3 fields are saved (about 3 x 2.5 dT = 30 evaluations),
    # then one per 2.5 dT: well under the 120 evaluations of 30 RK4 steps
    assert s.speedup_state()["convolutions"] < 60


def test_speedup_rkdp54_rejections_match_oracle(gpu):
    """Rejected steps leave the rejected state's field among the saved ones (time ahead of the
    restored step), as the reference does; the extrapolation must follow."""
    n, h, p = (64, 32, 4), (4e-9, 4e-9, 4e-9), bench_like_params()
    M0 = tilted_x(n[0] * n[1] * n[2], p.Ms, 9, tilt=0.2)
    s, o = _pair(api.BM_EVAL_RKDP54, n, h, p, M0, 3, Nz=8)
    s.run_to_iteration(12)
    o.run_to_iteration(12)
    assert o.state()["rejections"] >= 2
    assert s.state()["rejections"] == o.state()["rejections"]
    assert s.speedup_state()["convolutions"] == o.speedup_info()["convolutions"]
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-4


def test_speedup_off_again_equals_plain(gpu):
    """Mode 0 after a speedup run restores the plain path (fused row kernel) exactly."""
    n, h, p = SP4_N, SP4_H, sp4_params()
    M0 = tilted_x(n[0] * n[1] * n[2], p.Ms, 4)
    a = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    b = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    b.set_eval_speedup(2)
    b.set_eval_speedup(0)
    a.run_to_iteration(10)
    b.run_to_iteration(10)
    assert np.array_equal(a.M.cpu().numpy(), b.M.cpu().numpy())


@pytest.mark.parametrize("mode", [1, 3])
@pytest.mark.parametrize("method", [api.BM_EVAL_RK4, api.BM_EVAL_RKDP54])
def test_speedup_layer_stack_matches_oracle(gpu, method, mode):
    """A 3-layer stack (2/1/4 cells of 4 nm with 1-cell spacers, per-layer Ms, alpha, K1) with the
    speedup: SDemag's multilayered branch, each layer with its own self coefficient."""
    from test_gpu_layers import A, ALPHA, HZ, K1, MODS, MS, Z0, layer_avg, start
    nx, ny = 32, 24
    nl = nx * ny
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, ea=(0.0, 0.0, 1.0), Ha=(-2.0e4, 4.0e3, 0.0), modules=MODS)
    M0 = start(nl, seed=11)
    s = api.LayerStackSolver(method, nx, ny, 4e-9, 4e-9, HZ, Z0, p, Ms=MS, A=A, alpha=ALPHA, K1=K1, M0=M0.astype(np.float32))
    q = api.FMParams(**{**p.__dict__, "modules": MODS & ~api.BM_MOD_DEMAG})
    o = O.OracleSim(method, (nx, ny, 3), (4e-9, 4e-9, 4e-9), q)
    o.set_layers(4e-9, 4e-9, HZ, Z0, with_demag=True, Ms=MS, A=A, alpha=ALPHA, K1=K1)
    o.set_M(M0)
    s.set_eval_speedup(mode)
    o.set_speedup(mode)
    steps = 40
    s.run_to_iteration(steps)
    o.run_to_iteration(steps)
    assert s.speedup_state()["convolutions"] == o.speedup_info()["convolutions"]
    assert s.state()["rejections"] == o.state()["rejections"]
    Mg = s.M.cpu().numpy().astype(np.float64)
    assert np.abs(layer_avg(Mg, nl) - layer_avg(o.get_M(), nl)).max() <= 1e-4
    assert rel_l2(Mg, o.get_M()) <= 1e-4
