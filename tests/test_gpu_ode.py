"""GPU parity of the fused field + LLG stage path against the oracle."""
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


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("modules", [
    api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN,
    api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN | api.BM_MOD_ANIS_UNI])
def test_effective_field_matches_oracle(gpu, modules):
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
def test_short_run_matches_oracle(gpu, method):
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
