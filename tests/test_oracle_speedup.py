"""Pins the oracle's demag evaluation speedup (DemagCUDA::UpdateField with saved demag fields,
Boris/DemagCUDA.cpp:106-468) by properties the mathematics fixes: Lagrange extrapolation is
exact on polynomials of degree p-1, the slot rule keeps the p latest saves, a one-cell mesh
(whose demag field is the self term alone) integrates the same with and without the
speedup, and the speedup's deviation from the plain integration shrinks with the
extrapolation order p as dT shrinks."""
import numpy as np
import pytest

import oracle_lib as O
from boris2_amd.api import (BM_EVAL_RK4, BM_EVAL_RKDP54, BM_MOD_DEMAG, BM_MOD_EXCHANGE, BM_MOD_ZEEMAN, FMParams,
                            random_magnetisation)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_lagrange_coefficients_exact_on_polynomials(p):
    rng = np.random.default_rng(p)
    for trial in range(20):
        tdm = np.sort(rng.uniform(0, 1e-11, p))
        rng.shuffle(tdm)                         # slot order is cyclic, not sorted
        t = tdm.max() + rng.uniform(0, 3e-12)    # forward extrapolation, as the stages use
        a = O.speedup_lagrange(p, tdm, t)
        assert np.all(a[p:] == 0)
        coef = rng.normal(size=p)
        scale = 1e-11
        P = lambda x: np.polyval(coef, x / scale)
        got = float(np.dot(a[:p], P(tdm)))
        assert abs(got - P(t)) <= 1e-8 * max(1.0, np.abs(coef).sum() * 10 ** p), (p, trial, got, P(t))
        # degree p: not reproduced (the coefficients are of order p-1, not higher)
        if p < 6:
            Q = lambda x: (x / scale) ** p
            assert abs(float(np.dot(a[:p], Q(tdm))) - Q(t)) > 1e-6 * abs(np.prod(t - tdm) / scale ** p)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_slot_rule_keeps_latest_saves(p):
    tdm = list(np.arange(1, p + 1, dtype=float))   # fill phase: slots 0..p-1 in time order
    t = float(p)
    for k in range(40):
        t += 1.0
        s = O.speedup_next_slot(p, tdm)
        if p > 1:
            tdm[s] = t
            assert sorted(tdm) == [t - q for q in range(p - 1, -1, -1)], (p, k, tdm)
        else:
            assert s == 0


def _film(method, n, h, params, m0, speedup, dt, steps, fixed=True):
    s = O.OracleSim(method, n, h, params)
    s.set_M(m0)
    s.set_dt(dt)
    if fixed and method == BM_EVAL_RKDP54:
        s.set_astep(dtmin=dt, dtmax=dt)
    if speedup:
        s.set_speedup(speedup)
    s.run_to_iteration(steps)
    return s


def test_one_cell_speedup_equals_plain():
    """One cubic cell: the convolution is the self term alone (-1/3 on the diagonal, which the
    self coefficient also is), so the saved fields are ~0 and every extrapolated evaluation
    must add back exactly the self term of the current M.

    The reference computes the self coefficient with the cell size in metres
    (Boris/DemagCUDA.cpp:37), where the zero guards of the Newell f function drop terms;
    for non-cubic cells it then differs from the tensor's zero-offset value (pinned against
    the reference build in test_oracle_tensor.py).  A cube is exact either way."""
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, Ha=(-19576.058, 3421.831, 0.0),
                 modules=BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ZEEMAN)
    h = (4e-9, 4e-9, 4e-9)
    m0 = np.array([[0.6, 0.7, np.sqrt(1 - 0.85)]]) * p.Ms
    for method in (BM_EVAL_RK4, BM_EVAL_RKDP54):
        ref = _film(method, (1, 1, 1), h, p, m0, 0, 2e-13, 100, fixed=False)
        for mode in (1, 3, 6):
            s = _film(method, (1, 1, 1), h, p, m0, mode, 2e-13, 100, fixed=False)
            np.testing.assert_allclose(s.get_M(), ref.get_M(), rtol=0, atol=1e-9 * p.Ms)
            assert s.speedup_info()["convolutions"] < ref.speedup_info()["convolutions"]


def _deviation(method, mode, dt, t_end):
    n, h = (16, 8, 2), (4e-9, 4e-9, 4e-9)
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.1, Ha=(2e4, 1e4, 0.0), modules=BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ZEEMAN)
    m0 = random_magnetisation(n[0] * n[1] * n[2], p.Ms, seed=5)
    steps = int(round(t_end / dt))
    a = _film(method, n, h, p, m0, 0, dt, steps).get_M()
    b = _film(method, n, h, p, m0, mode, dt, steps).get_M()
    return np.linalg.norm(a - b) / np.linalg.norm(a)


@pytest.mark.parametrize("method", [BM_EVAL_RK4, BM_EVAL_RKDP54])
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_speedup_deviation_order(method, mode):
    """The extrapolated demag field is off by O(dT^p) at every stage, so the integrated state
    departs from the plain one by O(dT^p): halving dT divides the gap by about 2^p."""
    t_end = 6.4e-12
    d1 = _deviation(method, mode, 1e-13, t_end)
    d2 = _deviation(method, mode, 0.5e-13, t_end)
    rate = np.log2(d1 / d2)
    assert d1 > 0 and abs(rate - mode) < 0.45, (d1, d2, rate)


def test_speedup_convolution_count_and_saved_times():
    """Linked speedup step: a convolution at the first evaluation of every step, all
    evaluations convolve until p fields are saved; saved times are the latest step starts."""
    n, h = (8, 8, 1), (4e-9, 4e-9, 3e-9)
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.1, Ha=(2e4, 0.0, 0.0), modules=BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ZEEMAN)
    m0 = random_magnetisation(64, p.Ms, seed=2)
    dt, steps = 1e-13, 20
    for method, nst in ((BM_EVAL_RK4, 4), (BM_EVAL_RKDP54, 6)):
        for mode in (1, 3, 6):
            s = _film(method, n, h, p, m0, mode, dt, steps)
            info = s.speedup_info()
            assert info["num_saved"] == mode
            # steps 1..p-1 convolve at every stage, step p saves its last field and extrapolates
            assert info["convolutions"] == nst * (mode - 1) + (steps - mode + 1), (method, mode, info)
            if mode > 1:
                # the saves are the first evaluations of the last p steps, at their start times
                np.testing.assert_allclose(np.sort(info["tdm"][:mode]), [(steps - 1 - q) * dt for q in range(mode - 1, -1, -1)],
                                           rtol=1e-9)


def _stack(speedup, method=BM_EVAL_RK4, dt=1e-13, steps=0, modules=None, seed=3):
    """Three 2D layers of 12x10 cells, thicknesses 2/1/4 nm with 1 nm gaps, as a layer stack."""
    n, hx, hy = (12, 10, 3), 4e-9, 4e-9
    hz, z0 = [2e-9, 1e-9, 4e-9], [0.0, 3e-9, 5e-9]
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.1, Ha=(2e4, 1e4, 0.0),
                 modules=(BM_MOD_EXCHANGE | BM_MOD_ZEEMAN) if modules is None else modules)
    s = O.OracleSim(method, n, (hx, hy, max(hz)), p)
    s.set_layers(hx, hy, hz, z0, with_demag=True, Ms=[8e5, 6e5, 1.0e6])
    ms = np.repeat([8e5, 6e5, 1.0e6], n[0] * n[1])
    M0 = random_magnetisation(n[0] * n[1] * n[2], 1.0, seed) * ms[:, None]
    s.set_M(M0)
    s.set_dt(dt)
    if method == BM_EVAL_RKDP54:
        s.set_astep(dtmin=dt, dtmax=dt)
    if speedup:
        s.set_speedup(speedup)
    if steps:
        s.run_to_iteration(steps)
    return s, M0, (n, hx, hy, hz, z0)


def test_stack_step_mode_uses_each_layers_self_term():
    """STEP mode on a layer stack: the second evaluation's demag field must be the first
    convolution minus each layer's own self term of the first M, plus that layer's self term
    of the current M (SDemagCUDA: per layer mesh selfDemagCoeff from its own cell size)."""
    s, M0, (n, hx, hy, hz, z0) = _stack(1, modules=0)
    s.update_fields()
    s.iterate_stage()
    M1 = s.get_M()
    s.update_fields()
    H = s.heff()
    conv = O.OracleMLayer(n[0], n[1], hx, hy, hz, z0)(M0)
    nl = n[0] * n[1]
    c = np.zeros((n[0] * n[1] * n[2], 3))
    for l in range(3):
        sd = np.zeros(3)
        O.lib().or_self_demag(hx, hy, hz[l], sd)
        c[l * nl:(l + 1) * nl] = sd
    want = conv - c * M0 + c * M1
    np.testing.assert_allclose(H, want, rtol=0, atol=1e-9 * np.abs(want).max())
    # the layers' coefficients differ (1, 2 and 4 nm cells), so one shared value would fail
    assert np.ptp(c[:, 2]) > 0.1


@pytest.mark.parametrize("mode", [2, 3])
def test_stack_speedup_deviation_order(mode):
    t_end = 6.4e-12
    d = []
    for dt in (1e-13, 0.5e-13):
        steps = int(round(t_end / dt))
        a = _stack(0, dt=dt, steps=steps)[0].get_M()
        b = _stack(mode, dt=dt, steps=steps)[0].get_M()
        d.append(np.linalg.norm(a - b) / np.linalg.norm(a))
    rate = np.log2(d[0] / d[1])
    assert abs(rate - mode) < 0.45, (d, rate)
