"""Pins the oracle's AHeun stepper (error estimate, order-2 step control, defaults) and its
time-dependent Zeeman field (sinc pulse and ramp, held at the step's start time) against
closed forms and an independent numpy integrator written here.

Reference: Boris/DiffEqFM_Evals_AHeun.cpp:13-178 (predictor, corrector, lte = |M_corr -
M_pred| / |M_corr|), Boris/DiffEq_CommonBase_Control.cpp:32-41 with Boris/DiffEq_Defs.h:17-25
(AHeun defaults, eval_method_order = 2), Boris/DiffEq_CommonBase_Iterate.cpp:15-59
(SetAdaptiveTimeStep), Boris/Zeeman.cpp:270/397 (field equation at GetStageTime(), which
advances by dT only at the end of a step), BorisLib/TEquation_Function.h:247 (F_sinc:
sin(x)/x, 1 at x = 0).
"""
import numpy as np
import pytest

import oracle_lib as O
from boris2_amd.api import BM_EVAL_AHEUN, BM_EVAL_RK4, BM_EVAL_TEULER, BM_MOD_ZEEMAN, GAMMA, FMParams


# ---------------------------------------------------------------------------------------
# Independent single-cell LLG (numpy, written from the equation, not from the oracle):
# dM/dt = -gamma/(1+a^2) [M x H + (a/Ms) M x (M x H)]
def llg(M, H, Ms, alpha):
    mxh = np.cross(M, H)
    return -GAMMA / (1 + alpha ** 2) * (mxh + alpha / Ms * np.cross(M, mxh))


def field_at(p, t):
    """Applied field of FMParams p at time t: np.sinc is numpy's normalised sinc
    sin(pi x)/(pi x), so sin(x)/x = np.sinc(x / pi) (1 at 0) — an independent routine."""
    H = np.array(p.Ha, float)
    if p.wave == 1:
        H = H + p.amp * np.array(p.dir, float) * np.sinc(p.w * (t - p.t0) / np.pi)
    elif p.wave == 2:
        H = H + np.array(p.slope, float) * t
    return H


def rk4_macrospin(p, M0, dt, nsteps, substep_times=False):
    """Classical RK4 with renormalisation; H at the step start time (the reference), or at
    t + c_i dt (substep_times=True, what the reference does NOT do)."""
    M, t = np.array(M0, float), 0.0
    for _ in range(nsteps):
        Hs = [field_at(p, t + (c * dt if substep_times else 0.0)) for c in (0.0, 0.5, 0.5, 1.0)]
        k1 = llg(M, Hs[0], p.Ms, p.alpha)
        k2 = llg(M + dt / 2 * k1, Hs[1], p.Ms, p.alpha)
        k3 = llg(M + dt / 2 * k2, Hs[2], p.Ms, p.alpha)
        k4 = llg(M + dt * k3, Hs[3], p.Ms, p.alpha)
        M = M + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        M = M * p.Ms / np.linalg.norm(M)
        t += dt
    return M


def one_cell(method, p, M0):
    s = O.OracleSim(method, (1, 1, 1), (5e-9,) * 3, p)
    s.set_M(np.array([M0], float))
    return s


# ---------------------------------------------------------------------------------------
# Zeeman waveforms
@pytest.mark.parametrize("wave", [1, 2])
def test_waveform_field_at_step_start_times(wave):
    """The oracle's Zeeman field after k fixed steps is the closed-form waveform at the
    accumulated step time (not at a stage time), including t = t0 for the pulse."""
    dt = 2.5e-13
    p = FMParams(Ms=8e5, alpha=0.05, Ha=(100.0, -50.0, 30.0), wave=wave, amp=1.5e4, w=2 * np.pi / 1e-10, t0=1e-12,
                 dir=(0.6, 0.0, 0.8), slope=(1e14, -2e14, 5e13), modules=BM_MOD_ZEEMAN)
    s = one_cell(BM_EVAL_RK4, p, (p.Ms, 0.0, 0.0))
    s.set_dt(dt)
    for k in (0, 1, 4, 7, 13):
        s.run_to_iteration(k)
        st = s.state()
        assert st["stagetime"] == st["time"]
        s.update_fields()
        np.testing.assert_allclose(s.heff()[0], field_at(p, st["stagetime"]), rtol=1e-13, atol=1e-9)
    if wave == 1:
        # t = t0 exactly: F_sinc(0) = 1
        s0 = one_cell(BM_EVAL_RK4, FMParams(**{**p.__dict__, "t0": 0.0}), (p.Ms, 0.0, 0.0))
        s0.update_fields()
        np.testing.assert_allclose(s0.heff()[0], np.array(p.Ha) + p.amp * np.array(p.dir), rtol=1e-15)


@pytest.mark.parametrize("wave", [1, 2])
def test_waveform_run_holds_field_at_step_start(wave):
    """A one-cell Zeeman-only RK4 run under a pulse or a ramp equals an independent numpy RK4
    whose field is held at each step's start time, and differs from one that moves the field
    to the stage times (so the test can tell the two apart)."""
    dt, nsteps = 5e-13, 400
    p = FMParams(Ms=8e5, alpha=0.05, Ha=(0.0, 0.0, 2e4), wave=wave, amp=4e4, w=2 * np.pi / 4e-11, t0=3e-11,
                 dir=(0.0, 0.8, 0.6), slope=(2e14, 0.0, -3e14), modules=BM_MOD_ZEEMAN)
    M0 = np.array([p.Ms * 0.8, 0.0, p.Ms * 0.6])
    s = one_cell(BM_EVAL_RK4, p, M0)
    s.set_dt(dt)
    s.run_to_iteration(nsteps)
    got = s.get_M()[0]
    held = rk4_macrospin(p, M0, dt, nsteps)
    moved = rk4_macrospin(p, M0, dt, nsteps, substep_times=True)
    assert np.linalg.norm(got - held) / p.Ms <= 1e-10
    assert np.linalg.norm(moved - held) / p.Ms >= 1e-5


# ---------------------------------------------------------------------------------------
# AHeun
def aheun_step(M, H, Ms, alpha, dt):
    """One AHeun step written here: predictor, corrector with renormalisation, and the
    error estimate |M_corr - M_pred| / |M_corr|."""
    pred = M + dt * llg(M, H, Ms, alpha)
    corr = (M + pred + dt * llg(pred, H, Ms, alpha)) / 2
    corr = corr * Ms / np.linalg.norm(corr)
    return corr, np.linalg.norm(corr - pred) / np.linalg.norm(corr)


def test_aheun_defaults():
    """Boris/DiffEq_Defs.h:17-25: dT0 0.5e-13; after a step the controller keeps dT within
    [1e-15, 1e-12] and grows it at most 2x."""
    p = FMParams(Ms=8e5, alpha=0.1, Ha=(0.0, 0.0, 8e3), modules=BM_MOD_ZEEMAN)
    s = one_cell(BM_EVAL_AHEUN, p, (p.Ms, 0.0, 0.0))
    assert s.state()["dT"] == 0.5e-13
    s.run_to_iteration(1)
    assert s.state()["dT"] == pytest.approx(1e-13, rel=1e-15)    # tiny lte: capped at 2x
    s.run_to_iteration(60)
    assert s.state()["dT"] == 1e-12                              # capped at dT_max


def test_aheun_lte_and_order2_controller():
    """The oracle's lte of an accepted AHeun step equals the one computed here, and the next
    dT follows dT * clamp((0.8 err / lte)^(1/3), 0.01, 2) (order 2 -> exponent 1/3)."""
    p = FMParams(Ms=8e5, alpha=0.1, Ha=(0.0, 3e4, 8e4), modules=BM_MOD_ZEEMAN)
    M0 = np.array([p.Ms, 0.0, 0.0])
    s = one_cell(BM_EVAL_AHEUN, p, M0)
    dt = 4e-13
    s.set_dt(dt)
    s.set_astep(err=1e-3)
    s.run_to_iteration(1)
    st = s.state()
    Mx, lte = aheun_step(M0, np.array(p.Ha), p.Ms, p.alpha, dt)
    assert st["rejections"] == 0
    np.testing.assert_allclose(s.get_M()[0], Mx, rtol=1e-13, atol=1e-7)
    assert st["lte"] == pytest.approx(lte, rel=1e-9)
    c = min(2.0, max(0.01, (0.8 * 1e-3 / lte) ** (1.0 / 3.0)))
    assert st["dT"] == pytest.approx(min(1e-12, max(1e-15, dt * c)), rel=1e-12)
    # a wrong controller order (1/6, the RKDP54 exponent) would give a visibly different dT
    c6 = min(2.0, max(0.01, (0.8 * 1e-3 / lte) ** (1.0 / 6.0)))
    assert abs(c6 - c) / c > 1e-3


def test_aheun_forced_rejection():
    """lte above err_fail: time and iteration undone, M back at the step start (sM1), dT
    reduced by max(0.01, (0.8 err / lte)^(1/3)); the redone step is accepted."""
    p = FMParams(Ms=8e5, alpha=0.1, Ha=(0.0, 3e4, 8e4), modules=BM_MOD_ZEEMAN)
    M0 = np.array([p.Ms, 0.0, 0.0])
    s = one_cell(BM_EVAL_AHEUN, p, M0)
    dt = 1e-12
    s.set_dt(dt)
    _, lte = aheun_step(M0, np.array(p.Ha), p.Ms, p.alpha, dt)
    err = lte / 4
    s.set_astep(err=err)
    for _ in range(2):                     # predictor, corrector: computed, then rejected
        s.update_fields()
        s.iterate_stage()
    st = s.state()
    assert st["rejections"] == 1 and st["iteration"] == 0
    assert st["time"] == 0.0 and st["stagetime"] == 0.0
    np.testing.assert_array_equal(s.get_M()[0], M0)
    c = max(0.01, (0.8 * err / lte) ** (1.0 / 3.0))
    assert st["dT"] == pytest.approx(dt * c, rel=1e-12)
    s.run_to_iteration(1)
    st2 = s.state()
    assert st2["iteration"] == 1 and st2["rejections"] >= 1
    assert st2["lte"] <= err


def test_aheun_macrospin_order2():
    """Fixed-step AHeun (dT_min == dT_max) on the macrospin: error vs the closed form falls as
    dt^2 (Heun), and it equals TEuler step for step (same predictor / corrector)."""
    H, alpha, t_end = 8e4, 0.1, 2e-10

    def exact(t):
        gp = GAMMA / (1 + alpha ** 2)
        th = 2 * np.arctan(np.exp(-alpha * gp * H * t))
        ph = gp * H * t
        return np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])

    p = FMParams(Ms=8e5, alpha=alpha, Ha=(0.0, 0.0, H), modules=BM_MOD_ZEEMAN)
    errs = []
    for dt in (4e-13, 2e-13):
        s = one_cell(BM_EVAL_AHEUN, p, (p.Ms, 0.0, 0.0))
        s.set_dt(dt)
        s.set_astep(dtmin=dt, dtmax=dt)
        steps = int(round(t_end / dt))
        s.run_to_iteration(steps)
        te = one_cell(BM_EVAL_TEULER, p, (p.Ms, 0.0, 0.0))
        te.set_dt(dt)
        te.run_to_iteration(steps)
        np.testing.assert_array_equal(s.get_M(), te.get_M())
        errs.append(np.linalg.norm(s.get_M()[0] / p.Ms - exact(steps * dt)))
    assert errs[1] < 1e-3
    rate = np.log2(errs[0] / errs[1])
    assert 1.7 < rate < 2.3, (errs, rate)
