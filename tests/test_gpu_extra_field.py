"""The integrator's input for the fields of other modules (bm_ode_set_extra_field): before
every stage a caller function computes a field from that stage's magnetisation and the stage
adds it to the effective field, as the reference runs every module's UpdateField before each
stage (Boris/MeshBaseModules.cpp:98-111).  Checked by moving a fused module out: uniaxial
anisotropy computed by the caller (torch) with the built-in anisotropy off must follow the
FP64 oracle run with its own anisotropy module."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def tilted_x(ncell, Ms, seed, tilt=0.3):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, tilt, ncell)
    ph = rng.uniform(0, 2 * np.pi, ncell)
    m = np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], axis=1)
    return (m * Ms).astype(np.float32)


def params(modules):
    return api.FMParams(Ms=1e6, A=1.5e-11, alpha=0.1, K1=5e5, K2=1e5, ea=(0.0, 0.6, 0.8),
                        Ha=(2e3, -1e3, 39788.736), modules=modules)


def uniaxial_field(p):
    """The caller's module: H = 2/(mu0 Ms) (m.e) (K1 + 2 K2 (1 - (m.e)^2)) e, m = M/Ms."""
    e = torch.tensor(p.ea, dtype=torch.float32, device="cuda")

    def fn(M):
        d = (M @ e) / p.Ms
        return ((2.0 / (api.MU0 * p.Ms)) * d * (p.K1 + 2.0 * p.K2 * (1.0 - d * d)))[:, None] * e[None, :]
    return fn


@pytest.fixture(params=["row", "unfused"])
def path(request, monkeypatch):
    monkeypatch.setenv("BM_ROW_KERNEL", "1" if request.param == "row" else "0")
    return request.param


@pytest.mark.parametrize("method", [api.BM_EVAL_RK4, api.BM_EVAL_RKDP54])
def test_anisotropy_from_the_caller_matches_oracle(gpu, path, method):
    n, h = (64, 32, 4), (4e-9, 4e-9, 4e-9)
    full = api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN | api.BM_MOD_ANIS_UNI
    p_full, p_rest = params(full), params(full & ~api.BM_MOD_ANIS_UNI)
    ncell = n[0] * n[1] * n[2]
    M0 = tilted_x(ncell, p_full.Ms, 21)
    s = api.FMSolver(method, n, h, p_rest, M0)
    s.set_extra_field(uniaxial_field(p_full))
    o = O.OracleSim(method, n, h, p_full, Nz_override=8)
    o.set_M(M0.astype(np.float64))
    steps = 30
    s.run_to_iteration(steps)
    o.run_to_iteration(steps)
    st, so = s.state(), o.state()
    assert st["rejections"] == so["rejections"]
    assert abs(st["time"] - so["time"]) <= 1e-4 * so["time"]
    dm = float(np.abs(s.average_m() - o.average_m()).max())
    e = rel_l2(s.M.cpu().numpy(), o.get_M())
    print(f"{path} method {method}: <m> {dm:.2e}, per-cell {e:.2e}")
    assert dm <= 1e-4
    assert e <= 1e-4
    s.close()


def test_extra_field_enters_the_effective_field_and_detaches(gpu, path):
    """A constant field from the caller adds to the effective field the stage uses (keep_heff
    copy); detaching it restores the plain run bit for bit."""
    n, h = (32, 24, 2), (4e-9, 4e-9, 4e-9)
    p = params(api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)
    ncell = n[0] * n[1] * n[2]
    M0 = tilted_x(ncell, p.Ms, 22)
    Hc = torch.tensor([3e3, -4e3, 5e3], dtype=torch.float32, device="cuda").expand(ncell, 3).contiguous()
    seen = []

    def const(M):
        seen.append(M.data_ptr())
        return Hc

    a = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    b = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    for s in (a, b):
        s.keep_heff(True)
    a.set_extra_field(const)
    a.iterate_stage(); b.iterate_stage()
    Ha, Hb = a.heff().double(), b.heff().double()
    assert torch.allclose(Ha - Hb, Hc.double(), rtol=0, atol=1e-6 * Hb.abs().max().item())
    assert seen == [a.M.data_ptr()]                     # stage 0 evaluates the caller's M
    a.iterate_stage()
    assert len(seen) == 2 and seen[1] != a.M.data_ptr()   # stage 1: the solver's second buffer
    # detached: a solver that had a field function steps as one that never had it
    c = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    a2 = api.FMSolver(api.BM_EVAL_RK4, n, h, p, M0)
    a2.set_extra_field(const)
    a2.set_extra_field(None)
    a2.run_to_iteration(3); c.run_to_iteration(3)
    assert torch.equal(a2.M, c.M)
    for s in (a, b, c, a2):
        s.close()
