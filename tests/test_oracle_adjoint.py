"""Pins the oracle's discrete RK4 adjoint against central finite differences of the same
FP64 forward rollout: per-cell Ms gradients and the pulse-amplitude gradient, with demag,
exchange and a sinc Zeeman pulse.  The central difference converges quadratically onto the
adjoint (relative gap 1e-2, 1e-4, 1e-6, 1e-8 at h = 1e-2 ... 1e-5 relative), so the check
uses h = 1e-5 and a 1e-6 relative bound."""
import numpy as np
import pytest

import oracle_lib as O
from boris2_amd import api


def unit_field(rng, n, smooth=False):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def setup(n=(8, 6, 2), steps=30, dt=250e-15, seed=1):
    rng = np.random.default_rng(seed)
    ncell = n[0] * n[1] * n[2]
    Ms = rng.uniform(5e5, 1.1e6, ncell)
    d = rng.normal(size=3); d /= np.linalg.norm(d)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.05, modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN,
                     wave=1, amp=20e-3 / api.MU0, w=2 * np.pi / 1e-9, t0=2e-12, dir=tuple(d))
    m0 = unit_field(rng, ncell)
    mt = unit_field(rng, ncell)
    return n, p, Ms, m0, mt, dt, steps


def loss_at(n, p, Ms, m0, mt, dt, steps):
    a = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps, Nz_override=4 if n[2] > 1 else 0)
    return a.forward(m0, mt)[0]


def test_adjoint_matches_central_fd():
    n, p, Ms, m0, mt, dt, steps = setup()
    a = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps, Nz_override=4)
    L0, _ = a.forward(m0, mt)
    g, ga = a.backward(m0, mt)
    rng = np.random.default_rng(9)
    for i in rng.choice(len(Ms), 6, replace=False):
        hs = 1e-5 * Ms[i]
        Mp, Mm = Ms.copy(), Ms.copy()
        Mp[i] += hs; Mm[i] -= hs
        fd = (loss_at(n, p, Mp, m0, mt, dt, steps) - loss_at(n, p, Mm, m0, mt, dt, steps)) / (2 * hs)
        assert abs(g[i] - fd) <= 1e-6 * abs(fd) + 1e-18, (i, g[i], fd)
    ha = 1e-5 * p.amp
    pp = api.FMParams(**{**p.__dict__, "amp": p.amp + ha})
    pm = api.FMParams(**{**p.__dict__, "amp": p.amp - ha})
    fd = (loss_at(n, pp, Ms, m0, mt, dt, steps) - loss_at(n, pm, Ms, m0, mt, dt, steps)) / (2 * ha)
    assert abs(ga - fd) <= 1e-6 * abs(fd), (ga, fd)


@pytest.mark.parametrize("ckpt", [1, 3, 7, 39])
def test_checkpointed_adjoint_equals_whole_tape(ckpt):
    """Checkpoints every ckpt steps with segment recomputation in the backward pass give the
    same bits as keeping the whole tape (same FP64 arithmetic, recomputed in the same order),
    including a last segment shorter than ckpt."""
    n, p, Ms, m0, mt, dt, steps = setup()
    a = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps, Nz_override=4)
    L0, M0 = a.forward(m0, mt)
    g0, ga0 = a.backward(m0, mt)
    c = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps, Nz_override=4, ckpt=ckpt)
    L1, M1 = c.forward(m0, mt)
    g1, ga1 = c.backward(m0, mt)
    assert L1 == L0 and np.array_equal(M1, M0)
    assert np.array_equal(g1, g0) and ga1 == ga0


def test_multilayer_transpose_is_exact():
    """<v, N u> = <N^T v, u> for the multilayered demag, irregular layers included (the
    far-field series is not volume-symmetric, so the transpose uses the pair kernels)."""
    for nx, ny in ((16, 12), (48, 48)):
        ml = O.OracleMLayer(nx, ny, 4e-9, 4e-9, [8e-9, 4e-9, 16e-9], [0.0, 12e-9, 20e-9])
        rng = np.random.default_rng(4)
        u = rng.standard_normal((3 * nx * ny, 3)); v = rng.standard_normal((3 * nx * ny, 3))
        a, b = (v * ml(u)).sum(), (ml.transposed(v) * u).sum()
        assert abs(a - b) <= 1e-12 * abs(a), (nx, a, b)
        # with equal layers the operator is symmetric: N^T = N
        eq = O.OracleMLayer(nx, ny, 4e-9, 4e-9, [4e-9, 4e-9], [0.0, 8e-9])
        w = rng.standard_normal((2 * nx * ny, 3))
        np.testing.assert_allclose(eq.transposed(w), eq(w), rtol=0, atol=1e-10 * np.abs(eq(w)).max())


HZ, Z0 = [8e-9, 4e-9, 16e-9], [0.0, 12e-9, 20e-9]
LA, LAL, LK1 = [1.3e-11, 1.0e-11, 1.6e-11], [0.02, 0.08, 0.03], [0.0, 3e5, 1e5]


def layered_setup(nx=6, ny=5, steps=25, dt=200e-15, seed=3):
    rng = np.random.default_rng(seed)
    ncell = 3 * nx * ny
    Ms = np.repeat(rng.uniform(6e5, 1.2e6, 3), nx * ny) * rng.uniform(0.9, 1.1, ncell)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, K1=0.0, K2=2e4, ea=(0.3, 0.0, 0.95),
                     modules=api.BM_MOD_EXCHANGE | api.BM_MOD_ANIS_UNI | api.BM_MOD_ZEEMAN,
                     Ha=(-1e4, 5e3, 2e3), wave=2, slope=(4e14, -2e14, 1e14))
    return (nx, ny, 3), p, Ms, unit_field(rng, ncell), unit_field(rng, ncell), dt, steps


def layered_adj(n, p, Ms, dt, steps):
    a = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps)
    a.set_layers(4e-9, 4e-9, HZ, Z0, with_demag=True, A=LA, alpha=LAL, K1=LK1)
    return a


def test_layered_adjoint_with_anisotropy_and_ramp_matches_central_fd():
    n, p, Ms, m0, mt, dt, steps = layered_setup()
    a = layered_adj(n, p, Ms, dt, steps)
    a.forward(m0, mt)
    g, _ = a.backward(m0, mt)
    gz = a.gz.copy()

    def L(pp, MM):
        return layered_adj(n, pp, MM, dt, steps).forward(m0, mt)[0]

    rng = np.random.default_rng(2)
    for i in rng.choice(len(Ms), 6, replace=False):
        hs = 1e-5 * Ms[i]
        Mp, Mm = Ms.copy(), Ms.copy()
        Mp[i] += hs; Mm[i] -= hs
        fd = (L(p, Mp) - L(p, Mm)) / (2 * hs)
        assert abs(g[i] - fd) <= 1e-6 * abs(fd) + 1e-18, (i, g[i], fd)
    for c in range(3):
        hsl = 1e-5 * abs(p.slope[c])
        sp, sm = list(p.slope), list(p.slope)
        sp[c] += hsl; sm[c] -= hsl
        fd = (L(api.FMParams(**{**p.__dict__, "slope": tuple(sp)}), Ms) - L(api.FMParams(**{**p.__dict__, "slope": tuple(sm)}), Ms)) / (2 * hsl)
        assert abs(gz[1 + c] - fd) <= 1e-6 * abs(fd), (c, gz[1 + c], fd)


PBC_CASES = [((8, 8, 1), (3, 2, 0)), ((8, 6, 4), (2, 0, 2))]


@pytest.mark.parametrize("n,pbc", PBC_CASES, ids=["2d_xy", "3d_xz"])
def test_periodic_adjoint_matches_central_fd(n, pbc):
    """Periodic meshes: the wrapped exchange stencil and the image-summed demag tensor are
    both their own transposes, so the same adjoint sweep applies; central differences of the
    periodic forward pin it, at cells on the wrapped ends and inside."""
    _, p, Ms, m0, mt, dt, steps = setup(n=n, steps=20, seed=5)
    h = (4e-9,) * 3
    loss_p = lambda Ms_, p_: O.OracleAdjoint(n, h, p_, Ms_, dt, steps, pbc=pbc).forward(m0, mt)[0]
    a = O.OracleAdjoint(n, h, p, Ms, dt, steps, pbc=pbc)
    L0, _ = a.forward(m0, mt)
    g, ga = a.backward(m0, mt)
    # the images change the rollout (else this would test the open mesh again)
    assert abs(L0 - loss_at(n, p, Ms, m0, mt, dt, steps)) > 1e-7
    nx, ny = n[0], n[1]
    cells = [0, nx - 1, (ny - 1) * nx, nx * ny - 1, len(Ms) - 1, nx + 3]
    for i in cells:
        hs = 1e-5 * Ms[i]
        Mp, Mm = Ms.copy(), Ms.copy()
        Mp[i] += hs; Mm[i] -= hs
        fd = (loss_p(Mp, p) - loss_p(Mm, p)) / (2 * hs)
        assert abs(g[i] - fd) <= 1e-6 * abs(fd) + 1e-18, (i, g[i], fd)
    ha = 1e-5 * p.amp
    fd = (loss_p(Ms, api.FMParams(**{**p.__dict__, "amp": p.amp + ha}))
          - loss_p(Ms, api.FMParams(**{**p.__dict__, "amp": p.amp - ha}))) / (2 * ha)
    assert abs(ga - fd) <= 1e-6 * abs(fd), (ga, fd)


@pytest.mark.parametrize("n,pbc", PBC_CASES, ids=["2d_xy", "3d_xz"])
def test_periodic_adjoint_forward_equals_periodic_sim(n, pbc):
    """The adjoint's periodic forward rollout is the oracle integrator's periodic RK4 run
    (set_pbc: image tensor and wrapped exchange) on the same Ms map."""
    _, p, Ms, m0, mt, dt, steps = setup(n=n, steps=12, seed=6)
    a = O.OracleAdjoint(n, (4e-9,) * 3, p, Ms, dt, steps, pbc=pbc)
    _, Mf = a.forward(m0, mt)
    s = O.OracleSim(api.BM_EVAL_RK4, n, (4e-9,) * 3, p)
    s.set_pbc(pbc)
    s.set_ms_map(Ms)
    s.set_M(m0 * Ms[:, None])
    s.set_dt(dt)
    s.run_to_iteration(steps)
    M = s.get_M()
    assert np.max(np.abs(Mf - M)) <= 1e-9 * np.max(Ms)
