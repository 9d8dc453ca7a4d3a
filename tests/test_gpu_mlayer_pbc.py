"""GPU tests of the multilayered demag with periodic images along x and y
(bm_mlayer_create_pbc): the convolution against the FP64 oracle's periodic stack (itself pinned
against the reference's DemagTFunc_Shifted_PBC / Irregular_PBC and a cyclic real-space sum),
the split per-layer calls, the integrator (wrapped in-plane exchange) and the adjoint of a
periodic stack against the oracle, and the configuration rules."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import _lib, api
from boris2_amd.api import BM_EVAL_RK4, BM_MOD_ANIS_UNI, BM_MOD_DEMAG, BM_MOD_EXCHANGE, BM_MOD_ZEEMAN, FMParams

pytestmark = pytest.mark.gpu

HZ, Z0 = [8e-9, 4e-9, 16e-9], [0.0, 12e-9, 20e-9]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,pbc", [((32, 24), (2, 0)), ((16, 32), (0, 3)), ((32, 32), (4, 4)), ((64, 16), (1, 2))],
                         ids=["x", "y", "xy_far", "xy"])
def test_periodic_stack_convolution_matches_oracle(gpu, n, pbc):
    nx, ny = n
    M = api.random_magnetisation(3 * nx * ny, 8e5, 3 + nx).astype(np.float32)
    ref = O.OracleMLayer(nx, ny, 4e-9, 4e-9, HZ, Z0, pbc=pbc)(M.astype(np.float64))
    ml = api.MultiLayerDemag(nx, ny, 4e-9, 4e-9, HZ, Z0, pbc=pbc)
    H = ml.convolute(torch.from_numpy(M).cuda()).cpu().numpy()
    e = rel_l2(H, ref)
    # split per-layer calls (ForwardFFT / multiply / InverseFFT per mesh) equal the monolithic call
    Ms = [torch.from_numpy(M[l * nx * ny:(l + 1) * nx * ny]).cuda() for l in range(3)]
    Hs = [torch.empty_like(m) for m in Ms]
    ml.convolute_split(Ms, Hs)
    Hsplit = torch.cat(Hs).cpu().numpy()
    print(f"{n} pbc {pbc}: rel L2 {e:.2e}, split vs monolithic {np.abs(Hsplit - H).max():.2e}")
    assert e <= 1e-6
    assert np.abs(Hsplit - H).max() <= 1e-6 * np.abs(H).max()
    # the images matter at these sizes
    H0 = O.OracleMLayer(nx, ny, 4e-9, 4e-9, HZ, Z0)(M.astype(np.float64))
    assert rel_l2(H0, ref) > 1e-3
    ml.close()


def test_periodic_stack_rk4_matches_oracle(gpu):
    nx, ny, pbc = 32, 16, (2, 3)
    MS, A, ALPHA, K1 = [6e5, 1.2e6, 9e5], [1.3e-11] * 3, [0.02, 0.05, 0.02], [0.0, 3e5, 0.0]
    mods = BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ANIS_UNI | BM_MOD_ZEEMAN
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, ea=(0.0, 0.0, 1.0), Ha=(-2.0e4, 4.0e3, 0.0), modules=mods)
    rng = np.random.default_rng(12)
    th, ph = 0.4 * rng.random(3 * nx * ny), 2 * np.pi * rng.random(3 * nx * ny)
    M0 = np.stack([np.cos(th), np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)], 1) * np.repeat(MS, nx * ny)[:, None]
    s = api.LayerStackSolver(BM_EVAL_RK4, nx, ny, 4e-9, 4e-9, HZ, Z0, p, Ms=MS, A=A, alpha=ALPHA, K1=K1,
                             M0=M0.astype(np.float32), pbc=pbc)
    q = FMParams(**{**p.__dict__, "modules": mods & ~BM_MOD_DEMAG})
    o = O.OracleSim(BM_EVAL_RK4, (nx, ny, 3), (4e-9, 4e-9, 4e-9), q)
    o.set_pbc((pbc[0], pbc[1], 0))
    o.set_layers(4e-9, 4e-9, HZ, Z0, with_demag=True, Ms=MS, A=A, alpha=ALPHA, K1=K1)
    o.set_M(M0)
    s.set_dt(2e-13); o.set_dt(2e-13)
    s.run_to_iteration(100)
    o.run_to_iteration(100)
    torch.cuda.synchronize()
    Mg, Mo = s.M.cpu().numpy().astype(np.float64), o.get_M()
    nl = nx * ny
    avg = lambda M: np.array([(M[l * nl:(l + 1) * nl] / MS[l]).mean(0) for l in range(3)])
    print(f"periodic stack: <m> {np.abs(avg(Mg) - avg(Mo)).max():.2e}, cell {np.abs(Mg - Mo).max() / max(MS):.2e}")
    assert np.abs(avg(Mg) - avg(Mo)).max() <= 1e-4
    assert np.abs(Mg - Mo).max() <= 2e-3 * max(MS)
    s.close()


def test_periodic_stack_adjoint_matches_oracle(gpu):
    nx, ny, pbc = 16, 12, (2, 2)
    LAY = dict(hz=HZ, z0=Z0, A=[1.3e-11, 1.0e-11, 1.6e-11], alpha=[0.02, 0.08, 0.03], K1=[0.0, 3e5, 1e5])
    rng = np.random.default_rng(9)
    Ms = np.repeat(rng.uniform(6e5, 1.2e6, 3), nx * ny)
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, K2=2e4, ea=(0.3, 0.0, 0.95),
                     modules=BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ANIS_UNI | BM_MOD_ZEEMAN,
                     Ha=(-1e4, 5e3, 2e3), wave=2, slope=(4e14, -2e14, 1e14))
    m0 = rng.normal(size=(3 * nx * ny, 3)); m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    mt = rng.normal(size=(3 * nx * ny, 3)); mt /= np.linalg.norm(mt, axis=1, keepdims=True)
    dt, steps = 200e-15, 30
    q = api.FMParams(**{**p.__dict__, "modules": p.modules & ~BM_MOD_DEMAG})
    o = O.OracleAdjoint((nx, ny, 3), (4e-9,) * 3, q, Ms, dt, steps, pbc=(pbc[0], pbc[1], 0))
    o.set_layers(4e-9, 4e-9, HZ, Z0, with_demag=True, A=LAY["A"], alpha=LAY["alpha"], K1=LAY["K1"])
    L0, _ = o.forward(m0, mt)
    g0, _ = o.backward(m0, mt)
    r = api.RK4Adjoint((nx, ny, 3), (4e-9, 4e-9, 16e-9), p, torch.tensor(Ms, dtype=torch.float32), dt, steps, 7,
                       layers=LAY, pbc=(pbc[0], pbc[1], 0))
    L, _ = r.forward(torch.tensor(m0, dtype=torch.float32), torch.tensor(mt, dtype=torch.float32))
    g, _ = r.backward()
    g = g.cpu().numpy().astype(np.float64)
    eg = np.linalg.norm(g - g0) / np.linalg.norm(g0)
    print(f"periodic stack adjoint: loss {abs(L - L0):.2e}, dL/dMs rel L2 {eg:.2e}")
    assert abs(L - L0) <= 1e-6 and eg <= 5e-5
    r.close()


def test_periodic_stack_configuration_rules(gpu):
    import ctypes as C
    L = _lib.load()
    h = C.c_void_p()
    hz, z0 = np.array(HZ), np.array(Z0)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    # a periodic nx must be even with plans for nx / 2 and nx
    assert L.bm_mlayer_create_pbc(3, 6, 8, 4e-9, 4e-9, ptr(hz), ptr(z0), 1, 0, 0, C.byref(h)) == 2
    assert L.bm_mlayer_create_pbc(3, 8, 8, 4e-9, 4e-9, ptr(hz), ptr(z0), -1, 0, 0, C.byref(h)) == 1
    p = FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, modules=BM_MOD_DEMAG | BM_MOD_EXCHANGE)
    s = api.LayerStackSolver(BM_EVAL_RK4, 16, 16, 4e-9, 4e-9, HZ, Z0, p, pbc=(2, 0))
    # the mesh takes the stack's images, none along z
    with pytest.raises(RuntimeError, match="error 5"):
        _lib.check(L.bm_ode_set_pbc(s.handle, 1, 0, 0))
    with pytest.raises(RuntimeError, match="error 5"):
        _lib.check(L.bm_ode_set_pbc(s.handle, 2, 0, 1))
    with pytest.raises(RuntimeError, match="error 5"):
        s.set_eval_speedup(3)                                         # no speedup on a periodic stack
    s.close()
