"""GPU demag convolution with periodic boundary conditions against the FP64 oracle (whose
periodic tensors equal the reference's Boris/DemagTFunc_PBC.cpp bit for bit): 2D and 3D
meshes, images along each axis and all three, far-field images, radix-3 sizes, the set/add
and energy outputs; plus a size-independent check at 512 x 512 x 1 with 10 x 10 images."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def nz_open(nz):
    N = 1
    while nz > 1 and N < 2 * nz:
        N *= 2
    return N


CASES = [((32, 24, 1), (4e-9, 4e-9, 4e-9), (3, 0, 0)),
         ((24, 16, 1), (4e-9, 3e-9, 6e-9), (2, 4, 0)),
         ((16, 32, 1), (4e-9, 4e-9, 5e-9), (0, 5, 2)),
         ((48, 48, 1), (5e-9, 5e-9, 10e-9), (10, 10, 0)),
         ((16, 12, 4), (4e-9, 4e-9, 4e-9), (2, 0, 0)),
         ((12, 16, 8), (3e-9, 4e-9, 2e-9), (0, 2, 3)),
         ((8, 8, 4), (4e-9, 4e-9, 4e-9), (1, 1, 1)),
         ((24, 6, 3), (4e-9, 4e-9, 4e-9), (0, 3, 0))]


@pytest.mark.parametrize("n,h,pbc", CASES)
def test_pbc_convolution_matches_oracle(gpu, n, h, pbc):
    ncell = n[0] * n[1] * n[2]
    M = api.random_magnetisation(ncell, 8e5, ncell + sum(pbc)).astype(np.float32)
    d = api.Demag(n, h, pbc)
    N, _ = d.info()
    assert N[0] == (n[0] if pbc[0] else 2 * n[0]) and N[1] == (n[1] if pbc[1] else 2 * n[1])
    if n[2] > 1:
        assert N[2] == (n[2] if pbc[2] else nz_open(n[2]))
    H = d.convolute(torch.from_numpy(M).cuda()).cpu().numpy()
    ref = O.OracleDemag(n, h, pbc=pbc)(M.astype(np.float64))
    assert rel_l2(H, ref) <= 1e-6
    # the images change the field: the open-boundary result is far off
    H0 = O.OracleDemag(n, h, Nz_override=nz_open(n[2]))(M.astype(np.float64))
    assert rel_l2(H0, ref) > 1e-3


def test_pbc_add_and_energy(gpu):
    n, h, pbc = (32, 16, 4), (4e-9, 4e-9, 4e-9), (4, 4, 0)
    ncell = n[0] * n[1] * n[2]
    M = api.random_magnetisation(ncell, 1e6, 5).astype(np.float32)
    d = api.Demag(n, h, pbc)
    Mt = torch.from_numpy(M).cuda()
    base = torch.full_like(Mt, 3.0)
    e = torch.zeros(1, dtype=torch.float32, device="cuda")
    H = d.convolute(Mt, base.clone(), clear_out=False, energy=e).cpu().numpy()
    ref = O.OracleDemag(n, h, pbc=pbc)(M.astype(np.float64))
    assert rel_l2(H - 3.0, ref) <= 1e-6
    e_ref = -api.MU0 / (2 * ncell) * float((M.astype(np.float64) * ref).sum())
    assert abs(e.item() - e_ref) <= 1e-5 * abs(e_ref)


def test_pbc_thin_film_limit(gpu):
    """512 x 512 x 1 cells with 10 x 10 images: a uniform out-of-plane state sees close to
    the infinite-film field -Mz, and the in-plane state close to zero; the field is the
    same in every cell (translation invariance along periodic axes)."""
    n, h = (512, 512, 1), (4e-9, 4e-9, 4e-9)
    d = api.Demag(n, h, (10, 10, 0))
    ncell = n[0] * n[1]
    Mz = torch.zeros((ncell, 3), device="cuda"); Mz[:, 2] = 1.0
    Hz = d.convolute(Mz)
    assert float((Hz[:, 2] - Hz[0, 2]).abs().max()) <= 1e-5
    assert -1.0 < float(Hz[0, 2]) < -0.99
    Mx = torch.zeros((ncell, 3), device="cuda"); Mx[:, 0] = 1.0
    Hx = d.convolute(Mx)
    assert float((Hx[:, 0] - Hx[0, 0]).abs().max()) <= 1e-5
    assert -0.01 < float(Hx[0, 0]) < 0.0


# ---------------------------------------------------------------------------------------
# The integrator with periodic boundaries (bm_ode_set_pbc): demag with images and wrap-around
# exchange, against the oracle's Sim with the same images.

from test_gpu_ode import bench_like_params, sp4_params, tilted_x    # noqa: E402


def _ode_pair(method, n, h, p, M0, pbc, speedup=0):
    s = api.FMSolver(method, n, h, p, M0, pbc=pbc)
    o = O.OracleSim(method, n, h, p, Nz_override=nz_open(n[2]) if n[2] > 1 else 0)
    o.set_pbc(pbc)
    o.set_M(M0.astype(np.float64))
    if speedup:
        s.set_eval_speedup(speedup)
        o.set_speedup(speedup)
    return s, o


@pytest.mark.parametrize("path", ["row", "unfused"])
@pytest.mark.parametrize("n,pbc", [((64, 32, 1), (0, 4, 0)), ((32, 16, 4), (0, 2, 3)), ((32, 16, 1), (3, 3, 0)),
                                   ((32, 24, 4), (2, 0, 0))])
def test_pbc_rkdp54_matches_oracle(gpu, monkeypatch, path, n, pbc):
    """Periodic meshes take the unfused evaluation whatever the row-kernel switch says (the
    stage kernel wraps exchange around); both settings must match the oracle."""
    monkeypatch.setenv("BM_ROW_KERNEL", "1" if path == "row" else "0")
    h, p = (4e-9, 4e-9, 4e-9), bench_like_params()
    M0 = tilted_x(n[0] * n[1] * n[2], p.Ms, 31, tilt=0.3)
    s, o = _ode_pair(api.BM_EVAL_RKDP54, n, h, p, M0, pbc)
    s.run_to_iteration(20)
    o.run_to_iteration(20)
    st, so = s.state(), o.state()
    assert st["rejections"] == so["rejections"]
    assert abs(st["time"] - so["time"]) <= 1e-3 * so["time"]
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-4


def test_pbc_rk4_film_and_exchange_only(gpu):
    n, h = (64, 32, 1), (3.90625e-9, 3.90625e-9, 3e-9)
    M0 = tilted_x(n[0] * n[1], 8e5, 5, tilt=0.5)
    for p in (sp4_params(), api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, Ha=(0.0, 2e4, 0.0),
                                         modules=api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)):
        s, o = _ode_pair(api.BM_EVAL_RK4, n, h, p, M0, (5, 5, 0))
        s.set_dt(1e-13); o.set_dt(1e-13)
        s.run_to_iteration(50)
        o.run_to_iteration(50)
        assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-5


@pytest.mark.parametrize("pbc", [(4, 4, 0), (2, 0, 0)])
def test_pbc_speedup_matches_oracle(gpu, pbc):
    """Evaluation speedup with periodic images: the saved fields exclude SelfDemag_PBC's self
    term (the images' far-field terms included)."""
    n, h, p = (64, 32, 1), (3.90625e-9, 3.90625e-9, 3e-9), sp4_params()
    M0 = tilted_x(n[0] * n[1], p.Ms, 12)
    s, o = _ode_pair(api.BM_EVAL_RKDP54, n, h, p, M0, pbc, speedup=3)
    s.run_to_iteration(40)
    o.run_to_iteration(40)
    st, so = s.state(), o.state()
    assert st["rejections"] == so["rejections"]
    assert abs(st["time"] - so["time"]) <= 1e-3 * so["time"]
    assert s.speedup_state()["convolutions"] == o.speedup_info()["convolutions"]
    assert rel_l2(s.M.cpu().numpy(), o.get_M()) <= 1e-4


def test_pbc_images_must_match_the_demag_handle(gpu):
    import ctypes as C
    n, h, p = (32, 16, 1), (4e-9, 4e-9, 4e-9), sp4_params()
    s = api.FMSolver(api.BM_EVAL_RK4, n, h, p, tilted_x(512, p.Ms, 1), pbc=(2, 0, 0))
    assert s.L.bm_ode_set_pbc(s.handle, 3, 0, 0) == 5          # BM_ERR_STATE
    assert s.L.bm_ode_set_pbc(s.handle, 2, 0, 0) == 0


def test_pbc_averaged_inputs_duplicated_outputs(gpu):
    """Two-sublattice meshes on a periodic handle: both outputs hold the periodic field of the
    averaged input."""
    n, h, pbc = (32, 16, 4), (4e-9, 4e-9, 4e-9), (3, 2, 0)
    ncell = n[0] * n[1] * n[2]
    M1 = api.random_magnetisation(ncell, 8e5, 41).astype(np.float32)
    M2 = api.random_magnetisation(ncell, 8e5, 42).astype(np.float32)
    d = api.Demag(n, h, pbc)
    H2 = torch.empty((ncell, 3), dtype=torch.float32, device="cuda")
    H1 = d.convolute_averaged(torch.from_numpy(M1).cuda(), torch.from_numpy(M2).cuda(), H2=H2)
    ref = O.OracleDemag(n, h, pbc=pbc)(0.5 * (M1.astype(np.float64) + M2.astype(np.float64)))
    assert rel_l2(H1.cpu().numpy(), ref) <= 1e-6
    assert torch.equal(H1, H2)


def test_set_params_demag_on_with_other_images_refused(gpu):
    """A periodic mesh whose demag is switched on later by bm_ode_set_params needs a demag
    handle with the same images (bm_ode_set_pbc checks this only while demag is on)."""
    import ctypes as C
    from boris2_amd import _lib
    L = _lib.load()
    n, h = (32, 16, 1), (4e-9, 4e-9, 4e-9)
    p_off = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, modules=api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)
    p_on = api.FMParams(**{**p_off.__dict__, "modules": p_off.modules | api.BM_MOD_DEMAG})
    d_open = api.Demag(n, h)
    M = torch.from_numpy(api.random_magnetisation(512, 8e5, 1).astype(np.float32)).cuda()
    pc_off, pc_on = p_off.to_c(), p_on.to_c()
    hd = C.c_void_p()
    assert L.bm_ode_create(api.BM_EVAL_RK4, *n, *h, C.byref(pc_off), d_open.handle, api._ptr(M), C.byref(hd)) == 0
    assert L.bm_ode_set_pbc(hd, 2, 0, 0) == 0                 # demag off: only exchange wraps
    assert L.bm_ode_set_params(hd, C.byref(pc_on)) == 5       # BM_ERR_STATE: the handle is open
    assert L.bm_ode_set_pbc(hd, 0, 0, 0) == 0
    assert L.bm_ode_set_params(hd, C.byref(pc_on)) == 0       # open mesh, open handle
    assert L.bm_ode_destroy(hd) == 0
    d_open.close()
