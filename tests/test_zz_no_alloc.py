"""The allocation-free averaged convolution (bm_demag_convolute_averaged_static) against the
allocating one (bm_demag_convolute_averaged), both called through the C-ABI: bit-identical
fields and energies on two calls in a row, the second on a smaller mesh (so it runs over what
the first call left in the static scratch), with and without duplicated outputs, in set and
add mode; meshes beyond the static scratch are refused with BM_ERR_SIZE."""
import ctypes as C

import numpy as np
import pytest
import torch

from boris2_amd import _lib, api

pytestmark = pytest.mark.gpu

STATIC_MAX_CELLS = 1 << 20


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _run(L, fn, d, M1, M2, H1, H2, clear_out, energy):
    rc = getattr(L, fn)(d.handle, _vp(M1), _vp(M2), _vp(H1), _vp(H2), int(clear_out), _vp(energy), api._stream())
    assert rc == 0, L.bm_last_error().decode()


def _pair(n, seed, Ms=8e5):
    ncell = n[0] * n[1] * n[2]
    M1 = api.random_magnetisation(ncell, Ms, seed).astype(np.float32)
    M2 = api.random_magnetisation(ncell, Ms, seed + 1).astype(np.float32)
    return torch.from_numpy(M1).cuda(), torch.from_numpy(M2).cuda()


def _compare(L, n, h, seed):
    d = api.Demag(n, h)
    M1, M2 = _pair(n, seed)
    base = torch.randn_like(M1) * 1e4
    out = {}
    for fn in ("bm_demag_convolute_averaged", "bm_demag_convolute_averaged_static"):
        H1 = torch.empty_like(M1)
        H2 = torch.empty_like(M1)
        e = torch.zeros(1, dtype=torch.float32, device="cuda")
        _run(L, fn, d, M1, M2, H1, H2, 1, e)              # set, duplicated outputs, energy
        A = base.clone()
        _run(L, fn, d, M1, M2, A, None, 0, None)          # add, one output
        B = torch.empty_like(M1)
        _run(L, fn, d, M1, M2, B, None, 1, None)          # set, one output
        torch.cuda.synchronize()
        out[fn] = (H1.cpu(), H2.cpu(), e.cpu(), A.cpu(), B.cpu())
    d.close()
    ref, got = out["bm_demag_convolute_averaged"], out["bm_demag_convolute_averaged_static"]
    for r, g in zip(ref, got):
        assert torch.equal(r, g)
    assert torch.equal(ref[0], ref[1])
    assert float(ref[0].abs().max()) > 0.0
    return ref


def test_static_averaged_matches_allocating_twice_second_smaller(gpu):
    L = _lib.load()
    big = _compare(L, (96, 80, 6), (4e-9, 4e-9, 3e-9), seed=11)
    small = _compare(L, (40, 24, 1), (5e-9, 5e-9, 4e-9), seed=23)
    assert big[0].shape[0] == 96 * 80 * 6 and small[0].shape[0] == 40 * 24


def test_static_averaged_refuses_meshes_beyond_its_scratch(gpu):
    L = _lib.load()
    n = (1100, 1000, 1)
    assert n[0] * n[1] * n[2] > STATIC_MAX_CELLS
    d = api.Demag(n, (5e-9, 5e-9, 5e-9))
    M = torch.zeros((n[0] * n[1], 3), dtype=torch.float32, device="cuda")
    H = torch.empty_like(M)
    rc = L.bm_demag_convolute_averaged_static(d.handle, _vp(M), _vp(M), _vp(H), None, 1, None, api._stream())
    assert rc == 2
    assert "static scratch" in L.bm_last_error().decode()
    d.close()
