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
