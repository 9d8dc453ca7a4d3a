"""GPU parity of the multilayered (stacked 2D layers) demag against the oracle."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from boris2_amd import api

pytestmark = pytest.mark.gpu

# (nx, ny, hx, hy, hz per layer, z0 per layer)
CASES = [
    # three layers of different thickness with gaps (irregular kernels, near field only)
    (16, 12, 4e-9, 4e-9, [8e-9, 4e-9, 16e-9], [0.0, 9e-9, 15e-9]),
    # equal touching layers (regular z-shifted kernels)
    (24, 8, 5e-9, 5e-9, [5e-9, 5e-9], [0.0, 5e-9]),
    # the 3-layer stack of 2/1/4 cells of 4 nm with 1-cell spacers (the multilayer
    # benchmark geometry) at 64 x 64 cells, large enough for the far-field series
    (64, 64, 4e-9, 4e-9, [8e-9, 4e-9, 16e-9], [0.0, 12e-9, 20e-9]),
    # one layer: the plain 2D convolution
    (32, 16, 4e-9, 4e-9, [3e-9], [0.0]),
    # 24 layers (above the 16 the register product holds: the general product, out of place),
    # uneven thicknesses and spacings with repeats, so pairs share kernel slots
    (16, 12, 4e-9, 4e-9, [2e-9, 4e-9, 3e-9] * 8, list(np.cumsum([0.0] + [3e-9, 5e-9, 4e-9] * 7 + [3e-9, 5e-9]))),
    # 20 equal touching layers: a regular stack, 39 distinct kernels for 400 pairs
    (24, 8, 4e-9, 4e-9, [4e-9] * 20, [4e-9 * k for k in range(20)]),
]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES, ids=["irregular3", "regular2", "bench3_64", "single", "uneven24", "regular20"])
def test_mlayer_field_matches_oracle(gpu, case):
    nx, ny, hx, hy, hz, z0 = case
    ncell = len(hz) * nx * ny
    M = api.random_magnetisation(ncell, 8e5, 3).astype(np.float32)
    d = api.MultiLayerDemag(nx, ny, hx, hy, hz, z0)
    H = d.convolute(torch.from_numpy(M).to("cuda")).cpu().numpy()
    ref = O.OracleMLayer(nx, ny, hx, hy, hz, z0)(M.astype(np.float64))
    assert rel_l2(H, ref) <= 1e-6


def test_mlayer_add_mode(gpu):
    nx, ny, hx, hy, hz, z0 = CASES[0]
    ncell = len(hz) * nx * ny
    M = torch.from_numpy(api.random_magnetisation(ncell, 8e5, 4).astype(np.float32)).to("cuda")
    d = api.MultiLayerDemag(nx, ny, hx, hy, hz, z0)
    base = torch.full_like(M, 123.0)
    H = d.convolute(M, base.clone(), clear_out=False)
    H0 = d.convolute(M)
    assert torch.allclose(H, H0 + 123.0, rtol=1e-6, atol=1e-2)


def test_mlayer_zshift_kernels_match_oracle(gpu):
    nx, ny, hx, hy, hz, z0 = CASES[0]
    d = api.MultiLayerDemag(nx, ny, hx, hy, hz, z0)
    K = d.kernels()
    hmax = max(hx, hy, max(hz))
    N = (2 * nx, 2 * ny)
    for dst in range(len(hz)):
        for src in range(len(hz)):
            if dst == src:
                continue
            sc = (hx / hmax, hy / hmax, hz[src] / hmax)
            dc = (hx / hmax, hy / hmax, hz[dst] / hmax)
            Dd, Do = O.tensor2d_zshift((nx, ny), N, sc, dc, (z0[dst] - z0[src]) / hmax)
            Kd, Ko = O.kernels2d_zshift(N, Dd, Do)
            ref = np.concatenate([Kd, Ko], axis=-1)          # (ky, kx, 6)
            got = K[dst, src].transpose(2, 1, 0)             # (6, kx, ky) -> (ky, kx, 6)
            assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("case", CASES[:3] + CASES[4:5], ids=["irregular3", "regular2", "bench3_64", "uneven24"])
def test_mlayer_split_per_layer_calls_equal_monolithic(gpu, case):
    """Forward per layer array, one product, inverse per layer array (the reference's per-mesh
    ForwardFFT / KernelMultiplication_MultipleInputs / InverseFFT): the same field as the
    one-call convolution bit for bit, set and add, and each layer's energy density
    -mu0/(2 n_l) sum M.H against the oracle field."""
    nx, ny, hx, hy, hz, z0 = case
    L, nl = len(hz), nx * ny
    M = torch.from_numpy(api.random_magnetisation(L * nl, 8e5, 6).astype(np.float32)).cuda()
    d = api.MultiLayerDemag(nx, ny, hx, hy, hz, z0)
    H_all = d.convolute(M)
    Ms = [M[l * nl:(l + 1) * nl].clone() for l in range(L)]
    Hs = [torch.full((nl, 3), 5.0, dtype=torch.float32, device="cuda") for _ in range(L)]
    en = torch.zeros(L, dtype=torch.float32, device="cuda")
    d.convolute_split(Ms, Hs, clear_out=True, energies=en)
    assert torch.equal(torch.cat(Hs), H_all)
    Ha = [torch.full((nl, 3), 5.0, dtype=torch.float32, device="cuda") for _ in range(L)]
    d.convolute_split(Ms, Ha, clear_out=False)
    assert torch.allclose(torch.cat(Ha), H_all + 5.0, rtol=1e-6, atol=1e-2)
    ref = O.OracleMLayer(nx, ny, hx, hy, hz, z0)(M.cpu().numpy().astype(np.float64))
    Mn = M.cpu().numpy().astype(np.float64)
    for l in range(L):
        e_ref = -api.MU0 / (2 * nl) * float(np.sum(Mn[l * nl:(l + 1) * nl] * ref[l * nl:(l + 1) * nl]))
        assert abs(en[l].item() - e_ref) <= 1e-5 * abs(e_ref), (l, en[l].item(), e_ref)
