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


# Layers as meshes of their own: (nx, ny, hx, hy, hz, x0, y0, z0) per layer
H4 = 4e-9
MESH_CASES = {
    # a smaller layer inside the larger one's extent, offset (same cell)
    "offset": [(32, 24, H4, H4, 4e-9, 0.0, 0.0, 0.0), (12, 10, H4, H4, 2e-9, 5 * H4, 7 * H4, 6e-9)],
    # 2x coarser cells over the same extent (transfer weights 1/4)
    "coarse2": [(32, 16, H4, H4, 3e-9, 0.0, 0.0, 0.0), (16, 8, 2 * H4, 2 * H4, 6e-9, 0.0, 0.0, 5e-9)],
    # 3x coarser cells (reciprocal-distance weights of a 3x3 block), a third layer offset
    "coarse3": [(24, 18, H4, H4, 3e-9, 0.0, 0.0, 0.0), (8, 6, 3 * H4, 3 * H4, 4e-9, 0.0, 0.0, 4e-9),
                (6, 6, H4, H4, 2e-9, 9 * H4, 3 * H4, 9e-9)],
    # side by side: the convolution rectangles at different places (the reference's shifted kernels)
    "side": [(16, 12, H4, H4, 4e-9, 0.0, 0.0, 0.0), (16, 12, H4, H4, 4e-9, 20 * H4, 3 * H4, 2e-9)],
    # cells that are no multiple of each other: 20 cells of 5 nm over 16 of 6.25 nm
    "ratio": [(20, 20, 5e-9, 5e-9, 3e-9, 0.0, 0.0, 0.0), (16, 16, 6.25e-9, 6.25e-9, 3e-9, 0.0, 0.0, 5e-9)],
    # rectangles offset by fractions of a common cell (the reference's Complex_Full kernels)
    "frac_side": [(16, 12, H4, H4, 4e-9, 0.0, 0.0, 0.0), (16, 12, H4, H4, 4e-9, 20.5 * H4, 3.25 * H4, 2e-9)],
    "frac_three": [(16, 12, H4, H4, 8e-9, 0.0, 0.0, 0.0), (16, 12, H4, H4, 4e-9, 1.3e-9, -2.2e-9, 10e-9),
                   (16, 12, H4, H4, 4e-9, 5.1e-9, 0.6e-9, 16e-9)],
    # a 2x coarser layer offset by a fraction of the common cell, and a far one (far-field pairs)
    "frac_coarse": [(32, 16, H4, H4, 3e-9, 0.0, 0.0, 0.0), (16, 8, 2 * H4, 2 * H4, 6e-9, 0.7 * H4, 0.0, 5e-9)],
    "frac_far": [(24, 20, H4, H4, 4e-9, 0.0, 0.0, 0.0), (6, 4, H4, H4, 4e-9, 40.4 * H4, 44.6 * H4, 30e-9)],
}


@pytest.mark.parametrize("name", list(MESH_CASES))
def test_mesh_layers_match_oracle(gpu, name):
    """The common grid and mesh transfer of the reference (n_common, enlarged rectangles,
    reciprocal-distance weights) against the oracle's numpy restatement: fields of every
    layer's own cells (1e-6), and the split per-layer calls equal to the one-call form."""
    meshes = MESH_CASES[name]
    counts = [m[0] * m[1] for m in meshes]
    M = torch.from_numpy(api.random_magnetisation(sum(counts), 8e5, 12).astype(np.float32)).cuda()
    d = api.MultiLayerMeshDemag(meshes)
    H = d.convolute(M)
    ref = O.OracleMLayerMeshes(meshes)(M.cpu().numpy().astype(np.float64))
    e = rel_l2(H.cpu().numpy(), ref)
    print(f"{name}: rel L2 {e:.2e}")
    assert e <= 1e-6
    off = np.concatenate([[0], np.cumsum(counts)])
    Ms = [M[off[l]:off[l + 1]].clone() for l in range(len(meshes))]
    Hs = [torch.full((c, 3), 7.0, dtype=torch.float32, device="cuda") for c in counts]
    en = torch.zeros(len(meshes), dtype=torch.float32, device="cuda")
    d.convolute_split(Ms, Hs, clear_out=True, energies=en)
    assert torch.equal(torch.cat(Hs), H)
    Mn = M.cpu().numpy().astype(np.float64)
    for l in range(len(meshes)):
        e_ref = -api.MU0 / (2 * counts[l]) * float(np.sum(Mn[off[l]:off[l + 1]] * ref[off[l]:off[l + 1]]))
        assert abs(en[l].item() - e_ref) <= 1e-5 * abs(e_ref), (l, en[l].item(), e_ref)
    Ha = d.convolute(M, torch.full_like(M, 3.0), clear_out=False)
    assert torch.allclose(Ha, H + 3.0, rtol=1e-6, atol=1e-2)
    d.close()


def test_layer_size_without_a_plan(gpu):
    """A stack whose nx has no FFT plan (50): the layers on a padded common grid, the same
    field as the oracle's stack at N = 2n."""
    nx, ny, hz, z0 = 50, 14, [4e-9, 2e-9], [0.0, 6e-9]
    M = api.random_magnetisation(2 * nx * ny, 8e5, 13).astype(np.float32)
    d = api.MultiLayerDemag(nx, ny, 4e-9, 4e-9, hz, z0)
    H = d.convolute(torch.from_numpy(M).cuda()).cpu().numpy()
    ref = O.OracleMLayer(nx, ny, 4e-9, 4e-9, hz, z0)(M.astype(np.float64))
    assert rel_l2(H, ref) <= 1e-6
    d.close()


def test_fractional_offset_layers_refused_by_the_integrator(gpu):
    """Mesh layers (common grid, transfers, complex kernels) serve the convolution calls only:
    the integrator and the adjoint refuse such a handle with BM_ERR_SIZE."""
    d = api.MultiLayerMeshDemag([(16, 12, H4, H4, 4e-9, 0.0, 0.0, 0.0), (16, 12, H4, H4, 4e-9, 20.5 * H4, 0.0, 2e-9)])
    p = api.FMParams(Ms=8e5, A=1.3e-11, alpha=0.02, modules=api.BM_MOD_EXCHANGE)
    s = api.FMSolver(api.BM_EVAL_RK4, (36, 12, 2), (H4, H4, 4e-9), p)
    with pytest.raises(RuntimeError, match="error 2"):
        api._lib.check(s.L.bm_ode_set_layers(s.handle, d.handle, None, None, None, None))
    s.close(); d.close()
