"""Pins the oracle's demag convolution (kernels from the pinned tensor, FFT path).

* FFT convolution == direct O(n^2) real-space sum with the reference-pinned tensor;
* uniformly magnetised prism: cell-averaged field == Aharoni's closed-form demag factors;
* FP32 oracle within 1e-6 of FP64 on a realistic input (the reference-held vortex state);
* kernel DC term == sum of the tensor.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O


def direct_sum(n, N, h, M):
    nx, ny, nz = n
    if nz == 1:
        Dd, Dxy = O.tensor2d((nx, ny), (N[0], N[1]), h)
        Dd = Dd[None]
        Do = np.zeros(Dd.shape)
        Do[..., 0] = Dxy[None]
    else:
        Dd, Do = O.tensor3d(n, N, h)
    Mr = M.reshape(nz, ny, nx, 3)
    H = np.zeros_like(Mr)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                di = (i - np.arange(nx)) % N[0]
                dj = (j - np.arange(ny)) % N[1]
                dk = (k - np.arange(nz)) % N[2]
                d = Dd[np.ix_(dk, dj, di)]
                o = Do[np.ix_(dk, dj, di)]
                mx, my, mz = Mr[..., 0], Mr[..., 1], Mr[..., 2]
                H[k, j, i, 0] = (d[..., 0] * mx + o[..., 0] * my + o[..., 1] * mz).sum()
                H[k, j, i, 1] = (o[..., 0] * mx + d[..., 1] * my + o[..., 2] * mz).sum()
                H[k, j, i, 2] = (o[..., 1] * mx + o[..., 2] * my + d[..., 2] * mz).sum()
    return H.reshape(-1, 3)


@pytest.mark.parametrize("n,h,Nz", [((8, 4, 1), (4e-9, 4e-9, 3e-9), 1), ((6, 5, 3), (4e-9, 3e-9, 5e-9), 6),
                                    ((6, 4, 3), (4e-9, 4e-9, 4e-9), 8), ((4, 6, 5), (2e-9, 3e-9, 2.5e-9), 16)])
def test_fft_convolution_equals_direct_sum(n, h, Nz):
    rng = np.random.default_rng(sum(n))
    M = rng.normal(size=(n[0] * n[1] * n[2], 3)) * 8e5
    N = (2 * n[0], 2 * n[1], Nz)
    ref = direct_sum(n, N, h, M)
    H = O.OracleDemag(n, h, Nz_override=Nz if n[2] > 1 else 0)(M)
    assert np.linalg.norm(H - ref) / np.linalg.norm(ref) <= 1e-13
    H32 = O.OracleDemag(n, h, fp32=True, Nz_override=Nz if n[2] > 1 else 0)(M)
    assert np.linalg.norm(H32 - ref) / np.linalg.norm(ref) <= 5e-7


def aharoni_Dz(a, b, c):
    """Demag factor along z of a 2a x 2b x 2c prism (A. Aharoni, J. Appl. Phys. 83, 3432 (1998))."""
    R = np.sqrt(a * a + b * b + c * c)
    ab, bc, ac = np.sqrt(a * a + b * b), np.sqrt(b * b + c * c), np.sqrt(a * a + c * c)
    t = ((b * b - c * c) / (2 * b * c)) * np.log((R - a) / (R + a)) + ((a * a - c * c) / (2 * a * c)) * np.log((R - b) / (R + b)) \
        + (b / (2 * c)) * np.log((ab + a) / (ab - a)) + (a / (2 * c)) * np.log((ab + b) / (ab - b)) \
        + (c / (2 * a)) * np.log((bc - b) / (bc + b)) + (c / (2 * b)) * np.log((ac - a) / (ac + a)) \
        + 2 * np.arctan(a * b / (c * R)) + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c) \
        + ((a * a + b * b - 2 * c * c) / (3 * a * b * c)) * R + (c / (a * b)) * (ac + bc) - (ab ** 3 + bc ** 3 + ac ** 3) / (3 * a * b * c)
    return t / np.pi


@pytest.mark.parametrize("n,h", [((16, 8, 4), (4e-9, 4e-9, 4e-9)), ((20, 10, 1), (5e-9, 5e-9, 3e-9)),
                                 ((50, 12, 3), (2e-9, 3e-9, 2.5e-9)), ((128, 32, 1), (3.90625e-9, 3.90625e-9, 3e-9))])
def test_uniform_prism_matches_aharoni(n, h):
    a, b, c = n[0] * h[0] / 2, n[1] * h[1] / 2, n[2] * h[2] / 2
    N = {"z": aharoni_Dz(a, b, c), "x": aharoni_Dz(b, c, a), "y": aharoni_Dz(c, a, b)}
    d = O.OracleDemag(n, h)
    ncell = n[0] * n[1] * n[2]
    for comp, name in enumerate("xyz"):
        M = np.zeros((ncell, 3))
        M[:, comp] = 1.0
        H = d(M)
        assert abs(H[:, comp].mean() + N[name]) <= 1e-10, (name, H[:, comp].mean(), N[name])


def test_fp32_oracle_on_vortex_state():
    z = np.load(Path(__file__).parent / "golden" / "vortex_hh_clk_up.npz")
    M = z["M"].astype(np.float64)
    n, h = tuple(int(v) for v in z["n"]), tuple(float(v) for v in z["h"])
    H64 = O.OracleDemag(n, h)(M)
    H32 = O.OracleDemag(n, h, fp32=True)(M)
    assert np.linalg.norm(H32 - H64) / np.linalg.norm(H64) <= 1e-6


def test_kernel_dc_term_is_tensor_sum():
    n, N, h = (5, 4, 3), (10, 8, 6), (1.0, 1.0, 0.5)
    Dd, Do = O.tensor3d(n, N, h)
    Kd, Ko = O.kernels(n, N, h)
    np.testing.assert_allclose(Kd[0, 0, 0], Dd.sum(axis=(0, 1, 2)), rtol=1e-12)
    # odd components vanish at k = 0 (the kernel keeps only the symmetric part)
    assert np.abs(Ko[0, 0, 0]).max() <= 1e-12 * np.abs(Kd[0, 0, 0]).max()
    # full complex 3D FFT of each tensor component: real, and equal to the kept values
    for q in range(3):
        F = np.fft.fftn(Dd[..., q])
        assert np.abs(F.imag).max() <= 1e-12 * np.abs(F.real).max()
        np.testing.assert_allclose(Kd[..., q], F.real[:N[2] // 2 + 1, :N[1] // 2 + 1, :N[0] // 2 + 1], rtol=0,
                                   atol=1e-12 * np.abs(F).max())
        G = np.fft.fftn(Do[..., q])
        np.testing.assert_allclose(Ko[..., q], G.real[:N[2] // 2 + 1, :N[1] // 2 + 1, :N[0] // 2 + 1], rtol=0,
                                   atol=1e-12 * np.abs(G).max())
