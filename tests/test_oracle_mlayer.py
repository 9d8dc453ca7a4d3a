"""Pins the oracle's multilayered (stacked 2D layers) demag.

* z-shifted source->destination tensors, regular and irregular, near and far field:
  bit-identical to the reference's own Boris/DemagTFunc_Shifted.cpp / DemagTFunc_Irregular.cpp
  (golden vectors written by oracle/_ref/ref_dump, tests/golden/zshift_golden.npz);
* geometry, independent of the reference: a thick source layer equals the sum of thin
  slices and a thick destination averages thin slices (xx, yy at the slices' positions; zz
  and off-diagonal at the offset the reference's construction implies, see below); the
  regular z-shifted tensor of touching layers equals the 3D tensor's kz = +-1 planes;
* kernels: equal to the full complex 2D FFT of the tensor (the parts the product keeps,
  and zero where it assumes zero);
* multilayer convolution: equal to the direct real-space sum with the pinned tensors, and a
  stack of equal touching layers equal to the 3D oracle convolution of the same cells.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

GOLD = Path(__file__).resolve().parent / "golden" / "zshift_golden.npz"


def test_zshift_tensor_bit_identical_to_reference_golden():
    g = np.load(GOLD)
    for name in g["names"]:
        n, N = tuple(g[name + "_n"]), tuple(g[name + "_N"])
        Dd, Do = O.tensor2d_zshift(n, N, g[name + "_s"], g[name + "_d"], float(g[name + "_shift"]))
        idx = g[name + "_idx"]
        assert np.array_equal(Dd[idx[:, 0], idx[:, 1]], g[name + "_diag"]), name
        assert np.array_equal(Do[idx[:, 0], idx[:, 1]], g[name + "_off"]), name


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


# The reference's irregular tensor (Boris/DemagTFunc_Irregular.cpp with the grids of
# Boris/DemagTFunc_fg.cpp:507-620) places the xx and yy basis-function planes at
# {-s.z, 0, +d.z} + shift but the zz and off-diagonal planes at {-d.z, 0, +s.z} + shift
# (del = s.z - d.z enters with opposite signs).  With shift = destination bottom - source
# bottom, xx and yy are exactly the sum (average) over thin slices of a thick source
# (destination) at their true positions; zz, xy, xz and yz equal the same sums taken at
# shift + del.  Both are exact identities, so they pin the 36-term stencils of the
# restatement independently of the golden vectors.
@pytest.mark.parametrize("S,sh", [(1.0, 0.5), (1.0, 1.0), (1.5, -1.5), (1.5, 2.25)])
def test_thick_source_equals_sum_of_thin_slices(S, sh):
    n, N = (6, 5), (12, 10)
    c = 0.5
    thin, thick = (0.8, 0.6, c), (0.8, 0.6, S)
    m, dl = int(round(S / c)), S - c
    Dd, Do = O.tensor2d_zshift(n, N, thick, thin, sh)
    A = sum(O.tensor2d_zshift(n, N, thin, thin, sh - q * c)[0] for q in range(m))
    B = [sum(O.tensor2d_zshift(n, N, thin, thin, sh + dl - q * c)[p] for q in range(m)) for p in (0, 1)]
    for c_ in (0, 1):
        assert _rel(Dd[..., c_], A[..., c_]) <= 1e-9
    assert _rel(Dd[..., 2], B[0][..., 2]) <= 1e-9
    assert _rel(Do, B[1]) <= 1e-9


@pytest.mark.parametrize("S,sh", [(1.0, -0.5), (1.0, 1.0), (1.5, -2.0), (1.5, 3.5)])
def test_thick_destination_averages_thin_slices(S, sh):
    n, N = (6, 5), (12, 10)
    c = 0.5
    thin, thick = (0.8, 0.6, c), (0.8, 0.6, S)
    m, dl = int(round(S / c)), c - S
    Dd, Do = O.tensor2d_zshift(n, N, thin, thick, sh)
    A = sum(O.tensor2d_zshift(n, N, thin, thin, sh + q * c)[0] for q in range(m)) / m
    B = [sum(O.tensor2d_zshift(n, N, thin, thin, sh + dl + q * c)[p] for q in range(m)) / m for p in (0, 1)]
    for c_ in (0, 1):
        assert _rel(Dd[..., c_], A[..., c_]) <= 1e-9
    assert _rel(Dd[..., 2], B[0][..., 2]) <= 1e-9
    assert _rel(Do, B[1]) <= 1e-9


@pytest.mark.parametrize("n", [(7, 5), (45, 6)])
def test_regular_zshift_equals_3d_tensor_planes(n):
    h = (4e-9, 3e-9, 5e-9)
    hm = max(h)
    hr = tuple(v / hm for v in h)
    N3 = (2 * n[0], 2 * n[1], 4)
    Dd3, Do3 = O.tensor3d((n[0], n[1], 2), N3, h)
    # the +1 plane evaluates the same basis-function values in the same order: identical;
    # the -1 plane is the 3D tensor's mirror image, here evaluated at negative z: equal up
    # to the cancellation of the 27-term stencil far from the source (a few 1e-12 of max)
    Dd, Do = O.tensor2d_zshift(n, N3[:2], hr, hr, hr[2])
    assert np.array_equal(Dd, Dd3[1]) and np.array_equal(Do, Do3[1])
    Dd, Do = O.tensor2d_zshift(n, N3[:2], hr, hr, -hr[2])
    assert _rel(Dd, Dd3[3]) <= 1e-11 and _rel(Do, Do3[3]) <= 1e-11


def test_zshift_kernels_equal_full_fft_of_tensor():
    n, N = (9, 6), (18, 12)
    Dd, Do = O.tensor2d_zshift(n, N, (0.5, 0.5, 1.0), (0.5, 0.5, 0.5), 0.75)
    Kd, Ko = O.kernels2d_zshift(N, Dd, Do)
    nc, nyh = N[0] // 2 + 1, N[1] // 2 + 1
    F = [np.fft.fft2(Dd[..., c])[:nyh, :nc] for c in range(3)]
    G = [np.fft.fft2(Do[..., c])[:nyh, :nc] for c in range(3)]
    scale = max(np.abs(x).max() for x in F + G)
    for c in range(3):
        assert np.abs(Kd[..., c] - F[c].real).max() <= 1e-13 * scale
        assert np.abs(F[c].imag).max() <= 1e-13 * scale
    # xy is real; xz and yz are i times the stored values
    assert np.abs(Ko[..., 0] - G[0].real).max() <= 1e-13 * scale
    assert np.abs(Ko[..., 1] - G[1].imag).max() <= 1e-13 * scale
    assert np.abs(Ko[..., 2] - G[2].imag).max() <= 1e-13 * scale
    for c in range(3):
        assert np.abs(G[c].imag if c == 0 else G[c].real).max() <= 1e-13 * scale


# parity of the stored components along x and y: diagonal even; xy odd, odd; xz odd, even;
# yz even, odd
_PAR_D = np.array([[1, 1], [1, 1], [1, 1]])
_PAR_O = np.array([[-1, -1], [-1, 1], [1, -1]])


def _even_odd_parts(Dd, Do, pbc):
    """Along a periodic axis the image window k n + b (k = -p..p) is not symmetric under
    b -> n - b, and the reference's real kernels (Boris/DemagKernelCollection_Calc.cpp:288-420
    keeps the real or imaginary part of the tensor's FFT per axis) hold only the even or odd
    part of the tensor along it: (D(b) + s D(n - b)) / 2."""
    Dd, Do = Dd.copy(), Do.copy()
    for ax, npx in ((1, pbc[0]), (0, pbc[1])):          # array axes: (y, x, component)
        if not npx:
            continue
        col = 0 if ax == 1 else 1
        for D, par in ((Dd, _PAR_D), (Do, _PAR_O)):
            n = D.shape[ax]
            flip = np.take(D, (-np.arange(n)) % n, axis=ax)
            D[...] = (D + par[:, col] * flip) / 2
    return Dd, Do


def direct_sum_mlayer(nx, ny, hx, hy, hz, z0, M, pbc=(0, 0), offsets=None):
    """Real-space sum over the pinned tensors: cyclic over N (N = n along a periodic axis,
    where the tensors carry the images; 2n otherwise, where the padding keeps it linear)."""
    L = len(hz)
    hmax = max(max(hx, hy), max(hz))
    N = (nx if pbc[0] else 2 * nx, ny if pbc[1] else 2 * ny)
    periodic = any(pbc)
    Mr = M.reshape(L, ny, nx, 3)
    H = np.zeros_like(Mr)
    for d in range(L):
        for s in range(L):
            if d == s:
                if periodic:
                    hr = (hx / hmax, hy / hmax, hz[d] / hmax)
                    Dd, Dxy = O.tensor_pbc((N[0], N[1], 1), hr, (pbc[0], pbc[1], 0))
                else:
                    Dd, Dxy = O.tensor2d((nx, ny), N, (hx, hy, hz[d]))
                Do = np.zeros(Dd.shape)
                Do[..., 0] = Dxy
            else:
                sc = (hx / hmax, hy / hmax, hz[s] / hmax)
                dc = (hx / hmax, hy / hmax, hz[d] / hmax)
                off = (0.0, 0.0) if offsets is None else ((offsets[d][0] - offsets[s][0]) / hmax, (offsets[d][1] - offsets[s][1]) / hmax)
                if off != (0.0, 0.0):
                    Dd, Do = O.tensor2d_shifted((nx, ny), N, sc, dc, (off[0], off[1], (z0[d] - z0[s]) / hmax))
                elif periodic:
                    Dd, Do = O.tensor2d_zshift_pbc(N, sc, dc, (z0[d] - z0[s]) / hmax, *pbc)
                else:
                    Dd, Do = O.tensor2d_zshift((nx, ny), N, sc, dc, (z0[d] - z0[s]) / hmax)
            if periodic:
                Dd, Do = _even_odd_parts(Dd, Do, pbc)
            mx, my, mz = Mr[s, ..., 0], Mr[s, ..., 1], Mr[s, ..., 2]
            for j in range(ny):
                for i in range(nx):
                    di = (i - np.arange(nx)) % N[0]
                    dj = (j - np.arange(ny)) % N[1]
                    dd = Dd[np.ix_(dj, di)]
                    oo = Do[np.ix_(dj, di)]
                    H[d, j, i, 0] += (dd[..., 0] * mx + oo[..., 0] * my + oo[..., 1] * mz).sum()
                    H[d, j, i, 1] += (oo[..., 0] * mx + dd[..., 1] * my + oo[..., 2] * mz).sum()
                    H[d, j, i, 2] += (oo[..., 1] * mx + oo[..., 2] * my + dd[..., 2] * mz).sum()
    return H.reshape(-1, 3)


def test_mlayer_convolution_equals_direct_sum():
    nx, ny, hx, hy = 6, 5, 4e-9, 4e-9
    hz, z0 = [8e-9, 4e-9, 16e-9], [0.0, 9e-9, 15e-9]
    rng = np.random.default_rng(5)
    M = rng.normal(size=(3 * nx * ny, 3)) * 8e5
    ref = direct_sum_mlayer(nx, ny, hx, hy, hz, z0, M)
    H = O.OracleMLayer(nx, ny, hx, hy, hz, z0)(M)
    assert np.linalg.norm(H - ref) / np.linalg.norm(ref) <= 1e-13


def test_stack_of_touching_layers_equals_3d_convolution():
    nx, ny, L, h = 8, 6, 4, 4e-9
    rng = np.random.default_rng(6)
    M = rng.normal(size=(L * nx * ny, 3)) * 8e5
    H = O.OracleMLayer(nx, ny, h, h, [h] * L, [l * h for l in range(L)])(M)
    ref = O.OracleDemag((nx, ny, L), (h, h, h), Nz_override=8)(M)
    assert np.linalg.norm(H - ref) / np.linalg.norm(ref) <= 1e-12


# ---------------------------------------------------------------------------------------
# Layers as meshes of their own (the reference's common grid and mesh transfer).

def test_mesh_layers_on_the_common_grid_are_the_plain_multilayer():
    """Layers that already are the common grid: identity transfers, the plain multilayer."""
    meshes = [(12, 8, 4e-9, 4e-9, 4e-9, 0.0, 0.0, 0.0), (12, 8, 4e-9, 4e-9, 2e-9, 0.0, 0.0, 6e-9)]
    o = O.OracleMLayerMeshes(meshes)
    for T in o.Tin + o.Tout:
        assert np.array_equal(T, np.eye(T.shape[0]))
    M = np.random.default_rng(1).normal(size=(2 * 96, 3)) * 8e5
    ref = O.OracleMLayer(12, 8, 4e-9, 4e-9, [4e-9, 2e-9], [0.0, 6e-9])(M)
    assert np.array_equal(o(M), ref)


def test_transfer_weights_partition_and_closed_forms():
    """Out weights of every layer cell sum to 1; in weights of a common cell sum to its
    covered volume ratio; a 2x2 block of common cells takes weights 1/4 each, a 3x3 block
    weights (d_max - d) / sum with d in {0, h, sqrt(2) h} and d_max = sqrt(8 h^2 + hz^2)."""
    h, hz = 4e-9, 3e-9
    # layer 1: 4 x 4 cells of 3h over layer 0's 12 x 12 of h (same extent); layer 2: 6 x 6 of 2h
    meshes = [(12, 12, h, h, hz, 0.0, 0.0, 0.0), (4, 4, 3 * h, 3 * h, hz, 0.0, 0.0, 2 * hz), (6, 6, 2 * h, 2 * h, hz, 0.0, 0.0, 4 * hz)]
    o = O.OracleMLayerMeshes(meshes)
    for T in o.Tout:
        np.testing.assert_allclose(T.sum(axis=1), 1.0, rtol=1e-14)
    for T in o.Tin:
        s = T.sum(axis=1)
        assert np.all((np.abs(s - 1.0) < 1e-12) | (s == 0.0))
    # 2 x 2 (layer 2): exactly 1/4
    w2 = o.Tout[2][0][o.Tout[2][0] > 0]
    assert len(w2) == 4 and np.allclose(w2, 0.25, rtol=0, atol=1e-15)
    # 3 x 3 (layer 1), first coarse cell: centre, edges, corners
    dmax = np.sqrt(8 * h * h + hz * hz)
    raw = np.array([[dmax - np.hypot(a * h, b * h) for a in (-1, 0, 1)] for b in (-1, 0, 1)])
    T = o.Tout[1][0].reshape(12, 12)[:3, :3]
    np.testing.assert_allclose(T, raw / raw.sum(), rtol=1e-13)


def test_coarse_layer_field_is_the_average_of_the_fine_computation():
    """A layer of 2x coarser cells: its field is the mean over each 2x2 block of the field the
    same layer gets when given at the fine cell with its M repeated (Newell tensors are cell
    averages, so this is exact physics, not an approximation of the transfer)."""
    h, hz = 4e-9, 4e-9
    meshes = [(8, 6, h, h, hz, 0.0, 0.0, 0.0), (4, 3, 2 * h, 2 * h, 2 * hz, 0.0, 0.0, 6e-9)]
    rng = np.random.default_rng(5)
    M0, M1 = rng.normal(size=(48, 3)) * 8e5, rng.normal(size=(12, 3)) * 6e5
    H = O.OracleMLayerMeshes(meshes)(np.concatenate([M0, M1]))
    M1f = np.repeat(np.repeat(M1.reshape(3, 4, 3), 2, axis=0), 2, axis=1).reshape(-1, 3)
    Hf = O.OracleMLayer(8, 6, h, h, [hz, 2 * hz], [0.0, 6e-9])(np.concatenate([M0, M1f]))
    H1f = Hf[48:].reshape(6, 8, 3)
    avg = H1f.reshape(3, 2, 4, 2, 3).mean(axis=(1, 3)).reshape(-1, 3)
    np.testing.assert_allclose(H[48:], avg, rtol=1e-12, atol=1e-9 * np.abs(avg).max())
    np.testing.assert_allclose(H[:48], Hf[:48], rtol=1e-12, atol=1e-9 * np.abs(Hf[:48]).max())


def test_mesh_layers_translation_invariance():
    """Moving every layer by whole common cells changes nothing; a small layer inside a large
    one and side by side (convolution rectangles at different places)."""
    h = 4e-9
    base = [(10, 8, h, h, 3e-9, 0.0, 0.0, 0.0), (4, 4, h, h, 3e-9, 3 * h, 2 * h, 5e-9), (6, 4, h, h, 3e-9, 14 * h, 1 * h, 5e-9)]
    M = np.random.default_rng(7).normal(size=(80 + 16 + 24, 3)) * 8e5
    H0 = O.OracleMLayerMeshes(base)(M)
    moved = [(a, b, c, d, e, x + 5 * h, y - 3 * h, z) for (a, b, c, d, e, x, y, z) in base]
    H1 = O.OracleMLayerMeshes(moved)(M)
    np.testing.assert_allclose(H1, H0, rtol=1e-9, atol=1e-9 * np.abs(H0).max())


GOLD_PBC = Path(__file__).resolve().parent / "golden" / "zshift_pbc_golden.npz"


def test_periodic_zshift_tensor_bit_identical_to_reference_golden():
    """z-shifted tensors with x and y images against the reference's
    Boris/DemagTFunc_Shifted_PBC.cpp / DemagTFunc_Irregular_PBC.cpp (oracle/_ref/ref_dump)."""
    g = np.load(GOLD_PBC)
    for name in g["names"]:
        N, (px, py) = tuple(g[name + "_N"]), g[name + "_img"]
        Dd, Do = O.tensor2d_zshift_pbc(N, g[name + "_s"], g[name + "_d"], float(g[name + "_shift"]), px, py)
        idx = g[name + "_idx"]
        assert np.array_equal(Dd[idx[:, 0], idx[:, 1]], g[name + "_diag"]), name
        assert np.array_equal(Do[idx[:, 0], idx[:, 1]], g[name + "_off"]), name


@pytest.mark.parametrize("n,img", [((8, 6), (2, 0)), ((6, 10), (0, 3)), ((12, 12), (2, 2)), ((48, 8), (1, 1))])
def test_periodic_regular_zshift_equals_3d_periodic_tensor_plane(n, img):
    """Touching layers of one thickness: the periodic z-shifted tensor is the kz = 1 plane of
    the 3D periodic tensor with the same x, y images (N = n along a periodic axis)."""
    h = (4e-9, 3e-9, 5e-9)
    hm = max(h)
    hr = tuple(v / hm for v in h)
    N3 = (n[0] if img[0] else 2 * n[0], n[1] if img[1] else 2 * n[1], 4)
    Dd3, Do3 = O.tensor_pbc(N3, hr, (img[0], img[1], 0))
    Dd, Do = O.tensor2d_zshift_pbc(N3[:2], hr, hr, hr[2], *img)
    assert _rel(Dd, Dd3[1]) <= 1e-12 and _rel(Do, Do3[1]) <= 1e-12


def test_periodic_zshift_source_thickness_additivity():
    """A source layer of thickness 2h acts as two stacked layers of thickness h with the images
    summed (near field; the same plane placement as the open-boundary identity above: xx, yy
    at the true slice positions, zz and the off-diagonal terms at shift + del)."""
    N, img = (10, 8), (1, 2)            # every image within the near-field distance
    c, S, sh = 0.5, 1.0, 0.75
    thin, thick = (0.5, 0.5, c), (0.5, 0.5, S)
    m, dl = int(round(S / c)), S - c
    Dd, Do = O.tensor2d_zshift_pbc(N, thick, thin, sh, *img)
    A = sum(O.tensor2d_zshift_pbc(N, thin, thin, sh - q * c, *img)[0] for q in range(m))
    B = [sum(O.tensor2d_zshift_pbc(N, thin, thin, sh + dl - q * c, *img)[p] for q in range(m)) for p in (0, 1)]
    for c_ in (0, 1):
        assert _rel(Dd[..., c_], A[..., c_]) <= 1e-9
    assert _rel(Dd[..., 2], B[0][..., 2]) <= 1e-9
    assert _rel(Do, B[1]) <= 1e-9



@pytest.mark.parametrize("pbc", [(2, 0), (0, 3), (2, 2)])
def test_periodic_mlayer_convolution_equals_direct_sum(pbc):
    """Periodic stacks: the FFT convolution with the periodic self and pair kernels equals the
    cyclic real-space sum over the pinned periodic tensors (N = n along a periodic axis)."""
    nx, ny, hx, hy = 8, 6, 4e-9, 4e-9
    hz, z0 = [8e-9, 4e-9, 16e-9], [0.0, 9e-9, 15e-9]
    rng = np.random.default_rng(7)
    M = rng.normal(size=(3 * nx * ny, 3)) * 8e5
    ref = direct_sum_mlayer(nx, ny, hx, hy, hz, z0, M, pbc)
    H = O.OracleMLayer(nx, ny, hx, hy, hz, z0, pbc=pbc)(M)
    assert np.linalg.norm(H - ref) / np.linalg.norm(ref) <= 1e-13
    # and the images change the field (else this would test the open stack again)
    H0 = O.OracleMLayer(nx, ny, hx, hy, hz, z0)(M)
    assert np.linalg.norm(H - H0) / np.linalg.norm(H0) > 1e-3


@pytest.mark.parametrize("pbc", [(2, 0), (1, 3)])
def test_periodic_stack_of_touching_layers_equals_3d_periodic_convolution(pbc):
    """L touching layers of one thickness with x, y images = the 3D periodic convolution of an
    nz = L mesh with the same images (open along z)."""
    nx, ny, L, h = 8, 6, 4, 4e-9
    rng = np.random.default_rng(8)
    M = rng.normal(size=(L * nx * ny, 3)) * 8e5
    H = O.OracleMLayer(nx, ny, h, h, [h] * L, [l * h for l in range(L)], pbc=pbc)(M)
    ref = O.OracleDemag((nx, ny, L), (h, h, h), pbc=(pbc[0], pbc[1], 0))(M)
    # measured 1.4e-12 and 3.0e-12: the two stencils' roundoff summed over the images
    assert np.linalg.norm(H - ref) / np.linalg.norm(ref) <= 1e-11


GOLD_SHIFT = Path(__file__).resolve().parent / "golden" / "shift_golden.npz"


def test_inplane_shifted_tensor_bit_identical_to_reference_golden():
    """Tensors shifted in-plane (by fractions of a cell) and along z against the reference's
    general branch of DemagTFunc_Shifted.cpp / DemagTFunc_Irregular.cpp (oracle/_ref)."""
    g = np.load(GOLD_SHIFT)
    for name in g["names"]:
        n, N = tuple(g[name + "_n"]), tuple(g[name + "_N"])
        Dd, Do = O.tensor2d_shifted(n, N, g[name + "_s"], g[name + "_d"], g[name + "_shift"])
        idx = g[name + "_idx"]
        assert np.array_equal(Dd[idx[:, 0], idx[:, 1]], g[name + "_diag"]), name
        assert np.array_equal(Do[idx[:, 0], idx[:, 1]], g[name + "_off"]), name


@pytest.mark.parametrize("cells", [((0.5, 0.5, 1.0), (0.5, 0.5, 1.0)), ((0.5, 0.5, 1.0), (0.5, 0.5, 0.5))], ids=["reg", "irr"])
def test_whole_cell_shift_is_a_translated_zshift_tensor(cells):
    """An in-plane shift of whole cells is the z-shifted tensor read at translated offsets (the
    embedding of offset layers in one box), to the rounding of the shifted positions."""
    s, d = cells
    n, N = (12, 10), (24, 20)
    Dz, Oz = O.tensor2d_zshift(n, N, s, d, 0.75)
    di, dj = 3, -2
    Ds, Os = O.tensor2d_shifted((9, 8), N, s, d, (di * d[0], dj * d[1], 0.75))
    J, I = np.meshgrid(np.arange(-7, 8), np.arange(-8, 9), indexing="ij")
    for X, Z in ((Ds, Dz), (Os, Oz)):
        A = X[J % N[1], I % N[0]]
        B = Z[(J + dj) % N[1], (I + di) % N[0]]
        # the 27/36-term stencils cancel: 1e-11 of the tensor's scale where positions round apart
        assert np.abs(A - B).max() <= 1e-9 * np.abs(Z).max()


def test_box_residual_shift_equals_full_offset_tensor():
    """Layers offset by Delta + r cells: the box embedding evaluates the tensor with the residual
    shift r at box offsets J, the reference with the full shift at mesh offsets J - Delta; the
    same physical positions, so equal to the rounding of the positions (near and far field)."""
    s = d = (0.5, 0.5, 0.5)
    N = (100, 90)
    r, D = (0.15, -0.2), (7, 4)
    Tb, Ob = O.tensor2d_shifted((50, 45), N, s, d, (r[0] * 0.5, r[1] * 0.5, 1.0))
    Tf, Of = O.tensor2d_shifted((43, 41), N, s, d, ((D[0] + r[0]) * 0.5, (D[1] + r[1]) * 0.5, 1.0))
    J, I = np.meshgrid(np.arange(-40, 41), np.arange(-42, 43), indexing="ij")
    for X, Z in ((Tb, Tf), (Ob, Of)):
        A = X[(J + D[1]) % N[1], (I + D[0]) % N[0]]
        B = Z[J % N[1], I % N[0]]
        # measured 1.2e-9 (diagonal) and 3.0e-8 (off-diagonal) of the scale: the near-field
        # stencils' cancellation 25-40 cells out, evaluated at positions that round apart
        assert np.abs(A - B).max() <= 1e-7 * np.abs(Z).max()


def test_complex_kernels_are_the_2d_fft_of_the_tensor():
    N = (16, 12)
    Dd, Do = O.tensor2d_shifted((8, 6), N, (1.0, 0.8, 1.0), (1.0, 0.8, 0.5), (0.4, -0.25, 0.75))
    K = O.kernels2d_complex(N, Dd, Do)
    D = np.concatenate([Dd, Do], axis=-1)
    F = np.fft.fft2(D, axes=(0, 1))[:, :N[0] // 2 + 1]
    assert np.abs(K - F).max() <= 1e-13 * np.abs(F).max()



def test_offset_layers_convolution_equals_direct_sum():
    """Layers whose grids are offset in-plane by fractions of a cell: pairs couple through the
    in-plane shifted tensor's complex kernels (no symmetry); the FFT convolution equals the
    real-space sum over the pinned shifted tensors, and its transpose is the adjoint."""
    nx, ny, hx, hy = 8, 6, 4e-9, 4e-9
    hz, z0 = [8e-9, 4e-9, 16e-9], [0.0, 9e-9, 15e-9]
    offsets = [(0.0, 0.0), (1.3e-9, -0.7e-9), (0.0, 2.1e-9)]
    rng = np.random.default_rng(11)
    M = rng.normal(size=(3 * nx * ny, 3)) * 8e5
    ref = direct_sum_mlayer(nx, ny, hx, hy, hz, z0, M, offsets=offsets)
    ml = O.OracleMLayer(nx, ny, hx, hy, hz, z0, offsets=offsets)
    H = ml(M)
    assert np.linalg.norm(H - ref) / np.linalg.norm(ref) <= 1e-13
    H0 = O.OracleMLayer(nx, ny, hx, hy, hz, z0)(M)
    assert np.linalg.norm(H - H0) / np.linalg.norm(H0) > 1e-4
    v = rng.normal(size=M.shape)
    assert abs(np.vdot(v, H) - np.vdot(ml.transposed(v), M)) <= 1e-12 * abs(np.vdot(v, H))


def test_fractional_mesh_offsets_equal_reference_formulation():
    """Meshes of equal cells offset in-plane by fractions of a cell (the transfer is then the
    identity): the box embedding with residual shifts equals the reference's formulation, every
    pair convolved on its own n_common grid through the tensor shifted by the full rectangle
    offset (Boris/DemagKernelCollectionCUDA_Calc.cpp:70-104, Complex_Full kernels)."""
    h = 4e-9
    meshes = [(8, 6, h, h, 8e-9, 0.0, 0.0, 0.0), (8, 6, h, h, 4e-9, 1.3e-9, -2.2e-9, 10e-9),
              (8, 6, h, h, 4e-9, 5.1e-9, 0.6e-9, 16e-9)]
    rng = np.random.default_rng(12)
    M = rng.normal(size=(3 * 48, 3)) * 8e5
    H = O.OracleMLayerMeshes(meshes)(M)
    hmax = 8e-9
    nx, ny = 8, 6
    N = (2 * nx, 2 * ny)
    Mr = M.reshape(3, ny, nx, 3)
    ref = np.zeros_like(Mr)
    for d, md in enumerate(meshes):
        for s, ms in enumerate(meshes):
            sc, dc = (h / hmax, h / hmax, ms[4] / hmax), (h / hmax, h / hmax, md[4] / hmax)
            if d == s:
                Dd, Dxy = O.tensor2d((nx, ny), N, (h, h, md[4]))
                Do = np.zeros(Dd.shape); Do[..., 0] = Dxy
            else:
                sh = ((md[5] - ms[5]) / hmax, (md[6] - ms[6]) / hmax, (md[7] - ms[7]) / hmax)
                Dd, Do = O.tensor2d_shifted((nx, ny), N, sc, dc, sh)
            mx, my, mz = Mr[s, ..., 0], Mr[s, ..., 1], Mr[s, ..., 2]
            for j in range(ny):
                for i in range(nx):
                    dd = Dd[np.ix_((j - np.arange(ny)) % N[1], (i - np.arange(nx)) % N[0])]
                    oo = Do[np.ix_((j - np.arange(ny)) % N[1], (i - np.arange(nx)) % N[0])]
                    ref[d, j, i, 0] += (dd[..., 0] * mx + oo[..., 0] * my + oo[..., 1] * mz).sum()
                    ref[d, j, i, 1] += (oo[..., 0] * mx + dd[..., 1] * my + oo[..., 2] * mz).sum()
                    ref[d, j, i, 2] += (oo[..., 1] * mx + oo[..., 2] * my + dd[..., 2] * mz).sum()
    ref = ref.reshape(-1, 3)
    assert np.linalg.norm(H - ref) / np.linalg.norm(ref) <= 1e-10
