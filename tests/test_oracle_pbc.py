"""Oracle periodic boundary conditions (Boris/DemagTFunc_PBC.cpp, N = n along periodic axes,
Boris/ConvolutionDataCUDA.cpp:56-90):

* the periodic tensors equal, bit for bit, the reference's own DemagTFunc_PBC.cpp compiled
  here (committed golden vectors: 2D and 3D, images along each axis, far-field images);
* the FFT convolution with N = n equals a direct real-space sum over source cells, with the
  periodic tensor rebuilt from the pinned open-boundary tensor as the sum over the image
  offsets b + k n (k = -p..p).  That window is not symmetric under b -> n - b, and the
  reference's real kernel (real or imaginary part of the tensor's FFT per axis,
  Boris/DemagKernel.cpp:388-468) keeps only its even or odd part along each periodic axis:
  the direct sum therefore uses (D(b) + s D(n - b)) / 2, s = +1 for a component even along
  that axis and -1 for an odd one.  This pins the wrap-around packing, the kernel extraction
  and the scaling;
* trace identity: the images only add mutual terms, whose trace is zero, so the trace of the
  periodic tensor is -1 at offset 0 and 0 elsewhere;
* SelfDemag_PBC (the evaluation speedup's self term with images) equals the reference's bit
  for bit, and reduces to the plain self term without images;
* the integrator with periodic boundaries: exchange of a uniform state is zero and a
  periodic cosine is an eigenvector with eigenvalue (2 cos(2 pi / n) - 2) / h^2 at every
  cell (boundary cells included); a periodic mesh advanced from a state shifted by whole
  cells equals the shifted result (translation invariance, demag and exchange on).
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = np.load(Path(__file__).parent / "golden" / "pbc_golden.npz")
CASES = sorted(k[:-2] for k in GOLDEN.files if k.endswith("_N"))


@pytest.mark.parametrize("name", CASES)
def test_pbc_tensor_matches_reference(name):
    N = tuple(int(v) for v in GOLDEN[name + "_N"])
    h = GOLDEN[name + "_h"]
    hr = tuple(float(v) for v in h / h.max())
    img = tuple(int(v) for v in GOLDEN[name + "_img"])
    Dd, Do = O.tensor_pbc(N, hr, img)
    idx = GOLDEN[name + "_idx"]
    if N[2] == 1:
        Dd, Do = Dd[None], Do[None, ..., None]
    assert np.array_equal(Dd[idx[:, 0], idx[:, 1], idx[:, 2]], GOLDEN[name + "_diag"])
    assert np.array_equal(Do[idx[:, 0], idx[:, 1], idx[:, 2]], GOLDEN[name + "_off"])


# parity of each stored component (xx, yy, zz, xy, xz, yz) along x, y, z: +1 even, -1 odd
PARITY = np.array([[1, 1, 1], [1, 1, 1], [1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]])


def periodic_tensor_from_images(n, h, pbc):
    """(6, Nz, Ny, Nx) periodic tensor on the FFT grid: image sums of the open-boundary tensor
    along periodic axes, then the even/odd part per periodic axis."""
    N = [n[a] if pbc[a] else 2 * n[a] for a in range(3)]
    if n[2] == 1:
        N[2] = 1
    big = [(pbc[a] + 1) * n[a] if pbc[a] else n[a] for a in range(3)]
    if n[2] == 1:
        big[2] = pbc[2] + 1                                    # 2D: z images one cell thick
    hm = max(h)
    Nb = [2 * b for b in big]
    Dd, Do = O.tensor3d(tuple(big), tuple(Nb), tuple(v / hm for v in h))
    T = np.concatenate([np.moveaxis(Dd, -1, 0), np.moveaxis(Do, -1, 0)])   # (6, Nbz, Nby, Nbx)
    axes_off = []
    for a in range(3):
        if n[a] > 1 or a < 2:
            if pbc[a]:
                axes_off.append([np.arange(N[a])[:, None] + k * n[a] for k in range(-pbc[a], pbc[a] + 1)])
            else:
                o = np.arange(N[a]); o = np.where(o < N[a] // 2, o, o - N[a]); o[N[a] // 2] = 10 ** 6
                axes_off.append([o[:, None]])
        else:
            axes_off.append([np.array([[k]]) for k in range(-pbc[2], pbc[2] + 1)])
    D = np.zeros((6, N[2], N[1], N[0]))
    for ox in axes_off[0]:
        for oy in axes_off[1]:
            for oz in axes_off[2]:
                ix, iy, iz = ox[:, 0], oy[:, 0], oz[:, 0]
                valid = (np.abs(iz)[:, None, None] < big[2]) & (np.abs(iy)[None, :, None] < big[1]) & (np.abs(ix)[None, None, :] < big[0])
                vals = T[:, np.mod(iz, Nb[2])][:, :, np.mod(iy, Nb[1])][:, :, :, np.mod(ix, Nb[0])]
                D += np.where(valid[None], vals, 0.0)
    for a in range(3):
        if pbc[a] and (a < 2 or n[2] > 1):
            ax = 3 - a
            rev = np.take(D, np.mod(-np.arange(N[a]), N[a]), axis=ax)
            D = 0.5 * (D + PARITY[:, a].reshape(6, 1, 1, 1) * rev)
    return D, N


@pytest.mark.parametrize("n,h,pbc", [((8, 6, 1), (4e-9, 4e-9, 4e-9), (3, 0, 0)),
                                     ((6, 8, 1), (4e-9, 3e-9, 6e-9), (2, 2, 0)),
                                     ((8, 4, 1), (4e-9, 4e-9, 5e-9), (0, 3, 2)),
                                     ((4, 6, 4), (4e-9, 4e-9, 4e-9), (2, 0, 0)),
                                     ((6, 4, 4), (3e-9, 4e-9, 2e-9), (0, 2, 3)),
                                     ((4, 4, 4), (4e-9, 4e-9, 4e-9), (1, 1, 1))])
def test_pbc_convolution_equals_direct_sum(n, h, pbc):
    ncell = n[0] * n[1] * n[2]
    M = np.random.default_rng(sum(n) + sum(pbc)).normal(size=(ncell, 3)) * 8e5
    H = O.OracleDemag(n, h, pbc=pbc)(M)
    D, N = periodic_tensor_from_images(n, h, pbc)
    cz, cy, cx = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    pos = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    d = pos[:, None, :] - pos[None, :, :]                       # target - source
    ix, iy, iz = np.mod(d[..., 0], N[0]), np.mod(d[..., 1], N[1]), np.mod(d[..., 2], N[2])
    xx, yy, zz, xy, xz, yz = (D[c][iz, iy, ix] for c in range(6))
    if n[2] == 1:
        xz = yz = np.zeros_like(xx)                             # 2D: only Dxy off the diagonal
    Mx, My, Mz = M[None, :, 0], M[None, :, 1], M[None, :, 2]
    ref = np.stack([(xx * Mx + xy * My + xz * Mz).sum(1), (xy * Mx + yy * My + yz * Mz).sum(1),
                    (xz * Mx + yz * My + zz * Mz).sum(1)], axis=1)
    e = np.linalg.norm(H - ref) / np.linalg.norm(ref)
    assert e <= 1e-12, e
    # the periodic images matter at this size: the open-boundary field is far off
    H0 = O.OracleDemag(n, h)(M)
    assert np.linalg.norm(H0 - ref) / np.linalg.norm(ref) > 1e-3


@pytest.mark.parametrize("N,h,pbc", [((8, 12, 1), (1.0, 1.0, 1.0), (3, 0, 0)),
                                     ((16, 6, 1), (0.8, 0.6, 1.0), (0, 4, 2)),
                                     ((8, 10, 4), (1.0, 1.0, 1.0), (2, 2, 0)),
                                     ((12, 6, 8), (0.75, 1.0, 0.5), (0, 3, 4))])
def test_pbc_tensor_trace(N, h, pbc):
    Dd, _ = O.tensor_pbc(N, h, pbc)
    tr = Dd.sum(-1)
    if N[2] == 1:
        tr = tr[None]
    assert abs(tr[0, 0, 0] + 1.0) <= 1e-10     # cancellation in the Newell sums: ~1e-12
    tr[0, 0, 0] = 0.0
    # offsets that no image reaches (the zero-padded half of an open axis) are zero anyway
    assert np.abs(tr).max() <= 1e-10


def test_self_demag_pbc_matches_reference():
    for row, val in zip(GOLDEN["selfpbc_in"], GOLDEN["selfpbc_val"]):
        h, n, img = tuple(row[:3]), tuple(int(v) for v in row[3:6]), tuple(int(v) for v in row[6:])
        assert np.array_equal(O.self_demag_pbc(h, n, img), val), (h, n, img)
    h = (4e-9, 3e-9, 2e-9)
    plain = np.zeros(3)
    O.lib().or_self_demag(*h, plain)
    assert np.array_equal(O.self_demag_pbc(h, (8, 8, 1), (0, 0, 0)), plain)


def sim_params(**kw):
    from boris2_amd import api
    base = dict(Ms=8e5, A=1.3e-11, alpha=0.02, modules=api.BM_MOD_EXCHANGE)
    base.update(kw)
    return api.FMParams(**base)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_periodic_exchange_eigenvector(axis):
    from boris2_amd import api
    n, h = (8, 6, 4), (4e-9, 3e-9, 5e-9)
    p = sim_params(Ha=(0.0, 0.0, 0.0))
    o = O.OracleSim(api.BM_EVAL_RK4, n, h, p)
    pbc = [0, 0, 0]; pbc[axis] = 1
    o.set_pbc(pbc)
    idx = np.indices((n[2], n[1], n[0]))[::-1].reshape(3, -1)     # (i, j, k) per cell
    ph = 2 * np.pi * idx[axis] / n[axis]
    M = np.zeros((idx.shape[1], 3)); M[:, 0] = p.Ms * np.cos(ph); M[:, 1] = p.Ms * np.sin(ph)
    o.set_M(M)
    o.update_fields()
    lam = (2 * np.cos(2 * np.pi / n[axis]) - 2) / h[axis] ** 2
    expect = 2 * p.A / (api.MU0 * p.Ms ** 2) * lam * M
    assert np.abs(o.heff() - expect).max() <= 1e-9 * np.abs(expect).max()
    M[:] = (p.Ms, 0.0, 0.0)
    o.set_M(M); o.update_fields()
    assert np.abs(o.heff()).max() == 0.0


def test_periodic_translation_invariance():
    from boris2_amd import api
    n, h, pbc = (8, 8, 1), (4e-9, 4e-9, 4e-9), (2, 3, 0)
    p = sim_params(Ha=(2e4, 0.0, 1e4), modules=api.BM_MOD_DEMAG | api.BM_MOD_EXCHANGE | api.BM_MOD_ZEEMAN)
    M0 = api.random_magnetisation(64, p.Ms, 9).reshape(n[1], n[0], 3)
    out = []
    for sx, sy in ((0, 0), (3, 5)):
        o = O.OracleSim(api.BM_EVAL_RK4, n, h, p)
        o.set_pbc(pbc)
        o.set_M(np.roll(M0, (sy, sx), axis=(0, 1)).reshape(-1, 3))
        o.set_dt(1e-13)
        o.run_to_iteration(20)
        out.append(o.get_M().reshape(n[1], n[0], 3))
    assert np.abs(np.roll(out[0], (5, 3), axis=(0, 1)) - out[1]).max() <= 1e-9 * p.Ms
