"""ctypes access to the oracle (oracle/liboracle.so) for tests and bench.py's CPU baseline.

The oracle is the checker: it is never part of the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
LIB = ORACLE_DIR / "liboracle.so"
REF_DUMP = ORACLE_DIR / "_ref" / "ref_dump"

_L = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def lib() -> C.CDLL:
    global _L
    if _L is not None:
        return _L
    if not LIB.exists():
        build()
    L = C.CDLL(str(LIB))
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.or_tensor3d.argtypes = [i] * 6 + [d] * 3 + [dp, dp]
    L.or_tensor2d.argtypes = [i] * 4 + [d] * 3 + [dp, dp]
    L.or_self_demag.argtypes = [d, d, d, dp]
    L.or_newell_f.argtypes = [d, d, d]; L.or_newell_f.restype = d
    L.or_newell_g.argtypes = [d, d, d]; L.or_newell_g.restype = d
    L.or_asympt_diag.argtypes = [d] * 6; L.or_asympt_diag.restype = d
    L.or_asympt_offdiag.argtypes = [d] * 6; L.or_asympt_offdiag.restype = d
    L.or_kernels.argtypes = [i] * 6 + [d] * 3 + [dp, dp]
    L.or_fft.argtypes = [i, dp, i]
    L.or_demag_create.argtypes = [i, i, i, d, d, d, i, i]; L.or_demag_create.restype = vp
    L.or_demag_create_pbc.argtypes = [i, i, i, d, d, d, ip, i]; L.or_demag_create_pbc.restype = vp
    L.or_tensor_pbc.argtypes = [i, i, i, d, d, d, i, i, i, dp, dp]; L.or_tensor_pbc.restype = i
    L.or_self_demag_pbc.argtypes = [d, d, d, i, i, i, i, i, i, dp]
    L.or_sim_set_pbc.argtypes = [vp, i, i, i]; L.or_sim_set_pbc.restype = i
    L.or_sim_set_layer_meshes.argtypes = [vp, i, vp, vp]
    L.or_demag_destroy.argtypes = [vp, i]
    L.or_demag_convolute.argtypes = [vp, i, dp, dp, i]
    L.or_sim_create.argtypes = [i, i, i, i, d, d, d, dp, i, i, i]; L.or_sim_create.restype = vp
    L.or_sim_destroy.argtypes = [vp]
    for nm in ("or_sim_set_M", "or_sim_get_M", "or_sim_get_Heff", "or_sim_state", "or_sim_average_m"):
        getattr(L, nm).argtypes = [vp, dp]
    L.or_sim_set_Msmap.argtypes = [vp, vp]
    L.or_sim_set_layers.argtypes = [vp, i, d, d, dp, dp, i, vp, vp, vp, vp]
    L.or_sim_set_dt.argtypes = [vp, d]
    L.or_sim_set_astep.argtypes = [vp, d, d, d, d]
    L.or_sim_update_fields.argtypes = [vp]
    L.or_sim_iterate_stage.argtypes = [vp]
    L.or_sim_advance_time.argtypes = [vp]
    L.or_sim_run_to_iteration.argtypes = [vp, i]
    L.or_sim_set_speedup.argtypes = [vp, i, d]; L.or_sim_set_speedup.restype = i
    L.or_sim_speedup_info.argtypes = [vp, dp]
    L.or_speedup_next_slot.argtypes = [i, dp]; L.or_speedup_next_slot.restype = i
    L.or_speedup_lagrange.argtypes = [i, dp, d, dp]
    L.or_tensor2d_zshift.argtypes = [i] * 4 + [dp, dp, d, dp, dp]
    L.or_kernels2d_zshift.argtypes = [i, i, dp, dp, dp, dp]
    L.or_tensor2d_zshift_pbc.argtypes = [i, i, dp, dp, d, i, i, dp, dp]
    L.or_tensor2d_shifted.argtypes = [i] * 4 + [dp, dp, dp, dp, dp]
    L.or_kernels2d_complex.argtypes = [i, i, dp, dp, dp]
    L.or_mlayer_create.argtypes = [i, i, i, d, d, dp, dp]; L.or_mlayer_create.restype = vp
    L.or_mlayer_create_pbc.argtypes = [i, i, i, d, d, dp, dp, i, i]; L.or_mlayer_create_pbc.restype = vp
    L.or_mlayer_create_offset.argtypes = [i, i, i, d, d, dp, dp, dp, dp]; L.or_mlayer_create_offset.restype = vp
    L.or_mlayer_destroy.argtypes = [vp]
    L.or_mlayer_convolute.argtypes = [vp, dp, dp]
    L.or_mlayer_convolute_transposed.argtypes = [vp, dp, dp]
    L.or_adj_create.argtypes = [i, i, i, dp, dp, i, dp, d, i, i]; L.or_adj_create.restype = vp
    L.or_adj_destroy.argtypes = [vp]
    L.or_adj_set_checkpoint.argtypes = [vp, i]
    L.or_adj_set_pbc.argtypes = [vp, i, i, i]
    L.or_adj_forward.argtypes = [vp, dp, dp, dp]; L.or_adj_forward.restype = d
    L.or_adj_backward.argtypes = [vp, dp, dp, dp, dp]
    L.or_adj_set_layers.argtypes = [vp, i, d, d, dp, dp, i, vp, vp, vp]
    L.or_set_random_mag.argtypes = [dp, C.c_longlong, C.c_uint]
    L.or_borisrand.argtypes = [C.c_uint, i, dp]
    _L = L
    return L


def tensor3d(n, N, h):
    L = lib()
    m = N[0] * N[1] * N[2] * 3
    a, b = np.zeros(m), np.zeros(m)
    assert L.or_tensor3d(*n, *N, *h, a, b) == 0
    return a.reshape(N[2], N[1], N[0], 3), b.reshape(N[2], N[1], N[0], 3)


def tensor_pbc(N, h, pbc):
    """Periodic tensor (h normalised): 3D (Dd, Do) each (Nz, Ny, Nx, 3); 2D (Nz = 1) Dd
    (Ny, Nx, 3) and Dxy (Ny, Nx)."""
    L = lib()
    m = N[0] * N[1] * N[2]
    a, b = np.zeros(3 * m), np.zeros(m if N[2] == 1 else 3 * m)
    assert L.or_tensor_pbc(*N, *h, *pbc, a, b) == 0
    if N[2] == 1:
        return a.reshape(N[1], N[0], 3), b.reshape(N[1], N[0])
    return a.reshape(N[2], N[1], N[0], 3), b.reshape(N[2], N[1], N[0], 3)


def self_demag_pbc(h, n, pbc):
    out = np.zeros(3)
    lib().or_self_demag_pbc(*h, *n, *pbc, out)
    return out


def tensor2d(n, N, h):
    L = lib()
    a, b = np.zeros(N[0] * N[1] * 3), np.zeros(N[0] * N[1])
    assert L.or_tensor2d(*n, *N, *h, a, b) == 0
    return a.reshape(N[1], N[0], 3), b.reshape(N[1], N[0])


def kernels(n, N, h):
    """(Kdiag, Kodiag): 3D (Nz/2+1, Ny/2+1, Nx/2+1, 3) each; 2D Kdiag (Ny/2+1, Nx/2+1, 3), Kxy (Ny/2+1, Nx/2+1)."""
    L = lib()
    nc, nyh, nzh = N[0] // 2 + 1, N[1] // 2 + 1, N[2] // 2 + 1
    if n[2] == 1:
        a, b = np.zeros(nc * nyh * 3), np.zeros(nc * nyh)
        assert L.or_kernels(*n, *N, *h, a, b) == 0
        return a.reshape(nyh, nc, 3), b.reshape(nyh, nc)
    a, b = np.zeros(nc * nyh * nzh * 3), np.zeros(nc * nyh * nzh * 3)
    assert L.or_kernels(*n, *N, *h, a, b) == 0
    return a.reshape(nzh, nyh, nc, 3), b.reshape(nzh, nyh, nc, 3)


def tensor2d_zshift(n, N, s, d, shift):
    """2D z-shifted tensor (normalised source cell s, destination cell d, shift): (Dd, Do),
    each (Ny, Nx, 3); Do = (xy, xz, yz)."""
    L = lib()
    a, b = np.zeros(N[0] * N[1] * 3), np.zeros(N[0] * N[1] * 3)
    assert L.or_tensor2d_zshift(n[0], n[1], N[0], N[1], np.array(s, float), np.array(d, float), float(shift), a, b) == 0
    return a.reshape(N[1], N[0], 3), b.reshape(N[1], N[0], 3)


def tensor2d_shifted(n, N, s, d, shift):
    """2D tensor shifted in-plane and along z (shift = (x, y, z), normalised): (Dd, Do), each
    (Ny, Nx, 3)."""
    L = lib()
    a, b = np.zeros(N[0] * N[1] * 3), np.zeros(N[0] * N[1] * 3)
    assert L.or_tensor2d_shifted(n[0], n[1], N[0], N[1], np.array(s, float), np.array(d, float), np.array(shift, float), a, b) == 0
    return a.reshape(N[1], N[0], 3), b.reshape(N[1], N[0], 3)


def kernels2d_complex(N, Dd, Do):
    """Complex kernels of a shifted tensor: (Ny, Nx/2+1, 6) complex (xx, yy, zz, xy, xz, yz)."""
    L = lib()
    nc = N[0] // 2 + 1
    k = np.zeros(nc * N[1] * 6 * 2)
    assert L.or_kernels2d_complex(N[0], N[1], np.ascontiguousarray(Dd, float).reshape(-1), np.ascontiguousarray(Do, float).reshape(-1), k) == 0
    return k.view(np.complex128).reshape(N[1], nc, 6)


def tensor2d_zshift_pbc(N, s, d, shift, px, py):
    """The 2D z-shifted tensor with px, py periodic images (N = n along a periodic axis):
    (Dd, Do), each (Ny, Nx, 3)."""
    L = lib()
    a, b = np.zeros(N[0] * N[1] * 3), np.zeros(N[0] * N[1] * 3)
    assert L.or_tensor2d_zshift_pbc(N[0], N[1], np.array(s, float), np.array(d, float), float(shift), int(px), int(py), a, b) == 0
    return a.reshape(N[1], N[0], 3), b.reshape(N[1], N[0], 3)


def kernels2d_zshift(N, Dd, Do):
    """Real kernels of a z-shifted 2D tensor: (Kd, Ko), each (Ny/2+1, Nx/2+1, 3)."""
    L = lib()
    nc, nyh = N[0] // 2 + 1, N[1] // 2 + 1
    a, b = np.zeros(nc * nyh * 3), np.zeros(nc * nyh * 3)
    assert L.or_kernels2d_zshift(N[0], N[1], np.ascontiguousarray(Dd, float).reshape(-1),
                                 np.ascontiguousarray(Do, float).reshape(-1), a, b) == 0
    return a.reshape(nyh, nc, 3), b.reshape(nyh, nc, 3)


class OracleMLayer:
    """The oracle's multilayered 2D demag: L layers of nx x ny cells (hx, hy, hz[l]) with
    bottom faces at z0[l]; M and H are (L*nx*ny, 3), layer-major.  pbc: (px, py) periodic
    images along x and y."""

    def __init__(self, nx, ny, hx, hy, hz, z0, pbc=(0, 0), offsets=None):
        """offsets: per layer (rx, ry) in-plane offsets of its grid (m); pairs offset against each
        other couple through the in-plane shifted tensor's complex kernels."""
        self.L = lib()
        self.nl = len(hz)
        if offsets is not None:
            o = np.asarray(offsets, float).reshape(self.nl, 2)
            self.h = self.L.or_mlayer_create_offset(self.nl, nx, ny, hx, hy, np.array(hz, float), np.array(z0, float),
                                                    np.ascontiguousarray(o[:, 0]), np.ascontiguousarray(o[:, 1]))
        elif any(pbc):
            self.h = self.L.or_mlayer_create_pbc(self.nl, nx, ny, hx, hy, np.array(hz, float), np.array(z0, float), int(pbc[0]), int(pbc[1]))
        else:
            self.h = self.L.or_mlayer_create(self.nl, nx, ny, hx, hy, np.array(hz, float), np.array(z0, float))
        assert self.h

    def __call__(self, M):
        M = np.ascontiguousarray(M, dtype=np.float64).reshape(-1)
        H = np.zeros_like(M)
        self.L.or_mlayer_convolute(self.h, M, H)
        return H.reshape(-1, 3)

    def transposed(self, M):
        """The transposed operator N^T applied to M."""
        M = np.ascontiguousarray(M, dtype=np.float64).reshape(-1)
        H = np.zeros_like(M)
        self.L.or_mlayer_convolute_transposed(self.h, M, H)
        return H.reshape(-1, 3)

    def __del__(self):
        try:
            self.L.or_mlayer_destroy(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
# Multilayered demag of 2D layers that are meshes of their own sizes, cells and positions:
# the reference's common grid (SDemag::set_default_n_common / set_Rect_collection,
# Boris/SDemag.cpp:173-300) and mesh transfer (SDemag_Demag::Initialize_Mesh_Transfer,
# Boris/SDemag_Demag.cpp:64-160; Boris/BorisLib/VEC_MeshTransfer.h), restated in numpy over
# the oracle's multilayer convolution of the rectangles' bounding box.

def _recip_weights(r, o, h, n):
    """Transfer::build_meshcells_weights / build_supermeshcells_weights
    (Boris/BorisLib/VEC_MeshTransfer.h:304-350, 396-440): the cells (i + j n.x) of the grid
    with origin o, cell h and n = (nx, ny) cells (one in z) that rectangle r = (s, e) covers,
    weighted d_max - |centre(r) - centre(cell)|, d_max = |size(r)/2 + h/2|; returns
    (cells, weights, their total) in the reference's loop order.  Corner indices by its
    floor / ceil of a DBL3, which round with a 1e-4 cell tolerance (floor_epsilon,
    ceil_epsilon, Boris/BorisLib/Funcs_Math.h:161-173)."""
    s, e = np.asarray(r[0], float), np.asarray(r[1], float)
    o, h = np.asarray(o, float), np.asarray(h, float)
    coord = (e + s) / 2 - o
    ll = np.floor((s - o) / h + 1e-4).astype(int)
    ur = np.ceil((e - o) / h - 1e-4).astype(int)
    dmax = float(np.sqrt(np.sum(((e - s) / 2 + h / 2) ** 2)))
    cells, wts, tot = [], [], 0.0
    for i in range(max(ll[0], 0), min(ur[0], n[0])):
        for j in range(max(ll[1], 0), min(ur[1], n[1])):
            for k in range(max(ll[2], 0), min(ur[2], 1)):
                d = dmax - float(np.sqrt(np.sum((coord - (np.array([i, j, k]) + 0.5) * h) ** 2)))
                tot += d
                cells.append(i + j * n[0]); wts.append(d)
    return cells, wts, tot


def _covered(c, m):
    """Covered volume ratio of cell c by rectangle m (Rect::get_intersection / intersects /
    IsPlane / IsPoint with the 1e-12 tolerance, Boris/BorisLib/Types_Rect.h)."""
    tol = 1e-12
    s = np.maximum(c[0], m[0]); e = np.minimum(c[1], m[1])
    d = e - s
    if not all(v > 0 or abs(v) < tol for v in d):
        return 0.0
    z = [abs(v) < tol for v in d]
    plane = (z[0] and not z[1] and not z[2]) or (z[1] and not z[0] and not z[2]) or (z[2] and not z[0] and not z[1])
    if plane or all(v == 0 for v in d):
        return 0.0
    size = np.asarray(c[1], float) - np.asarray(c[0], float)
    return float(np.prod(d / size))


class OracleMLayerMeshes:
    """Layers given as meshes (nx, ny, hx, hy, hz, x0, y0, z0): the reference's n_common (the
    largest cell counts), each layer's convolution rectangle (enlarged to the largest
    extents, starting as close to the smallest start as its end allows) and common cell
    (rectangle / n_common); M of a layer onto its rectangle's common cells by
    MESHTRANSFERTYPE_WEIGHTED (reciprocal distance x covered volume ratio,
    VEC_MeshTransfer.h:899-945), fields back by reciprocal distance (:1221-1260).  The
    rectangles are placed in their bounding box at the common cell and convolved there with
    the oracle's multilayered demag.  M and H: the layers' own cells concatenated."""

    def __init__(self, meshes):
        self.meshes = [tuple(m) for m in meshes]
        nc = (max(m[0] for m in meshes), max(m[1] for m in meshes))
        size = (max(m[0] * m[2] for m in meshes), max(m[1] * m[3] for m in meshes))
        lo = (min(m[5] for m in meshes), min(m[6] for m in meshes))
        self.Rm, self.Rc = [], []
        for (nx, ny, hx, hy, hz, x0, y0, z0) in meshes:
            s = [x0, y0, z0]; e = [x0 + nx * hx, y0 + ny * hy, z0 + hz]
            self.Rm.append((np.array(s), np.array(e)))
            cs, ce = list(s), list(e)
            for a in range(2):
                if ce[a] - size[a] < lo[a]:
                    cs[a] = lo[a]; ce[a] = cs[a] + size[a]
                else:
                    cs[a] = ce[a] - size[a]
            self.Rc.append((np.array(cs), np.array(ce)))
        hc = (size[0] / nc[0], size[1] / nc[1])
        bx, by = min(c[0][0] for c in self.Rc), min(c[0][1] for c in self.Rc)

        def split(v):
            # whole common cells (within 1e-6 of one) and the remaining fraction of a cell
            r = np.floor(v + 1e-6)
            return int(r), (0.0 if abs(v - r) <= 1e-6 else v - r)
        parts = [(split((c[0][0] - bx) / hc[0]), split((c[0][1] - by) / hc[1])) for c in self.Rc]
        self.off = [(px[0], py[0]) for px, py in parts]
        res = [(px[1] * hc[0], py[1] * hc[1]) for px, py in parts]
        self.nb = (max(o[0] for o in self.off) + nc[0], max(o[1] for o in self.off) + nc[1])
        self.nc, self.hc = nc, hc
        self.counts = [m[0] * m[1] for m in meshes]
        self.moff = np.concatenate([[0], np.cumsum(self.counts)])
        # rectangles offset by fractions of a common cell couple through shifted kernels
        fractional = any(r != (0.0, 0.0) for r in res)
        self.conv = OracleMLayer(self.nb[0], self.nb[1], hc[0], hc[1], [m[4] for m in meshes], [m[7] for m in meshes],
                                 offsets=res if fractional else None)
        # transfer matrices (dense, small cases): Tin[l] (box plane x layer cells), Tout[l]
        self.Tin, self.Tout = [], []
        for l, m in enumerate(meshes):
            nx, ny, hx, hy, hz = m[:5]
            c, mr = self.Rc[l], self.Rm[l]
            h_c = np.array([(c[1][0] - c[0][0]) / nc[0], (c[1][1] - c[0][1]) / nc[1], c[1][2] - c[0][2]])
            h_m = np.array([hx, hy, hz])
            plane = self.nb[0] * self.nb[1]
            Tin = np.zeros((plane, nx * ny)); Tout = np.zeros((nx * ny, plane))
            ox, oy = self.off[l]
            for j in range(nc[1]):
                for i in range(nc[0]):
                    rs = c[0] + np.array([i, j, 0]) * h_c
                    rc = (rs, rs + h_c)
                    cov = _covered(rc, mr)
                    if cov > 0:
                        cells, w, tot = _recip_weights(rc, mr[0], h_m, (nx, ny))
                        for cell, wt in zip(cells, w):
                            Tin[(oy + j) * self.nb[0] + ox + i, cell] += wt * cov / tot if tot > 0 else wt
            for J in range(ny):
                for I in range(nx):
                    rs = mr[0] + np.array([I, J, 0]) * h_m
                    cells, w, tot = _recip_weights((rs, rs + h_m), c[0], h_c, nc)
                    for cell, wt in zip(cells, w):
                        i, j = cell % nc[0], cell // nc[0]
                        Tout[I + J * nx, (oy + j) * self.nb[0] + ox + i] += wt / tot if tot > 0 else wt
            self.Tin.append(Tin); self.Tout.append(Tout)

    def __call__(self, M):
        M = np.asarray(M, float).reshape(-1, 3)
        planes = [self.Tin[l] @ M[self.moff[l]:self.moff[l + 1]] for l in range(len(self.meshes))]
        Hb = self.conv(np.concatenate(planes)).reshape(len(self.meshes), -1, 3)
        return np.concatenate([self.Tout[l] @ Hb[l] for l in range(len(self.meshes))])


def speedup_next_slot(p, tdm):
    """Slot the next computed demag field replaces once p slots hold saves at times tdm."""
    t = np.zeros(6); t[:len(tdm)] = tdm
    return lib().or_speedup_next_slot(int(p), t)


def speedup_lagrange(p, tdm, t):
    """Extrapolation coefficients of the p saved fields at time t (6 values, zeros past p)."""
    td = np.zeros(6); td[:len(tdm)] = tdm
    a = np.zeros(6)
    lib().or_speedup_lagrange(int(p), td, float(t), a)
    return a


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    L = lib()
    a = np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1).reshape(-1), dtype=np.float64)
    L.or_fft(len(x), a, sign)
    return a[0::2] + 1j * a[1::2]


class OracleDemag:
    """The oracle's demag convolution (FP64, or FP32 with fp32=True)."""

    def __init__(self, n, h, fp32=False, Nz_override=0, pbc=None):
        self.L = lib()
        self.n = tuple(n)
        self.fp32 = int(fp32)
        if pbc is not None and any(pbc):
            self.h = self.L.or_demag_create_pbc(*n, *h, np.ascontiguousarray(pbc, dtype=np.int32), self.fp32)
        else:
            self.h = self.L.or_demag_create(*n, *h, Nz_override, self.fp32)
        assert self.h

    def __call__(self, M: np.ndarray, H: np.ndarray | None = None, clear_out=True) -> np.ndarray:
        M = np.ascontiguousarray(M, dtype=np.float64).reshape(-1)
        out = np.zeros_like(M) if H is None else np.ascontiguousarray(H, dtype=np.float64).reshape(-1).copy()
        self.L.or_demag_convolute(self.h, self.fp32, M, out, int(clear_out))
        return out.reshape(-1, 3)

    def __del__(self):
        try:
            self.L.or_demag_destroy(self.h, self.fp32)
        except Exception:
            pass


def pack_params(p) -> tuple[np.ndarray, int]:
    """boris2_amd.api.FMParams -> the oracle C-ABI's 22-double vector and module flags."""
    v = [p.Ms, p.A, p.alpha, p.K1, p.K2, *p.ea, *p.Ha, float(p.wave), p.amp, p.w, p.t0, *p.dir, *p.slope]
    m = p.modules
    flags = (1 if m & 1 else 0) | (2 if m & 2 else 0) | (4 if m & 4 else 0) | (8 if m & 8 else 0)
    return np.array(v, dtype=np.float64), flags


class OracleSim:
    """The oracle's mesh + fields + ODE (mirror of the reference's AdvanceTime loop)."""

    def __init__(self, method, n, h, params, fp32=False, Nz_override=0):
        self.L = lib()
        self.n = tuple(n)
        self.ncell = n[0] * n[1] * n[2]
        prm, flags = pack_params(params)
        self.h = self.L.or_sim_create(method, *n, *h, prm, flags, Nz_override, int(fp32))
        assert self.h

    def set_M(self, M):
        self.L.or_sim_set_M(self.h, np.ascontiguousarray(M, dtype=np.float64).reshape(-1))

    def get_M(self):
        out = np.zeros(self.ncell * 3)
        self.L.or_sim_get_M(self.h, out)
        return out.reshape(-1, 3)

    def heff(self):
        out = np.zeros(self.ncell * 3)
        self.L.or_sim_get_Heff(self.h, out)
        return out.reshape(-1, 3)

    def set_ms_map(self, ms):
        self._ms = None if ms is None else np.ascontiguousarray(ms, dtype=np.float64)
        self.L.or_sim_set_Msmap(self.h, None if ms is None else self._ms.ctypes.data_as(C.c_void_p))

    def set_layers(self, hx, hy, hz, z0, with_demag=True, Ms=None, A=None, alpha=None, K1=None):
        """The sim's nz planes become a stack of 2D layers (create the sim with demag off)."""
        hz = np.ascontiguousarray(hz, dtype=np.float64)
        z0 = np.ascontiguousarray(z0, dtype=np.float64)
        self._lay = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (Ms, A, alpha, K1)]
        ptrs = [None if v is None else v.ctypes.data_as(C.c_void_p) for v in self._lay]
        self.L.or_sim_set_layers(self.h, len(hz), float(hx), float(hy), hz, z0, int(with_demag), *ptrs)

    def set_dt(self, dt):
        self.L.or_sim_set_dt(self.h, dt)

    def set_astep(self, err=-1.0, inc=-1.0, dtmin=-1.0, dtmax=-1.0):
        self.L.or_sim_set_astep(self.h, err, inc, dtmin, dtmax)

    def update_fields(self):
        self.L.or_sim_update_fields(self.h)

    def iterate_stage(self):
        self.L.or_sim_iterate_stage(self.h)

    def advance_time(self):
        self.L.or_sim_advance_time(self.h)

    def run_to_iteration(self, it):
        self.L.or_sim_run_to_iteration(self.h, int(it))

    def state(self):
        s = np.zeros(9)
        self.L.or_sim_state(self.h, s)
        keys = ["time", "dT", "lte", "iteration", "evalStep", "primed", "rejections", "evaluations", "stagetime"]
        return dict(zip(keys, s.tolist()))

    def average_m(self):
        out = np.zeros(3)
        self.L.or_sim_average_m(self.h, out)
        return out

    def set_layer_meshes(self, groups=None, Ha=None):
        """Stacks of 3D meshes: mesh index per layer (exchange across planes of a mesh) and a
        constant applied field (L, 3) per layer."""
        self._grp = None if groups is None else np.ascontiguousarray(groups, dtype=np.int32)
        self._lha = None if Ha is None else np.ascontiguousarray(Ha, dtype=np.float64).reshape(-1)
        L = self.n[2]
        self.L.or_sim_set_layer_meshes(self.h, L, None if self._grp is None else self._grp.ctypes.data_as(C.c_void_p),
                                       None if self._lha is None else self._lha.ctypes.data_as(C.c_void_p))

    def set_pbc(self, pbc):
        """Periodic images per axis: demag tensor and wrap-around exchange."""
        assert self.L.or_sim_set_pbc(self.h, *[int(v) for v in pbc]) == 0

    def set_speedup(self, mode, dTspeedup=0.0):
        """Demag evaluation speedup: mode 0..6 (none, step, linear ... quintic); dTspeedup > 0
        unlinks the speedup time step from dT."""
        assert self.L.or_sim_set_speedup(self.h, int(mode), float(dTspeedup)) == 0

    def speedup_info(self):
        out = np.zeros(8)
        self.L.or_sim_speedup_info(self.h, out)
        return {"num_saved": int(out[0]), "tdm": out[1:7].copy(), "convolutions": int(out[7])}

    def __del__(self):
        try:
            self.L.or_sim_destroy(self.h)
        except Exception:
            pass


class OracleAdjoint:
    """FP64 discrete adjoint of the fixed-step RK4 rollout: loss 1 - <(M/Ms) . m_target> of
    the final state from M0 = Ms m0; gradients dL/dMs (per cell) and dL/d amp."""

    def __init__(self, n, h, params, Ms_map, dt, nsteps, Nz_override=0, ckpt=0, pbc=(0, 0, 0)):
        """ckpt > 0: keep M every ckpt steps and recompute segments in backward() (same bits);
        pbc: periodic images per axis."""
        self.L = lib()
        self.n = tuple(n)
        self.ncell = n[0] * n[1] * n[2]
        prm, flags = pack_params(params)
        self.h = self.L.or_adj_create(*n, np.array(h, float), prm, flags, np.ascontiguousarray(Ms_map, float).reshape(-1),
                                      float(dt), int(nsteps), int(Nz_override))
        assert self.h
        if ckpt:
            self.L.or_adj_set_checkpoint(self.h, int(ckpt))
        if any(pbc):
            self.L.or_adj_set_pbc(self.h, *(int(v) for v in pbc))

    def forward(self, m0, mt):
        Mf = np.zeros(self.ncell * 3)
        loss = self.L.or_adj_forward(self.h, np.ascontiguousarray(m0, float).reshape(-1),
                                     np.ascontiguousarray(mt, float).reshape(-1), Mf)
        return loss, Mf.reshape(-1, 3)

    def set_layers(self, hx, hy, hz, z0, with_demag=True, A=None, alpha=None, K1=None):
        """The nz planes become a stack of 2D layers (create with demag off)."""
        self._lay = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (A, alpha, K1)]
        ptrs = [None if v is None else v.ctypes.data_as(C.c_void_p) for v in self._lay]
        self.L.or_adj_set_layers(self.h, len(hz), float(hx), float(hy), np.array(hz, float), np.array(z0, float),
                                 int(with_demag), *ptrs)

    def backward(self, m0, mt):
        """dL/dMs per cell and dL/d amp; self.gz holds (dL/d amp, dL/d slope xyz)."""
        g = np.zeros(self.ncell)
        gz = np.zeros(4)
        self.L.or_adj_backward(self.h, np.ascontiguousarray(m0, float).reshape(-1), np.ascontiguousarray(mt, float).reshape(-1), g, gz)
        self.gz = gz
        return g, float(gz[0])

    def __del__(self):
        try:
            self.L.or_adj_destroy(self.h)
        except Exception:
            pass


def borisrand(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count)
    lib().or_borisrand(seed, count, out)
    return out


def ref_available() -> bool:
    return REF_DUMP.exists()


def ref_dump(*args) -> np.ndarray:
    """Run oracle/_ref/ref_dump (the reference's own tensor sources) and read its output."""
    out = Path("/tmp") / f"ref_dump_{abs(hash(args))}.bin"
    subprocess.run([str(REF_DUMP), *[str(a) if not isinstance(a, float) else repr(a) for a in args], str(out)], check=True)
    v = np.fromfile(out)
    out.unlink()
    return v
