"""Host-side mirror of the reference's module / ODE interface for this hot path.

Objects here own device memory through torch (the allocator and stream provider only) and
call the C-ABI for every piece of compute:

* :class:`Demag` — ``DemagCUDA`` / ``ConvolutionCUDA::Convolute``
  (Boris/DemagCUDA.cpp:31-121, Boris/ConvolutionCUDA.h:217-223).
* :class:`FMSolver` — a ferromagnetic mesh with its Zeeman / Demag / Exchange / uniaxial
  anisotropy modules and ``DifferentialEquationFMCUDA`` + ``ODECommon_Base`` step control
  (Boris/DiffEqFMCUDA.h:38-137, Boris/DiffEq_CommonBase_IterateCUDA.cpp:13,
  Boris/SuperMeshSimulation.cpp:122-149).
* :func:`random_magnetisation` — ``FMesh::SetRandomMag`` with the ``BorisRand`` LCG
  (Boris/Mesh_Ferromagnetic_Control.cpp:230-255, BorisLib/BLib_prng.h:18-90).
* :func:`read_ovf2` / :func:`write_ovf2` — ``OVF2`` file handlers (Boris/OVF2_Handlers.cpp).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import (BM_EVAL_AHEUN, BM_EVAL_RK4, BM_EVAL_RKDP54, BM_EVAL_TEULER,  # noqa: F401
                   BM_MOD_ANIS_UNI, BM_MOD_DEMAG, BM_MOD_EXCHANGE, BM_MOD_ZEEMAN, check)

MU0 = 1.256637061435916e-6      # BorisLib/Funcs_Math_base.h:9
GAMMA = 2.212761569e5           # BorisLib/Funcs_Math_base.h:16


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor) -> C.c_void_p:
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError("expected a contiguous float32 CUDA tensor")
    return C.c_void_p(t.data_ptr())


class _DeviceArray:
    """A float32 device array owned by the library, exposed to torch through the CUDA array
    interface (torch.as_tensor makes a view, no copy)."""

    def __init__(self, ptr: int, count: int):
        self.__cuda_array_interface__ = {"shape": (count,), "typestr": "<f4", "data": (int(ptr), False),
                                         "version": 3, "strides": None}


def _device_view(ptr: int, count: int) -> torch.Tensor:
    return torch.as_tensor(_DeviceArray(ptr, count), device="cuda")


def require_device() -> None:
    """Raise unless the HIP library is loadable and a gfx950 device is present."""
    L = _lib.load()
    if not torch.cuda.is_available() or not L.bm_device_ok():
        raise RuntimeError("boris2_amd needs an MI355X (gfx950) device; no fallback path exists")


class Demag:
    """FFT demag convolution of one nx*ny*nz mesh of cells h (m).  pbc: periodic images
    along x, y, z (0 = open boundary), as the reference's `pbc` command sets them."""

    def __init__(self, n: tuple[int, int, int], h: tuple[float, float, float], pbc: tuple[int, int, int] = (0, 0, 0)):
        require_device()
        self.L = _lib.load()
        self.n = tuple(int(v) for v in n)
        self.h = tuple(float(v) for v in h)
        self.pbc = tuple(int(v) for v in pbc)
        self.handle = C.c_void_p()
        if any(self.pbc):
            check(self.L.bm_demag_create_pbc(*self.n, *self.h, *self.pbc, 0, C.byref(self.handle)))
        else:
            check(self.L.bm_demag_create(*self.n, *self.h, 0, C.byref(self.handle)))

    @property
    def ncells(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    def info(self) -> tuple[tuple[int, int, int], int]:
        N = (C.c_int * 3)()
        b = C.c_longlong()
        check(self.L.bm_demag_info(self.handle, N, C.byref(b)))
        return (N[0], N[1], N[2]), b.value

    def convolute(self, M: torch.Tensor, Heff: torch.Tensor | None = None, clear_out: bool = True,
                  energy: torch.Tensor | None = None) -> torch.Tensor:
        """Heff (=, +=) H_demag(M); M, Heff are (n, 3) float32 CUDA tensors."""
        if Heff is None:
            Heff = torch.empty_like(M)
            clear_out = True
        e = C.c_void_p(energy.data_ptr()) if energy is not None else None
        check(self.L.bm_demag_convolute(self.handle, _ptr(M), _ptr(Heff), int(clear_out), e, _stream()))
        return Heff

    def convolute_averaged(self, M1: torch.Tensor, M2: torch.Tensor, H1: torch.Tensor | None = None,
                           H2: torch.Tensor | None = None, clear_out: bool = True,
                           energy: torch.Tensor | None = None) -> torch.Tensor:
        """H1 (and H2 if given) (=, +=) H_demag((M1 + M2) / 2): the two-sublattice convolution
        (ConvolutionCUDA::Convolute_AveragedInputs[_DuplicatedOutputs])."""
        if H1 is None:
            H1 = torch.empty_like(M1)
            clear_out = True
        e = C.c_void_p(energy.data_ptr()) if energy is not None else None
        check(self.L.bm_demag_convolute_averaged(self.handle, _ptr(M1), _ptr(M2), _ptr(H1),
                                                 _ptr(H2) if H2 is not None else None, int(clear_out), e, _stream()))
        return H1

    def convolute_planar(self, M: torch.Tensor, H3n: torch.Tensor | None = None) -> torch.Tensor:
        if H3n is None:
            H3n = torch.empty((3, self.ncells), dtype=torch.float32, device=M.device)
        check(self.L.bm_demag_convolute_planar(self.handle, _ptr(M), _ptr(H3n), _stream()))
        return H3n

    def kernels(self) -> np.ndarray:
        (Nx, Ny, Nz), _ = self.info()
        nc, nyh, nzh = Nx // 2 + 1, Ny // 2 + 1, Nz // 2 + 1
        cnt = 4 * nc * nyh if self.n[2] == 1 else 6 * nzh * nc * nyh
        out = np.zeros(cnt, dtype=np.float32)
        check(self.L.bm_demag_copy_kernels(self.handle, out.ctypes.data_as(C.c_void_p), cnt))
        return out

    def close(self) -> None:
        if self.handle:
            self.L.bm_demag_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiLayerMeshDemag:
    """Multilayered demag of 2D layers that are meshes of their own: meshes is a list of
    (nx, ny, hx, hy, hz, x0, y0, z0) per layer (cells, cell size, bottom-left corner in m).
    M and H hold every layer's own cells concatenated, layer by layer ((sum nx*ny, 3) float32
    CUDA); convolute_split takes one (nx*ny, 3) array per layer."""

    def __init__(self, meshes):
        require_device()
        self.L = _lib.load()
        self.meshes = [tuple(m) for m in meshes]
        arr = (_lib.LayerMesh * len(self.meshes))()
        for k, (nx, ny, hx, hy, hz, x0, y0, z0) in enumerate(self.meshes):
            arr[k] = _lib.LayerMesh(int(nx), int(ny), float(hx), float(hy), float(hz), float(x0), float(y0), float(z0))
        self.handle = C.c_void_p()
        check(self.L.bm_mlayer_create_meshes(len(self.meshes), C.cast(arr, C.c_void_p), 0, C.byref(self.handle)))
        self.counts = [int(m[0]) * int(m[1]) for m in self.meshes]

    @property
    def ncells(self) -> int:
        return sum(self.counts)

    def convolute(self, M: torch.Tensor, Heff: torch.Tensor | None = None, clear_out: bool = True) -> torch.Tensor:
        if Heff is None:
            Heff = torch.empty_like(M)
            clear_out = True
        check(self.L.bm_mlayer_convolute(self.handle, _ptr(M), _ptr(Heff), int(clear_out), _stream()))
        return Heff

    def convolute_split(self, M_layers, H_layers, clear_out: bool = True, energies: torch.Tensor | None = None) -> None:
        st = _stream()
        for l, M in enumerate(M_layers):
            check(self.L.bm_mlayer_forward(self.handle, l, _ptr(M), st))
        check(self.L.bm_mlayer_multiply(self.handle, st))
        for l, (M, H) in enumerate(zip(M_layers, H_layers)):
            ep = None if energies is None else C.c_void_p(energies.data_ptr() + 4 * l)
            check(self.L.bm_mlayer_inverse(self.handle, l, _ptr(M), _ptr(H), int(clear_out), ep, st))

    def close(self) -> None:
        if self.handle:
            self.L.bm_mlayer_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiLayerDemag:
    """Multilayered demag of stacked 2D layers: nlayers layers of nx*ny cells (hx, hy, hz[l]),
    layer l's bottom face at height z0[l] (m).  M and H are (nlayers*nx*ny, 3) float32 CUDA
    tensors, layer-major.  pbc: (px, py) periodic images along x and y."""

    def __init__(self, nx: int, ny: int, hx: float, hy: float, hz, z0, pbc=(0, 0)):
        require_device()
        self.L = _lib.load()
        self.nx, self.ny = int(nx), int(ny)
        self.hz = np.ascontiguousarray(hz, dtype=np.float64)
        self.z0 = np.ascontiguousarray(z0, dtype=np.float64)
        if self.hz.shape != self.z0.shape or self.hz.ndim != 1:
            raise ValueError("hz and z0 must be 1-D of the same length")
        self.nlayers = len(self.hz)
        self.handle = C.c_void_p()
        self.pbc = tuple(int(v) for v in pbc)
        check(self.L.bm_mlayer_create_pbc(self.nlayers, self.nx, self.ny, float(hx), float(hy),
                                          self.hz.ctypes.data_as(C.c_void_p), self.z0.ctypes.data_as(C.c_void_p),
                                          *self.pbc, 0, C.byref(self.handle)))

    @property
    def ncells(self) -> int:
        return self.nlayers * self.nx * self.ny

    def convolute(self, M: torch.Tensor, Heff: torch.Tensor | None = None, clear_out: bool = True) -> torch.Tensor:
        """Heff (=, +=) the demag field of every layer from all layers' magnetisation."""
        if Heff is None:
            Heff = torch.empty_like(M)
            clear_out = True
        check(self.L.bm_mlayer_convolute(self.handle, _ptr(M), _ptr(Heff), int(clear_out), _stream()))
        return Heff

    def convolute_split(self, M_layers, H_layers, clear_out: bool = True, energies: torch.Tensor | None = None) -> None:
        """The split form over per-layer arrays (each (nx*ny, 3) float32 CUDA): forward of
        every layer, one kernel product, inverse of every layer into H_layers (set or add);
        energies (float32 CUDA, nlayers) receives each layer's demag energy density."""
        st = _stream()
        for l, M in enumerate(M_layers):
            check(self.L.bm_mlayer_forward(self.handle, l, _ptr(M), st))
        check(self.L.bm_mlayer_multiply(self.handle, st))
        for l, (M, H) in enumerate(zip(M_layers, H_layers)):
            ep = None if energies is None else C.c_void_p(energies.data_ptr() + 4 * l)
            check(self.L.bm_mlayer_inverse(self.handle, l, _ptr(M), _ptr(H), int(clear_out), ep, st))

    def kernels(self) -> np.ndarray:
        """(nlayers, nlayers, 6, nx+1, ny+1) float32: destination, source, component, kx, ky."""
        L, nc, nyh = self.nlayers, self.nx + 1, self.ny + 1
        out = np.zeros(L * L * 6 * nc * nyh, dtype=np.float32)
        check(self.L.bm_mlayer_copy_kernels(self.handle, out.ctypes.data_as(C.c_void_p), out.size))
        return out.reshape(L, L, 6, nc, nyh)

    def close(self) -> None:
        if self.handle:
            self.L.bm_mlayer_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class FMParams:
    """Material parameters and applied field of one ferromagnetic mesh (SI units)."""
    Ms: float = 8e5
    A: float = 1.3e-11
    alpha: float = 0.02
    K1: float = 0.0
    K2: float = 0.0
    ea: tuple = (1.0, 0.0, 0.0)
    grel: float = 1.0
    Ha: tuple = (0.0, 0.0, 0.0)
    wave: int = 0
    amp: float = 0.0
    w: float = 0.0
    t0: float = 0.0
    dir: tuple = (0.0, 0.0, 0.0)
    slope: tuple = (0.0, 0.0, 0.0)
    modules: int = BM_MOD_DEMAG | BM_MOD_EXCHANGE | BM_MOD_ZEEMAN

    def to_c(self) -> _lib.FmParams:
        p = _lib.FmParams()
        p.Ms, p.A, p.alpha, p.K1, p.K2, p.grel = self.Ms, self.A, self.alpha, self.K1, self.K2, self.grel
        p.ea[:] = list(self.ea)
        p.Ha[:] = list(self.Ha)
        p.wave, p.amp, p.w, p.t0 = int(self.wave), self.amp, self.w, self.t0
        p.dir[:] = list(self.dir)
        p.slope[:] = list(self.slope)
        p.modules = int(self.modules)
        return p


class FMSolver:
    """One ferromagnetic mesh integrated with the LLG equation."""

    def __init__(self, method: int, n: tuple[int, int, int], h: tuple[float, float, float],
                 params: FMParams, M0: np.ndarray | torch.Tensor | None = None, pbc: tuple[int, int, int] = (0, 0, 0)):
        """pbc: periodic images per axis (demag and exchange), as the reference's `pbc` command."""
        require_device()
        self.L = _lib.load()
        self.n = tuple(int(v) for v in n)
        self.h = tuple(float(v) for v in h)
        self.params = params
        ncell = self.n[0] * self.n[1] * self.n[2]
        self.M = torch.zeros((ncell, 3), dtype=torch.float32, device="cuda")
        if M0 is None:
            self.M[:, 0] = -params.Ms      # a new mesh starts at (-Ms, 0, 0), Boris/Mesh_Ferromagnetic.cpp:194
        else:
            self.M.copy_(torch.as_tensor(np.asarray(M0, dtype=np.float32).reshape(ncell, 3)))
        self.pbc = tuple(int(v) for v in pbc)
        self.demag = Demag(self.n, self.h, self.pbc) if params.modules & BM_MOD_DEMAG else None
        self._pc = params.to_c()
        self.handle = C.c_void_p()
        dh = self.demag.handle if self.demag else None
        check(self.L.bm_ode_create(method, *self.n, *self.h, C.byref(self._pc), dh, _ptr(self.M), C.byref(self.handle)))
        if any(self.pbc):
            check(self.L.bm_ode_set_pbc(self.handle, *self.pbc))
        self.Ms_map = None

    def set_ms_map(self, ms: np.ndarray | torch.Tensor | None) -> None:
        if ms is None:
            self.Ms_map = None
            check(self.L.bm_ode_set_ms_map(self.handle, None))
            return
        self.Ms_map = torch.as_tensor(np.asarray(ms, dtype=np.float32).reshape(-1)).cuda().contiguous()
        check(self.L.bm_ode_set_ms_map(self.handle, _ptr(self.Ms_map)))

    def set_params(self, params: FMParams) -> None:
        self.params = params
        self._pc = params.to_c()
        check(self.L.bm_ode_set_params(self.handle, C.byref(self._pc)))

    def set_dt(self, dT: float) -> None:
        check(self.L.bm_ode_set_dt(self.handle, float(dT)))

    def set_extra_field(self, fn) -> None:
        """Fields of other modules: before every stage fn(M) is called with the (ncell, 3)
        CUDA tensor view of the magnetisation that stage evaluates and must return the other
        modules' summed field as an (ncell, 3) float32 tensor (A/m) on the current stream;
        the stage adds it to the effective field (the reference's per-stage module
        UpdateField loop).  fn None detaches it."""
        if fn is None:
            check(self.L.bm_ode_set_extra_field(self.handle, _lib.FIELD_FN(), None, None))
            self._ext = None
            return
        ncell = self.M.shape[0]
        Hext = torch.zeros((ncell, 3), dtype=torch.float32, device="cuda")
        views = {}

        def tramp(user, m_ptr, h_ptr, stream):
            # the stage input is one of two arrays: the caller's M or the solver's second
            # buffer; wrap each once as a tensor view over its device memory
            if m_ptr not in views:
                views[m_ptr] = self.M if m_ptr == self.M.data_ptr() else _device_view(m_ptr, ncell * 3).view(ncell, 3)
            Hext.copy_(fn(views[m_ptr]))

        cb = _lib.FIELD_FN(tramp)
        self._ext = (cb, Hext, views)           # keep the callback and buffers alive
        check(self.L.bm_ode_set_extra_field(self.handle, cb, None, _ptr(Hext)))

    def set_astep(self, err_fail=-1.0, dT_increase=-1.0, dT_min=-1.0, dT_max=-1.0) -> None:
        check(self.L.bm_ode_set_astep(self.handle, err_fail, dT_increase, dT_min, dT_max))

    def set_eval_speedup(self, mode: int, dTspeedup: float = 0.0) -> None:
        """Demag evaluation speedup (the reference's ``evalspeedup`` setting): mode 0 none,
        1 step, 2 linear, 3 quadratic, 4 cubic, 5 quartic, 6 quintic extrapolation of saved
        demag fields; dTspeedup > 0 convolves every dTspeedup of simulated time instead of
        at the start of every step."""
        check(self.L.bm_ode_set_eval_speedup(self.handle, int(mode), float(dTspeedup)))

    def speedup_state(self) -> dict:
        n, conv = C.c_int(), C.c_longlong()
        tdm = (C.c_double * 6)()
        check(self.L.bm_ode_get_speedup_state(self.handle, C.byref(n), tdm, C.byref(conv)))
        return {"num_saved": n.value, "tdm": np.array(tdm[:]), "convolutions": conv.value}

    def advance_time(self) -> None:
        check(self.L.bm_ode_advance_time(self.handle, _stream()))

    def run_to_iteration(self, it: int) -> None:
        check(self.L.bm_ode_run_to_iteration(self.handle, int(it), _stream()))

    def iterate_stage(self) -> None:
        check(self.L.bm_ode_iterate_stage(self.handle, _stream()))

    def restore(self) -> None:
        check(self.L.bm_ode_restore(self.handle, _stream()))

    def state(self) -> dict:
        s = _lib.OdeState()
        check(self.L.bm_ode_get_state(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.OdeState._fields_}

    def keep_heff(self, on: bool = True) -> None:
        check(self.L.bm_ode_keep_heff(self.handle, int(on)))

    def heff(self) -> torch.Tensor:
        H = torch.empty_like(self.M)
        check(self.L.bm_ode_copy_heff(self.handle, _ptr(H), _stream()))
        return H

    def average_m(self) -> np.ndarray:
        out = (C.c_double * 3)()
        check(self.L.bm_ode_average_m(self.handle, out, _stream()))
        return np.array(out[:])

    def close(self) -> None:
        if self.handle:
            torch.cuda.synchronize()
            self.L.bm_ode_destroy(self.handle)
            self.handle = C.c_void_p()
        if self.demag is not None:
            self.demag.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LayerStackSolver(FMSolver):
    """A stack of 2D layers integrated together: layer l has nx*ny cells (hx, hy, hz[l]) with
    its bottom face at z0[l], the layers coupled by the multilayered demag and each with its
    own Ms, A, alpha, K1 (per-layer sequences; None keeps params' value).  The reference
    holds such a stack as one ferromagnetic mesh per layer under SDemag (Boris/SDemag.cpp).
    M is (L*nx*ny, 3), layer-major.  pbc: (px, py) periodic images along x and y (demag and
    in-plane exchange)."""

    def __init__(self, method: int, nx: int, ny: int, hx: float, hy: float, hz, z0, params: FMParams,
                 Ms=None, A=None, alpha=None, K1=None, M0=None, pbc=(0, 0)):
        hz = np.ascontiguousarray(hz, dtype=np.float64)
        nl = len(hz)
        pbc = tuple(int(v) for v in pbc)
        self.mlayer = MultiLayerDemag(nx, ny, hx, hy, hz, z0, pbc) if params.modules & BM_MOD_DEMAG else None
        base = FMParams(**{**params.__dict__, "modules": params.modules & ~BM_MOD_DEMAG})
        if M0 is None and Ms is not None:
            M0 = np.zeros((nl * nx * ny, 3), dtype=np.float32)
            M0[:, 0] = -np.repeat(np.asarray(Ms, dtype=np.float32), nx * ny)
        super().__init__(method, (nx, ny, nl), (hx, hy, float(hz.max())), base, M0, pbc=(pbc[0], pbc[1], 0))
        self._lay = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (Ms, A, alpha, K1)]
        for v in self._lay:
            if v is not None and v.shape != (nl,):
                raise ValueError("per-layer values need one entry per layer")
        ptrs = [None if v is None else v.ctypes.data_as(C.c_void_p) for v in self._lay]
        if self.mlayer is not None or any(v is not None for v in self._lay):
            # without demag the stack still needs its per-layer values: attach a handle anyway
            if self.mlayer is None:
                self.mlayer = MultiLayerDemag(nx, ny, hx, hy, hz, z0, pbc)
            check(self.L.bm_ode_set_layers(self.handle, self.mlayer.handle, *ptrs))
        self.set_params(params)

    def set_layer_meshes(self, mesh_of_layer=None, Ha=None) -> None:
        """Stacks of 3D meshes: consecutive layers with equal mesh_of_layer[l] are the z planes of
        one mesh (exchange couples them; they must touch and have one thickness); Ha (L, 3) A/m
        gives every layer its own constant applied field."""
        g = None if mesh_of_layer is None else np.ascontiguousarray(mesh_of_layer, dtype=np.int32)
        f = None if Ha is None else np.ascontiguousarray(Ha, dtype=np.float64).reshape(-1)
        self._groups = (g, f)
        check(self.L.bm_ode_set_layer_meshes(self.handle, None if g is None else g.ctypes.data_as(C.c_void_p),
                                             None if f is None else f.ctypes.data_as(C.c_void_p)))

    def close(self) -> None:
        super().close()
        if getattr(self, "mlayer", None) is not None:
            self.mlayer.close()


def stack_of_meshes(meshes):
    """Planes of a multilayered set-up of 3D meshes, as LayerStackSolver arguments: meshes is a
    list of (nz, hz, z_bottom); returns (hz per plane, z0 per plane, mesh index per plane)."""
    hz, z0, grp = [], [], []
    for m, (nz, h, zb) in enumerate(meshes):
        for k in range(int(nz)):
            hz.append(float(h)); z0.append(float(zb) + k * float(h)); grp.append(m)
    return np.array(hz), np.array(z0), np.array(grp, dtype=np.int32)


class BatchSolver:
    """Many independent meshes of uneven sizes integrated together, with a fixed-step method
    (RK4 or TEuler) or RKDP54 (each mesh with its own adaptive step, controlled on the device):
    one launch per pass size per field evaluation, a step replayed from a captured hipGraph.  meshes: list of (n, h, FMParams); M0: list of per-mesh (ncell, 3)
    arrays or None.  self.M holds every mesh's magnetisation concatenated; mesh(b) is a view.
    layers: None, or per mesh None or dict(hz=..., z0=..., Ms=None, A=None, alpha=None,
    K1=None) making that mesh's n[2] planes a layer stack (as LayerStackSolver).
    autotune: False, True (launch shapes timed before the first step; results unchanged bit
    for bit) or "radix" (the FFT passes' radix schedules timed too; results within FP32
    rounding of the default schedules).  pbc: None, or per mesh None or (px, py, pz) periodic
    images (as FMSolver's pbc)."""

    def __init__(self, method: int, meshes, M0=None, use_graph: bool = True, layers=None, autotune=False, pbc=None):
        require_device()
        self.L = _lib.load()
        self.meshes = [(tuple(int(v) for v in n), tuple(float(v) for v in h), p) for n, h, p in meshes]
        nb = len(self.meshes)
        counts = [n[0] * n[1] * n[2] for n, _, _ in self.meshes]
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.M = torch.zeros((int(self.offsets[-1]), 3), dtype=torch.float32, device="cuda")
        for b, (n, h, p) in enumerate(self.meshes):
            v = self.mesh(b)
            if M0 is None or M0[b] is None:
                ms = p.Ms
                if layers is not None and layers[b] is not None and layers[b].get("Ms") is not None:
                    ms = torch.as_tensor(np.repeat(np.asarray(layers[b]["Ms"], np.float32), n[0] * n[1]), device="cuda")
                v[:, 0] = -ms
            else:
                v.copy_(torch.as_tensor(np.asarray(M0[b], dtype=np.float32).reshape(-1, 3)))
        n3 = np.array([v for n, _, _ in self.meshes for v in n], dtype=np.int32)
        h3 = np.array([v for _, h, _ in self.meshes for v in h], dtype=np.float64)
        self._pc = (_lib.FmParams * nb)(*[p.to_c() for _, _, p in self.meshes])
        self.handle = C.c_void_p()
        ld = None
        if layers is not None:
            self._keep = []
            ld = (_lib.LayerDesc * nb)()
            for b, lay in enumerate(layers):
                if lay is None:
                    continue
                arrs = {k: np.ascontiguousarray(lay[k], dtype=np.float64).reshape(-1) for k in ("hz", "z0", "Ms", "A", "alpha", "K1", "Ha")
                        if lay.get(k) is not None}
                if lay.get("mesh_of_layer") is not None:
                    arrs["mesh_of_layer"] = np.ascontiguousarray(lay["mesh_of_layer"], dtype=np.int32)
                self._keep.append(arrs)
                ld[b].nlayers = len(arrs["hz"])
                for k, a in arrs.items():
                    setattr(ld[b], k, a.ctypes.data_as(C.c_void_p))
        check(self.L.bm_batch_create_layers(method, nb, n3.ctypes.data_as(C.c_void_p), h3.ctypes.data_as(C.c_void_p),
                                            C.cast(self._pc, C.c_void_p), C.cast(ld, C.c_void_p) if ld is not None else None,
                                            _ptr(self.M), (0 if use_graph else 1) | (6 if autotune == "radix" else 2 if autotune else 0),
                                            C.byref(self.handle)))
        for b, pb in enumerate(pbc or []):
            if pb is not None and any(pb):
                check(self.L.bm_batch_set_pbc(self.handle, b, *(int(v) for v in pb)))

    def mesh(self, b: int) -> torch.Tensor:
        return self.M[int(self.offsets[b]):int(self.offsets[b + 1])]

    def autotune_report(self) -> list[str]:
        """The autotuner's choices (after the first run with autotune=True), one line per pass."""
        buf = C.create_string_buffer(16384)
        check(self.L.bm_batch_autotune_report(self.handle, buf, len(buf)))
        return [l for l in buf.value.decode().splitlines() if l]

    def set_schedule(self, pass_: str, size: int, f: int = 0, per_cu: int = 0, radix: int = 0) -> None:
        """Pin one pass-size group's launch shape and radix schedule (pass_: x_fwd, x_inv, y_fwd,
        y_inv, z, y2d, ml_mult), replacing any pending autotuning."""
        check(self.L.bm_batch_set_schedule(self.handle, pass_.encode(), int(size), int(f), int(per_cu), int(radix)))

    def apply_report(self, lines) -> None:
        """Re-apply the choices of an earlier autotune_report() (e.g. saved from another run of the
        same batch), so a process need not tune again."""
        for l in lines:
            w = l.split()
            kv = dict(x.split("=") for x in w[2:5])
            self.set_schedule(w[0], int(w[1]), int(kv["f"]), int(kv["per_cu"]), int(kv["radix"]))

    def set_dt(self, dT: float) -> None:
        check(self.L.bm_batch_set_dt(self.handle, float(dT)))

    def run(self, nsteps: int) -> None:
        check(self.L.bm_batch_run(self.handle, int(nsteps), _stream()))

    def set_astep(self, err_fail=-1.0, dT_increase=-1.0, dT_min=-1.0, dT_max=-1.0) -> None:
        """RKDP54 batches: adaptive step controls of every mesh (<= 0 keeps the current value)."""
        check(self.L.bm_batch_set_astep(self.handle, err_fail, dT_increase, dT_min, dT_max))

    def state(self) -> dict:
        t, it = C.c_double(), C.c_longlong()
        check(self.L.bm_batch_get_state(self.handle, C.byref(t), C.byref(it)))
        return {"time": t.value, "iteration": it.value}

    def mesh_state(self, b: int) -> dict:
        """Mesh b's time, dT, accepted iterations and rejected attempts after the last run."""
        t, dt, it, rj = C.c_double(), C.c_double(), C.c_longlong(), C.c_longlong()
        check(self.L.bm_batch_get_mesh_state(self.handle, int(b), C.byref(t), C.byref(dt), C.byref(it), C.byref(rj)))
        return {"time": t.value, "dT": dt.value, "iteration": it.value, "rejections": rj.value}

    def close(self) -> None:
        if self.handle:
            torch.cuda.synchronize()
            self.L.bm_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RK4Adjoint:
    """Differentiable fixed-step RK4 rollout of one mesh: demag, exchange, uniaxial anisotropy
    and a Zeeman field (constant, sinc pulse params.wave = 1 or ramp wave = 2), per-cell Ms
    map.  loss = 1 - <(M_final/Ms) . m_target>; backward() gives dL/dMs (per cell) and
    dL/d amp by the discrete adjoint, zeeman_grads() (dL/d amp, dL/d slope xyz).
    layers = dict(hz=..., z0=..., A=None, alpha=None, K1=None) makes the n[2] planes a stack
    of 2D layers (hx, hy from h) coupled by the multilayered demag (params must enable demag).
    pbc: periodic images along x, y, z (exchange wraps around, demag sums the images)."""

    def __init__(self, n, h, params: "FMParams", Ms_map: torch.Tensor, dt: float, nsteps: int, ckpt_every: int = 100,
                 layers: dict | None = None, pbc: tuple[int, int, int] = (0, 0, 0)):
        require_device()
        self.L = _lib.load()
        self.n = tuple(int(v) for v in n)
        self.h = tuple(float(v) for v in h)
        self.params = params
        self.Ms_map = Ms_map.detach().to(device="cuda", dtype=torch.float32).contiguous().reshape(-1)
        self.demag = self.mlayer = None
        pc = params
        if layers is not None:
            if not params.modules & BM_MOD_DEMAG:
                raise ValueError("a layer stack needs the demag module")
            self.mlayer = MultiLayerDemag(self.n[0], self.n[1], self.h[0], self.h[1], layers["hz"], layers["z0"], (pbc[0], pbc[1]))
            pc = FMParams(**{**params.__dict__, "modules": params.modules & ~BM_MOD_DEMAG})
        elif params.modules & BM_MOD_DEMAG:
            self.demag = Demag(self.n, self.h, pbc)
        self._pc = pc.to_c()
        self.handle = C.c_void_p()
        check(self.L.bm_adj_create(*self.n, *self.h, C.byref(self._pc), self.demag.handle if self.demag else None,
                                   _ptr(self.Ms_map), float(dt), int(nsteps), int(ckpt_every), C.byref(self.handle)))
        if layers is not None:
            self._lay = [None if layers.get(k) is None else np.ascontiguousarray(layers[k], dtype=np.float64)
                         for k in ("A", "alpha", "K1")]
            ptrs = [None if v is None else v.ctypes.data_as(C.c_void_p) for v in self._lay]
            check(self.L.bm_adj_set_layers(self.handle, self.mlayer.handle, *ptrs))
        if any(pbc):
            check(self.L.bm_adj_set_pbc(self.handle, *(int(v) for v in pbc)))

    def forward(self, m0: torch.Tensor, mt: torch.Tensor) -> tuple[float, torch.Tensor]:
        self._m0 = m0.detach().to(device="cuda", dtype=torch.float32).contiguous()
        self._mt = mt.detach().to(device="cuda", dtype=torch.float32).contiguous()
        Mf = torch.empty_like(self._m0)
        loss = C.c_double()
        check(self.L.bm_adj_forward(self.handle, _ptr(self._m0), _ptr(self._mt), _ptr(Mf), C.byref(loss), _stream()))
        return loss.value, Mf

    def backward(self) -> tuple[torch.Tensor, float]:
        g = torch.empty_like(self.Ms_map)
        ga = C.c_double()
        check(self.L.bm_adj_backward(self.handle, _ptr(g), C.byref(ga), _stream()))
        return g, ga.value

    def zeeman_grads(self) -> np.ndarray:
        """(dL/d amp, dL/d slope_x, dL/d slope_y, dL/d slope_z) of the last backward()."""
        out = (C.c_double * 4)()
        check(self.L.bm_adj_zeeman_grads(self.handle, out))
        return np.array(out[:])

    def close(self) -> None:
        if self.handle:
            self.L.bm_adj_destroy(self.handle)
            self.handle = C.c_void_p()
        for obj in (self.demag, self.mlayer):
            if obj is not None:
                obj.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RK4AdjointBatch:
    """Several RK4Adjoint instances (same nsteps and checkpoint interval) run in lockstep, each
    on its own stream, so their kernels overlap on the device: losses and gradients of all
    instances from one forward() and one backward()."""

    def __init__(self, adjoints: list[RK4Adjoint]):
        self.L = _lib.load()
        self.adjoints = list(adjoints)
        self._h = (C.c_void_p * len(self.adjoints))(*[a.handle.value for a in self.adjoints])

    def forward(self, m0s, mts) -> tuple[list[float], list[torch.Tensor]]:
        nb = len(self.adjoints)
        self._m0 = [m.detach().to(device="cuda", dtype=torch.float32).contiguous() for m in m0s]
        self._mt = [m.detach().to(device="cuda", dtype=torch.float32).contiguous() for m in mts]
        Mf = [torch.empty_like(m) for m in self._m0]
        ptrs = lambda ts: (C.c_void_p * nb)(*[t.data_ptr() for t in ts])
        loss = (C.c_double * nb)()
        check(self.L.bm_adj_batch_forward(nb, self._h, ptrs(self._m0), ptrs(self._mt), ptrs(Mf), loss, _stream()))
        return list(loss[:]), Mf

    def backward(self) -> list[tuple[torch.Tensor, float]]:
        nb = len(self.adjoints)
        gs = [torch.empty_like(a.Ms_map) for a in self.adjoints]
        ga = (C.c_double * nb)()
        check(self.L.bm_adj_batch_backward(nb, self._h, (C.c_void_p * nb)(*[g.data_ptr() for g in gs]), ga, _stream()))
        return [(g, ga[b]) for b, g in enumerate(gs)]


class RK4RolloutLoss(torch.autograd.Function):
    """loss = RK4RolloutLoss.apply(Ms_map, amp, n, h, params, dt, nsteps, m0, mt, ckpt_every):
    differentiable in Ms_map (n floats, CUDA) and the pulse amplitude amp (0-d tensor, A/m)."""

    @staticmethod
    def forward(ctx, Ms_map, amp, n, h, params, dt, nsteps, m0, mt, ckpt_every=100):
        import dataclasses
        p = dataclasses.replace(params, amp=float(amp))
        roll = RK4Adjoint(n, h, p, Ms_map, dt, nsteps, ckpt_every)
        loss, _ = roll.forward(m0, mt)
        ctx.roll = roll
        return torch.tensor(loss, dtype=torch.float64, device=Ms_map.device)

    @staticmethod
    def backward(ctx, gout):
        g, ga = ctx.roll.backward()
        ctx.roll.close()
        s = gout.to(torch.float32)
        return (g * s).reshape(ctx.roll.Ms_map.shape), gout * ga, None, None, None, None, None, None, None, None


class LayerStackRolloutLoss(torch.autograd.Function):
    """loss = LayerStackRolloutLoss.apply(Ms_layers, slope, nx, ny, hx, hy, layers, params, dt,
    nsteps, m0, mt, ckpt_every): fixed-step RK4 rollout of a layer stack (layers: dict with
    hz, z0 and optional per-layer A, alpha, K1), differentiable in the per-layer Ms
    (Ms_layers, L values) and the Zeeman ramp slope (3 values, A/m/s; params.wave = 2)."""

    @staticmethod
    def forward(ctx, Ms_layers, slope, nx, ny, hx, hy, layers, params, dt, nsteps, m0, mt, ckpt_every=100):
        import dataclasses
        L = len(layers["hz"])
        p = dataclasses.replace(params, wave=2, slope=tuple(float(v) for v in slope.detach().cpu().reshape(3)))
        Ms_map = Ms_layers.detach().to(device="cuda", dtype=torch.float32).reshape(L, 1).expand(L, nx * ny).reshape(-1)
        roll = RK4Adjoint((nx, ny, L), (hx, hy, float(max(layers["hz"]))), p, Ms_map, dt, nsteps, ckpt_every, layers=layers)
        loss, _ = roll.forward(m0, mt)
        ctx.roll, ctx.L, ctx.nl = roll, L, nx * ny
        ctx.shapes = (Ms_layers.shape, Ms_layers.dtype, Ms_layers.device, slope.shape, slope.dtype, slope.device)
        return torch.tensor(loss, dtype=torch.float64, device=Ms_layers.device)

    @staticmethod
    def backward(ctx, gout):
        g, _ = ctx.roll.backward()
        gz = ctx.roll.zeeman_grads()
        ctx.roll.close()
        ms_shape, ms_dt, ms_dev, sl_shape, sl_dt, sl_dev = ctx.shapes
        g_layers = g.double().reshape(ctx.L, ctx.nl).sum(1).to(device=ms_dev, dtype=ms_dt).reshape(ms_shape)
        g_slope = torch.tensor(gz[1:4], dtype=sl_dt, device=sl_dev).reshape(sl_shape)
        s = gout.to(ms_dt)
        return g_layers * s, g_slope * gout.to(sl_dt), None, None, None, None, None, None, None, None, None, None, None


def random_magnetisation(ncell: int, Ms: float, seed: int, M_norm: np.ndarray | None = None) -> np.ndarray:
    """Reference ``random <seed>``: i.i.d. directions on the unit sphere from BorisRand.

    Serial LCG x <- 1664525 x + 1013904223 (mod 2^32) seeded with ``seed``;
    u = x / (2^32 - 1); theta = acos(1 - 2 u1), phi = 2 pi u2.  Returns (ncell, 3) float64.
    """
    a, c = np.uint64(1664525), np.uint64(1013904223)
    mask = np.uint64(0xFFFFFFFF)
    # draw 2*ncell numbers with a vectorised jump-ahead of the LCG
    cnt = 2 * ncell
    # affine composition: x_k = A_k x0 + C_k (mod 2^32), computed by doubling
    out = np.empty(cnt, dtype=np.uint64)
    A = np.uint64(1)
    Cc = np.uint64(0)
    # powers for a block of size B, then iterate blocks
    B = min(cnt, 4096)
    Ab = np.empty(B, dtype=np.uint64)
    Cb = np.empty(B, dtype=np.uint64)
    for k in range(B):
        A = (A * a) & mask
        Cc = (Cc * a + c) & mask
        Ab[k] = A
        Cb[k] = Cc
    x = np.uint64(seed) & mask
    for s in range(0, cnt, B):
        m = min(B, cnt - s)
        blk = (Ab[:m] * x + Cb[:m]) & mask
        out[s:s + m] = blk
        x = blk[-1]
    u = out.astype(np.float64) / 4294967295.0
    th = np.arccos(1.0 - 2.0 * u[0::2])
    ph = u[1::2] * 2.0 * np.pi
    norm = Ms if M_norm is None else M_norm
    m = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], axis=1)
    return m * np.asarray(norm).reshape(-1, 1)


@dataclass
class OVF2Data:
    """Contents of an OVF 2.0 file: values (cells, valuedim) float64, node counts n, cell size h
    and rectangle rect = (xmin, ymin, zmin, xmax, ymax, zmax), all in metres."""
    values: np.ndarray
    n: tuple
    h: tuple
    rect: tuple
    data_bytes: int


def read_ovf2(path) -> OVF2Data:
    """``OVF2::Read_OVF2`` (Boris/OVF2_Handlers.cpp:41-690): vector (valuedim 3) or scalar
    (valuedim 1) data, binary 4 / binary 8 / text, meshunit m or nm."""
    L = _lib.load()
    n3 = (C.c_int * 3)()
    h3 = (C.c_double * 3)()
    r6 = (C.c_double * 6)()
    vd, nb = C.c_int(), C.c_int()
    p = str(path).encode()
    check(L.bm_ovf2_info(p, n3, h3, r6, C.byref(vd), C.byref(nb)))
    ncell = n3[0] * n3[1] * n3[2]
    out = np.zeros(ncell * vd.value, dtype=np.float64)
    check(L.bm_ovf2_read(p, out.ctypes.data_as(C.c_void_p), out.size))
    return OVF2Data(out.reshape(ncell, vd.value), tuple(n3[:]), tuple(h3[:]), tuple(r6[:]), nb.value)


def write_ovf2(path, values, n, h, rect=None, data_type: str = "bin8", norm: float = 1.0) -> None:
    """``OVF2::Write_OVF2_VEC`` / ``Write_OVF2_SCA`` (Boris/OVF2_Handlers.cpp:692-902): values
    (cells, 3) or (cells,) / (cells, 1), divided by norm; data_type "bin4", "bin8" or "text";
    rect defaults to the mesh at the origin."""
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
    ncell = int(n[0]) * int(n[1]) * int(n[2])
    vd = 1 if v.size == ncell else 3
    if v.size != vd * ncell:
        raise ValueError("values must hold 1 or 3 components per cell")
    if rect is None:
        rect = (0.0, 0.0, 0.0, n[0] * h[0], n[1] * h[1], n[2] * h[2])
    codes = {"bin4": 4, "bin8": 8, "text": 0}
    if data_type not in codes:
        raise ValueError("data_type must be bin4, bin8 or text")
    L = _lib.load()
    n3 = (C.c_int * 3)(*[int(x) for x in n])
    h3 = (C.c_double * 3)(*[float(x) for x in h])
    r6 = (C.c_double * 6)(*[float(x) for x in rect])
    check(L.bm_ovf2_write(str(path).encode(), v.ctypes.data_as(C.c_void_p), vd, n3, h3, r6, codes[data_type], float(norm)))
