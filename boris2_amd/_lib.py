"""ctypes binding of libboris_mi355x.so (the C-ABI declared in include/boris_mi355x.h).

The library is the only compute path: if it is missing or no gfx950 device is visible the
calls raise instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libboris_mi355x.so"

BM_EVAL_TEULER, BM_EVAL_RK4, BM_EVAL_AHEUN, BM_EVAL_RKDP54 = 1, 2, 7, 9
BM_MOD_DEMAG, BM_MOD_EXCHANGE, BM_MOD_ANIS_UNI, BM_MOD_ZEEMAN = 1, 2, 4, 8

# every entry point declared in include/boris_mi355x.h
EXPORTS = [
    "bm_last_error", "bm_device_ok",
    "bm_demag_create", "bm_demag_create_pbc", "bm_demag_destroy", "bm_demag_info", "bm_demag_convolute",
    "bm_demag_convolute_planar", "bm_demag_copy_kernels", "bm_demag_convolute_averaged",
    "bm_demag_convolute_averaged_static",
    "bm_ode_create", "bm_ode_destroy", "bm_ode_set_ms_map", "bm_ode_set_params", "bm_ode_set_layers", "bm_ode_set_dt",
    "bm_ode_set_astep", "bm_ode_set_pbc", "bm_ode_set_layer_meshes", "bm_ode_set_eval_speedup", "bm_ode_get_speedup_state", "bm_ode_advance_time", "bm_ode_run_to_iteration", "bm_ode_iterate_stage",
    "bm_ode_restore", "bm_ode_get_state", "bm_ode_keep_heff", "bm_ode_copy_heff", "bm_ode_average_m",
    "bm_ode_set_extra_field",
    "bm_prof_begin", "bm_prof_end", "bm_prof_query",
    "bm_mlayer_create", "bm_mlayer_create_pbc", "bm_mlayer_create_meshes", "bm_mlayer_destroy", "bm_mlayer_convolute", "bm_mlayer_copy_kernels",
    "bm_mlayer_forward", "bm_mlayer_multiply", "bm_mlayer_inverse",
    "bm_batch_create", "bm_batch_create_layers", "bm_batch_destroy", "bm_batch_set_dt", "bm_batch_run", "bm_batch_get_state",
    "bm_batch_mesh_offset", "bm_batch_autotune_report", "bm_batch_set_schedule", "bm_batch_set_pbc", "bm_batch_set_astep", "bm_batch_get_mesh_state",
    "bm_adj_create", "bm_adj_destroy", "bm_adj_forward", "bm_adj_backward", "bm_adj_zeeman_grads", "bm_adj_set_layers", "bm_adj_set_pbc",
    "bm_ovf2_info", "bm_ovf2_read", "bm_ovf2_write", "bm_adj_batch_forward", "bm_adj_batch_backward",
]

PROF_KINDS = ["fwd_x_r2c", "y_fft", "z_fft_kernel_ifft", "y_ifft", "inv_x_c2r", "y2d_fft_kernel_ifft",
              "llg_stage", "rkdp_advance", "row_c2r_stage_r2c", "fwd_x_advance", "ml_kernel_product"]


# bm_field_fn: void (*)(void* user, const float* M, float* Hext, void* stream)
FIELD_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)


class FmParams(C.Structure):
    _fields_ = [
        ("Ms", C.c_double), ("A", C.c_double), ("alpha", C.c_double),
        ("K1", C.c_double), ("K2", C.c_double), ("ea", C.c_double * 3), ("grel", C.c_double),
        ("Ha", C.c_double * 3), ("wave", C.c_int), ("amp", C.c_double), ("w", C.c_double),
        ("t0", C.c_double), ("dir", C.c_double * 3), ("slope", C.c_double * 3), ("modules", C.c_int),
    ]


class LayerMesh(C.Structure):
    """bm_layer_mesh: one layer mesh of its own cells and position (bm_mlayer_create_meshes)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double)]


class LayerDesc(C.Structure):
    _fields_ = [("nlayers", C.c_int), ("hz", C.c_void_p), ("z0", C.c_void_p), ("Ms", C.c_void_p),
                ("A", C.c_void_p), ("alpha", C.c_void_p), ("K1", C.c_void_p), ("mesh_of_layer", C.c_void_p),
                ("Ha", C.c_void_p)]


class OdeState(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("stagetime", C.c_double), ("dT", C.c_double), ("dT_last", C.c_double),
        ("lte", C.c_double), ("iteration", C.c_longlong), ("stageiteration", C.c_longlong),
        ("rejections", C.c_longlong), ("evaluations", C.c_longlong),
        ("evalStep", C.c_int), ("primed", C.c_int), ("available", C.c_int),
    ]


_lib = None


def load() -> C.CDLL:
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m boris2_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    vp, i, d, ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
    sig = {
        "bm_last_error": (C.c_char_p, []),
        "bm_device_ok": (i, []),
        "bm_demag_create": (i, [i, i, i, d, d, d, i, C.POINTER(vp)]),
        "bm_demag_create_pbc": (i, [i, i, i, d, d, d, i, i, i, i, C.POINTER(vp)]),
        "bm_demag_destroy": (i, [vp]),
        "bm_demag_info": (i, [vp, C.POINTER(i), C.POINTER(ll)]),
        "bm_demag_convolute": (i, [vp, vp, vp, i, vp, vp]),
        "bm_demag_convolute_planar": (i, [vp, vp, vp, vp]),
        "bm_demag_convolute_averaged": (i, [vp, vp, vp, vp, vp, i, vp, vp]),
        "bm_demag_convolute_averaged_static": (i, [vp, vp, vp, vp, vp, i, vp, vp]),
        "bm_demag_copy_kernels": (i, [vp, vp, ll]),
        "bm_mlayer_create": (i, [i, i, i, d, d, vp, vp, i, C.POINTER(vp)]),
        "bm_mlayer_create_pbc": (i, [i, i, i, d, d, vp, vp, i, i, i, C.POINTER(vp)]),
        "bm_mlayer_create_meshes": (i, [i, vp, i, C.POINTER(vp)]),
        "bm_mlayer_destroy": (i, [vp]),
        "bm_mlayer_convolute": (i, [vp, vp, vp, i, vp]),
        "bm_mlayer_copy_kernels": (i, [vp, vp, ll]),
        "bm_mlayer_forward": (i, [vp, i, vp, vp]),
        "bm_mlayer_multiply": (i, [vp, vp]),
        "bm_mlayer_inverse": (i, [vp, i, vp, vp, i, vp, vp]),
        "bm_batch_create": (i, [i, i, vp, vp, vp, vp, i, C.POINTER(vp)]),
        "bm_batch_create_layers": (i, [i, i, vp, vp, vp, vp, vp, i, C.POINTER(vp)]),
        "bm_batch_destroy": (i, [vp]),
        "bm_batch_set_dt": (i, [vp, d]),
        "bm_batch_run": (i, [vp, ll, vp]),
        "bm_batch_get_state": (i, [vp, C.POINTER(d), C.POINTER(ll)]),
        "bm_batch_mesh_offset": (i, [vp, i, C.POINTER(ll), C.POINTER(ll)]),
        "bm_batch_autotune_report": (i, [vp, C.c_char_p, i]),
        "bm_batch_set_schedule": (i, [vp, C.c_char_p, i, i, i, i]),
        "bm_batch_set_pbc": (i, [vp, i, i, i, i]),
        "bm_batch_set_astep": (i, [vp, d, d, d, d]),
        "bm_batch_get_mesh_state": (i, [vp, i, C.POINTER(d), C.POINTER(d), C.POINTER(ll), C.POINTER(ll)]),
        "bm_adj_create": (i, [i, i, i, d, d, d, C.POINTER(FmParams), vp, vp, d, i, i, C.POINTER(vp)]),
        "bm_adj_destroy": (i, [vp]),
        "bm_adj_forward": (i, [vp, vp, vp, vp, C.POINTER(d), vp]),
        "bm_adj_backward": (i, [vp, vp, C.POINTER(d), vp]),
        "bm_adj_zeeman_grads": (i, [vp, C.POINTER(d)]),
        "bm_adj_set_layers": (i, [vp, vp, vp, vp, vp]),
        "bm_adj_set_pbc": (i, [vp, i, i, i]),
        "bm_adj_batch_forward": (i, [i, vp, vp, vp, vp, vp, vp]),
        "bm_adj_batch_backward": (i, [i, vp, vp, vp, vp]),
        "bm_ode_create": (i, [i, i, i, i, d, d, d, C.POINTER(FmParams), vp, vp, C.POINTER(vp)]),
        "bm_ode_destroy": (i, [vp]),
        "bm_ode_set_ms_map": (i, [vp, vp]),
        "bm_ode_set_params": (i, [vp, C.POINTER(FmParams)]),
        "bm_ode_set_layers": (i, [vp, vp, vp, vp, vp, vp]),
        "bm_ode_set_dt": (i, [vp, d]),
        "bm_ode_set_astep": (i, [vp, d, d, d, d]),
        "bm_ode_set_pbc": (i, [vp, i, i, i]),
        "bm_ode_set_layer_meshes": (i, [vp, vp, vp]),
        "bm_ode_set_eval_speedup": (i, [vp, i, d]),
        "bm_ode_get_speedup_state": (i, [vp, C.POINTER(i), C.POINTER(d), C.POINTER(ll)]),
        "bm_ode_advance_time": (i, [vp, vp]),
        "bm_ode_run_to_iteration": (i, [vp, ll, vp]),
        "bm_ode_iterate_stage": (i, [vp, vp]),
        "bm_ode_restore": (i, [vp, vp]),
        "bm_ode_get_state": (i, [vp, C.POINTER(OdeState)]),
        "bm_ode_keep_heff": (i, [vp, i]),
        "bm_ode_copy_heff": (i, [vp, vp, vp]),
        "bm_ode_average_m": (i, [vp, C.POINTER(d), vp]),
        "bm_ode_set_extra_field": (i, [vp, FIELD_FN, vp, vp]),
        "bm_ovf2_info": (i, [C.c_char_p, vp, vp, vp, C.POINTER(i), C.POINTER(i)]),
        "bm_ovf2_read": (i, [C.c_char_p, vp, ll]),
        "bm_ovf2_write": (i, [C.c_char_p, vp, i, vp, vp, vp, i, d]),
        "bm_prof_begin": (i, [i]),
        "bm_prof_end": (i, []),
        "bm_prof_query": (i, [i, C.POINTER(ll), C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = load().bm_last_error().decode(errors="replace")
        raise RuntimeError(f"boris_mi355x error {rc}: {msg}")
