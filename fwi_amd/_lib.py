"""ctypes binding of fwi_amd/libfwi.so (the C ABI declared in include/fwi_serialcalls.h and include/fwi.h).

Loading never falls back to anything: if the library is missing or cannot be loaded, every entry point
raises.  Device arrays are passed as raw pointers (torch tensors on the GPU: tensor.data_ptr()).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfwi.so")

c_int = ctypes.c_int
c_float = ctypes.c_float
c_double = ctypes.c_double
c_ll = ctypes.c_longlong
vp_ = ctypes.c_void_p


class FwiShot(ctypes.Structure):
    _fields_ = [("x0", c_int), ("nx", c_int), ("isx", c_int), ("isz", c_int), ("irz", c_int), ("ir0", c_int),
                ("nrec", c_int)]


class FwiSurvey(ctypes.Structure):
    _fields_ = [("nx", c_int), ("nz", c_int), ("nt", c_int), ("cpml", c_int), ("dh", c_float), ("dt", c_float),
                ("nshots", c_int), ("shots", ctypes.POINTER(FwiShot)), ("wavelet", ctypes.POINTER(c_float)),
                ("cpml_a", ctypes.POINTER(c_float)), ("cpml_b", ctypes.POINTER(c_float))]


class FwiLineSearch(ctypes.Structure):
    _fields_ = [("alpha", c_double * 3), ("count", c_int), ("max_failures", c_int)]


class FwiOptions(ctypes.Structure):
    _fields_ = [("grad_mode", c_int), ("arith", c_int), ("ckpt_interval", c_int), ("max_batch", c_int),
                ("mem_budget", c_double), ("rows_per_item", c_int), ("item_order", c_int),
                ("grad_group", c_int), ("ckpt_format", c_int), ("store_segments", c_int),
                ("use_graph", c_int), ("step_pairs", c_int), ("surface_rows", c_int),
                ("adj_rows_per_item", c_int), ("adj_item_order", c_int), ("adj_step_pairs", c_int)]


# Every exported symbol with its signature (argument types, return type).
_F = vp_  # device float*
LEGACY = {
    "pre_variables_h": ([_F] * 6 + [c_int, c_int], None),
    "CPML_x_h": ([_F, _F, c_int, c_float, c_int, c_float, c_float, c_float, c_float, c_float], None),
    "CPML_z_h": ([_F, _F, c_int, c_float, c_int, c_float, c_float, c_float, c_float, c_float], None),
    "PSI_F_h": ([_F] * 8 + [c_int, c_int, c_float, c_int], None),
    "ZETA_F_h": ([_F] * 14 + [c_int, c_int, c_float, c_float, c_int], None),
    "propagator_F_h": ([_F] * 14 + [c_float, c_float, c_int, c_int, c_int, c_int, c_int, c_int], None),
    "res_h": ([_F] * 3 + [c_int, c_int], None),
    "SecondDerivate_P_h": ([_F] * 4 + [c_float, c_int, c_int], None),
    "launchPropagation": ([vp_, vp_] + [_F] * 8 + [c_float, c_int, c_int] + [_F] * 6 + [c_float, c_int]
                          + [_F] * 9 + [c_float, c_int, c_int, c_int], None),
    "launchBackPropagation": ([vp_, vp_, c_int, c_int, c_int] + [_F] * 16 + [c_float] + [_F] * 10 + [c_float, c_int]
                              + [_F] * 9 + [c_float, _F, c_int, c_int, c_int, c_float] + [_F] * 3, None),
    "NORMAL_h": ([_F] * 3 + [c_int, c_int, c_float, c_float, c_float], None),
    "UpdateModel_h": ([_F] * 6 + [c_float, c_float, c_float, c_int, c_int], None),
    "compute_sy_h": ([_F] * 18 + [c_int, c_int, c_int], None),
    "compute_qi_h": ([_F] * 6 + [c_float, c_float, c_float, c_int, c_int, c_int], None),
    "compute_ri_h": ([_F] * 6 + [c_float, c_float, c_float, c_int, c_int], None),
    "update_r_h": ([_F] * 6 + [c_float] * 6 + [c_int, c_int, c_int], None),
    "update_LBFGS_h": ([_F] * 9 + [c_int, c_int, c_float, c_float, c_float] + [c_int] * 5, None),
    "fwi_last_error": ([], c_int),
}
BATCHED = {
    "fwi_options_default": ([ctypes.POINTER(FwiOptions)], None),
    "fwi_plan_create": ([ctypes.POINTER(FwiSurvey), ctypes.POINTER(FwiOptions), ctypes.POINTER(vp_)], c_int),
    "fwi_plan_destroy": ([vp_], None),
    "fwi_plan_set_wavelet": ([vp_, ctypes.POINTER(c_float)], c_int),
    "fwi_plan_set_cpml": ([vp_, ctypes.POINTER(c_float), ctypes.POINTER(c_float)], c_int),
    "fwi_plan_gather_size": ([vp_], c_ll),
    "fwi_plan_batch_size": ([vp_], c_int),
    "fwi_plan_batch_count": ([vp_], c_int),
    "fwi_plan_workspace_bytes": ([vp_], c_double),
    "fwi_forward": ([vp_, _F, _F, _F, _F, vp_], c_int),
    "fwi_misfit": ([vp_, _F, _F, _F, _F, ctypes.POINTER(c_double), vp_], c_int),
    "fwi_gradient": ([vp_, _F, _F, _F, _F, vp_, ctypes.POINTER(c_double), _F, vp_], c_int),
    "fwi_gradient_src": ([vp_, _F, _F, _F, _F, vp_, _F, vp_], c_int),
    "fwi_chain_rule": ([_F, _F, _F, vp_, vp_, c_ll, vp_], c_int),
    "fwi_cpml_table": ([c_int, c_float, c_float, c_float, c_float, c_float, c_float, ctypes.POINTER(c_float),
                        ctypes.POINTER(c_float)], None),
    "fwi_error_message": ([], ctypes.c_char_p),
    "fwi_lbfgs_create": ([c_int, c_int, c_int, ctypes.POINTER(vp_)], c_int),
    "fwi_lbfgs_destroy": ([vp_], None),
    "fwi_lbfgs_sd_step": ([vp_, _F, _F, _F, vp_, ctypes.POINTER(c_double), vp_], c_int),
    "fwi_lbfgs_direction": ([vp_, _F, _F, _F, vp_, vp_], c_int),
    "fwi_lbfgs_trial": ([vp_, _F, _F, _F, ctypes.POINTER(c_double), vp_], c_int),
    "fwi_lbfgs_restore": ([vp_, _F, _F, _F, vp_], c_int),
    "fwi_lbfgs_get_direction": ([vp_, vp_, vp_], c_int),
    "fwi_lbfgs_pairs": ([vp_], c_int),
    "fwi_model_clamp": ([vp_, vp_, vp_, c_ll, ctypes.c_float, ctypes.c_float, vp_], c_int),
    "fwi_dot3": ([vp_, vp_, c_ll, ctypes.POINTER(c_double), vp_], c_int),
    "fwi_dot3_noalloc": ([vp_, vp_, c_ll, vp_, vp_], c_int),
    "fwi_plan_set_timing": ([vp_, c_int], None),
    "fwi_plan_kernel_time": ([vp_, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_ll)], c_int),
    "fwi_plan_kernel_launches": ([vp_, c_int], c_ll),
    "fwi_plan_core_cells": ([vp_], c_ll),
    "fwi_plan_cells": ([vp_, ctypes.POINTER(c_ll)], c_ll),
    "fwi_plan_shot_misfits": ([vp_, ctypes.POINTER(c_double), vp_], c_int),
    "fwi_plan_ckpt_interval": ([vp_], c_int),
    "fwi_plan_rows_per_item": ([vp_], c_int),
    "fwi_plan_item_order": ([vp_], c_int),
    "fwi_plan_kept_steps": ([vp_], c_int),
    "fwi_plan_pair_rows": ([vp_], c_int),
    "fwi_plan_adj_pair_rows": ([vp_], c_int),
    "fwi_plan_adj_core_cells": ([vp_], c_ll),
    "fwi_plan_surface_rows": ([vp_], c_int),
    "fwi_plan_graph_stats": ([vp_, ctypes.POINTER(c_ll), ctypes.POINTER(c_ll)], c_int),
    "fwi_line_search_init": ([ctypes.POINTER(FwiLineSearch)], None),
    "fwi_line_search_step": ([ctypes.POINTER(FwiLineSearch), c_double, ctypes.POINTER(c_double)], c_int),
    "fwi_plan_graph_nodes": ([vp_, ctypes.POINTER(c_ll), ctypes.POINTER(c_ll)], c_int),
    "fwi_plan_adj_tiling": ([vp_, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int)], c_int),
}
SYMBOLS = {**LEGACY, **BATCHED}

_lib = None


class FwiError(RuntimeError):
    pass


def load():
    """Load libfwi.so (built by fwi_amd/build.py).  Raises if it is absent: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FwiError(f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                       " or `python fwi_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in SYMBOLS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def hip_runtime():
    """The HIP runtime already loaded in this process (libfwi.so links libamdhip64.so.7)."""
    load()
    return ctypes.CDLL("libamdhip64.so.7")


def create_blocking_stream() -> ctypes.c_void_p:
    """A stream created with hipStreamCreate: blocking with respect to the null stream, as the reference's
    cudaStreamCreate streams are (src/main.c:269-272).  torch's streams are non-blocking."""
    s = ctypes.c_void_p()
    rt = hip_runtime()
    rt.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    rc = rt.hipStreamCreate(ctypes.byref(s))
    if rc != 0:
        raise FwiError(f"hipStreamCreate failed ({rc})")
    return s


def destroy_stream(s: ctypes.c_void_p):
    rt = hip_runtime()
    rt.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    rt.hipStreamDestroy(s)


def check(rc: int, what: str):
    if rc != 0:
        msg = load().fwi_error_message()
        raise FwiError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def check_last(what: str):
    e = load().fwi_last_error()
    if e != 0:
        raise FwiError(f"{what}: HIP error {e}")
