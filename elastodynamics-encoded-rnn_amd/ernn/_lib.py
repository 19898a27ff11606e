"""ctypes binding of libernn_hip.so (C ABI in include/ernn.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime PyTorch-ROCm already loaded (both carry soname libamdhip64.so.7).
There is no fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import torch  # noqa: F401  (must precede loading the HIP library)

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "libernn_hip.so")

ERNN_ABI_VERSION = 7
PARAM_CONST6, PARAM_DFIELD, PARAM_FIELD6 = 0, 1, 2
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "capacity exceeded"}


class Instance(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nt", C.c_int32), ("set", C.c_int32),
                ("h", C.c_double), ("dt", C.c_double), ("rho", C.c_double),
                ("src_begin", C.c_int32), ("src_count", C.c_int32),
                ("rec_begin", C.c_int32), ("rec_count", C.c_int32)]


class Source(C.Structure):
    _fields_ = [("i", C.c_int32), ("j", C.c_int32), ("w", C.c_float), ("comp", C.c_int32)]


class Receiver(C.Structure):
    _fields_ = [("i", C.c_int32), ("j", C.c_int32)]


class Desc(C.Structure):
    _fields_ = [("n_inst", C.c_int32), ("n_sets", C.c_int32), ("param_mode", C.c_int32),
                ("rec_stride", C.c_int32), ("snapshot", C.c_int32), ("ckpt_every", C.c_int32),
                ("fwd_v", C.c_int32), ("fwd_r", C.c_int32), ("adj_v", C.c_int32),
                ("adj_r", C.c_int32), ("use_graph", C.c_int32), ("fwd_bpc", C.c_int32),
                ("adj_bpc", C.c_int32), ("fuse", C.c_int32),
                ("blk_v", C.c_int32), ("blk_rw", C.c_int32), ("blk_ks", C.c_int32),
                ("ablk_v", C.c_int32), ("ablk_rw", C.c_int32), ("ablk_ks", C.c_int32),
                ("ablk_ff_rw", C.c_int32), ("fixed_c0", C.c_int32), ("reverse_u", C.c_int32),
                ("ortho", C.c_int32),
                ("inst", C.POINTER(Instance)), ("src", C.POINTER(Source)), ("n_src", C.c_int64),
                ("rec", C.POINTER(Receiver)), ("n_rec", C.c_int64)]


class PlanInfo(C.Structure):
    _fields_ = [("workspace_bytes", C.c_int64), ("trace_floats", C.c_int64),
                ("grad_field_doubles", C.c_int64), ("field_floats", C.c_int64),
                ("n_frames_max", C.c_int32), ("ckpt_every", C.c_int32), ("n_ckpt", C.c_int32),
                ("fwd_v", C.c_int32), ("fwd_r", C.c_int32), ("adj_v", C.c_int32),
                ("adj_r", C.c_int32), ("fwd_tiles", C.c_int32), ("adj_tiles", C.c_int32),
                ("use_graph", C.c_int32), ("fwd_bpc", C.c_int32), ("adj_bpc", C.c_int32),
                ("fwd_blocks", C.c_int32), ("adj_blocks", C.c_int32), ("n_cu", C.c_int32),
                ("fuse", C.c_int32), ("blk_v", C.c_int32), ("blk_rw", C.c_int32), ("blk_ks", C.c_int32),
                ("blk_blocks", C.c_int32), ("ablk_v", C.c_int32), ("ablk_rw", C.c_int32),
                ("ablk_ks", C.c_int32), ("ablk_blocks", C.c_int32), ("ablk_ff_rw", C.c_int32),
                ("fixed_c0", C.c_int32), ("reverse_u", C.c_int32), ("ortho", C.c_int32)]


class Call(C.Structure):
    _fields_ = [("workspace", C.c_void_p), ("params", C.c_void_p), ("field", C.c_void_p),
                ("wavelets", C.c_void_p), ("traces", C.c_void_p), ("grad_traces", C.c_void_p),
                ("grad_params", C.c_void_p), ("grad_field", C.c_void_p)]


FEM_CONST, FEM_DAMAGE = 0, 1


class FemDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nt", C.c_int32), ("mode", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double), ("rho", C.c_double),
                ("rec_stride", C.c_int32), ("rec_scale", C.c_float), ("tile_v", C.c_int32),
                ("tile_r", C.c_int32), ("use_graph", C.c_int32), ("blk_rw", C.c_int32), ("blk_ks", C.c_int32),
                ("loads", C.POINTER(Source)), ("n_loads", C.c_int64)]


class FemInfo(C.Structure):
    _fields_ = [("workspace_bytes", C.c_int64), ("frame_floats", C.c_int64), ("n_frames", C.c_int32),
                ("tile_v", C.c_int32), ("tile_r", C.c_int32), ("n_tiles", C.c_int32), ("blocks", C.c_int32),
                ("n_loads", C.c_int32), ("use_graph", C.c_int32), ("blk_rw", C.c_int32), ("blk_ks", C.c_int32),
                ("blk_blocks", C.c_int32)]


class FemCall(C.Structure):
    _fields_ = [("workspace", C.c_void_p), ("params", C.c_void_p), ("field", C.c_void_p),
                ("wavelet", C.c_void_p), ("frames_x", C.c_void_p), ("frames_y", C.c_void_p)]


TV_MAX_FIELDS = 8


class TvAdamDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("n_fields", C.c_int32), ("log_params", C.c_int32),
                ("step", C.c_int32), ("pad", C.c_int32), ("lambda_tv", C.c_double), ("lr", C.c_double),
                ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double), ("weight_decay", C.c_double),
                ("lo", C.c_float * TV_MAX_FIELDS), ("hi", C.c_float * TV_MAX_FIELDS)]


class TvAdamCall(C.Structure):
    _fields_ = [("theta", C.c_void_p * TV_MAX_FIELDS), ("grad", C.c_void_p * TV_MAX_FIELDS),
                ("m", C.c_void_p * TV_MAX_FIELDS), ("v", C.c_void_p * TV_MAX_FIELDS),
                ("tv", C.c_void_p), ("workspace", C.c_void_p)]


EXPORTS = {
    "ernn_tv_adam_workspace_bytes": ([C.POINTER(TvAdamDesc), C.POINTER(C.c_int64)], C.c_int),
    "ernn_tv_adam_step": ([C.POINTER(TvAdamDesc), C.POINTER(TvAdamCall), C.c_void_p], C.c_int),
    "ernn_tv_adam_step_static": ([C.POINTER(TvAdamDesc), C.POINTER(TvAdamCall), C.c_void_p], C.c_int),
    "ernn_fem_plan_create": ([C.POINTER(FemDesc), C.POINTER(C.c_void_p)], C.c_int),
    "ernn_fem_plan_destroy": ([C.c_void_p], C.c_int),
    "ernn_fem_plan_info": ([C.c_void_p, C.POINTER(FemInfo)], C.c_int),
    "ernn_fem_run": ([C.c_void_p, C.POINTER(FemCall), C.c_void_p], C.c_int),
    "ernn_fem_set_tiling": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "ernn_fem_set_blocking": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "ernn_fem_time_steps": ([C.c_void_p, C.c_int, C.POINTER(FemCall), C.c_void_p, C.POINTER(C.c_float)],
                            C.c_int),
    "ernn_abi_version": ([], C.c_int),
    "ernn_last_error": ([], C.c_char_p),
    "ernn_plan_create": ([C.POINTER(Desc), C.POINTER(C.c_void_p)], C.c_int),
    "ernn_plan_destroy": ([C.c_void_p], C.c_int),
    "ernn_plan_info": ([C.c_void_p, C.POINTER(PlanInfo)], C.c_int),
    "ernn_trace_offsets": ([C.c_void_p, C.POINTER(C.c_int64)], C.c_int),
    "ernn_workspace_init": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "ernn_forward": ([C.c_void_p, C.POINTER(Call), C.c_void_p], C.c_int),
    "ernn_adjoint": ([C.c_void_p, C.POINTER(Call), C.c_void_p], C.c_int),
    "ernn_plan_set_tiling": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "ernn_plan_set_blocking": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "ernn_time_steps": ([C.c_void_p, C.c_int, C.c_int, C.POINTER(Call), C.c_void_p,
                         C.POINTER(C.c_float)], C.c_int),
}

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m ernn.build` "
                    "(or __graft_entry__.build()); there is no non-HIP fallback")
            h = C.CDLL(LIB_PATH)
            for name, (args, res) in EXPORTS.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = res
            v = h.ernn_abi_version()
            if v != ERNN_ABI_VERSION:
                raise RuntimeError(f"libernn_hip ABI {v}, expected {ERNN_ABI_VERSION}")
            _lib = h
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().ernn_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}")


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream
