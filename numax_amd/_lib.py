"""Loader and ctypes signatures for ``libnumax_hip.so`` (the C ABI in
``include/numax_hip.h``).

There is no fallback: if the library is missing or no GPU is visible, every
entry point raises.  Device memory and streams come from PyTorch (plumbing);
all compute runs in the HIP kernels of the library.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libnumax_hip.so")

_c_int = ctypes.c_int
_c_vp = ctypes.c_void_p
_c_dbl = ctypes.c_double
_c_flt = ctypes.c_float
_c_i64 = ctypes.c_int64
_c_u64 = ctypes.c_uint64

# name -> (restype, argtypes); every declaration in include/numax_hip.h
SIGNATURES = {
    "nmx_ctx_create": (_c_int, [_c_int, _c_vp, ctypes.POINTER(_c_vp)]),
    "nmx_ctx_destroy": (_c_int, [_c_vp]),
    "nmx_ctx_set_stream": (_c_int, [_c_vp, _c_vp]),
    "nmx_sync": (_c_int, [_c_vp]),
    "nmx_ctx_check": (_c_int, [_c_vp]),
    "nmx_status_string": (ctypes.c_char_p, [_c_int]),
    "nmx_version": (_c_int, []),
    "nmx_ctx_set_profiling": (_c_int, [_c_vp, _c_int]),
    "nmx_ctx_profile": (_c_int, [_c_vp, _c_int, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_dbl),
                                 ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_dbl)]),
    "nmx_kernel_name": (ctypes.c_char_p, [_c_int]),
    "nmx_getrf_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_getrf_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_getrs_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int]),
    "nmx_getrs_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int]),
    "nmx_laswp_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp]),
    "nmx_laswp_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp]),
    "nmx_diag_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp]),
    "nmx_diag_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp]),
    "nmx_trsm_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_trsm_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_potrf_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_potrf_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_potrs_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_potrs_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_gesv_mpir_nrhs": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_int, _c_vp,
                                    _c_vp]),
    "nmx_graph_begin": (_c_int, [_c_vp]),
    "nmx_graph_end": (_c_int, [_c_vp, ctypes.POINTER(_c_vp)]),
    "nmx_graph_launch": (_c_int, [_c_vp]),
    "nmx_graph_destroy": (_c_int, [_c_vp]),
    "nmx_batch_autotune": (_c_int, [_c_vp, _c_int, _c_vp]),
    "nmx_ir_autotune": (_c_int, [_c_vp, _c_int, _c_vp]),
    "nmx_geqrf_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_geqrf_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_ormqr_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_ormqr_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_orgqr_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_orgqr_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_qr_r_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_qr_r_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_int]),
    "nmx_gels_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_gels_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_int]),
    "nmx_gemm_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_dbl, _c_vp, _c_int, _c_vp,
                              _c_int, _c_dbl, _c_vp, _c_int]),
    "nmx_gemm_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_flt, _c_vp, _c_int, _c_vp,
                              _c_int, _c_flt, _c_vp, _c_int]),
    "nmx_fill_random_f64": (_c_int, [_c_vp, _c_int, _c_u64, _c_i64, _c_i64, _c_dbl, _c_dbl, _c_vp]),
    "nmx_batch_cfg_default": (None, [_c_vp]),
    "nmx_batched_syevj_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64, _c_vp,
                                       _c_int]),
    "nmx_batched_gels_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_vp, _c_vp, _c_vp]),
    "nmx_getri_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int]),
    "nmx_getri_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int]),
    "nmx_lu_unpack_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_int, _c_vp, _c_int, _c_vp,
                                   _c_int]),
    "nmx_lu_unpack_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_int, _c_vp, _c_int, _c_vp,
                                   _c_int]),
    "nmx_batched_lu_nopiv_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "nmx_batched_solve_nopiv_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64,
                                             _c_vp]),
    "nmx_batched_inverse_nopiv_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "nmx_batched_det_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_i64, _c_int, _c_vp]),
    "nmx_batched_slogdet_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_vp]),
    "nmx_batched_slogdet_cholesky_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "nmx_map_blocks_solve_f64": (_c_int, [_c_vp, _c_int, _c_int, _c_i64, _c_vp, _c_vp]),
    "nmx_map_blocks_solve_f32": (_c_int, [_c_vp, _c_int, _c_int, _c_i64, _c_vp, _c_vp]),
    "nmx_sytrf_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp]),
    "nmx_sytrf_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp]),
    "nmx_ldl_unpack_f64": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_vp, _c_int, _c_vp]),
    "nmx_ldl_unpack_f32": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_int, _c_vp, _c_int, _c_vp]),
    "nmx_tensor_file_info": (_c_int, [ctypes.c_char_p, _c_int, _c_vp, _c_vp, _c_vp]),
    "nmx_tensor_file_load": (_c_int, [_c_vp, ctypes.c_char_p, _c_int, _c_int, _c_int, _c_vp, _c_vp]),
    "nmx_tensor_file_save": (_c_int, [_c_vp, ctypes.c_char_p, _c_int, _c_int, _c_int, _c_vp, _c_vp]),
    "nmx_batched_gesv": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64, _c_vp,
                                  _c_vp, _c_vp, _c_vp, _c_vp]),
    "nmx_batched_gesv_noalloc": (_c_int, [_c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64,
                                          _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "nmx_ir_cfg_default": (None, [_c_vp]),
    "nmx_gesv_mpir": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_vp, _c_vp, _c_vp, _c_vp]),
    "nmx_ir_report_read": (_c_int, [_c_vp, _c_vp]),
    "nmx_mp_factor_debug": (_c_int, [_c_vp, _c_int, _c_vp, _c_int, _c_int, _c_int, _c_u64, _c_int, _c_int,
                                     _c_vp, _c_vp, _c_vp, _c_vp]),
}


class IrCfg(ctypes.Structure):
    """``nmx_ir_cfg`` (include/numax_hip.h)."""
    _fields_ = [("precision", _c_int), ("block", _c_int), ("max_iter", _c_int), ("solver", _c_int),
                ("gmres_restart", _c_int), ("tol", _c_dbl), ("seed", _c_u64), ("pivot", _c_int)]


class IrReport(ctypes.Structure):
    """``nmx_ir_report`` (include/numax_hip.h)."""
    _fields_ = [("iterations", _c_int), ("gmres_iterations", _c_int), ("converged", _c_int), ("info", _c_int),
                ("backward_error", _c_dbl), ("factor_ms", _c_dbl), ("refine_ms", _c_dbl),
                ("history", _c_dbl * 32)]

class BatchCfg(ctypes.Structure):
    """``nmx_batch_cfg`` (include/numax_hip.h)."""
    _fields_ = [("max_iter", _c_int), ("tol", _c_dbl), ("teams_per_cu", _c_int * 3)]


_lib = None
_lock = threading.Lock()


class NativeError(RuntimeError):
    """A nonzero status from libnumax_hip."""


def load() -> ctypes.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                    "(numax_amd has no CPU fallback)")
            # One HIP runtime per process: PyTorch ships its own libamdhip64, so
            # it is loaded first and the library binds to it (same soname);
            # loaded the other way round the two runtimes see no common device.
            import torch  # noqa: F401
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


# status codes (include/numax_hip.h)
NMX_OK = 0
NMX_ERR_HIP = 1
NMX_ERR_ALLOC = 2
NMX_ERR_UNSUPPORTED = 3
NMX_ERR_SYNC_TIMEOUT = 4
NMX_ERR_NOT_CONVERGED = 5
NMX_ERR_IO = 6
NMX_ERR_FORMAT = 7
NMX_ERR_DTYPE = 8
NMX_ERR_SHAPE = 9
NMX_ERR_LAYOUT = 10


def status_string(status: int) -> str:
    return load().nmx_status_string(int(status)).decode()


def check(status: int, what: str) -> None:
    if status != 0:
        if status < 0:
            raise ValueError(f"{what}: argument {-status} is invalid")
        raise NativeError(f"{what}: {status_string(status)} (status {status})")


class Context:
    """One ``nmx_ctx`` bound to a device and a stream handle."""

    def __init__(self, device: int = 0, stream: int | None = None):
        lib = load()
        handle = _c_vp()
        check(lib.nmx_ctx_create(int(device), _c_vp(stream) if stream else None, ctypes.byref(handle)),
              "nmx_ctx_create")
        self.handle = handle
        self.device = device
        self.stream = stream

    def set_stream(self, stream: int | None) -> None:
        if stream != self.stream:
            check(load().nmx_ctx_set_stream(self.handle, _c_vp(stream) if stream else None), "nmx_ctx_set_stream")
            self.stream = stream

    def sync(self) -> None:
        check(load().nmx_sync(self.handle), "nmx_sync")

    def check_device(self) -> None:
        check(load().nmx_ctx_check(self.handle), "device")

    def set_profiling(self, on: bool) -> None:
        check(load().nmx_ctx_set_profiling(self.handle, int(bool(on))), "nmx_ctx_set_profiling")

    def profile(self) -> dict:
        """Per kernel class: launches, milliseconds, algorithmic flops and bytes."""
        lib = load()
        out = {}
        for kind in range(14):
            n = _c_i64()
            ms, fl, by = _c_dbl(), _c_dbl(), _c_dbl()
            check(lib.nmx_ctx_profile(self.handle, kind, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl),
                                      ctypes.byref(by)), "nmx_ctx_profile")
            if n.value:
                out[lib.nmx_kernel_name(kind).decode()] = {"launches": n.value, "ms": ms.value,
                                                            "flops": fl.value, "bytes": by.value}
        return out

    def close(self) -> None:
        if self.handle:
            load().nmx_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts: dict[int, Context] = {}


def context_for(device: int, stream: int | None) -> Context:
    """The per-device context, pointed at ``stream`` (torch's current one)."""
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device, stream)
        _contexts[device] = ctx
    else:
        ctx.set_stream(stream)
    return ctx
