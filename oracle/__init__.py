"""CPU oracle for numax's dense factor-and-solve path -- TEST INFRASTRUCTURE.

ctypes bindings over ``oracle/lib/libnmx_oracle.so``, the C restatement of
numax's float64 host algorithms (``oracle/nmx_oracle.c`` cites the reference
file:line of every routine).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import this package, and only to check the
device results against or as the CPU baseline.  Nothing in ``numax_amd``
imports it.

Everything takes and returns row-major NumPy float64 arrays; the inputs are
copied, never modified.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libnmx_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i = ctypes.c_int
_vp = ctypes.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load() -> ctypes.CDLL:
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    sig = {
        "orc_array_gesv_batch": (None, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "orc_array_eigh_batch": (None, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
        "orc_array_lstsq_batch": (None, [_i, _i, _i, _vp, _vp, _vp]),
        "orc_set_threads": (_i, [_i]),
        "orc_gemm_sub": (None, [_i, _i, _i, _dp, _i, _dp, _i, _i, _dp, _i]),
        "orc_getrf": (_i, [_i, _dp, _i, _ip, _ip, _i, _i]),
        "orc_laswp": (None, [_i, _i, _dp, _i, _ip]),
        "orc_trsv": (None, [_i, _dp, _i, _dp, _i, _i, _i, _i]),
        "orc_trsm": (None, [_i, _i, _dp, _i, _dp, _i, _i, _i, _i, _i]),
        "orc_getrs": (None, [_i, _i, _dp, _i, _ip, _dp, _i, _i]),
        "orc_potrf": (_i, [_i, _dp, _i, _ip, _i, _i]),
        "orc_potrs": (None, [_i, _i, _dp, _i, _dp, _i, _i]),
        "orc_geqrf": (None, [_i, _i, _dp, _i, _dp, _i]),
        "orc_ormqr_t": (None, [_i, _i, _i, _dp, _i, _dp, _dp, _i, _i]),
        "orc_gels": (None, [_i, _i, _dp, _i, _dp, _dp, _dp, _i]),
        "orc_array_lu_factor": (_i, [_i, _dp, _ip]),
        "orc_array_lu_solve": (None, [_i, _dp, _ip, _dp, _dp]),
        "orc_array_cholesky": (None, [_i, _dp, _dp]),
        "orc_array_cholesky_solve": (None, [_i, _dp, _dp, _dp]),
        "orc_array_solve_nopivot": (None, [_i, _dp, _dp, _dp]),
        "orc_array_lu_nopivot": (None, [_i, _dp]),
        "orc_sytf2_lower": (_i, [_i, _dp, _ip]),
        "orc_ldl_unpack": (None, [_i, _dp, _ip, _dp, _dp, _vp]),
        "orc_array_inverse_nopivot": (None, [_i, _dp, _dp]),
        "orc_array_det": (ctypes.c_double, [_i, _dp, _i]),
        "orc_array_slogdet": (None, [_i, _dp, _dp, _dp]),
        "orc_array_slogdet_cholesky": (ctypes.c_double, [_i, _dp]),
        "orc_array_eigh": (None, [_i, _dp, _dp, _dp, _i]),
        "orc_array_lstsq": (None, [_i, _i, _dp, _dp, _dp]),
        "orc_backward_error": (ctypes.c_double, [_i, _i, _dp, _i, _dp, _i, _dp, _i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.array(x, dtype=np.float64, copy=True))


def _p(x: np.ndarray):
    if x.dtype == np.int32:
        return x.ctypes.data_as(_ip)
    return x.ctypes.data_as(_dp)


def _a(x: np.ndarray):
    """The address of a contiguous array (the batch entry points)."""
    return ctypes.c_void_p(x.ctypes.data)


def set_threads(t: int) -> int:
    """Threads for the row/column-parallel loops (results do not depend on it)."""
    return int(_lib.orc_set_threads(int(t)))


def getrf(a, block: int = 32, base: int = 16):
    """numax ``lu_factor`` host path: returns ``(lu, ipiv, info, sign)``."""
    lu = _f64(a)
    n = lu.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    info = np.zeros(1, dtype=np.int32)
    sign = _lib.orc_getrf(n, _p(lu), n, _p(ipiv), _p(info), block, base)
    return lu, ipiv, int(info[0]), int(sign)


def laswp(b, ipiv):
    x = _f64(b)
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    _lib.orc_laswp(x.shape[0], nrhs, _p(x), nrhs, _p(np.ascontiguousarray(ipiv, dtype=np.int32)))
    return x


def getrs(lu, ipiv, b, block: int = 32):
    """numax ``LU.solve``: vector form for a 1-D ``b``, matrix form otherwise."""
    lu = _f64(lu)
    x = _f64(b)
    n = lu.shape[0]
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    ldb = 1 if x.ndim == 1 else nrhs
    _lib.orc_getrs(n, nrhs, _p(lu), n, _p(np.ascontiguousarray(ipiv, dtype=np.int32)), _p(x), ldb, block)
    return x


def trsv(a, b, upper: bool, unit: bool = False, trans: bool = False, block: int = 16):
    a = _f64(a)
    x = _f64(b)
    _lib.orc_trsv(a.shape[0], _p(a), a.shape[0], _p(x), int(upper), int(unit), int(trans), block)
    return x


def trsm(a, b, upper: bool, unit: bool = False, trans: bool = False, block: int = 16):
    a = _f64(a)
    x = _f64(b)
    _lib.orc_trsm(a.shape[0], x.shape[1], _p(a), a.shape[0], _p(x), x.shape[1], int(upper), int(unit),
                  int(trans), block)
    return x


def solve_triangular(a, b, upper: bool = False, unit: bool = False, trans: bool = False, block: int = 16):
    """numax ``solve_triangular``: ``upper`` names the stored triangle and
    ``trans`` flips which one the substitution walks (triangular.mojo:290-296)."""
    return trsv(a, b, upper=(upper != trans), unit=unit, trans=trans, block=block)


def potrf(a, block: int = 64, tile: int = 128):
    """numax ``cholesky`` host path: returns ``(L, info)`` (L is ``tril``'d)."""
    l = _f64(a)
    n = l.shape[0]
    info = np.zeros(1, dtype=np.int32)
    _lib.orc_potrf(n, _p(l), n, _p(info), block, tile)
    return l, int(info[0])


def potrs(l, b, block: int = 16):
    l = _f64(l)
    x = _f64(b)
    n = l.shape[0]
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    ldb = 1 if x.ndim == 1 else nrhs
    _lib.orc_potrs(n, nrhs, _p(l), n, _p(x), ldb, block)
    return x


def geqrf(a, block: int = 16):
    """numax ``qr_factor``: returns ``(factored, taus)`` in LAPACK's packed form."""
    qr = _f64(a)
    m, n = qr.shape
    tau = np.zeros(n, dtype=np.float64)
    _lib.orc_geqrf(m, n, _p(qr), n, _p(tau), block)
    return qr, tau


def ormqr_t(qr, tau, b, block: int = 16):
    """numax ``QR.apply_q_transpose``: ``Q^T B`` for an ``m x k`` B."""
    qr = _f64(qr)
    tau = _f64(tau)
    x = _f64(b)
    if x.ndim == 1:
        x = x.reshape(-1, 1)
    m, n = qr.shape
    _lib.orc_ormqr_t(m, n, x.shape[1], _p(qr), n, _p(tau), _p(x), x.shape[1], block)
    return x


def gels(a, b, block: int = 16):
    """numax ``lstsq`` (method "qr"): least-squares ``x`` for ``m >= n``."""
    qr = _f64(a)
    m, n = qr.shape
    tau = np.zeros(n, dtype=np.float64)
    bb = _f64(b)
    x = np.zeros(n, dtype=np.float64)
    _lib.orc_gels(m, n, _p(qr), n, _p(tau), _p(bb), _p(x), block)
    return x


def array_lu_factor(a):
    """Register tier ``lu_factor``: ``(lu, permutation, sign)``."""
    lu = _f64(a)
    n = lu.shape[0]
    perm = np.zeros(n, dtype=np.int32)
    sign = _lib.orc_array_lu_factor(n, _p(lu), _p(perm))
    return lu, perm, int(sign)


def array_lu_solve(lu, perm, b):
    lu = _f64(lu)
    bb = _f64(b)
    x = np.zeros_like(bb)
    _lib.orc_array_lu_solve(lu.shape[0], _p(lu), _p(np.ascontiguousarray(perm, dtype=np.int32)), _p(bb), _p(x))
    return x


def array_cholesky(a):
    a = _f64(a)
    l = np.zeros_like(a)
    _lib.orc_array_cholesky(a.shape[0], _p(a), _p(l))
    return l


def array_cholesky_solve(l, b):
    l = _f64(l)
    bb = _f64(b)
    x = np.zeros_like(bb)
    _lib.orc_array_cholesky_solve(l.shape[0], _p(l), _p(bb), _p(x))
    return x


def array_solve(a, b):
    """Register tier ``solve``: unpivoted LU + substitutions."""
    a = _f64(a)
    bb = _f64(b)
    x = np.zeros_like(bb)
    _lib.orc_array_solve_nopivot(a.shape[0], _p(a), _p(bb), _p(x))
    return x


def array_lu(a):
    """Register tier ``lu``: unpivoted Doolittle, packed L\\U."""
    f = _f64(a).copy()
    _lib.orc_array_lu_nopivot(f.shape[0], _p(f))
    return f


def array_inverse(a):
    """Register tier ``inverse`` (unpivoted)."""
    a = _f64(a)
    out = np.zeros_like(a)
    _lib.orc_array_inverse_nopivot(a.shape[0], _p(a), _p(out))
    return out


def array_det(a, pivoted: bool = False) -> float:
    """Register tier ``det`` (unpivoted) or ``PivotedLU.det``."""
    a = _f64(a)
    return float(_lib.orc_array_det(a.shape[0], _p(a), int(pivoted)))


def array_slogdet(a) -> tuple[float, float]:
    """``PivotedLU.slogdet`` of ``lu_factor(a)``: ``(sign, ln|det|)``."""
    a = _f64(a)
    s = np.zeros(1)
    la = np.zeros(1)
    _lib.orc_array_slogdet(a.shape[0], _p(a), _p(s), _p(la))
    return float(s[0]), float(la[0])


def array_slogdet_cholesky(lower) -> float:
    """``slogdet_cholesky``: ``2 sum ln L_ii``."""
    lower = _f64(lower)
    return float(_lib.orc_array_slogdet_cholesky(lower.shape[0], _p(lower)))


def sytf2_lower(a):
    """``sytf2_lower``: ``(packed factor, ipiv 1-based, info)``."""
    f = _f64(a).copy()
    n = f.shape[0]
    ipiv = np.zeros(max(n, 1), dtype=np.int32)
    info = _lib.orc_sytf2_lower(n, _p(f), ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return f, ipiv[:n], int(info)


def ldl(a):
    """numax's ``ldl``: SciPy's ``(lu, d, perm)``."""
    f, ipiv, _ = sytf2_lower(a)
    n = f.shape[0]
    lu = np.zeros_like(f)
    d = np.zeros_like(f)
    perm = np.zeros(max(n, 1), dtype=np.int64)
    _lib.orc_ldl_unpack(n, _p(f), ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _p(lu), _p(d),
                        perm.ctypes.data_as(ctypes.c_void_p))
    return lu, d, perm[:n]


def array_eigh(a, sweeps: int = 12):
    """Register tier ``eigh``: ``(values unsorted, vectors as columns)``."""
    a = _f64(a)
    n = a.shape[0]
    w = np.zeros(n, dtype=np.float64)
    v = np.zeros((n, n), dtype=np.float64)
    _lib.orc_array_eigh(n, _p(a), _p(w), _p(v), sweeps)
    return w, v


def array_lstsq(a, b):
    a = _f64(a)
    bb = _f64(b)
    m, n = a.shape
    x = np.zeros(n, dtype=np.float64)
    _lib.orc_array_lstsq(m, n, _p(a), _p(bb), _p(x))
    return x


def array_gesv_batch(a, b, n, kind, offa, offb):
    """``count`` uneven systems (kind 0: ``lu_factor`` + ``PivotedLU.solve``,
    1: ``cholesky`` + ``cholesky_solve``), one problem per thread."""
    aa = _f64(a)
    bb = _f64(b)
    x = np.zeros_like(bb)
    nn = np.ascontiguousarray(n, dtype=np.int32)
    oa = np.ascontiguousarray(offa, dtype=np.int64)
    ob = np.ascontiguousarray(offb, dtype=np.int64)
    kd = np.ascontiguousarray(kind, dtype=np.uint8)
    _lib.orc_array_gesv_batch(len(nn), _a(nn), _a(oa), _a(ob), _a(kd), _a(aa), _a(bb), _a(x))
    return x


def array_eigh_batch(a, n, offa, offw, sweeps: int = 12):
    """``count`` register-tier ``eigh`` problems packed back to back, one per
    thread: ``(w, v)`` laid out like ``a`` (values at offw, vectors at offa)."""
    aa = _f64(a)
    nn = np.ascontiguousarray(n, dtype=np.int32)
    ow = np.ascontiguousarray(offw, dtype=np.int64)
    w = np.zeros(int((ow + nn).max()) if len(nn) else 1, dtype=np.float64)
    v = np.zeros_like(aa)
    oa = np.ascontiguousarray(offa, dtype=np.int64)
    _lib.orc_array_eigh_batch(len(nn), _a(nn), _a(oa), _a(ow), _a(aa), _a(w), _a(v), int(sweeps))
    return w, v


def array_lstsq_batch(a, b):
    """``count`` register-tier ``lstsq`` problems (a: count x m x n, b: count
    x m), one per thread: x (count x n)."""
    aa = _f64(a)
    bb = _f64(b)
    count, m, n = aa.shape
    x = np.zeros((count, n), dtype=np.float64)
    _lib.orc_array_lstsq_batch(count, m, n, _a(aa), _a(bb), _a(x))
    return x


def backward_error(a, x, b) -> float:
    """``max_c ||b - A x||_inf / (||A||_inf ||x||_inf)``."""
    a = _f64(a)
    xx = _f64(x)
    bb = _f64(b)
    if xx.ndim == 1:
        xx = xx.reshape(-1, 1)
        bb = bb.reshape(-1, 1)
    n = a.shape[0]
    return float(_lib.orc_backward_error(n, xx.shape[1], _p(a), n, _p(xx), xx.shape[1], _p(bb), bb.shape[1]))
