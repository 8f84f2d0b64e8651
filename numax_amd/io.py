"""numax.io's tensor files, with the payload going straight to or from device
memory through the library (``nmx_tensor_file_*``):

* ``numpy.save(a, path)`` / ``numpy.load(path, dtype, shape)`` /
  ``numpy.load_dyn(path, dtype)`` -- NumPy's ``.npy`` (numax/io/npy.mojo:262-453);
* ``nmx.save(a, path)`` / ``nmx.load(path, dtype, shape)`` -- numax's own NMX1
  (numax/io/io.mojo:59-194).

Arrays are torch tensors on a ROCm device (row-major, contiguous); a host or
NumPy array given to ``save`` is copied to the device first.  Errors raise
``NumaxError`` with numax's message texts.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from . import _lib
from .linalg import NumaxError, _ctx, _device_tensor, _ptr

__all__ = ["numpy", "nmx", "DTYPES"]

# dtype codes of the C ABI
DTYPES = {torch.float16: 0, torch.float32: 1, torch.float64: 2, torch.int8: 3, torch.int16: 4, torch.int32: 5,
          torch.int64: 6, torch.uint8: 7, torch.uint16: 8, torch.uint32: 9, torch.uint64: 10, torch.bool: 11}
_BY_CODE = {v: k for k, v in DTYPES.items()}
_NPY_DESCR = {0: "<f2", 1: "<f4", 2: "<f8", 3: "|i1", 4: "<i2", 5: "<i4", 6: "<i8", 7: "|u1", 8: "<u2", 9: "<u4",
              10: "<u8", 11: "|b1"}


def _code(dtype, who: str) -> int:
    if isinstance(dtype, str) or isinstance(dtype, np.dtype) or (isinstance(dtype, type) and issubclass(dtype, np.generic)):
        dtype = getattr(torch, np.dtype(dtype).name)
    if dtype not in DTYPES:
        raise NumaxError(f"numax.io.{who}: {dtype} has no NumPy dtype, so no .npy file can hold it")
    return DTYPES[dtype]


def _raise(status: int, who: str, path: str, fmt: int, dtype_code: int) -> None:
    if status == 0:
        return
    if status == _lib.NMX_ERR_IO:
        raise NumaxError(f"numax.io.{who}: cannot open, read or write {path}")
    if status == _lib.NMX_ERR_FORMAT:
        if fmt == 0:
            raise NumaxError(f"numax.io.{who}: bad magic bytes -- not a .npy file. A .npz archive is a zip container, "
                             "not an .npy file: unzip it, or re-save each array with numpy.save")
        raise NumaxError(f"numax.io.{who}: bad magic bytes -- not a numax tensor file")
    if status == _lib.NMX_ERR_DTYPE:
        if fmt == 0:
            raise NumaxError(f"numax.io.{who}: dtype mismatch -- the file holds another dtype but "
                             f"'{_NPY_DESCR[dtype_code]}' was requested")
        raise NumaxError(f"numax.io.{who}: dtype mismatch -- file was saved with a different dtype")
    if status == _lib.NMX_ERR_SHAPE:
        raise NumaxError(f"numax.io.{who}: shape mismatch")
    if status == _lib.NMX_ERR_LAYOUT:
        raise NumaxError(f"numax.io.{who}: Fortran-ordered or big-endian .npy -- numax tensors are row-major "
                         "little-endian; re-save with numpy.ascontiguousarray(arr) (and arr.astype(arr.dtype."
                         "newbyteorder('<')))")
    _lib.check(status, f"numax.io.{who}")


def _info(path: str, fmt: int, who: str):
    code = ctypes.c_int()
    rank = ctypes.c_int()
    dims = (ctypes.c_int64 * 8)()
    st = _lib.load().nmx_tensor_file_info(os.fsencode(path), fmt, ctypes.byref(code), ctypes.byref(rank), dims)
    _raise(st, who, path, fmt, -1)
    return code.value, tuple(int(dims[i]) for i in range(rank.value))


def _save(a, path: str, fmt: int, who: str) -> None:
    t = _device_tensor(a)
    code = _code(t.dtype, who)
    dims = (ctypes.c_int64 * 8)(*t.shape)
    st = _lib.load().nmx_tensor_file_save(_ctx(t).handle, os.fsencode(path), fmt, code, t.dim(), dims, _ptr(t))
    _raise(st, who, path, fmt, code)


def _load(path: str, fmt: int, dtype, shape, who: str, device="cuda:0") -> torch.Tensor:
    code = _code(dtype, who)
    shape = tuple(int(s) for s in shape)
    out = torch.empty(shape, dtype=_BY_CODE[code], device=device)
    dims = (ctypes.c_int64 * 8)(*shape)
    st = _lib.load().nmx_tensor_file_load(_ctx(out).handle, os.fsencode(path), fmt, code, len(shape), dims,
                                         _ptr(out) if out.numel() else None)
    _raise(st, who, path, fmt, code)
    return out


class numpy:  # noqa: N801 -- numax's name (numax.io.numpy)
    """NumPy's ``.npy`` (numax/io/npy.mojo:262-453)."""

    @staticmethod
    def save(a, path: str) -> None:
        """Write ``a`` as ``.npy`` format 1.0, byte-identical to ``numpy.save``."""
        _save(a, path, 0, "numpy.save")

    @staticmethod
    def load(path: str, dtype, shape, device="cuda:0") -> torch.Tensor:
        """Read an ``.npy`` file whose dtype and shape are exactly ``dtype`` /
        ``shape`` onto ``device``."""
        return _load(path, 0, dtype, shape, "numpy.load", device)

    @staticmethod
    def load_dyn(path: str, dtype, rank: int | None = None, device="cuda:0") -> torch.Tensor:
        """Read an ``.npy`` file of ``dtype`` taking its extents from the file
        (a file of another rank than ``rank``, when given, raises)."""
        code, shape = _info(path, 0, "numpy.load_dyn")
        if rank is not None and len(shape) != rank:
            raise NumaxError(f"numax.io.numpy.load_dyn: rank mismatch -- the file has rank {len(shape)}, "
                             f"not {rank}")
        return _load(path, 0, dtype, shape, "numpy.load_dyn", device)


class nmx:  # noqa: N801 -- numax's name (numax.io.nmx)
    """numax's own NMX1 format (numax/io/io.mojo:59-194)."""

    @staticmethod
    def save(a, path: str) -> None:
        _save(a, path, 1, "nmx.save")

    @staticmethod
    def load(path: str, dtype, shape, device="cuda:0") -> torch.Tensor:
        return _load(path, 1, dtype, shape, "nmx.load", device)
