"""Seeded synthetic inputs for the factor-and-solve workloads.

Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
as 1, 2, 3", SC'11) keyed by the seed, counter = element offset.  The word
to uniform mapping and the Box--Muller form follow numax's own sampler
(numax/stats/random.mojo:112-147): word ``i`` of the stream is half of the
Philox block at offset ``i // 2``, the two 32-bit words packed high-first
into a 64-bit integer ``u``, and ``uniform = (u >> 11) * 2**-53``; a normal
is ``sqrt(-2 ln(1 - u1)) cos(2 pi u2)`` on the uniforms at words ``2i`` and
``2i + 1``.  (numax's Philox counter/key layout lives in the Mojo stdlib and
is not restated; these streams are this library's own.)

Everything here is exact integer work plus, for normals, NumPy's ``log``,
``sqrt`` and ``cos``; a matrix is generated once on the host and handed
bit-identically to every solver that consumes it.
"""
from __future__ import annotations

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)
_TWO_PI = 6.283185307179586


def philox4x32_10(ctr: np.ndarray, key: tuple[int, int]) -> np.ndarray:
    """Philox4x32-10 on an ``(N, 4)`` uint32 counter array; returns ``(N, 4)``."""
    c = [ctr[:, i].astype(np.uint64) for i in range(4)]
    k0 = np.uint32(key[0] & 0xFFFFFFFF)
    k1 = np.uint32(key[1] & 0xFFFFFFFF)
    for r in range(10):
        if r:
            k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return np.stack([x.astype(np.uint32) for x in c], axis=1)


def _blocks(seed: int, first: int, count: int) -> np.ndarray:
    off = np.arange(first, first + count, dtype=np.uint64)
    ctr = np.zeros((count, 4), dtype=np.uint32)
    ctr[:, 0] = (off & _MASK).astype(np.uint32)
    ctr[:, 1] = (off >> np.uint64(32)).astype(np.uint32)
    return philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Words ``offset .. offset+count`` of the stream as float64 in [0, 1)."""
    if count <= 0:
        return np.zeros(0, dtype=np.float64)
    first = offset // 2
    last = (offset + count - 1) // 2
    w = _blocks(seed, first, last - first + 1).astype(np.uint64)
    pairs = np.empty((w.shape[0], 2), dtype=np.uint64)
    pairs[:, 0] = (w[:, 0] << np.uint64(32)) | w[:, 1]
    pairs[:, 1] = (w[:, 2] << np.uint64(32)) | w[:, 3]
    flat = pairs.reshape(-1)[offset - 2 * first: offset - 2 * first + count]
    return (flat >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(seed: int, shape, low: float = -1.0, high: float = 1.0, offset: int = 0) -> np.ndarray:
    n = int(np.prod(shape))
    return (low + (high - low) * uniform01(seed, n, offset)).reshape(shape)


def normal(seed: int, shape, offset: int = 0) -> np.ndarray:
    """N(0, 1) via Box--Muller on the uniform pair ``(2i, 2i+1)``."""
    n = int(np.prod(shape))
    u = uniform01(seed, 2 * n, 2 * offset).reshape(-1, 2)
    return (np.sqrt(-2.0 * np.log(1.0 - u[:, 0])) * np.cos(_TWO_PI * u[:, 1])).reshape(shape)


def lu_matrix(seed: int, n: int) -> np.ndarray:
    """``A_ij ~ U[-1, 1]``, the LU oracle workload."""
    return uniform(seed, (n, n), -1.0, 1.0)


def spd_matrix(seed: int, n: int) -> np.ndarray:
    """``A = G G^T / n + I`` with ``G ~ N(0, 1)``, symmetric by construction."""
    g = normal(seed, (n, n))
    a = g @ g.T / n + np.eye(n)
    return (a + a.T) * 0.5


def _device_fill(kind: int, seed: int, shape, lo: float, hi: float, offset: int, device):
    import ctypes

    import torch

    from . import _lib
    from .linalg import _ctx

    out = torch.empty(tuple(shape), dtype=torch.float64, device=device)
    ctx = _ctx(out)
    _lib.check(_lib.load().nmx_fill_random_f64(ctx.handle, kind, seed, offset, out.numel(), lo, hi,
                                               ctypes.c_void_p(out.data_ptr())), "nmx_fill_random_f64")
    return out


def device_uniform(seed: int, shape, low: float = -1.0, high: float = 1.0, offset: int = 0, device="cuda:0"):
    """``uniform`` generated on the GPU (bit-identical to the host stream)."""
    return _device_fill(0, seed, shape, low, high, offset, device)


def device_normal(seed: int, shape, offset: int = 0, device="cuda:0"):
    """``normal`` generated on the GPU: the same Philox words as the host
    stream; the Box--Muller ``log``/``cos`` are the device's, so values can
    differ from the host stream in the last bits."""
    return _device_fill(1, seed, shape, 0.0, 0.0, offset, device)
