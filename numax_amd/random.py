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


# ---- the benchmark workloads, generated on the GPU ---------------------------
def device_haar(seed: int, n: int, device="cuda:0"):
    """Haar-distributed orthogonal n x n: Q of the QR of a seeded N(0,1)
    matrix with Mezzadri's sign fix (Q diag(sign(diag R)))."""
    import torch
    g = device_normal(seed, (n, n), device=device)
    q, r = torch.linalg.qr(g)
    d = torch.sign(torch.diagonal(r))
    d[d == 0] = 1.0
    del g, r
    return q * d


def device_svd_matrix(n: int, kappa: float, seed: int = 3, factors=None, device="cuda:0"):
    """The mixed-precision workload: ``A = U diag(sigma) V^T`` with U, V Haar
    (seeds ``seed`` and ``seed + 100``) and ``sigma_i = kappa^(-i/(n-1))``.
    ``factors`` may pass ``(U, V)`` already generated, to build several kappas
    from one pair."""
    import torch
    u, v = factors if factors is not None else (device_haar(seed, n, device), device_haar(seed + 100, n, device))
    i = torch.arange(n, dtype=torch.float64, device=u.device)
    sigma = kappa ** (-i / max(n - 1, 1))
    return ((u * sigma) @ v.T).contiguous()


def device_uneven_batch(count: int, seed: int, device="cuda:0"):
    """The uneven batch: n ~ U{8..256}; even problems LU with
    ``A ~ U[-1,1] + 0.1 n I``, odd problems Cholesky with ``A = G G^T / n + I``;
    b ~ N(0,1).  Packed back to back (row-major); returns
    ``(a, b, n, kind, offsets_a, offsets_b)`` with a, b on the device."""
    import torch
    rng = np.random.default_rng(seed)
    n = rng.integers(8, 257, size=count).astype(np.int32)
    kind = (np.arange(count) % 2).astype(np.uint8)
    offa = np.zeros(count, dtype=np.int64)
    offb = np.zeros(count, dtype=np.int64)
    if count:
        offa[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
        offb[1:] = np.cumsum(n.astype(np.int64))[:-1]
    a = torch.empty(max(int((n.astype(np.int64) ** 2).sum()), 1), dtype=torch.float64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    for size in np.unique(n):
        size = int(size)
        for kd in (0, 1):
            idx = np.flatnonzero((n == size) & (kind == kd))
            if idx.size == 0:
                continue
            eye = torch.eye(size, dtype=torch.float64, device=device)
            if kd == 0:
                m = torch.rand((idx.size, size, size), dtype=torch.float64, device=device, generator=gen) * 2 - 1
                m += 0.1 * size * eye
            else:
                g = torch.randn((idx.size, size, size), dtype=torch.float64, device=device, generator=gen)
                m = torch.bmm(g, g.transpose(1, 2).contiguous()) / size + eye
            dst = torch.as_tensor(offa[idx], device=device)[:, None] + torch.arange(size * size, device=device)
            a[dst.reshape(-1)] = m.reshape(-1)
    b = torch.randn(max(int(n.sum()), 1), dtype=torch.float64, device=device, generator=gen)
    return a, b, n, kind, offa, offb


def hamiltonian(n: int, omega: float) -> np.ndarray:
    """Finite-difference harmonic well on [-4, 4] (numax's quantum_well
    example, examples/advanced/quantum_well.mojo:66-80): diagonal
    ``1/h^2 + omega^2 x^2 / 2``, off-diagonal ``-1/(2 h^2)``, ``h = 8/(n-1)``."""
    h = 8.0 / (n - 1)
    x = np.linspace(-4.0, 4.0, n)
    a = np.diag(1.0 / h ** 2 + 0.5 * omega ** 2 * x ** 2)
    off = -0.5 / h ** 2
    return a + np.diag(np.full(n - 1, off), 1) + np.diag(np.full(n - 1, off), -1)


def eigh_batch(count: int, seed: int, sizes=(24, 32, 48, 64)):
    """The batched eigensolver workload: ``count`` symmetric float32 matrices,
    equal shares of each size; in each share half are finite-difference
    Hamiltonians (omega ~ U[0.5, 2]) and half ``(G + G^T)/2`` with
    ``G ~ N(0,1)``.  Returns ``(a, n, offsets_a, offsets_w)`` on the host."""
    rng = np.random.default_rng(seed)
    per = count // len(sizes)
    ns = np.repeat(np.asarray(sizes, dtype=np.int32), per)
    ns = np.concatenate([ns, np.full(count - ns.size, sizes[-1], dtype=np.int32)])
    offa = np.zeros(count, dtype=np.int64)
    offw = np.zeros(count, dtype=np.int64)
    offa[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    offw[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    a = np.empty(int((ns.astype(np.int64) ** 2).sum()), dtype=np.float32)
    for size in sizes:
        idx = np.flatnonzero(ns == size)
        half = idx.size // 2
        blk = np.empty((idx.size, size, size), dtype=np.float64)
        h = 8.0 / (size - 1)
        x = np.linspace(-4.0, 4.0, size)
        omega = rng.uniform(0.5, 2.0, size=half)
        blk[:half] = 0.0
        di = np.arange(size)
        blk[:half, di, di] = 1.0 / h ** 2 + 0.5 * (omega[:, None] ** 2) * x[None, :] ** 2
        blk[:half, di[:-1], di[1:]] = -0.5 / h ** 2
        blk[:half, di[1:], di[:-1]] = -0.5 / h ** 2
        g = rng.standard_normal((idx.size - half, size, size))
        blk[half:] = 0.5 * (g + g.transpose(0, 2, 1))
        dst = offa[idx][:, None] + np.arange(size * size)
        a[dst.reshape(-1)] = blk.reshape(-1).astype(np.float32)
    return a, ns, offa, offw


def lstsq_batch(count: int, m: int, n: int, seed: int):
    """The batched least-squares workload: ``count`` problems with
    ``A ~ N(0,1)`` (m x n) and ``b ~ N(0,1)``, float32, on the host."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((count, m, n), dtype=np.float32)
    b = rng.standard_normal((count, m), dtype=np.float32)
    return a, b
