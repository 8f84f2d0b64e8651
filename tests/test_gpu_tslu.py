"""The mixed-precision factor with tournament pivoting (tslu.hip), through the
C ABI's factor-stage test hook (nmx_mp_factor_debug) and nmx_gesv_mpir.

Tournament pivoting picks different pivot rows than partial pivoting, so
there is no pivot-by-pivot oracle: the factor is checked by what defines an
LU factorization, independently of the kernels --

* ipiv is a LAPACK interchange sequence (ipiv[j] >= j) and the permutation it
  makes is one;
* P A = L U: with P from ipiv (applied on the host), L unit lower and U upper
  read back from the FP32 factor, || P A - L U ||_max <= 64 n u32 max|A|
  (FP32 arithmetic, trailing updates in FP32 here), computed in FP64;
* |L| <= a modest bound (tournament pivoting's multipliers stay O(1) on these
  matrices; GEPP's are <= 1);
* the refined solve reaches FP64 backward error and agrees with the FP64
  oracle solve;
* an exactly zero column gives info = its 1-based index.

The sizes cover one chunk (n <= 256: the tournament is plain partial
pivoting), one merge level (G <= 8 chunks), two and three levels (n > 2048,
n > 16384 at the first leaf), ragged widths (n not a multiple of 32).
"""
import ctypes

import numpy as np
import pytest

import oracle
from numax_amd import _lib
from numax_amd import linalg as nl
from numax_amd import random as R

pytestmark = pytest.mark.gpu

U32 = np.finfo(np.float32).eps


def _factor(a, precision=3, block=256):
    import torch
    A = torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    n = A.shape[0]
    F = torch.empty((n, n), dtype=torch.float32, device="cuda")
    diags = torch.empty(4 * n, dtype=torch.float32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    lib = _lib.load()
    ctx = nl._ctx(A)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _lib.check(lib.nmx_mp_factor_debug(ctx.handle, n, P(A), n, precision, block, 0, 1, 2, P(F), P(diags), P(info),
                                       P(ipiv)), "nmx_mp_factor_debug")
    return A, F, ipiv, int(info.item())


def _perm(ipiv):
    n = len(ipiv)
    p = np.arange(n)
    for j, q in enumerate(ipiv):
        p[j], p[q] = p[q], p[j]
    return p


def _check_factor(a, F, ipiv, growth=8.0):
    """P A = L U in FP64 on the GPU (torch matmul, not this library)."""
    import torch
    n = a.shape[0]
    piv = ipiv.cpu().numpy()
    assert np.all(piv >= np.arange(n)) and np.all(piv < n), "ipiv is not a LAPACK interchange sequence"
    perm = _perm(piv)
    assert np.array_equal(np.sort(perm), np.arange(n))
    F64 = F.double()
    L = torch.tril(F64, -1) + torch.eye(n, dtype=torch.float64, device="cuda")
    U = torch.triu(F64)
    A = torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    PA = A[torch.as_tensor(perm, device="cuda")]
    err = float((PA - L @ U).abs().max())
    amax = float(A.abs().max())
    assert err <= 64 * n * U32 * amax, f"|PA - LU| = {err:.3e} (bound {64 * n * U32 * amax:.3e})"
    lmax = float(L.abs().max())
    assert lmax <= growth, f"max |L| = {lmax}"
    return err / amax, lmax


@pytest.mark.parametrize("n", [1, 5, 31, 33, 256, 257, 700, 2048, 2049, 5000])
def test_tournament_factor_is_an_lu(n):
    a = R.lu_matrix(40 + n, n)
    _, F, ipiv, info = _factor(a)
    assert info == 0
    _check_factor(a, F, ipiv)


def test_tournament_factor_three_merge_levels():
    # the first leaves span more than 64 chunks of 256 rows: three merge levels
    import torch
    n = 16900
    A = R.device_normal(77, (n, n))
    a = A.cpu().numpy()
    del A
    torch.cuda.empty_cache()
    _, F, ipiv, info = _factor(a)
    assert info == 0
    _check_factor(a, F, ipiv, growth=16.0)


def test_tournament_one_chunk_is_partial_pivoting():
    # n <= 256 is one chunk: the tournament is partial pivoting -- every
    # multiplier is at most 1 in magnitude (up to the pivot key's 2^-15
    # resolution) and the first pivot is the largest entry of column 0
    n = 200
    a = R.lu_matrix(9, n)
    _, F, ipiv, info = _factor(a)
    assert int(ipiv[0].item()) == int(np.argmax(np.abs(a[:, 0])))
    _check_factor(a, F, ipiv, growth=1.0 + 1e-3)


def test_tournament_zero_column_info():
    n = 300
    a = R.lu_matrix(12, n)
    a[:, 37] = 0.0
    _, F, ipiv, info = _factor(a)
    assert info == 38


@pytest.mark.parametrize("n,precision", [(1001, "fp16s"), (2048, "fp16s"), (3000, "fp32"), (4097, "fp16s"),
                                         (2048, "bf16")])
def test_tournament_mpir_solves_to_fp64(n, precision):
    a = R.lu_matrix(60 + n, n)
    b = R.normal(61 + n, (n,))
    x, rep = nl.solve_mpir(a, b, precision=precision, pivot="tournament")
    x = x.cpu().numpy()
    assert rep["converged"], rep
    eta = oracle.backward_error(a, x, b)
    assert eta <= 1e-15, eta
    xo = np.linalg.solve(a, b)
    assert np.max(np.abs(x - xo)) <= 1e-10 * np.max(np.abs(xo))


def test_tournament_mpir_multi_rhs():
    n, k = 1500, 8
    a = R.lu_matrix(5, n)
    B = R.normal(6, (n, k))
    X, reps = nl.solve_mpir(a, B, pivot="tournament")
    X = X.cpu().numpy()
    for j in range(k):
        assert reps[j]["converged"]
        assert oracle.backward_error(a, X[:, j], B[:, j]) <= 1e-15
