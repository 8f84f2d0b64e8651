"""GPU parity: blocked LU, triangular solves, GEMM and Cholesky through the
C ABI, against the CPU oracle (numax's float64 host algorithm restated).

Bar (the numax float64 contract): identical pivot sequences, and a normwise
backward error ||b - A x|| / (||A|| ||x||) within 10 n eps of the oracle's.
"""
import numpy as np
import pytest
import torch

import oracle
from numax_amd import linalg as nl
from numax_amd import random as R

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("transa,transb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 129, 17), (128, 128, 16), (300, 257, 256), (513, 64, 33)])
def test_gemm_matches_numpy(dtype, transa, transb, m, n, k):
    rng = np.random.default_rng(m * 1000 + n + k)
    a = rng.standard_normal((k, m) if transa else (m, k))
    b = rng.standard_normal((n, k) if transb else (k, n))
    c = rng.standard_normal((m, n))
    npd = np.float64 if dtype == torch.float64 else np.float32
    a, b, c = a.astype(npd), b.astype(npd), c.astype(npd)
    out = _np(nl.gemm(a, b, c, alpha=-1.0, beta=1.0, transa=bool(transa), transb=bool(transb)))
    ref = c.astype(np.float64) - (a.T if transa else a).astype(np.float64) @ (b.T if transb else b).astype(np.float64)
    tol = 1e-12 if dtype == torch.float64 else 2e-4
    assert np.max(np.abs(out - ref)) <= tol * max(1.0, np.max(np.abs(ref))) * np.sqrt(k)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("transa,transb", [(0, 0), (1, 1)])
def test_gemm_tile_shapes_agree_bitwise(dtype, transa, transb):
    # a 2944 x 2944 update is 23 x 23 = 529 tiles of 128 x 128 (the large-tile
    # path); its leading 384 x 320 block on its own is too small to fill the
    # device with those and takes 64 x 64 tiles.  Every element sums its k
    # slabs in the same order either way, so the two agree bit for bit.
    m = n = 2944
    k = 72
    rng = np.random.default_rng(2944 + transa)
    npd = np.float64 if dtype == torch.float64 else np.float32
    a = rng.standard_normal((k, m) if transa else (m, k)).astype(npd)
    b = rng.standard_normal((n, k) if transb else (k, n)).astype(npd)
    c = rng.standard_normal((m, n)).astype(npd)
    full = _np(nl.gemm(a, b, c, alpha=-1.0, beta=1.0, transa=bool(transa), transb=bool(transb)))
    a_s = np.ascontiguousarray(a[:, :384] if transa else a[:384])
    b_s = np.ascontiguousarray(b[:320] if transb else b[:, :320])
    part = _np(nl.gemm(a_s, b_s, np.ascontiguousarray(c[:384, :320]), alpha=-1.0, beta=1.0, transa=bool(transa),
                       transb=bool(transb)))
    assert np.array_equal(full[:384, :320], part)
    ref = c.astype(np.float64) - (a.T if transa else a).astype(np.float64) @ (b.T if transb else b).astype(np.float64)
    tol = 1e-12 if dtype == torch.float64 else 2e-4
    assert np.max(np.abs(full - ref)) <= tol * max(1.0, np.max(np.abs(ref))) * np.sqrt(k)


@pytest.mark.parametrize("n", [1, 7, 64, 65, 200, 300])
@pytest.mark.parametrize("upper,unit,trans", [(0, 0, 0), (0, 1, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)])
def test_solve_triangular_matches_oracle(n, upper, unit, trans):
    rng = np.random.default_rng(n + 10 * upper + 100 * unit + 1000 * trans)
    t = rng.standard_normal((n, n)) * 0.1 + np.eye(n) * 2.0
    b = rng.standard_normal((n, 3))
    x = _np(nl.solve_triangular(t, b, upper=bool(upper), unit=bool(unit), trans=bool(trans)))
    stored = np.triu(t) if upper else np.tril(t)
    if unit:
        stored = stored - np.diag(np.diag(stored)) + np.eye(n)
    op = stored.T if trans else stored
    np.testing.assert_allclose(op @ x, b, atol=1e-11 * n, rtol=0)
    for c in range(3):  # the oracle's vector form, column by column
        xo = oracle.solve_triangular(t, b[:, c], upper=bool(upper), unit=bool(unit), trans=bool(trans))
        np.testing.assert_allclose(x[:, c], xo, atol=1e-10, rtol=1e-10)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (5, 3), (300, 1), (4097, 16), (16384, 1), (16385, 2), (20000, 1)])
def test_laswp_matches_oracle(n, nrhs):
    # n <= 16384 forms the permutation in LDS, larger n in device memory
    rng = np.random.default_rng(n + nrhs)
    rows = np.arange(n)
    ipiv = (rows + (rng.random(n) * (n - rows))).astype(np.int32)  # ipiv[j] in [j, n)
    keep = rng.random(n) < 0.3
    ipiv[keep] = rows[keep]  # steps without an interchange
    b = rng.standard_normal((n, nrhs)) if nrhs > 1 else rng.standard_normal(n)
    out = _np(nl.laswp(b, ipiv))
    assert np.array_equal(out, oracle.laswp(b, ipiv))


def _lu_parity(a, b, block=256, base=32):
    lu = nl.lu_factor(a, block=block, base=base)
    olu, opiv, oinfo, osign = oracle.getrf(a, block=32, base=16)
    piv = _np(lu.pivots)
    assert np.array_equal(piv, opiv), f"first pivot mismatch at {np.argmax(piv != opiv)}"
    assert lu.sign == osign
    assert lu.info == oinfo
    x = _np(lu.solve(b))
    xo = oracle.getrs(olu, opiv, b, block=32)
    n = a.shape[0]
    eta = oracle.backward_error(a, x, b)
    eta_o = oracle.backward_error(a, xo, b)
    assert eta <= eta_o + 10 * n * EPS, (eta, eta_o)
    return lu, x


def test_lu_c1_oracle_case():
    """C1: n=512, A ~ U[-1,1], b ~ N(0,1), seed 1."""
    a = R.lu_matrix(1, 512)
    b = R.normal(1, (512,))
    lu, x = _lu_parity(a, b)
    assert np.isfinite(x).all()


@pytest.mark.parametrize("n", [1, 2, 5, 31, 33, 100, 257, 1000, 1025, 2100, 4097])
def test_lu_sizes_and_ragged(n):
    a = R.normal(100 + n, (n, n))
    b = R.normal(200 + n, (n, 4))
    _lu_parity(a, b)


@pytest.mark.parametrize("block,base", [(16, 16), (32, 8), (64, 32), (100, 25), (512, 32)])
def test_lu_block_invariance(block, base):
    """numax pins blocked == unblocked (docs/features.md:285-287); pivots
    must not depend on the blocking."""
    a = R.normal(7, (700, 700))
    b = R.normal(8, (700,))
    _lu_parity(a, b, block=block, base=base)


def test_lu_exchange_matrix_known_answer():
    """test_tensor_linalg.mojo:816-832: [[0,1],[1,0]] has det -1, x = (5,3)
    for b = (3,5)."""
    lu = nl.lu_factor(np.array([[0.0, 1.0], [1.0, 0.0]]))
    assert lu.det() == -1.0
    np.testing.assert_array_equal(_np(lu.solve(np.array([3.0, 5.0]))), [5.0, 3.0])


def test_lu_textbook_3x3():
    """test_linalg.mojo:134-139: x = (2, 3, -1)."""
    a = np.array([[2.0, 1.0, -1.0], [-3.0, -1.0, 2.0], [-2.0, 1.0, 2.0]])
    b = np.array([8.0, -11.0, -3.0])
    np.testing.assert_allclose(_np(nl.solve(a, b)), [2.0, 3.0, -1.0], atol=1e-14)


def test_lu_pivot_known_answer():
    """test_panel.mojo:258-274: the first pivot of [[0,1,2],[4,5,6],[7,8,10]]
    is row 2."""
    lu = nl.lu_factor(np.array([[0.0, 1.0, 2.0], [4.0, 5.0, 6.0], [7.0, 8.0, 10.0]]))
    assert int(_np(lu.pivots)[0]) == 2


def test_lu_singular_reports_info_not_error():
    a = np.array([[1.0, 2.0], [2.0, 4.0]])
    lu = nl.lu_factor(a)
    assert lu.singular() and lu.info == 2
    assert lu.slogdet() == (0.0, -np.inf)
    olu, opiv, oinfo, _ = oracle.getrf(a)
    assert oinfo == lu.info


def test_lu_zero_column_info_first_only():
    a = np.zeros((40, 40))
    lu = nl.lu_factor(a)
    assert lu.info == 1
    np.testing.assert_array_equal(_np(lu.pivots), np.arange(40))


def test_lu_scipy_plu_known_answer():
    """tests/linalg/test_lu_plu.mojo:15-109: SciPy's (P, L, U) of a 4x4 whose
    first pivot is zero, to 1e-14."""
    from tests.golden.kat import PLU_A, PLU_L, PLU_P, PLU_U
    lu = nl.lu_factor(PLU_A)
    f = _np(lu.factored)
    np.testing.assert_allclose(np.tril(f, -1) + np.eye(4), PLU_L, atol=1e-14)
    np.testing.assert_allclose(np.triu(f), PLU_U, atol=1e-14)
    perm = np.arange(4)
    for j, q in enumerate(_np(lu.pivots)):
        perm[[j, q]] = perm[[q, j]]
    p = np.eye(4)[perm].T  # A = P L U with row i of P^T A = row perm[i] of A
    np.testing.assert_allclose(p, PLU_P, atol=0)


def test_lu_solve_shape_error_text():
    lu = nl.lu_factor(np.eye(3))
    with pytest.raises(nl.NumaxError, match="lu_solve: b has 2 entries for an order-3 factorization"):
        lu.solve(np.ones(2))
    with pytest.raises(nl.NumaxError, match="square"):
        nl.lu_factor(np.ones((2, 3)))


def test_lu_det_slogdet():
    a = R.normal(11, (60, 60))
    lu = nl.lu_factor(a)
    sign, logdet = np.linalg.slogdet(a)
    s, l = lu.slogdet()
    assert s == sign and abs(l - logdet) < 1e-10 * abs(logdet)
    spd3 = np.array([[4.0, 2.0, 0.6], [2.0, 5.0, 1.0], [0.6, 1.0, 3.0]])
    assert abs(nl.lu_factor(spd3).det() - np.linalg.det(spd3)) < 1e-12


def test_lu_float32_matches_reference_gpu_tolerances():
    """tests_gpu/linalg/test_lu_plu_gpu.mojo:17-44 at float32 on its n=48
    sin-generated matrix: the pivots of the float64 host algorithm, and
    P L U == A to 1e-5 as the reference checks.  (The matrix has condition
    number ~4.5e7, so float32 L/U entries are compared through the
    reconstruction, not entry by entry against float64.)"""
    from tests.golden.kat import sin48
    a = sin48()
    lu32 = nl.lu_factor(a)
    _, opiv, _, _ = oracle.getrf(a.astype(np.float64), block=16)
    np.testing.assert_array_equal(_np(lu32.pivots), opiv)
    f = _np(lu32.factored).astype(np.float64)
    l = np.tril(f, -1) + np.eye(48)
    u = np.triu(f)
    perm = np.arange(48)
    for j, q in enumerate(_np(lu32.pivots)):
        perm[[j, q]] = perm[[q, j]]
    np.testing.assert_allclose(l @ u, a.astype(np.float64)[perm], atol=1e-5)


def test_lu_c2_scale_pivots_and_backward_error():
    """C2: n=8192, A ~ N(0,1), 16 right-hand sides, block 256: identical pivots
    with the oracle and backward error within 10 n eps of it."""
    oracle.set_threads(16)
    n = 8192
    a = R.device_normal(2, (n, n)).cpu().numpy()
    b = R.device_normal(3, (n, 16)).cpu().numpy()
    _lu_parity(a, b, block=256, base=32)


def test_lu_many_part_panels_match_oracle():
    """n = 9216: the panels of the first 1024 columns span 18 workgroups of
    512 rows, where the pivot exchange stages every part's key first and then
    only the winner's row; pivots identical to the oracle's."""
    oracle.set_threads(16)
    n = 9216
    a = R.device_normal(5, (n, n)).cpu().numpy()
    b = R.device_normal(6, (n, 4)).cpu().numpy()
    _lu_parity(a, b, block=256, base=32)


@pytest.mark.parametrize("n", [1, 5, 64, 65, 300, 1000, 1300])
def test_cholesky_matches_oracle(n):
    a = R.spd_matrix(300 + n, n)
    l = _np(nl.cholesky(a))
    lo, info = oracle.potrf(a)
    assert info == 0
    np.testing.assert_allclose(l, lo, atol=1e-12 * np.sqrt(n), rtol=0)
    assert np.array_equal(np.triu(l, 1), np.zeros((n, n)))
    b = R.normal(400 + n, (n, 3))
    x = _np(nl.cholesky_solve(l, b))
    eta = oracle.backward_error(a, x, b)
    eta_o = oracle.backward_error(a, oracle.potrs(lo, b), b)
    assert eta <= eta_o + 10 * n * EPS


def test_cholesky_reads_lower_triangle_only():
    a = R.spd_matrix(5, 150)
    junk = a + np.triu(np.full_like(a, 7.0), 1)
    np.testing.assert_allclose(_np(nl.cholesky(junk)), _np(nl.cholesky(a)), atol=0, rtol=0)


def test_cholesky_not_pd_raises_with_index():
    """test_panel.mojo:146-161: [[4,2,2],[2,1,0],[2,0,9]] fails at index 1
    (info 2)."""
    with pytest.raises(nl.NumaxError, match=r"non-positive pivot at index 1\)"):
        nl.cholesky(np.array([[4.0, 2.0, 2.0], [2.0, 1.0, 0.0], [2.0, 0.0, 9.0]]))
    a = R.spd_matrix(9, 500)
    a[321, 321] = -50.0
    _, info = oracle.potrf(a)
    with pytest.raises(nl.NumaxError, match=f"index {info - 1}\\)"):
        nl.cholesky(a)


def test_cholesky_c2_scale():
    """C2 Cholesky: A = G G^T / n + I, n = 8192, 16 right-hand sides."""
    n = 8192
    g = R.device_normal(2, (n, n))
    a = g @ g.T / n + torch.eye(n, dtype=torch.float64, device="cuda")
    a = (a + a.T) * 0.5
    l = nl.cholesky(a)
    b = R.device_normal(3, (n, 16))
    x = nl.cholesky_solve(l, b)
    r = b - a @ x
    eta = (r.abs().max(0).values / (a.abs().sum(1).max() * x.abs().max(0).values)).max().item()
    assert eta < 10 * n * EPS
