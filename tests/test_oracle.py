"""Pins the CPU oracle before anything is checked against it.

Each oracle routine is held to something independent of its own formula:
LAPACK via the committed SciPy/NumPy goldens (tests/golden/goldens.npz, made
by tests/golden/make_goldens.py), the reference's own known-answer tests
(tests/golden/kat.py), exact rational arithmetic on small integer matrices,
closed forms, and invariants (P A = L U, L L^T = A, Q^T Q = I, trace and
Frobenius identities).  A planted-mistake test shows the pivot checks have
teeth.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from numax_amd import random as R
from tests.golden import kat

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "goldens.npz"))
EPS = np.finfo(np.float64).eps


def perm_from_ipiv(ipiv):
    perm = np.arange(len(ipiv))
    for j, q in enumerate(ipiv):
        perm[[j, q]] = perm[[q, j]]
    return perm


# ---------------------------------------------------------------- generator
def test_philox_known_answers():
    """Random123's published Philox4x32-10 vectors."""
    kats = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
            ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
            ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
             (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kats:
        got = R.philox4x32_10(np.array([ctr], dtype=np.uint32), key)[0]
        assert tuple(int(x) for x in got) == want


def test_uniform_stream_offsets_agree():
    full = R.uniform01(9, 101)
    np.testing.assert_array_equal(R.uniform01(9, 50, offset=51), full[51:])
    assert full.min() >= 0 and full.max() < 1


# ---------------------------------------------------------------- LU
def test_getrf_c1_matches_lapack():
    a = R.lu_matrix(1, 512)
    b = R.normal(1, (512,))
    lu, piv, info, sign = oracle.getrf(a)
    np.testing.assert_array_equal(piv, G["c1_piv"])
    x = oracle.getrs(lu, piv, b)
    np.testing.assert_allclose(x, G["c1_x"], rtol=1e-10, atol=1e-12)
    assert oracle.backward_error(a, x, b) < 10 * 512 * EPS


@pytest.mark.parametrize("n", [4, 17, 64, 130])
def test_getrf_small_matches_lapack(n):
    a = R.normal(1000 + n, (n, n))
    lu, piv, info, _ = oracle.getrf(a, block=8, base=4)
    np.testing.assert_array_equal(piv, G[f"lu{n}_piv"])
    np.testing.assert_allclose(lu, G[f"lu{n}_lu"], atol=1e-12 * n)


def test_getrf_plu_known_answer():
    lu, piv, _, _ = oracle.getrf(kat.PLU_A)
    np.testing.assert_array_equal(piv, G["plu_piv"])
    np.testing.assert_allclose(np.tril(lu, -1) + np.eye(4), kat.PLU_L, atol=1e-14)
    np.testing.assert_allclose(np.triu(lu), kat.PLU_U, atol=1e-14)
    p = np.eye(4)[perm_from_ipiv(piv)].T
    np.testing.assert_array_equal(p, kat.PLU_P)


def exact_gepp(a):
    """Partial pivoting in exact rational arithmetic: first maximum of |a|,
    full-row swaps -- the mathematical object every pivot rule restates."""
    n = len(a)
    m = [[Fraction(int(v)) for v in row] for row in a]
    piv = []
    for k in range(n):
        best = max(range(k, n), key=lambda i: (abs(m[i][k]), -i))
        piv.append(best)
        m[k], m[best] = m[best], m[k]
        if m[k][k] == 0:
            continue
        for i in range(k + 1, n):
            f = m[i][k] / m[k][k]
            m[i][k] = f
            for j in range(k + 1, n):
                m[i][j] -= f * m[k][j]
    return np.array(piv), np.array([[float(v) for v in row] for row in m])


@pytest.mark.parametrize("seed", range(6))
def test_getrf_exact_rational_small_integer(seed):
    rng = np.random.default_rng(seed)
    n = 9
    a = rng.integers(-9, 10, size=(n, n)).astype(np.float64)
    piv_e, lu_e = exact_gepp(a)
    lu, piv, _, _ = oracle.getrf(a, block=4, base=2)
    np.testing.assert_array_equal(piv, piv_e)
    np.testing.assert_allclose(lu, lu_e, atol=1e-12)


def test_getrf_ties_take_the_first_row():
    """Planted mistake: a last-maximum pivot rule picks row 1 here; the oracle
    (numax's strict '>' scan, LAPACK idamax) must pick row 0."""
    a = np.array([[1.0, 2.0], [-1.0, 5.0]])
    _, piv, _, _ = oracle.getrf(a)
    assert piv[0] == 0

    def last_max(col):
        return max(range(len(col)), key=lambda i: (abs(col[i]), i))
    assert last_max(a[:, 0]) == 1  # the planted rule would fail the assertion above


@pytest.mark.parametrize("block,base", [(1, 1), (2, 1), (4, 3), (5, 5), (32, 16), (70, 70)])
def test_getrf_block_invariance(block, base):
    a = kat.dynamic_lu_entry(70)
    ref, rpiv, _, _ = oracle.getrf(a, block=70, base=70)
    lu, piv, _, _ = oracle.getrf(a, block=block, base=base)
    np.testing.assert_array_equal(piv, rpiv)
    np.testing.assert_allclose(lu, ref, atol=1e-12)


def test_getrf_invariant_pa_eq_lu():
    a = R.normal(5, (200, 200))
    lu, piv, _, _ = oracle.getrf(a)
    l = np.tril(lu, -1) + np.eye(200)
    u = np.triu(lu)
    np.testing.assert_allclose(a[perm_from_ipiv(piv)], l @ u, atol=1e-12)


def test_known_answers_lu():
    lu, piv, _, sign = oracle.getrf(kat.EXCHANGE)
    assert sign * np.prod(np.diag(lu)) == kat.EXCHANGE_DET
    np.testing.assert_array_equal(oracle.getrs(lu, piv, kat.EXCHANGE_B), kat.EXCHANGE_X)
    lu, piv, _, _ = oracle.getrf(kat.TEXTBOOK_A)
    np.testing.assert_allclose(oracle.getrs(lu, piv, kat.TEXTBOOK_B), kat.TEXTBOOK_X, atol=1e-14)
    _, piv, _, _ = oracle.getrf(kat.PIVOT3)
    assert piv[0] == kat.PIVOT3_FIRST


def test_singular_is_info_not_error():
    lu, piv, info, _ = oracle.getrf(np.array([[1.0, 2.0], [2.0, 4.0]]))
    assert info == 2 and lu[1, 1] == 0.0
    _, _, info, _ = oracle.getrf(np.zeros((5, 5)))
    assert info == 1


def test_getrs_matrix_and_vector_forms_agree():
    a = R.normal(12, (90, 90))
    b = R.normal(13, (90, 5))
    lu, piv, _, _ = oracle.getrf(a)
    xm = oracle.getrs(lu, piv, b)
    for c in range(5):
        np.testing.assert_allclose(oracle.getrs(lu, piv, b[:, c]), xm[:, c], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a @ xm, b, atol=1e-10)


@pytest.mark.parametrize("upper,unit,trans", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1)])
def test_trsv_trsm_against_numpy(upper, unit, trans):
    rng = np.random.default_rng(3)
    n = 45
    t = rng.standard_normal((n, n)) * 0.2 + 3 * np.eye(n)
    b = rng.standard_normal((n, 4))
    stored = np.triu(t) if upper else np.tril(t)
    if unit:
        stored = stored - np.diag(np.diag(stored)) + np.eye(n)
    op = stored.T if trans else stored
    walk_upper = bool(upper) != bool(trans)
    x = oracle.trsm(t, b, upper=walk_upper, unit=bool(unit), trans=bool(trans), block=8)
    np.testing.assert_allclose(op @ x, b, atol=1e-12)
    xv = oracle.trsv(t, b[:, 0], upper=walk_upper, unit=bool(unit), trans=bool(trans), block=7)
    np.testing.assert_allclose(xv, x[:, 0], atol=1e-12)


# ---------------------------------------------------------------- Cholesky
@pytest.mark.parametrize("n", [5, 40, 150])
def test_potrf_matches_lapack(n):
    a = R.spd_matrix(2000 + n, n)
    l, info = oracle.potrf(a, block=16, tile=7)
    assert info == 0
    np.testing.assert_allclose(l, G[f"chol{n}"], atol=1e-13 * n)


def test_potrf_known_answers():
    l, info = oracle.potrf(kat.SPD5, block=2, tile=2)
    np.testing.assert_allclose(l, G["spd5_chol"], atol=1e-14)
    _, info = oracle.potrf(kat.NOT_PD3)
    assert info == kat.NOT_PD3_INFO
    _, info = oracle.potrf(kat.LATE_NOT_PD, block=2)
    assert info - 1 == kat.LATE_NOT_PD_INDEX


@pytest.mark.parametrize("block,tile", [(1, 2), (2, 3), (3, 5), (5, 7), (64, 128)])
def test_potrf_block_tile_invariance(block, tile):
    a = kat.spd_n(40)
    ref, _ = oracle.potrf(a, block=40, tile=40)
    l, _ = oracle.potrf(a, block=block, tile=tile)
    np.testing.assert_allclose(l, ref, atol=1e-13)
    np.testing.assert_allclose(l @ l.T, a, atol=1e-12)


def test_potrs():
    a = kat.spd_n(33)
    b = np.arange(33.0)
    l, _ = oracle.potrf(a)
    np.testing.assert_allclose(a @ oracle.potrs(l, b), b, atol=1e-12)


# ---------------------------------------------------------------- QR
def _fix_signs(r):
    return r * np.sign(np.diag(r))[:, None]


def test_geqrf_matches_lapack_r():
    a = R.normal(50, (50, 20))
    qr, tau = oracle.geqrf(a, block=16)
    np.testing.assert_allclose(np.triu(qr[:20]), G["qr50x20_r"], atol=1e-12)  # dlarfg signs
    x = oracle.gels(a, R.normal(51, (50,)))
    np.testing.assert_allclose(x, G["qr50x20_x"], atol=1e-12)


def test_geqrf_tall_6x3_and_block_invariance():
    qr1, _ = oracle.geqrf(kat.TALL_6x3, block=1)
    qr2, _ = oracle.geqrf(kat.TALL_6x3, block=2)
    np.testing.assert_allclose(np.triu(qr2[:3]), G["tall6x3_r"], atol=1e-13)
    np.testing.assert_allclose(qr1, qr2, atol=1e-13)
    np.testing.assert_allclose(oracle.gels(kat.TALL_6x3, np.arange(1.0, 7.0)), G["tall6x3_x"], atol=1e-12)


def test_ormqr_orthogonality():
    a = R.normal(70, (40, 12))
    qr, tau = oracle.geqrf(a, block=5)
    qt = oracle.ormqr_t(qr, tau, np.eye(40), block=5)  # Q^T
    np.testing.assert_allclose(qt @ qt.T, np.eye(40), atol=1e-13)
    np.testing.assert_allclose((qt @ a)[:12], np.triu(qr[:12]), atol=1e-12)
    np.testing.assert_allclose((qt @ a)[12:], 0, atol=1e-12)


# ---------------------------------------------------------------- register tier
def test_array_lu_matches_exact_pivoting():
    rng = np.random.default_rng(11)
    a = rng.integers(-9, 10, size=(8, 8)).astype(np.float64)
    lu, perm, sign = oracle.array_lu_factor(a)
    piv_e, lu_e = exact_gepp(a)
    np.testing.assert_array_equal(perm, perm_from_ipiv(piv_e))
    np.testing.assert_allclose(lu, lu_e, atol=1e-12)
    b = np.arange(8.0)
    np.testing.assert_allclose(a @ oracle.array_lu_solve(lu, perm, b), b, atol=1e-11)


def test_array_cholesky_and_solve_match_map_blocks_batch():
    """tests/core/test_map_blocks.mojo:192-218: per-lane solve vs host solve
    at 1e-12 over the hashed SPD batch."""
    A, B = kat.map_blocks_systems(batch=60)
    for p in range(60):
        l = oracle.array_cholesky(A[p])
        x = oracle.array_cholesky_solve(l, B[p])
        np.testing.assert_allclose(x, np.linalg.solve(A[p], B[p]), atol=1e-12)
        np.testing.assert_allclose(oracle.array_solve(A[p], B[p]), x, atol=1e-12)


def test_array_cholesky_negative_pivot_is_nan():
    l = oracle.array_cholesky(np.array([[1.0, 2.0], [2.0, 1.0]]))
    assert np.isnan(l[1, 1])


def hamiltonian(n, omega):
    h = 8.0 / (n - 1)
    x = -4.0 + h * np.arange(n)
    a = np.diag(1.0 / h**2 + 0.5 * omega**2 * x**2)
    return a + np.diag(np.full(n - 1, -0.5 / h**2), 1) + np.diag(np.full(n - 1, -0.5 / h**2), -1)


def test_array_eigh_matches_lapack_and_invariants():
    a = hamiltonian(24, 1.0)
    w, v = oracle.array_eigh(a)
    np.testing.assert_allclose(np.sort(w), G["ham24_w"], atol=1e-10)
    assert abs(np.sort(w)[0] - 0.496) < 0.01  # quantum_well.mojo:29-31
    np.testing.assert_allclose(v.T @ v, np.eye(24), atol=1e-12)
    np.testing.assert_allclose(a @ v, v * w, atol=1e-9)
    g = R.normal(60, (32, 32))
    s = (g + g.T) / 2
    w, _ = oracle.array_eigh(s)
    np.testing.assert_allclose(np.sort(w), G["sym32_w"], atol=1e-10)
    assert abs(w.sum() - np.trace(s)) < 1e-11
    assert abs((w ** 2).sum() - (s ** 2).sum()) < 1e-9


def test_array_eigh_closed_form_free_particle():
    """omega = 0: tridiagonal(-1/(2h^2), 1/h^2, -1/(2h^2)) has
    lambda_k = h^-2 (1 - cos(k pi / (n + 1)))."""
    n = 24
    h = 8.0 / (n - 1)
    w, _ = oracle.array_eigh(hamiltonian(n, 0.0))
    k = np.arange(1, n + 1)
    np.testing.assert_allclose(np.sort(w), (1 - np.cos(k * np.pi / (n + 1))) / h**2, rtol=1e-12, atol=1e-10)


def test_array_lstsq_matches_lapack():
    np.testing.assert_allclose(oracle.array_lstsq(kat.TALL_6x3, np.arange(1.0, 7.0)), G["tall6x3_x"], atol=1e-12)
    a = R.normal(61, (1024, 32)).astype(np.float32).astype(np.float64)
    b = R.normal(62, (1024,)).astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(oracle.array_lstsq(a, b), G["ts1024x32_x"], atol=1e-12)
    np.testing.assert_allclose(oracle.gels(a, b), G["ts1024x32_x"], atol=1e-12)


def test_backward_error_closed_form():
    a = np.diag([2.0, 4.0])
    x = np.array([1.0, 1.0])
    b = np.array([2.0, 4.5])  # residual 0.5 on row 2
    assert abs(oracle.backward_error(a, x, b) - 0.5 / (4.0 * 1.0)) < 1e-16


def test_batch_entry_points_equal_the_single_problem_functions():
    # the CPU baselines of the batched configurations run these: one problem
    # per thread, each exactly the single-problem function
    from numax_amd import random as R
    rng = np.random.default_rng(4)
    sizes = [int(s) for s in rng.integers(1, 60, size=30)]
    mats, rhss = [], []
    kinds = (np.arange(30) % 2).astype(np.uint8)
    for i, n in enumerate(sizes):
        if kinds[i] == 0:
            mats.append(rng.uniform(-1, 1, (n, n)) + 0.1 * n * np.eye(n))
        else:
            g = rng.standard_normal((n, n))
            mats.append(g @ g.T / n + np.eye(n))
        rhss.append(rng.standard_normal(n))
    n = np.array(sizes, dtype=np.int32)
    oa = np.zeros(30, dtype=np.int64)
    ob = np.zeros(30, dtype=np.int64)
    oa[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    ob[1:] = np.cumsum(n)[:-1]
    oracle.set_threads(4)
    x = oracle.array_gesv_batch(np.concatenate([m.ravel() for m in mats]), np.concatenate(rhss), n, kinds, oa, ob)
    for p in range(30):
        if kinds[p] == 0:
            lu, perm, _ = oracle.array_lu_factor(mats[p])
            xo = oracle.array_lu_solve(lu, perm, rhss[p])
        else:
            xo = oracle.array_cholesky_solve(oracle.array_cholesky(mats[p]), rhss[p])
        assert np.array_equal(x[ob[p]:ob[p] + sizes[p]], xo)
    a2, n2, oa2, ow2 = R.eigh_batch(16, 3, sizes=(5, 9))
    w, v = oracle.array_eigh_batch(a2.astype(np.float64), n2, oa2, ow2)
    for p in range(16):
        k = int(n2[p])
        wo, vo = oracle.array_eigh(a2[oa2[p]:oa2[p] + k * k].reshape(k, k).astype(np.float64))
        assert np.array_equal(w[ow2[p]:ow2[p] + k], wo)
        assert np.array_equal(v[oa2[p]:oa2[p] + k * k].reshape(k, k), vo)
    A, B = R.lstsq_batch(6, 40, 5, 2)
    X = oracle.array_lstsq_batch(A, B)
    for p in range(6):
        assert np.array_equal(X[p], oracle.array_lstsq(A[p].astype(np.float64), B[p].astype(np.float64)))
    oracle.set_threads(1)


def test_register_blocked_gemm_matches_numpy_and_is_thread_invariant():
    # c - a b (and a b^T) against numpy, with ragged sizes around the 4 x 8
    # register blocks; the result may not depend on the thread count
    rng = np.random.default_rng(5)
    for (m, n, k, tb) in [(1, 1, 1, 0), (4, 8, 3, 0), (5, 9, 17, 1), (37, 70, 33, 0), (64, 64, 64, 1)]:
        a = rng.standard_normal((m, k))
        b = rng.standard_normal((n, k) if tb else (k, n))
        c = rng.standard_normal((m, n))
        ref = c - a @ (b.T if tb else b)
        outs = []
        for t in (1, 3):
            oracle.set_threads(t)
            cc = c.copy()
            oracle._lib.orc_gemm_sub(m, n, k, oracle._p(a), k, oracle._p(b), b.shape[1], tb, oracle._p(cc), n)
            outs.append(cc)
        np.testing.assert_allclose(outs[0], ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()) * k)
        assert np.array_equal(outs[0], outs[1])
    oracle.set_threads(1)


# ------------------------------------------- register tier: lu / inverse / det
def _dyadic_lu(seed, n):
    """A = L0 U0 with small integer entries and power-of-two pivots: every
    step of the unpivoted Doolittle is exact, so it must return L0 \\ U0."""
    rng = np.random.default_rng(seed)
    l0 = np.tril(rng.integers(-3, 4, size=(n, n)).astype(np.float64), -1) + np.eye(n)
    u0 = np.triu(rng.integers(-3, 4, size=(n, n)).astype(np.float64), 1)
    u0[np.arange(n), np.arange(n)] = 2.0 ** rng.integers(-2, 3, size=n)
    return l0, u0


@pytest.mark.parametrize("seed,n", [(1, 3), (2, 6), (3, 9)])
def test_array_lu_nopivot_recovers_exact_factors(seed, n):
    l0, u0 = _dyadic_lu(seed, n)
    f = oracle.array_lu(l0 @ u0)
    np.testing.assert_array_equal(np.tril(f, -1) + np.eye(n), l0)
    np.testing.assert_array_equal(np.triu(f), u0)


def test_array_lu_nopivot_zero_pivot_uses_the_floor():
    # the exchange matrix: the zero pivot is divided as 1e-30 and kept as 0
    f = oracle.array_lu(kat.EXCHANGE)
    assert f[0, 0] == 0.0 and f[1, 0] == 1.0 / 1e-30 and f[1, 1] == -(1.0 / 1e-30)
    # so the unpivoted det is 0 where the pivoted one is -1
    assert oracle.array_det(kat.EXCHANGE) == 0.0
    assert oracle.array_det(kat.EXCHANGE, pivoted=True) == kat.EXCHANGE_DET


def test_array_det_known_answers():
    assert abs(oracle.array_det(kat.SPD3) - kat.SPD3_DET) <= 1e-12
    assert abs(oracle.array_det(kat.SLOGDET_M3, pivoted=True) - kat.SLOGDET_M3_DET) <= 1e-12
    rng = np.random.default_rng(4)
    for n in (1, 5, 17, 40):
        a = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        want = np.linalg.det(a)
        assert abs(oracle.array_det(a) - want) <= 1e-12 * abs(want)
        assert abs(oracle.array_det(a, pivoted=True) - want) <= 1e-12 * abs(want)


def test_array_slogdet_known_answers():
    s, la = oracle.array_slogdet(kat.SLOGDET_M3)
    assert s == -1.0 and abs(la - np.log(18.0)) <= 1e-12
    s, la = oracle.array_slogdet(kat.SLOGDET_SINGULAR)
    assert s == 0.0 and la == -np.inf
    # test_slogdet.mojo:92-118's diagonal of 40s, at the register tier's 64
    s, la = oracle.array_slogdet(40.0 * np.eye(64))
    assert s == 1.0 and abs(la - 64 * np.log(40.0)) <= 1e-12
    rng = np.random.default_rng(5)
    for n in (2, 9, 33, 64):
        a = rng.standard_normal((n, n))
        ws, wl = np.linalg.slogdet(a)
        s, la = oracle.array_slogdet(a)
        assert s == ws and abs(la - wl) <= 1e-12 * max(1.0, abs(wl))


def test_array_slogdet_cholesky():
    l = oracle.array_cholesky(kat.SPD3)
    assert abs(oracle.array_slogdet_cholesky(l) - np.log(kat.SPD3_DET)) <= 1e-14
    a = kat.spd_n(30)
    assert abs(oracle.array_slogdet_cholesky(oracle.array_cholesky(a)) - np.linalg.slogdet(a)[1]) <= 1e-12


def test_array_inverse():
    # test_linalg.mojo's general3 (leading minors nonzero)
    inv = oracle.array_inverse(kat.GENERAL3)
    np.testing.assert_allclose(kat.GENERAL3 @ inv, np.eye(3), atol=1e-13)
    # a diagonal matrix inverts to the rounded reciprocals exactly
    d = np.array([3.0, -7.0, 0.1, 11.0])
    np.testing.assert_array_equal(oracle.array_inverse(np.diag(d)), np.diag(1.0 / d))
    rng = np.random.default_rng(6)
    for n in (1, 8, 31, 64):
        a = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        np.testing.assert_allclose(oracle.array_inverse(a), np.linalg.inv(a), rtol=0, atol=1e-13)
        # inverse = solve against each column of the identity
        np.testing.assert_array_equal(oracle.array_inverse(a)[:, n // 2],
                                      oracle.array_solve(a, np.eye(n)[:, n // 2]))


# ----------------------------------------------------------------- ldl (sytf2)
def test_ldl_known_answers():
    """tests/linalg/test_ldl.mojo:49-111 (SciPy's factors)."""
    lu, d, perm = oracle.ldl(kat.LDL_A)
    np.testing.assert_allclose(lu, kat.LDL_LU, atol=1e-14)
    np.testing.assert_allclose(d, kat.LDL_D, atol=1e-14)
    np.testing.assert_array_equal(perm, kat.LDL_PERM)
    lu, d, perm = oracle.ldl(kat.LDL_Z)
    np.testing.assert_allclose(lu, kat.LDL_Z_LU, atol=1e-14)
    np.testing.assert_allclose(d, kat.LDL_Z_D, atol=1e-14)
    np.testing.assert_array_equal(perm, kat.LDL_Z_PERM)


@pytest.mark.parametrize("n,seed", [(5, 1), (17, 2), (40, 3), (97, 4)])
def test_ldl_matches_scipy(n, seed):
    """SciPy's ldl calls LAPACK's blocked sytrf, whose pivot choices are
    sytf2's: same permutation and block structure, factors to rounding."""
    from scipy import linalg
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n))
    a = (g + g.T) / 2  # indefinite
    lu, d, perm = oracle.ldl(a)
    lu_s, d_s, perm_s = linalg.ldl(a, lower=True)
    np.testing.assert_array_equal(perm, perm_s)
    np.testing.assert_array_equal(d != 0, d_s != 0)  # the same 1 x 1 / 2 x 2 blocks
    np.testing.assert_allclose(lu, lu_s, rtol=0, atol=1e-10)
    np.testing.assert_allclose(d, d_s, rtol=0, atol=1e-10 * np.abs(d_s).max())
    np.testing.assert_allclose(lu @ d @ lu.T, a, rtol=0, atol=1e-11 * n)
    # lu[perm] is unit lower triangular
    t = lu[perm]
    np.testing.assert_array_equal(np.diag(t), 1.0)
    assert np.all(np.triu(t, 1) == 0.0)


def test_sytf2_singular_stays_inside_ipiv():
    """A zero first column: info = 1, the NaNs that follow the zero pivot
    never open a 2 x 2 block at the last column, and nothing is stored past
    ipiv's n entries (guard words after them keep their value)."""
    import ctypes
    for n in (2, 3, 4, 7):
        a = np.zeros((n, n))
        if n > 1:
            g = np.arange(1.0, n) + 1.0
            a[1:, 1:] = np.diag(g) + 0.5
        f = np.ascontiguousarray(a)
        ipiv = np.full(n + 4, -777, dtype=np.int32)
        info = oracle._lib.orc_sytf2_lower(n, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        assert info == 1
        np.testing.assert_array_equal(ipiv[n:], -777)
        assert ipiv[n - 1] > 0 or (n >= 2 and ipiv[n - 2] == ipiv[n - 1] < 0)
        assert np.all(np.abs(ipiv[:n]) >= 1) and np.all(np.abs(ipiv[:n]) <= n)


def test_ldl_definite_takes_only_scalar_pivots():
    """tests/linalg/test_ldl.mojo:126-140."""
    s = np.array([[4.0, 1.0, 0.5], [1.0, 3.0, 0.2], [0.5, 0.2, 2.0]])
    lu, d, perm = oracle.ldl(s)
    assert np.all(d == np.diag(np.diag(d)))
    np.testing.assert_allclose(lu @ d @ lu.T, s, atol=1e-14)
