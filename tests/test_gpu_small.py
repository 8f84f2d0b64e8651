"""GPU parity of the register-tier small-matrix operations (small.hip) and of
the Tensor-tier inverse / lu -> PLU (drivers.hip) against the oracle.

The small-matrix kernels keep numax's loop orders and round every multiply
and subtraction separately (no fused multiply-add), as the oracle does, so
lu / solve / inverse / det / the map_blocks solves must equal it bit for bit
in float64; slogdet goes through log, whose last bit may differ between the
device library and glibc (1e-14 relative).  float32 map_blocks results are
compared with the float64 oracle at 2e-5 relative (well-conditioned inputs).
The Tensor-tier inverse solves through blocked MFMA substitutions, so it is
compared with NumPy / the oracle to rounding (1e-12 relative at the sizes
used), and lu() against SciPy's factors from the reference's known answer.
"""
import numpy as np
import pytest

import oracle
from numax_amd import linalg as nl
from tests.golden import kat

pytestmark = pytest.mark.gpu


def _dominant(rng, n):
    return rng.uniform(-1.0, 1.0, (n, n)) + 0.5 * n * np.eye(n)


def _spd(rng, n):
    g = rng.standard_normal((n, n))
    return g @ g.T / n + np.eye(n)


def _uneven(seed, count=160):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 65, size=count)
    sizes[:4] = [1, 2, 63, 64]
    mats = [_dominant(rng, int(n)) for n in sizes]
    rhss = [rng.standard_normal(int(n)) for n in sizes]
    return nl.pack_systems(mats, rhss), mats, rhss


def test_batched_lu_solve_inverse_det_bit_exact():
    (a, b, n, oa, ob), mats, rhss = _uneven(21)
    lu = nl.batched_lu_nopiv(a, n, oa).cpu().numpy()
    x = nl.batched_solve_nopiv(a, b, n, oa, ob).cpu().numpy()
    inv = nl.batched_inverse(a, n, oa).cpu().numpy()
    det = nl.batched_det(a, n, oa).cpu().numpy()
    detp = nl.batched_det(a, n, oa, pivoted=True).cpu().numpy()
    sg, la = (t.cpu().numpy() for t in nl.batched_slogdet(a, n, oa))
    for p, m in enumerate(mats):
        k = int(n[p])
        sa, sb = slice(int(oa[p]), int(oa[p]) + k * k), slice(int(ob[p]), int(ob[p]) + k)
        np.testing.assert_array_equal(lu[sa].reshape(k, k), oracle.array_lu(m), err_msg=f"lu p={p} n={k}")
        np.testing.assert_array_equal(x[sb], oracle.array_solve(m, rhss[p]), err_msg=f"solve p={p} n={k}")
        np.testing.assert_array_equal(inv[sa].reshape(k, k), oracle.array_inverse(m), err_msg=f"inverse p={p}")
        assert det[p] == oracle.array_det(m), (p, k)
        assert detp[p] == oracle.array_det(m, pivoted=True), (p, k)
        ws, wl = oracle.array_slogdet(m)
        assert sg[p] == ws and abs(la[p] - wl) <= 1e-14 * max(1.0, abs(wl)), (p, k)


def test_batched_slogdet_cholesky():
    rng = np.random.default_rng(22)
    sizes = [1, 3, 17, 40, 64]
    factors = [oracle.array_cholesky(_spd(rng, k)) for k in sizes]
    l, _, n, ol, _ = nl.pack_systems(factors, [np.zeros(k) for k in sizes])
    got = nl.batched_slogdet_cholesky(l, n, ol).cpu().numpy()
    for p, f in enumerate(factors):
        want = oracle.array_slogdet_cholesky(f)
        assert abs(got[p] - want) <= 1e-14 * max(1.0, abs(want)), p
    # spd3 (tests/linalg/test_linalg.mojo:170-173): ln 67
    l3, _, n3, o3, _ = nl.pack_systems([oracle.array_cholesky(kat.SPD3)], [np.zeros(3)])
    assert abs(float(nl.batched_slogdet_cholesky(l3, n3, o3)[0]) - np.log(kat.SPD3_DET)) <= 1e-14


def test_batched_known_answers():
    mats = [kat.EXCHANGE, kat.SLOGDET_M3, kat.SLOGDET_SINGULAR, 40.0 * np.eye(64), kat.SPD3, kat.GENERAL3]
    a, _, n, oa, _ = nl.pack_systems(mats, [np.zeros(m.shape[0]) for m in mats])
    det = nl.batched_det(a, n, oa).cpu().numpy()
    detp = nl.batched_det(a, n, oa, pivoted=True).cpu().numpy()
    sg, la = (t.cpu().numpy() for t in nl.batched_slogdet(a, n, oa))
    # the exchange matrix: unpivoted 0 (zero pivot floored), pivoted -1
    assert det[0] == 0.0 and detp[0] == kat.EXCHANGE_DET
    assert sg[1] == -1.0 and abs(la[1] - np.log(-kat.SLOGDET_M3_DET)) <= 1e-14
    assert abs(detp[1] - kat.SLOGDET_M3_DET) <= 1e-13
    assert sg[2] == 0.0 and la[2] == -np.inf
    assert sg[3] == 1.0 and abs(la[3] - 64 * np.log(40.0)) <= 1e-12
    assert abs(det[4] - kat.SPD3_DET) <= 1e-12
    inv = nl.batched_inverse(a, n, oa).cpu().numpy()
    g = inv[int(oa[5]):int(oa[5]) + 9].reshape(3, 3)
    np.testing.assert_allclose(kat.GENERAL3 @ g, np.eye(3), atol=1e-13)
    x = nl.batched_solve_nopiv(kat.TEXTBOOK_A.ravel(), kat.TEXTBOOK_B, [3], [0], [0]).cpu().numpy()
    np.testing.assert_allclose(x, kat.TEXTBOOK_X, atol=1e-14)


def test_batched_small_skips_invalid_problems():
    import torch
    rng = np.random.default_rng(23)
    m = _dominant(rng, 5)
    a = np.concatenate([m.ravel(), np.zeros(10)])
    n = np.array([5, 0, 65, 5, 5], dtype=np.int32)
    oa = np.array([0, 0, 0, 30, -1], dtype=np.int64)  # the 4th overruns a, the 5th is negative
    det = torch.full((5,), 7.0, dtype=torch.float64, device="cuda")
    A = torch.as_tensor(a, device="cuda")
    from numax_amd import _lib
    ctx = nl._ctx(A)
    N = torch.as_tensor(n, device="cuda")
    OA = torch.as_tensor(oa, device="cuda")
    _lib.check(_lib.load().nmx_batched_det_f64(ctx.handle, 5, nl._ptr(N), nl._ptr(OA), nl._ptr(A), A.numel(), 0,
                                               nl._ptr(det)), "det")
    d = det.cpu().numpy()
    assert d[0] == oracle.array_det(m)
    np.testing.assert_array_equal(d[1:], 7.0)


def _soa(mats, rhss):
    n = mats[0].shape[0]
    xs = np.empty((n * n + n, len(mats)))
    for p, (m, r) in enumerate(zip(mats, rhss)):
        xs[:n * n, p] = m.ravel()
        xs[n * n:, p] = r
    return xs


def _oracle_step(kind, m, r):
    if kind == "lu":
        f, perm, _ = oracle.array_lu_factor(m)
        return oracle.array_lu_solve(f, perm, r)
    if kind == "cholesky":
        return oracle.array_cholesky_solve(oracle.array_cholesky(m), r)
    return oracle.array_solve(m, r)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 9, 16, 33, 64])
@pytest.mark.parametrize("kind", ["lu", "cholesky", "nopivot"])
def test_map_blocks_solve_f64_bit_exact(n, kind):
    rng = np.random.default_rng(100 + n)
    batch = 300 if n <= 8 else 70
    mats = [(_spd(rng, n) if kind == "cholesky" else rng.uniform(-1, 1, (n, n)) + (0.5 * n if kind == "nopivot" else 0)
             * np.eye(n)) for _ in range(batch)]
    rhss = [rng.standard_normal(n) for _ in range(batch)]
    ys = nl.map_blocks_solve(_soa(mats, rhss), n, kind).cpu().numpy()
    for p in range(batch):
        np.testing.assert_array_equal(ys[:, p], _oracle_step(kind, mats[p], rhss[p]), err_msg=f"{kind} n={n} p={p}")


def test_map_blocks_known_answer():
    """tests/core/test_map_blocks.mojo:192-218: one Array-tier solve per lane
    over the hashed SPD batch agrees with the host solve at 1e-12."""
    a, b = kat.map_blocks_systems(1003, 4)
    xs = _soa(list(a), list(b))
    for kind in ("nopivot", "cholesky", "lu"):
        ys = nl.map_blocks_solve(xs, 4, kind).cpu().numpy()
        np.testing.assert_allclose(ys.T, np.linalg.solve(a, b[:, :, None])[:, :, 0], rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [4, 8, 12, 48])
def test_map_blocks_solve_f32(n):
    import torch
    rng = np.random.default_rng(200 + n)
    batch = 257
    mats = [_spd(rng, n) for _ in range(batch)]
    rhss = [rng.standard_normal(n) for _ in range(batch)]
    xs = _soa(mats, rhss).astype(np.float32)
    for kind in ("lu", "cholesky", "nopivot"):
        ys = nl.map_blocks_solve(torch.as_tensor(xs), n, kind).cpu().numpy()
        assert ys.dtype == np.float32
        for p in range(0, batch, 16):
            want = _oracle_step(kind, mats[p].astype(np.float32).astype(np.float64),
                                rhss[p].astype(np.float32).astype(np.float64))
            assert np.max(np.abs(ys[:, p] - want)) <= 2e-5 * max(1.0, np.abs(want).max()), (kind, n, p)


def test_map_blocks_rejects_bad_shapes():
    with pytest.raises(nl.NumaxError):
        nl.map_blocks_solve(np.zeros((19, 4)), 4)
    with pytest.raises(nl.NumaxError):
        nl.map_blocks_solve(np.zeros((20, 4)), 4, kind="qr")


# ------------------------------------------------------------- Tensor tier
def test_lu_plu_known_answer():
    """tests/linalg/test_lu_plu.mojo:44-109: SciPy's (P, L, U) of a 4x4 whose
    first pivot is zero; P L U = A; det agrees (:112-124)."""
    f = nl.lu(kat.PLU_A)
    p, l, u = (t.cpu().numpy() for t in f)
    np.testing.assert_allclose(p, kat.PLU_P, atol=1e-14)
    np.testing.assert_allclose(l, kat.PLU_L, atol=1e-14)
    np.testing.assert_allclose(u, kat.PLU_U, atol=1e-14)
    np.testing.assert_allclose(p @ l @ u, kat.PLU_A, atol=1e-14)
    # P is a 4-cycle (odd): det(A) = -prod(diag U)
    assert abs(nl.lu_factor(kat.PLU_A).det() - (-np.prod(np.diag(u)))) <= 1e-12


@pytest.mark.parametrize("n,dtype", [(1, "float64"), (37, "float64"), (300, "float64"), (130, "float32")])
def test_lu_plu_reconstructs(n, dtype):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, (n, n)).astype(dtype)
    f = nl.lu(a)
    p, l, u = (t.cpu().numpy().astype(np.float64) for t in f)
    assert np.array_equal(np.tril(l, -1) + np.eye(n), l) and np.array_equal(np.triu(u), u)
    assert np.array_equal(np.sort(p.argmax(0)), np.arange(n)) and np.array_equal(p.sum(0), np.ones(n))
    tol = (1e-12 if dtype == "float64" else 2e-5) * n
    np.testing.assert_allclose(p @ l @ u, a.astype(np.float64), atol=tol)
    if dtype == "float64":
        # the same pivots as the oracle's lu_factor: perm replays its ipiv
        _, piv, _, _ = oracle.getrf(a)
        perm = np.arange(n)
        for j, q in enumerate(piv):
            perm[j], perm[q] = perm[q], perm[j]
        np.testing.assert_array_equal(f.perm.cpu().numpy(), perm)


@pytest.mark.parametrize("n,dtype", [(1, "float64"), (5, "float64"), (300, "float64"), (777, "float64"),
                                     (100, "float32")])
def test_inverse(n, dtype):
    rng = np.random.default_rng(300 + n)
    a = (rng.uniform(-1, 1, (n, n)) + 0.3 * np.sqrt(n) * np.eye(n)).astype(dtype)
    inv = nl.inverse(a).cpu().numpy().astype(np.float64)
    a64 = a.astype(np.float64)
    want = np.linalg.inv(a64)
    tol = 1e-12 if dtype == "float64" else 1e-4
    assert np.max(np.abs(inv - want)) <= tol * np.abs(want).max() * max(1.0, np.linalg.cond(a64) / 100)
    np.testing.assert_allclose(a64 @ inv, np.eye(n), atol=50 * tol)
    if dtype == "float64" and n <= 300:
        lu, piv, _, _ = oracle.getrf(a64)
        np.testing.assert_allclose(inv, oracle.getrs(lu, piv, np.eye(n)), rtol=0, atol=1e-12 * np.abs(want).max())


def test_inverse_of_a_singular_matrix_is_not_finite():
    inv = nl.inverse(kat.SLOGDET_SINGULAR).cpu().numpy()
    assert not np.all(np.isfinite(inv)) or np.abs(inv).max() > 1e14
