"""GPU parity of the device Householder QR (nmx_geqrf/ormqr/orgqr/qr_r/gels,
FP64) with the oracle's restatement of numax's qr_factor / apply_q_transpose
/ lstsq (numax/linalg/qr.mojo:123-821, panel.mojo:1975-2169), through the C ABI.

Bar: the reflectors, taus and R agree with the oracle to 1e-12 relative to
the matrix scale (the blocked products sum in another order; the dlarfg sign
convention makes the factorization unique), Q is orthonormal to 1e-13, and
least-squares solutions agree with the oracle and LAPACK to 1e-10 relative.
"""
import numpy as np
import pytest
import torch

import oracle
from numax_amd import linalg as nl
from numax_amd import random as R
from tests.golden import kat

pytestmark = pytest.mark.gpu

G = np.load("tests/golden/goldens.npz")


def _scale(a):
    return max(1.0, float(np.abs(a).max()))


@pytest.mark.parametrize("m,n,block", [(50, 20, 16), (6, 3, 2), (6, 3, 1), (40, 12, 5), (128, 128, 16),
                                       (300, 170, 64), (257, 33, 16), (1000, 64, 32), (65, 1, 16)])
def test_geqrf_matches_oracle(m, n, block):
    a = R.normal(m * 7 + n, (m, n))
    f = nl.qr_factor(a, block=block)
    qr_o, tau_o = oracle.geqrf(a, block=block)
    qr = f.factored.cpu().numpy()
    tau = f.taus.cpu().numpy()
    np.testing.assert_allclose(tau, tau_o, atol=1e-12)
    np.testing.assert_allclose(qr, qr_o, atol=1e-12 * _scale(qr_o))
    np.testing.assert_allclose(f.r().cpu().numpy(), np.triu(qr_o[:n]), atol=1e-12 * _scale(qr_o))


def test_geqrf_lapack_goldens():
    # dlarfg signs, the same as numax's geqr2_panel (test_tensor_linalg.mojo _tall_6x3)
    a = R.normal(50, (50, 20))
    np.testing.assert_allclose(nl.qr_factor(a, block=16).r().cpu().numpy(), G["qr50x20_r"], atol=1e-12)
    np.testing.assert_allclose(nl.lstsq(a, R.normal(51, (50,))).cpu().numpy(), G["qr50x20_x"], atol=1e-12)
    np.testing.assert_allclose(nl.qr_factor(kat.TALL_6x3, block=2).r().cpu().numpy(), G["tall6x3_r"], atol=1e-13)
    np.testing.assert_allclose(nl.lstsq(kat.TALL_6x3, np.arange(1.0, 7.0)).cpu().numpy(), G["tall6x3_x"],
                               atol=1e-12)


@pytest.mark.parametrize("m,n,block", [(40, 12, 5), (200, 96, 16), (513, 130, 64)])
def test_q_orthonormal_and_reconstructs(m, n, block):
    a = R.normal(m + 3 * n, (m, n))
    f = nl.qr_factor(a, block=block)
    q = f.q().cpu().numpy()
    r = f.r().cpu().numpy()
    np.testing.assert_allclose(q.T @ q, np.eye(n), atol=1e-13 * max(1, m / 64))
    np.testing.assert_allclose(q @ r, a, atol=1e-12 * _scale(a))


@pytest.mark.parametrize("nrhs", [1, 3, 40])
def test_apply_q_transpose_matches_oracle(nrhs):
    m, n, block = 150, 70, 16
    a = R.normal(91, (m, n))
    b = R.normal(92, (m, nrhs)) if nrhs > 1 else R.normal(92, (m,))
    f = nl.qr_factor(a, block=block)
    qr_o, tau_o = oracle.geqrf(a, block=block)
    y = f.apply_q_transpose(b).cpu().numpy().reshape(m, -1)
    y_o = oracle.ormqr_t(qr_o, tau_o, b, block=block)
    np.testing.assert_allclose(y, y_o, atol=1e-12 * _scale(y_o))


@pytest.mark.parametrize("m,n", [(50, 20), (400, 200), (1024, 32), (77, 77)])
def test_lstsq_matches_oracle_and_lapack(m, n):
    a = R.normal(m + n, (m, n))
    b = R.normal(m + n + 1, (m,))
    x = nl.lstsq(a, b).cpu().numpy()
    x_o = oracle.gels(a, b)
    x_np = np.linalg.lstsq(a, b, rcond=None)[0]
    np.testing.assert_allclose(x, x_o, atol=1e-10 * _scale(x_o))
    np.testing.assert_allclose(x, x_np, atol=1e-10 * _scale(x_np))


def test_solve_reuses_factorization_multi_rhs():
    m, n = 300, 120
    a = R.normal(5, (m, n))
    f = nl.qr_factor(a)
    for k, nrhs in enumerate([1, 5]):
        b = R.normal(6 + k, (m, nrhs)) if nrhs > 1 else R.normal(6 + k, (m,))
        x = f.solve(b).cpu().numpy().reshape(n, -1)
        x_np = np.linalg.lstsq(a, np.asarray(b).reshape(m, -1), rcond=None)[0]
        np.testing.assert_allclose(x, x_np, atol=1e-10 * _scale(x_np))


def test_zero_column_gives_identity_reflector():
    a = R.normal(8, (20, 6))
    a[:, 2] = 0.0
    f = nl.qr_factor(a, block=4)
    qr_o, tau_o = oracle.geqrf(a, block=4)
    np.testing.assert_allclose(f.taus.cpu().numpy(), tau_o, atol=1e-13)
    assert f.taus.cpu().numpy()[2] == 0.0 and tau_o[2] == 0.0
    np.testing.assert_allclose(f.factored.cpu().numpy(), qr_o, atol=1e-12)


def test_shape_errors():
    with pytest.raises(nl.NumaxError):
        nl.qr_factor(np.ones((3, 5)))
    with pytest.raises(nl.NumaxError):
        nl.lstsq(np.ones((6, 3)), np.ones(5))
    with pytest.raises(nl.NumaxError):
        nl.qr_factor(torch.ones((6, 3), dtype=torch.float16))


# ---- FP32: numax's device QR runs at float32 (qr.mojo:615-743, lstsq :746-821).
# Checked against the FP64 oracle at FP32 tolerances, tolerances stated here:
# R and the reflectors within 2e-5 * max|A| (scaled by sqrt(m) for the
# panel's reductions), Q^T Q = I and Q R = A within 1e-5 relative, the
# least-squares x within 1e-4 relative of the oracle's on well-conditioned
# N(0,1) problems, and its residual orthogonal to range(A).
@pytest.mark.parametrize("m,n,block", [(64, 64, 16), (300, 120, 16), (1000, 257, 32), (513, 64, 64), (37, 5, 4)])
def test_qr_factor_f32_matches_oracle(m, n, block):
    a = R.normal(m + n, (m, n))
    f = nl.qr_factor(a.astype(np.float32), block=block)
    assert f.factored.dtype == torch.float32
    qr_o, tau_o = oracle.geqrf(a.astype(np.float32).astype(np.float64), block=block)
    tol = 2e-5 * np.sqrt(m) * np.abs(a).max()
    np.testing.assert_allclose(f.factored.cpu().numpy(), qr_o, atol=tol)
    np.testing.assert_allclose(f.taus.cpu().numpy(), tau_o, atol=1e-5)
    q = f.q().cpu().numpy().astype(np.float64)
    r = f.r().cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(q.T @ q, np.eye(n), atol=1e-5 * np.sqrt(m))
    np.testing.assert_allclose(q @ r, a, atol=1e-5 * np.sqrt(m) * np.abs(a).max())


@pytest.mark.parametrize("m,n,nrhs", [(400, 100, 1), (2000, 64, 3), (128, 128, 2)])
def test_lstsq_f32_matches_oracle(m, n, nrhs):
    a = R.normal(3 * m + n, (m, n)).astype(np.float32)
    b = R.normal(5 * m + n, (m, nrhs) if nrhs > 1 else (m,)).astype(np.float32)
    x = nl.lstsq(a, b).cpu().numpy().astype(np.float64)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    cols = [b64] if nrhs == 1 else [b64[:, j] for j in range(nrhs)]
    xs = [x] if nrhs == 1 else [x[:, j] for j in range(nrhs)]
    for bj, xj in zip(cols, xs):
        xo = oracle.gels(a64, bj)
        assert np.max(np.abs(xj - xo)) <= 1e-4 * max(1.0, np.abs(xo).max())
        rr = bj - a64 @ xj
        assert np.max(np.abs(a64.T @ rr)) <= 1e-3 * np.linalg.norm(a64) * np.linalg.norm(bj)
    # QR.solve and apply_q_transpose at float32 agree with lstsq
    f = nl.qr_factor(a)
    np.testing.assert_allclose(f.solve(b).cpu().numpy(), x, rtol=0, atol=1e-5 * max(1.0, np.abs(x).max()))
    qtb = f.apply_q_transpose(b).cpu().numpy().astype(np.float64)
    q = f.q().cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(qtb[:n], q.T @ b64, atol=1e-4 * np.abs(b64).max() * np.sqrt(m))


def test_laswp_f32_is_exact():
    n, k = 777, 5
    rng = np.random.default_rng(3)
    ipiv = np.array([rng.integers(j, n) for j in range(n)], dtype=np.int32)
    b = rng.standard_normal((n, k)).astype(np.float32)
    got = nl.laswp(b, ipiv).cpu().numpy()
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got, oracle.laswp(b.astype(np.float64), ipiv).astype(np.float32))
