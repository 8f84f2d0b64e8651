"""GPU parity of the batched Jacobi eigensolver and the batched tall-skinny
least squares against the FP64 oracle (numax's register-tier eigh,
numax/linalg/_array/eigen.mojo:316-392, and Householder lstsq,
numax/linalg/qr.mojo:746-821).

Bars (FP32 kernels, FP64 checks, tolerances stated here):
- eigh: sorted eigenvalues within 2e-5 * ||A||_2 of the oracle's and of
  LAPACK's; ||A V - V diag(w)|| <= 1e-4 ||A||, ||V^T V - I|| <= 1e-4; the
  closed form of the free-particle tridiagonal; trace and Frobenius identities.
- lstsq: x within 1e-4 relative of the FP64 oracle on well-conditioned
  Gaussian problems; the least-squares residual is orthogonal to range(A).
"""
import numpy as np
import pytest

import oracle
from numax_amd import linalg as nl

pytestmark = pytest.mark.gpu


def _pack(mats):
    n = np.array([m.shape[0] for m in mats], dtype=np.int32)
    oa = np.zeros(len(mats), dtype=np.int64)
    ow = np.zeros(len(mats), dtype=np.int64)
    oa[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    ow[1:] = np.cumsum(n.astype(np.int64))[:-1]
    a = np.concatenate([m.astype(np.float32).ravel() for m in mats])
    return a, n, oa, ow


def _hamiltonian(n, omega):
    # finite-difference well on [-4, 4] (numax's quantum_well example, lines 66-80)
    h = 8.0 / (n - 1)
    x = np.linspace(-4.0, 4.0, n)
    a = np.diag(1.0 / h ** 2 + 0.5 * omega ** 2 * x ** 2)
    off = -0.5 / h ** 2
    a += np.diag(np.full(n - 1, off), 1) + np.diag(np.full(n - 1, off), -1)
    return a


def _mats(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(sizes):
        if i % 2:
            out.append(_hamiltonian(n, rng.uniform(0.5, 2.0)))
        else:
            g = rng.standard_normal((n, n))
            out.append((g + g.T) / 2)
    return out


@pytest.mark.parametrize("sizes", [[1, 2, 3, 5, 8, 16, 24, 31, 32], [33, 40, 48, 63, 64], [24, 32, 48, 64] * 8])
def test_eigh_batched_matches_oracle(sizes):
    mats = _mats(sizes, seed=len(sizes))
    a, n, oa, ow = _pack(mats)
    w, v = nl.eigh_batched(a, n, oa, ow)
    w, v = w.cpu().numpy(), v.cpu().numpy()
    for p, m in enumerate(mats):
        k = m.shape[0]
        wp = w[ow[p]:ow[p] + k].astype(np.float64)
        vp = v[oa[p]:oa[p] + k * k].reshape(k, k).astype(np.float64)
        m32 = m.astype(np.float32).astype(np.float64)
        scale = max(np.abs(np.linalg.eigvalsh(m32)).max(), 1e-30)
        wo, _ = oracle.array_eigh(m32)
        assert np.max(np.abs(np.sort(wp) - np.sort(wo))) <= 2e-5 * scale, f"problem {p} n={k}"
        assert np.max(np.abs(np.sort(wp) - np.linalg.eigvalsh(m32))) <= 2e-5 * scale
        assert np.max(np.abs(m32 @ vp - vp * wp)) <= 1e-4 * scale
        assert np.max(np.abs(vp.T @ vp - np.eye(k))) <= 1e-4
        assert abs(wp.sum() - np.trace(m32)) <= 1e-4 * scale * k
        assert abs((wp ** 2).sum() - (m32 ** 2).sum()) <= 1e-4 * scale ** 2 * k


def test_eigh_free_particle_closed_form():
    # omega = 0: tridiagonal (-1/2h^2, 1/h^2, -1/2h^2), lambda_k = h^-2 (1 - cos(k pi / (n + 1)))
    mats = [_hamiltonian(n, 0.0) for n in (24, 32, 48, 64)]
    a, n, oa, ow = _pack(mats)
    w, _ = nl.eigh_batched(a, n, oa, ow, vectors=False)
    w = w.cpu().numpy()
    for p, m in enumerate(mats):
        k = m.shape[0]
        h = 8.0 / (k - 1)
        exact = (1.0 - np.cos(np.arange(1, k + 1) * np.pi / (k + 1))) / h ** 2
        assert np.max(np.abs(np.sort(w[ow[p]:ow[p] + k]) - exact)) <= 2e-5 * exact.max()


def test_eigh_ground_energy_quantum_well():
    # quantum_well.mojo:29-31: n = 24, omega = 1 ground energy ~ 0.496
    a, n, oa, ow = _pack([_hamiltonian(24, 1.0)])
    w, _ = nl.eigh_batched(a, n, oa, ow, vectors=False)
    assert abs(float(np.min(w.cpu().numpy()[:24])) - 0.496) < 0.01


@pytest.mark.parametrize("m,n,count", [(1024, 32, 64), (100, 7, 9), (32, 32, 3), (33, 1, 5)])
def test_lstsq_batched_matches_oracle(m, n, count):
    rng = np.random.default_rng(m + n)
    a = rng.standard_normal((count, m, n)).astype(np.float32)
    b = rng.standard_normal((count, m)).astype(np.float32)
    x = nl.lstsq_batched(a, b).cpu().numpy().astype(np.float64)
    for p in range(count):
        a64, b64 = a[p].astype(np.float64), b[p].astype(np.float64)
        xo = oracle.array_lstsq(a64, b64)
        assert np.max(np.abs(x[p] - xo)) <= 1e-4 * max(1.0, np.abs(xo).max()), f"problem {p}"
        r = b64 - a64 @ x[p]
        assert np.max(np.abs(a64.T @ r)) <= 1e-3 * np.linalg.norm(a64) * np.linalg.norm(b64)


def test_lstsq_tall_6x3_known_answer():
    from tests.golden import kat
    g = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "goldens.npz"))
    b = np.array([1.0, -2.0, 3.0, 0.5, 4.0, -1.0])
    x = nl.lstsq_batched(kat.TALL_6x3[None].astype(np.float32), b[None].astype(np.float32)).cpu().numpy()[0]
    xo = oracle.array_lstsq(kat.TALL_6x3, b)
    assert np.allclose(x, xo, atol=1e-5, rtol=1e-5)
    assert "tall6x3_x" in g.files


def test_lstsq_shape_errors():
    with pytest.raises(nl.NumaxError):
        nl.lstsq_batched(np.zeros((2, 8, 4), np.float32), np.zeros((2, 7), np.float32))
    with pytest.raises(ValueError):
        nl.lstsq_batched(np.zeros((1, 8, 40), np.float32), np.zeros((1, 8), np.float32))
