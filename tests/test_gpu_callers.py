"""The callers one step up (numax_amd/callers.py) against SciPy and closed
forms: their factorizations and solves run in the library, so these check
the boundary under the call patterns of multivariate_normal (Cholesky +
triangular solve of all points), least_squares (one QR solve per damped
step) and solve_bvp (a run-time-order LU per Newton Jacobian)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_multivariate_normal_matches_scipy():
    import torch
    from scipy import stats

    from numax_amd.callers import multivariate_normal as mvn
    rng = np.random.default_rng(3)
    d, m = 40, 300
    g = rng.standard_normal((d, d))
    cov = g @ g.T / d + 0.5 * np.eye(d)
    mean = rng.standard_normal(d)
    x = rng.standard_normal((m, d)) * 1.5 + mean
    want = stats.multivariate_normal(mean, cov).logpdf(x)
    got = mvn.logpdf(x, mean, cov).cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(mvn.pdf(x[:5], mean, cov).cpu().numpy(), np.exp(want[:5]), rtol=1e-11)
    assert abs(mvn.entropy(cov) - stats.multivariate_normal(mean, cov).entropy()) <= 1e-11
    # draws: the sample mean and covariance of 200,000 draws
    gen = torch.Generator(device="cuda").manual_seed(5)
    draws = mvn.rvs(mean, cov, 200_000, gen).cpu().numpy()
    assert np.max(np.abs(draws.mean(0) - mean)) < 0.02
    assert np.max(np.abs(np.cov(draws.T) - cov)) < 0.05


def test_multivariate_normal_rejects_a_non_pd_covariance():
    from numax_amd.callers import multivariate_normal as mvn
    from numax_amd.linalg import NumaxError
    with pytest.raises(NumaxError, match="not positive definite"):
        mvn.logpdf(np.zeros((2, 2)), np.zeros(2), np.array([[1.0, 2.0], [2.0, 1.0]]))


def test_least_squares_matches_scipy():
    import torch
    from scipy import optimize

    from numax_amd.callers import curve_fit
    rng = np.random.default_rng(4)
    t = np.linspace(0, 4, 200)
    y = 2.5 * np.exp(-1.3 * t) + 0.5 + 0.01 * rng.standard_normal(t.size)

    def model(ts, p):
        return p[0] * torch.exp(p[1] * ts) + p[2]

    def jac(ts, p):
        e = torch.exp(p[1] * ts)
        return torch.stack([e, p[0] * ts * e, torch.ones_like(ts)], dim=1)

    fit = curve_fit(model, jac, t, y, np.array([1.0, -0.5, 0.0]))
    assert fit.converged, fit
    ref = optimize.least_squares(lambda p: p[0] * np.exp(p[1] * t) + p[2] - y, [1.0, -0.5, 0.0], method="lm",
                                 xtol=1e-15, ftol=1e-15, gtol=1e-15)
    np.testing.assert_allclose(fit.x.cpu().numpy(), ref.x, rtol=1e-7, atol=1e-9)
    assert abs(fit.cost - ref.cost) <= 1e-10 * max(1.0, ref.cost)


def _bratu_exact(x):
    # y'' + exp(y) = 0, y(0) = y(1) = 0: y = -2 ln(cosh((x - 1/2) th / 2) / cosh(th / 4)),
    # th the smaller root of th = sqrt(2) cosh(th / 4)
    th = 1.0
    for _ in range(100):
        th = math.sqrt(2.0) * math.cosh(th / 4)
    return -2 * np.log(np.cosh((x - 0.5) * th / 2) / np.cosh(th / 4))


def test_solve_bvp_newton_bratu():
    import torch
    from scipy import integrate

    from numax_amd.callers import solve_bvp_newton

    def fun(x, y):
        return torch.stack([y[1], -torch.exp(y[0])])

    def bc(ya, yb):
        return torch.stack([ya[0], yb[0]])

    x = np.linspace(0.0, 1.0, 41)
    res = solve_bvp_newton(fun, bc, x, np.zeros((2, x.size)), tol=1e-8)
    assert res.converged and not res.singular, res
    y = res.y.cpu().numpy()
    assert np.max(np.abs(y[0] - _bratu_exact(x))) < 1e-7
    ref = integrate.solve_bvp(lambda xx, yy: np.vstack([yy[1], -np.exp(yy[0])]),
                              lambda ya, yb: np.array([ya[0], yb[0]]), x, np.zeros((2, x.size)), tol=1e-8)
    assert np.max(np.abs(y[0] - ref.sol(x)[0])) < 1e-6


def test_solve_bvp_newton_reports_a_singular_jacobian():
    import torch

    from numax_amd.callers import solve_bvp_newton

    def fun(x, y):
        return torch.stack([y[1], -y[0]])

    def bc(ya, yb):  # both conditions on y(0): the system has no unique solution
        return torch.stack([ya[0], 2 * ya[0]])

    res = solve_bvp_newton(fun, bc, np.linspace(0, 1, 6), np.zeros((2, 6)))
    assert res.singular and not res.converged
