"""The callers one step above numax.linalg's dense path, with their
factorizations and solves going through the library (the C ABI):

* ``multivariate_normal`` -- ``logpdf`` / ``pdf`` / ``rvs`` / ``entropy``
  (numax/stats/multivariate.mojo:36-260): one Cholesky of the covariance
  (``nmx_potrf``), one blocked triangular solve of all centered points
  (``nmx_trsm``), the log-determinant from the factor's diagonal;
* ``least_squares`` / ``curve_fit`` -- Levenberg-Marquardt through the
  augmented system ``[J; sqrt(mu) I] d = [-r; 0]`` handed to the QR
  least-squares solve (numax/optimize/least_squares.mojo:119-163, :218-428;
  ``nmx_gels``), the gradient ``J^T r`` on the MFMA GEMM (``nmx_gemm``);
* ``solve_bvp_newton`` -- the damped Newton iteration of ``solve_bvp`` on a
  fixed mesh (numax/integrate/bvp.mojo:1100-1200, SciPy's ``solve_newton``):
  fourth-order Lobatto IIIA collocation residuals, forward-difference
  Jacobian assembled dense, factored by ``lu_factor`` at a run-time order
  (``nmx_getrf``) with numax's ``singular()`` check, steps by ``LU.solve``.

The element-wise glue (centering, sums, residual functions) runs as torch
operations on the same device; every factorization, triangular solve, QR
solve and GEMM is the library's.  Mesh refinement, unknown parameters and
singular terms of ``solve_bvp`` are not restated here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import linalg as nl

__all__ = ["multivariate_normal", "least_squares", "curve_fit", "FitResult", "solve_bvp_newton", "BvpResult"]

_LN_2PI = 1.8378770664093453


def _log_det_from_cholesky(lower: torch.Tensor) -> float:
    """``ln det C = 2 sum ln L_ii`` (numax/stats/multivariate.mojo:36-51)."""
    d = torch.diagonal(lower).double().cpu()
    if bool((d <= 0).any()):
        raise nl.NumaxError("multivariate_normal: the covariance is not positive definite")
    return float(2.0 * torch.log(d).sum())


class multivariate_normal:  # noqa: N801 -- numax's name
    """``scipy.stats.multivariate_normal`` over device tensors."""

    @staticmethod
    def logpdf(x, mean, cov) -> torch.Tensor:
        """The log density of each row of ``x`` (m, d)."""
        X = nl._device_tensor(x)
        dt = X.dtype
        mu = nl._device_tensor(mean, dt)
        C = nl._device_tensor(cov, dt)
        lower = nl.cholesky(C)
        log_det = _log_det_from_cholesky(lower)
        centered = (X - mu).T.contiguous()          # column j is x[j] - mean
        z = nl.solve_triangular(lower, centered)     # L z = x - mean, all points at once
        return -(X.shape[1] * _LN_2PI + log_det + (z * z).sum(0)) / 2

    @staticmethod
    def pdf(x, mean, cov) -> torch.Tensor:
        return torch.exp(multivariate_normal.logpdf(x, mean, cov))

    @staticmethod
    def rvs(mean, cov, count: int, generator: torch.Generator | None = None) -> torch.Tensor:
        """``count`` draws, one per row: ``mean + L z`` with z ~ N(0, I)."""
        mu = nl._device_tensor(mean)
        lower = nl.cholesky(nl._device_tensor(cov, mu.dtype))
        z = torch.randn((mu.shape[0], count), dtype=mu.dtype, device=mu.device, generator=generator)
        return (nl.gemm(lower, z).T + mu).contiguous()

    @staticmethod
    def entropy(cov) -> float:
        C = nl._device_tensor(cov)
        return (C.shape[0] * (1.0 + _LN_2PI) + _log_det_from_cholesky(nl.cholesky(C))) / 2.0


@dataclass
class FitResult:
    """numax's ``FitResult`` (numax/optimize/least_squares.mojo:90-117)."""
    x: torch.Tensor
    cost: float
    grad_norm: float
    iterations: int
    converged: bool


def _damped_step(jac: torch.Tensor, r: torch.Tensor, damping: float) -> torch.Tensor:
    """``[J; sqrt(mu) I] d = [-r; 0]`` by QR least squares
    (numax/optimize/least_squares.mojo:119-163)."""
    n_resid, n_params = jac.shape
    a = torch.cat([jac, math.sqrt(damping) * torch.eye(n_params, dtype=jac.dtype, device=jac.device)])
    b = torch.cat([-r, torch.zeros(n_params, dtype=r.dtype, device=r.device)])
    return nl.lstsq(a, b)


def least_squares(residuals, jacobian, x0, tol: float = 1e-10, max_iter: int = 100, block: int = 16) -> FitResult:
    """``scipy.optimize.least_squares`` by Levenberg-Marquardt, numax's loop
    (numax/optimize/least_squares.mojo:218-428): convergence is
    ``max|J^T r| < tol``; each iteration tries up to 30 dampings (x3 on a
    rejected step, /3 down to 1e-12 on an accepted one); a fit that runs out
    of iterations or cannot find a downhill step returns ``converged=False``
    with the best parameters reached."""
    x = nl._device_tensor(x0).clone()
    cost = grad_norm = 0.0
    damping = 1e-3
    for iteration in range(max_iter):
        r = residuals(x)
        jac = jacobian(x).contiguous()
        cost = float((r * r).sum()) / 2
        g = nl.gemm(jac, r[:, None], transa=True)    # J^T r on the MFMA GEMM
        grad_norm = float(g.abs().max())
        if grad_norm < tol:
            return FitResult(x, cost, grad_norm, iteration, True)
        accepted = False
        candidate = x
        for _ in range(30):
            candidate = x + _damped_step(jac, r, damping)
            rc = residuals(candidate)
            if float((rc * rc).sum()) / 2 < cost:
                accepted = True
                damping = max(damping / 3, 1e-12)
                break
            damping = damping * 3
        if not accepted:
            return FitResult(x, cost, grad_norm, iteration + 1, False)
        x = candidate
    return FitResult(x, cost, grad_norm, max_iter, False)


def curve_fit(model, model_jac, xdata, ydata, p0, tol: float = 1e-10, max_iter: int = 100) -> FitResult:
    """``scipy.optimize.curve_fit``'s unweighted fit through ``least_squares``:
    residuals ``model(x, p) - y``, Jacobian ``model_jac(x, p)``
    (n_points x n_params)."""
    xs = nl._device_tensor(xdata)
    ys = nl._device_tensor(ydata, xs.dtype)
    return least_squares(lambda p: model(xs, p) - ys, lambda p: model_jac(xs, p), p0, tol, max_iter)


@dataclass
class BvpResult:
    y: torch.Tensor       # (n, m)
    iterations: int
    jacobians: int
    converged: bool
    singular: bool
    max_residual: float


def _collocation(fun, x, h, y):
    """SciPy's collocation residuals (Lobatto IIIA, fourth order): f at the
    nodes, y and f at the interval midpoints, and the (n, m-1) residuals."""
    f = fun(x, y)
    y_mid = 0.5 * (y[:, 1:] + y[:, :-1]) - 0.125 * h * (f[:, 1:] - f[:, :-1])
    f_mid = fun(x[:-1] + 0.5 * h, y_mid)
    res = y[:, 1:] - y[:, :-1] - h / 6 * (f[:, :-1] + f[:, 1:] + 4 * f_mid)
    return res, f_mid


def _residual(fun, bc, x, h, y):
    res, f_mid = _collocation(fun, x, h, y)
    return torch.cat([res.T.reshape(-1), bc(y[:, 0], y[:, -1])]), f_mid


def solve_bvp_newton(fun, bc, x, y, tol: float = 1e-3, bc_tol: float | None = None, max_njev: int = 4,
                     max_iter: int = 8) -> BvpResult:
    """The damped Newton iteration of numax's ``solve_bvp`` on the fixed mesh
    ``x`` (m nodes) from the guess ``y`` (n, m): ``fun(x, y) -> (n, k)`` and
    ``bc(ya, yb) -> (n,)``, vectorized over nodes.  Unknowns are ordered
    node by node; the residual is the collocation residuals, then the
    boundary residual.  The Jacobian is by forward differences of the whole
    residual at step ``sqrt(eps) (1 + |y|)``, factored dense by the
    library's LU (``lu_factor``); a singular factor stops the iteration.
    Steps are accepted by SciPy's Armijo test on ``|J^-1 r|^2`` with up to
    four halvings, and a full step reuses the Jacobian (at most ``max_njev``
    Jacobians, ``max_iter`` iterations).  Converged when every collocation
    residual is below ``(2/3) h 0.05 tol (1 + |f_mid|)`` and the boundary
    residual below ``bc_tol`` (default ``tol``)."""
    X = nl._device_tensor(x)
    Y = nl._device_tensor(y, X.dtype).clone()
    n, m = Y.shape
    h = (X[1:] - X[:-1])[None, :]
    btol = tol if bc_tol is None else bc_tol
    eps = torch.finfo(X.dtype).eps
    sigma, tau, n_trial = 0.2, 0.5, 4

    def residual_of(yv):
        return _residual(fun, bc, X, h, yv)

    def done(r, f_mid):
        col = r[:n * (m - 1)].reshape(m - 1, n).T
        ok = (col.abs() < (2.0 / 3.0) * h * 0.05 * tol * (1 + f_mid.abs())).all()
        return bool(ok) and bool((r[n * (m - 1):].abs() < btol).all())

    r, f_mid = residual_of(Y)
    njev, recompute, factor, step, cost = 0, True, None, None, 0.0
    it = 0
    for it in range(max_iter):
        if recompute:
            if njev == max_njev:
                break
            flat = Y.T.reshape(-1)                          # node-major unknowns
            step_h = math.sqrt(eps) * (1 + flat.abs())
            cols = []
            for j in range(n * m):
                yp = flat.clone()
                yp[j] += step_h[j]
                rp, _ = residual_of(yp.reshape(m, n).T)
                cols.append((rp - r) / step_h[j])
            jac = torch.stack(cols, dim=1).contiguous()
            njev += 1
            factor = nl.lu_factor(jac)
            if factor.singular():
                return BvpResult(Y, it, njev, False, True, float(r.abs().max()))
            step = factor.solve(r)
            cost = float((step * step).sum())
        alpha = 1.0
        for trial in range(n_trial + 1):
            y_new = Y - alpha * step.reshape(m, n).T
            r_new, f_mid_new = residual_of(y_new)
            step_new = factor.solve(r_new)
            cost_new = float((step_new * step_new).sum())
            if cost_new < (1 - 2 * alpha * sigma) * cost or trial == n_trial:
                break
            alpha *= tau
        Y, r, f_mid = y_new, r_new, f_mid_new
        if done(r, f_mid):
            return BvpResult(Y, it + 1, njev, True, False, float(r.abs().max()))
        # a full step keeps the Jacobian; a damped one asks for a new one
        recompute = alpha != 1.0
        step, cost = step_new, cost_new
    return BvpResult(Y, it + 1, njev, done(r, f_mid), False, float(r.abs().max()))
