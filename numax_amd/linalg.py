"""Host-side mirror of numax.linalg's dense factor-and-solve surface.

Same names, argument meaning and error behaviour as numax's Tensor tier
(numax/linalg/lu.mojo, cholesky.mojo, triangular.mojo, basic.mojo), with the
device work done by libnumax_hip through its C ABI:

* ``lu_factor(a)`` -> ``LU`` (``.factored``, ``.pivots``, ``.sign``, ``.solve``,
  ``.det``, ``.slogdet``, ``.singular``)            -- lu.mojo:75-701, :998-1361
* ``solve(a, b)``                                    -- basic.mojo:55-146
* ``cholesky(a)``, ``cholesky_solve(lower, b)``      -- cholesky.mojo:233-580
* ``solve_triangular(a, b, upper, unit, trans)``     -- triangular.mojo:228-425

Arrays are ``torch`` tensors on a ROCm device (row-major, contiguous); NumPy
inputs are copied to ``cuda:0``.  Inputs are never modified (the reference
copies ``a`` into a work tensor too, lu.mojo:593).  There is no CPU path:
without the built library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

__all__ = ["LU", "PLU", "LDL", "NumaxError", "lu_factor", "lu", "ldl", "inverse", "solve", "solve_mpir", "solve_mpir_graph", "cholesky", "cholesky_solve", "solve_triangular",
           "gemm", "PRECISIONS"]

# trailing-update precisions of the mixed-precision solver (nmx_ir_cfg.precision)
# how the low-precision factor pivots: per-column partial pivoting (numax's
# pivots), tournament pivoting (one reduction per 32-column leaf), or random
# butterflies with no pivoting
PIVOTS = {"partial": 1, "tournament": 2, "butterfly": 0}
PRECISIONS = {"fp16": 0, "bf16": 1, "fp16s": 2, "fp32": 3}


class NumaxError(ValueError):
    """The reference's raised ``Error`` (shape problems, non-PD Cholesky)."""


_SUFFIX = {torch.float64: "f64", torch.float32: "f32"}


def _device_tensor(x, dtype=None) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if t.device.type != "cuda":
        t = t.to("cuda:0")
    return t.contiguous()


def _ctx(t: torch.Tensor) -> _lib.Context:
    dev = t.device.index or 0
    stream = torch.cuda.current_stream(dev).cuda_stream
    return _lib.context_for(dev, stream)


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def _fn(name: str, dtype: torch.dtype):
    if dtype not in _SUFFIX:
        raise NumaxError(f"{name}: dtype {dtype} is not a supported floating-point type (float32, float64)")
    return getattr(_lib.load(), f"nmx_{name}_{_SUFFIX[dtype]}")


def _square(a: torch.Tensor, who: str) -> int:
    if a.dim() != 2 or a.shape[0] != a.shape[1]:
        raise NumaxError(f"{who}: matrix must be square, got shape {tuple(a.shape)}")
    return int(a.shape[0])


class LU:
    """A blocked LU factorization with partial pivoting (numax ``LU`` /
    ``DynamicLU``, lu.mojo:75-451, :998-1197).

    ``factored`` holds L strictly below the diagonal (unit diagonal implicit)
    and U on and above it; ``pivots[j]`` is the row column j interchanged with
    (LAPACK ipiv, 0-based); ``sign`` is the interchange parity.
    """

    def __init__(self, factored: torch.Tensor, pivots: torch.Tensor, sign: int, info: int):
        self.factored = factored
        self.pivots = pivots
        self.sign = sign
        self.info = info

    @property
    def n(self) -> int:
        return int(self.factored.shape[0])

    def singular(self) -> bool:
        """``DynamicLU.singular`` (lu.mojo:1047-1056): an exactly zero pivot
        was met (it was divided as 1e-30, not raised)."""
        return self.info != 0

    def solve(self, b) -> torch.Tensor:
        """``x`` with ``A x = b`` (a vector) or ``X`` with ``A X = B`` (an
        ``n x k`` matrix), reusing this factorization (lu.mojo:181-354)."""
        x = _device_tensor(b, self.factored.dtype).clone()
        n = self.n
        if x.dim() == 1:
            if x.shape[0] != n:
                raise NumaxError(f"lu_solve: b has {x.shape[0]} entries for an order-{n} factorization")
            nrhs, ldb = 1, 1
        elif x.dim() == 2:
            if x.shape[0] != n:
                raise NumaxError(f"lu_solve: b has {x.shape[0]} rows for an order-{n} factorization")
            nrhs = int(x.shape[1])
            ldb = max(nrhs, 1)
        else:
            raise NumaxError("lu_solve: b must be a vector or a matrix")
        ctx = _ctx(self.factored)
        _lib.check(_fn("getrs", self.factored.dtype)(ctx.handle, n, nrhs, _ptr(self.factored), n,
                                                       _ptr(self.pivots), _ptr(x), ldb), "lu_solve")
        return x

    def _diagonal(self) -> np.ndarray:
        d = torch.empty(self.n, dtype=self.factored.dtype, device=self.factored.device)
        ctx = _ctx(self.factored)
        _lib.check(_fn("diag", self.factored.dtype)(ctx.handle, self.n, _ptr(self.factored), self.n, _ptr(d)),
                   "det")
        return d.cpu().numpy()

    def det(self) -> float:
        """Product of U's diagonal times the swap parity (lu.mojo:382-404)."""
        prod = float(self.sign)
        for v in self._diagonal():
            prod *= float(v)
        return prod

    def slogdet(self) -> tuple[float, float]:
        """``(sign, ln|det|)``; ``(0, -inf)`` on a zero diagonal entry
        (lu.mojo:406-451)."""
        sign = float(self.sign)
        total = 0.0
        for v in self._diagonal():
            v = float(v)
            if v == 0.0:
                return 0.0, -math.inf
            if v < 0:
                sign = -sign
            total += math.log(abs(v))
        return sign, total


def lu_factor(a, block: int = 16, base: int = 16) -> LU:
    """Factor ``a`` into ``P L U`` (numax ``lu_factor``, lu.mojo:454-701).

    ``block`` is the panel width and ``base`` the recursive panel's leaf width
    (the reference's parameters of the same names, with its device defaults
    16 / 16, lu.mojo:454-459); the pivots do not depend on them, the speed
    does: 256 / 32 is the fastest shape on MI355X at large n.  A singular
    matrix is not reported as an error (see ``LU.singular``).
    """
    work = _device_tensor(a).clone()
    n = _square(work, "lu_factor")
    piv = torch.zeros(max(n, 1), dtype=torch.int32, device=work.device)
    info = torch.zeros(1, dtype=torch.int32, device=work.device)
    ctx = _ctx(work)
    _lib.check(_fn("getrf", work.dtype)(ctx.handle, n, _ptr(work), max(n, 1), _ptr(piv), _ptr(info), block, base),
               "lu_factor")
    host_piv = piv[:n].cpu().numpy()
    ctx.check_device()
    sign = 1
    for j, p in enumerate(host_piv):
        if p != j:
            sign = -sign
    return LU(work, piv[:n], sign, int(info.item()))


class PLU:
    """What numax's ``lu`` returns (lu.mojo:849-874): ``p``, ``l``, ``u`` with
    ``A = P @ L @ U``, SciPy's ``(p, l, u)``; ``perm[i]`` is the row of ``A``
    that row ``i`` of ``L @ U`` came from."""

    def __init__(self, p: torch.Tensor, l: torch.Tensor, u: torch.Tensor, perm: torch.Tensor):
        self.p = p
        self.l = l
        self.u = u
        self.perm = perm

    def __iter__(self):
        return iter((self.p, self.l, self.u))


def lu(a, block: int = 16) -> PLU:
    """numax ``lu`` (lu.mojo:877-975): ``lu_factor``, then the packed factor
    and interchanges unpacked on the device into ``P``, ``L``, ``U``."""
    f = lu_factor(a, block=block)
    n = f.n
    dev, dt = f.factored.device, f.factored.dtype
    p = torch.empty((n, n), dtype=dt, device=dev)
    l = torch.empty_like(p)
    u = torch.empty_like(p)
    perm = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ld = max(n, 1)
    _lib.check(_fn("lu_unpack", dt)(_ctx(f.factored).handle, n, _ptr(f.factored), ld, _ptr(f.pivots), _ptr(perm),
                                     _ptr(p), ld, _ptr(l), ld, _ptr(u), ld), "lu")
    return PLU(p, l, u, perm[:n])


def inverse(a, block: int = 16) -> torch.Tensor:
    """numax ``inverse`` (basic.mojo:149-195): one ``lu_factor`` and one
    ``LU.solve`` against the whole identity.  A singular ``a`` does not raise;
    it yields non-finite entries."""
    f = lu_factor(a, block=block)
    n = f.n
    out = torch.empty((n, n), dtype=f.factored.dtype, device=f.factored.device)
    ld = max(n, 1)
    _lib.check(_fn("getri", f.factored.dtype)(_ctx(f.factored).handle, n, _ptr(f.factored), ld, _ptr(f.pivots),
                                               _ptr(out), ld), "inverse")
    return out


class LDL:
    """What numax's ``ldl`` returns (numax/linalg/ldl.mojo:30-56): SciPy's
    ``(lu, d, perm)`` with ``a = lu @ d @ lu^T`` and ``lu[perm]`` unit lower
    triangular; ``factored`` / ``ipiv`` are sytf2's packed factor and pivots."""

    def __init__(self, lu, d, perm, factored, ipiv, info):
        self.lu = lu
        self.d = d
        self.perm = perm
        self.factored = factored
        self.ipiv = ipiv
        self.info = info

    def __iter__(self):
        return iter((self.lu, self.d, self.perm))


def ldl(a) -> LDL:
    """numax ``ldl`` (numax/linalg/ldl.mojo:59-148), ``scipy.linalg.ldl(a)``
    with ``lower=True``: Bunch-Kaufman ``L D L^T`` of the symmetric ``a``
    (lower triangle read).  A singular ``a`` does not raise; its zero pivot
    yields non-finite entries."""
    work = _device_tensor(a).clone()
    n = _square(work, "ldl")
    dev, dt = work.device, work.dtype
    ld = max(n, 1)
    ipiv = torch.zeros(ld, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    ctx = _ctx(work)
    _lib.check(_fn("sytrf", dt)(ctx.handle, n, _ptr(work), ld, _ptr(ipiv), _ptr(info)), "ldl")
    lu = torch.empty((n, n), dtype=dt, device=dev)
    d = torch.empty_like(lu)
    perm = torch.empty(ld, dtype=torch.int64, device=dev)
    _lib.check(_fn("ldl_unpack", dt)(ctx.handle, n, _ptr(work), ld, _ptr(ipiv), _ptr(lu), ld, _ptr(d), ld, _ptr(perm)),
               "ldl")
    return LDL(lu, d, perm[:n], work, ipiv[:n], int(info.item()))


def laswp(b, pivots) -> torch.Tensor:
    """laswp / laswp_matrix (numax/linalg/panel.mojo:1824-1852, :2279-2313):
    apply the interchanges ``pivots`` (LAPACK order, 0-based) to the rows of a
    float64 or float32 vector or n x k matrix ``b``; returns the permuted copy."""
    x = _device_tensor(b).clone()
    n = x.shape[0]
    nrhs = 1 if x.dim() == 1 else x.shape[1]
    piv = _device_tensor(pivots, torch.int32)
    if piv.numel() != n:
        raise NumaxError(f"laswp: {piv.numel()} pivots for {n} rows")
    _lib.check(_fn("laswp", x.dtype)(_ctx(x).handle, n, nrhs, _ptr(x), max(nrhs, 1), _ptr(piv)), "laswp")
    return x


def solve(a, b, block: int = 16) -> torch.Tensor:
    """``lu_factor`` then the solve at the same device (basic.mojo:55-146)."""
    return lu_factor(a, block=block).solve(b)


def cholesky(a, block: int = 32) -> torch.Tensor:
    """Lower ``L`` with ``L L^T = a`` (cholesky.mojo:233-396).  Reads only the
    lower triangle; raises naming the 0-based index of the first non-positive
    pivot, after the whole factorization has run.  ``block`` defaults to
    numax's device default (32, cholesky.mojo:233-237); 256 is the fastest
    panel width on MI355X at large n."""
    work = _device_tensor(a).clone()
    n = _square(work, "cholesky")
    info = torch.zeros(1, dtype=torch.int32, device=work.device)
    ctx = _ctx(work)
    _lib.check(_fn("potrf", work.dtype)(ctx.handle, n, _ptr(work), max(n, 1), _ptr(info), block), "cholesky")
    flag = int(info.item())
    if flag != 0:
        raise NumaxError(f"cholesky: matrix is not positive definite (non-positive pivot at index {flag - 1})")
    return work


def cholesky_solve(lower, b) -> torch.Tensor:
    """``x`` with ``L L^T x = b`` from the Cholesky factor (cholesky.mojo:471-580)."""
    l = _device_tensor(lower)
    n = _square(l, "cholesky_solve")
    x = _device_tensor(b, l.dtype).clone()
    if x.shape[0] != n:
        raise NumaxError(f"cholesky_solve: b has {x.shape[0]} entries for an order-{n} factor")
    nrhs = 1 if x.dim() == 1 else int(x.shape[1])
    ctx = _ctx(l)
    _lib.check(_fn("potrs", l.dtype)(ctx.handle, n, nrhs, _ptr(l), n, _ptr(x), max(nrhs, 1)), "cholesky_solve")
    return x


def solve_triangular(a, b, upper: bool = False, unit: bool = False, trans: bool = False) -> torch.Tensor:
    """``solve_triangular`` (triangular.mojo:228-425): ``upper`` names the
    stored triangle (SciPy's ``lower`` inverted), ``trans`` solves against
    ``a^T`` without transposing, ``unit`` ignores the diagonal."""
    t = _device_tensor(a)
    n = _square(t, "solve_triangular")
    x = _device_tensor(b, t.dtype).clone()
    if x.shape[0] != n:
        raise NumaxError(f"solve_triangular: b has {x.shape[0]} entries for an order-{n} matrix")
    nrhs = 1 if x.dim() == 1 else int(x.shape[1])
    walk_upper = bool(upper) != bool(trans)
    ctx = _ctx(t)
    _lib.check(_fn("trsm", t.dtype)(ctx.handle, int(walk_upper), int(bool(unit)), int(bool(trans)), n, nrhs, _ptr(t),
                                    n, _ptr(x), max(nrhs, 1)), "solve_triangular")
    return x


class QR:
    """``QR`` (numax/linalg/qr.mojo:330-612): a blocked Householder QR held on
    the device -- ``factored`` has R on and above the diagonal and the
    reflectors below it (unit diagonal implicit), ``taus`` one scale per
    column, ``block`` the panel width it ran at.  ``r()`` and ``q()`` form
    the factors; ``apply_q_transpose`` and ``solve`` never form Q."""

    def __init__(self, factored: torch.Tensor, taus: torch.Tensor, block: int):
        self.factored = factored
        self.taus = taus
        self.block = int(block)

    @property
    def m(self) -> int:
        return int(self.factored.shape[0])

    @property
    def n(self) -> int:
        return int(self.factored.shape[1])

    def r(self) -> torch.Tensor:
        """The ``n x n`` upper-triangular R, zeros below (qr.mojo:371-403)."""
        n = self.n
        out = torch.empty((n, n), dtype=self.factored.dtype, device=self.factored.device)
        _lib.check(_fn("qr_r", out.dtype)(_ctx(out).handle, n, _ptr(self.factored), max(n, 1), _ptr(out), max(n, 1)),
                   "QR.r")
        return out

    def q(self) -> torch.Tensor:
        """The thin ``m x n`` Q with orthonormal columns (qr.mojo:405-468)."""
        m, n = self.m, self.n
        out = torch.empty((m, n), dtype=self.factored.dtype, device=self.factored.device)
        _lib.check(_fn("orgqr", out.dtype)(_ctx(out).handle, m, n, _ptr(self.factored), max(n, 1), _ptr(self.taus),
                                           _ptr(out), max(n, 1), self.block), "QR.q")
        return out

    def _apply(self, b, trans: int, who: str) -> torch.Tensor:
        x = _device_tensor(b, self.factored.dtype).clone()
        if x.shape[0] != self.m:
            raise NumaxError(f"{who}: b has {x.shape[0]} rows for an m = {self.m} factorization")
        nrhs = 1 if x.dim() == 1 else int(x.shape[1])
        _lib.check(_fn("ormqr", x.dtype)(_ctx(x).handle, trans, self.m, self.n, nrhs, _ptr(self.factored),
                                         max(self.n, 1), _ptr(self.taus), _ptr(x), max(nrhs, 1), self.block), who)
        return x

    def apply_q_transpose(self, b) -> torch.Tensor:
        """``Q^T b`` for an ``m``-row ``b`` (qr.mojo:470-534)."""
        return self._apply(b, 1, "QR.apply_q_transpose")

    def solve(self, b) -> torch.Tensor:
        """Least-squares ``x`` of ``A x ~= b`` (qr.mojo:536-612): ``Q^T b``, then
        the back substitution with R; Q is never formed."""
        y = self._apply(b, 1, "QR.solve")
        n = self.n
        nrhs = 1 if y.dim() == 1 else int(y.shape[1])
        x = y[:n].clone()
        _lib.check(_fn("trsm", x.dtype)(_ctx(x).handle, 1, 0, 0, n, nrhs, _ptr(self.factored), max(n, 1), _ptr(x),
                                        max(nrhs, 1)), "QR.solve")
        return x


def _qr_input(a, who: str) -> torch.Tensor:
    t = _device_tensor(a)
    if t.dtype not in (torch.float64, torch.float32):
        raise NumaxError(f"{who}: dtype {t.dtype} is not supported by the device QR (float32, float64)")
    if t.dim() != 2 or t.shape[0] < t.shape[1]:
        raise NumaxError(f"{who}: needs an m x n matrix with m >= n, got shape {tuple(t.shape)}")
    return t


def qr_factor(a, block: int = 16) -> QR:
    """``qr_factor`` (numax/linalg/qr.mojo:615-743): blocked Householder QR of
    an ``m x n`` float32 (numax's device precision) or float64 matrix with
    ``m >= n``; ``block`` is the panel width (1..64, numax's default 16)."""
    work = _qr_input(a, "qr_factor").clone()
    m, n = work.shape
    taus = torch.zeros(n, dtype=work.dtype, device=work.device)
    _lib.check(_fn("geqrf", work.dtype)(_ctx(work).handle, m, n, _ptr(work), max(n, 1), _ptr(taus), block),
               "qr_factor")
    return QR(work, taus, block)


def lstsq(a, b, block: int = 16, method: str = "qr") -> torch.Tensor:
    """``lstsq`` (numax/linalg/qr.mojo:746-821), method "qr": the ``x``
    minimizing ``||A x - b||`` for ``m >= n`` -- ``qr_factor`` then
    ``QR.solve`` in one call."""
    if method != "qr":
        raise NumaxError(f"lstsq: method {method!r} is not on this device path (\"qr\")")
    work = _qr_input(a, "lstsq").clone()
    m, n = work.shape
    x = _device_tensor(b, work.dtype).clone()
    if x.shape[0] != m:
        raise NumaxError(f"lstsq: b has {x.shape[0]} rows for an {m} x {n} matrix")
    nrhs = 1 if x.dim() == 1 else int(x.shape[1])
    taus = torch.zeros(n, dtype=work.dtype, device=work.device)
    _lib.check(_fn("gels", work.dtype)(_ctx(work).handle, m, n, nrhs, _ptr(work), max(n, 1), _ptr(taus), _ptr(x),
                                       max(nrhs, 1), block), "lstsq")
    return x[:n].clone()


def gemm(a, b, c=None, alpha: float = 1.0, beta: float = 0.0, transa: bool = False, transb: bool = False):
    """``c = beta c + alpha op(a) op(b)`` on the MFMA kernel the drivers use."""
    A = _device_tensor(a)
    B = _device_tensor(b, A.dtype)
    m, k = (A.shape[1], A.shape[0]) if transa else (A.shape[0], A.shape[1])
    kb, n = (B.shape[1], B.shape[0]) if transb else (B.shape[0], B.shape[1])
    if k != kb:
        raise NumaxError(f"gemm: inner dimensions {k} and {kb} differ")
    C = torch.zeros((m, n), dtype=A.dtype, device=A.device) if c is None else _device_tensor(c, A.dtype).clone()
    ctx = _ctx(A)
    _lib.check(_fn("gemm", A.dtype)(ctx.handle, int(transa), int(transb), m, n, k, alpha, _ptr(A), A.shape[1],
                                    _ptr(B), B.shape[1], beta, _ptr(C), max(n, 1)), "gemm")
    return C


SOLVERS = {"auto": 0, "classical": 1, "gmres": 2}


def _report(rep) -> dict:
    return {"iterations": rep.iterations, "gmres_iterations": rep.gmres_iterations, "converged": bool(rep.converged),
            "backward_error": rep.backward_error, "factor_ms": rep.factor_ms, "refine_ms": rep.refine_ms,
            "info": rep.info, "history": [rep.history[i] for i in range(min(rep.iterations + 1, 32))]}


def solve_mpir(a, b, precision: str = "fp16s", tol: float = 1e-15, max_iter: int = 30, block=None,
               seed: int = 0x5EED, raise_on_failure: bool = True, pivot: str = "tournament", solver: str = "auto",
               gmres_restart: int = 30):
    """``x`` with ``A x = b`` to FP64 backward error from a low-precision
    factorization (no reference counterpart; numax's ``solve`` is the FP64
    path).  ``precision`` picks the trailing-update arithmetic: ``"fp16"``,
    ``"bf16"``, ``"fp16s"`` (FP16 hi + lo split, ~FP32 accuracy) or ``"fp32"``.
    ``block``: the panel width (a multiple of 64; None: the autotuned width
    cached for this order and precision, else 256).  ``pivot``:
    ``"tournament"`` (every 32-column leaf's pivots from one reduction tree),
    ``"partial"`` (column by column, numax's pivot sequence) or
    ``"butterfly"`` (random butterflies, no pivoting).  ``solver``:
    ``"classical"`` refinement, ``"gmres"`` (GMRES-IR: FP64 GMRES
    on the correction equation, preconditioned by the low-precision factor)
    or ``"auto"`` (classical while each step at least halves the backward
    error, GMRES-IR from the first step that does not).  Returns ``(x, report)``; ``report`` holds the
    refinement history, the GMRES iterations and the final normwise
    backward error."""
    A = _device_tensor(a, torch.float64)
    n = _square(A, "solve_mpir")
    B = _device_tensor(b, torch.float64)
    if B.dim() == 1:
        cols = [B]
    elif B.dim() == 2:
        cols = [B[:, j].contiguous() for j in range(B.shape[1])]
    else:
        raise NumaxError("solve_mpir: b must be a vector or a matrix")
    if B.shape[0] != n:
        raise NumaxError(f"solve_mpir: b has {B.shape[0]} entries for an order-{n} matrix")
    if precision not in PRECISIONS:
        raise NumaxError(f"solve_mpir: precision must be one of {sorted(PRECISIONS)}")
    lib = _lib.load()
    cfg = _lib.IrCfg()
    lib.nmx_ir_cfg_default(ctypes.byref(cfg))
    cfg.precision = PRECISIONS[precision]
    cfg.tol = tol
    cfg.max_iter = max_iter
    if block is None:
        from numax_amd import autotune
        block = autotune.cached("mpir", A.device.index or 0, n=n, precision=precision) or 256
    cfg.block = int(block)
    cfg.seed = seed
    if pivot not in PIVOTS:
        raise NumaxError(f"solve_mpir: pivot must be one of {sorted(PIVOTS)}")
    cfg.pivot = PIVOTS[pivot]
    if solver not in SOLVERS:
        raise NumaxError(f"solve_mpir: solver must be one of {sorted(SOLVERS)}")
    cfg.solver = SOLVERS[solver]
    cfg.gmres_restart = gmres_restart
    ctx = _ctx(A)
    if B.dim() == 2 and B.shape[1] > 1:
        # one factorization, all columns refined together
        if cfg.solver == SOLVERS["gmres"]:
            raise NumaxError("solve_mpir: GMRES-IR takes one right-hand side at a time")
        nrhs = int(B.shape[1])
        X = torch.empty_like(B)
        reps = (_lib.IrReport * nrhs)()
        st = lib.nmx_gesv_mpir_nrhs(ctx.handle, n, nrhs, _ptr(A), n, _ptr(B), nrhs, _ptr(X), nrhs,
                                    ctypes.byref(cfg), reps)
        if st != 0 and not (st == 5 and not raise_on_failure):
            _lib.check(st, "solve_mpir")
        return X, [_report(r) for r in reps]
    xs, reports = [], []
    for col in cols:
        x = torch.empty_like(col)
        rep = _lib.IrReport()
        st = lib.nmx_gesv_mpir(ctx.handle, n, _ptr(A), n, _ptr(col), _ptr(x), ctypes.byref(cfg), ctypes.byref(rep))
        info = _report(rep)
        if st != 0 and not (st == 5 and not raise_on_failure):
            _lib.check(st, "solve_mpir")
        xs.append(x)
        reports.append(info)
    x = xs[0] if B.dim() == 1 else torch.stack(xs, dim=1)
    return x, (reports[0] if B.dim() == 1 else reports)


class MpirGraph:
    """A captured :func:`solve_mpir` (one right-hand side, classical
    refinement): :meth:`replay` solves with whatever ``a`` and ``b`` hold at
    that moment into ``x`` with one graph launch and returns ``(x, report)``,
    the report read back from the device (no timings)."""

    def __init__(self, handle, ctx, a: torch.Tensor, b: torch.Tensor, x: torch.Tensor, eager_report: dict, keep):
        self._handle = handle
        self._ctx = ctx
        self._keep = keep
        self.a = a
        self.b = b
        self.x = x
        self.eager_report = eager_report

    def replay(self):
        lib = _lib.load()
        _lib.check(lib.nmx_graph_launch(self._handle), "graph replay")
        rep = _lib.IrReport()
        _lib.check(lib.nmx_ir_report_read(self._ctx.handle, ctypes.byref(rep)), "nmx_ir_report_read")
        return self.x, _report(rep)

    def close(self) -> None:
        if self._handle:
            _lib.load().nmx_graph_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_mpir_graph(a, b, precision: str = "fp16s", tol: float = 1e-15, max_iter: int = 30, block: int = 256,
                     pivot: str = "tournament") -> MpirGraph:
    """:func:`solve_mpir` with classical refinement run once eagerly (which
    sizes every buffer and creates the look-ahead stream), then captured into
    a graph on the same stream; the eager solution is in ``graph.x`` and its
    report in ``graph.eager_report`` until the first replay.  The graph holds
    ``max_iter`` refinement steps, the ones after the device-side stop test
    returning at once."""
    A = _device_tensor(a, torch.float64).clone()
    n = _square(A, "solve_mpir_graph")
    Bv = _device_tensor(b, torch.float64).clone()
    if Bv.dim() != 1 or Bv.shape[0] != n:
        raise NumaxError("solve_mpir_graph: b must be a vector of the matrix's order")
    if precision not in PRECISIONS:
        raise NumaxError(f"solve_mpir_graph: precision must be one of {sorted(PRECISIONS)}")
    if pivot not in PIVOTS:
        raise NumaxError(f"solve_mpir_graph: pivot must be one of {sorted(PIVOTS)}")
    lib = _lib.load()
    cfg = _lib.IrCfg()
    lib.nmx_ir_cfg_default(ctypes.byref(cfg))
    cfg.precision = PRECISIONS[precision]
    cfg.tol = tol
    cfg.max_iter = max_iter
    cfg.block = int(block)
    cfg.pivot = PIVOTS[pivot]
    cfg.solver = SOLVERS["classical"]
    ctx = _ctx(A)
    x = torch.empty_like(Bv)
    rep = _lib.IrReport()
    st = lib.nmx_gesv_mpir(ctx.handle, n, _ptr(A), n, _ptr(Bv), _ptr(x), ctypes.byref(cfg), ctypes.byref(rep))
    if st not in (0, 5):
        _lib.check(st, "solve_mpir_graph (eager)")
    eager = _report(rep)
    torch.cuda.synchronize(A.device)
    _lib.check(lib.nmx_graph_begin(ctx.handle), "graph begin")
    st = lib.nmx_gesv_mpir(ctx.handle, n, _ptr(A), n, _ptr(Bv), _ptr(x), ctypes.byref(cfg), None)
    handle = ctypes.c_void_p()
    st_end = lib.nmx_graph_end(ctx.handle, ctypes.byref(handle))
    _lib.check(st, "solve_mpir (captured)")
    _lib.check(st_end, "graph end")
    return MpirGraph(handle, ctx, A, Bv, x, eager, cfg)


# ---- batched uneven-size solves ---------------------------------------------
class BatchResult:
    """Per-problem results of :func:`batched_solve`: the FP64 solutions
    (packed like the right-hand sides), ``info`` (first zero / non-positive
    pivot, 1-based; -1 for a problem that was not taken), refinement steps and
    final backward error."""

    def __init__(self, x, info, iters, berr, offsets_b, n):
        self.x, self.info, self.iters, self.berr = x, info, iters, berr
        self._offb, self._n = offsets_b, n

    def solution(self, p: int) -> torch.Tensor:
        o = int(self._offb[p])
        return self.x[o:o + int(self._n[p])]


def pack_systems(mats, rhss):
    """Pack square matrices and vectors of uneven sizes into the flat layout
    of :func:`batched_solve` (row-major, back to back)."""
    n = np.array([np.asarray(m).shape[0] for m in mats], dtype=np.int32)
    offa = np.zeros(len(mats), dtype=np.int64)
    offb = np.zeros(len(mats), dtype=np.int64)
    if len(mats):
        offa[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
        offb[1:] = np.cumsum(n.astype(np.int64))[:-1]
    a = np.concatenate([np.asarray(m, dtype=np.float64).ravel() for m in mats]) if len(mats) else np.zeros(1)
    b = np.concatenate([np.asarray(r, dtype=np.float64).ravel() for r in rhss]) if len(mats) else np.zeros(1)
    return a, b, n, offa, offb


def batched_solve(a, b, n, kind, offsets_a, offsets_b, tol: float = 1e-15, max_iter: int = 10,
                  teams_per_cu=None) -> BatchResult:
    """Solve ``len(n)`` independent systems of order 1..256 in one call: the
    register-tier ``lu_factor`` + ``PivotedLU.solve`` (kind 0) and
    ``cholesky`` + ``cholesky_solve`` (kind 1) of numax
    (numax/linalg/_array/lu.mojo:165-216, numax/linalg/_array/cholesky.mojo:26-101)
    that ``map_blocks`` runs one problem per thread
    (numax/core/functional.mojo:691-834), past its n <= 64, one-n-per-launch
    limit.  FP32 factorization, FP64 residual and refinement to backward error
    ``tol``.  ``a``, ``b``: flat float64 arrays; ``offsets_*``: element offsets
    of each problem's matrix / vector."""
    A = _device_tensor(a, torch.float64)
    B = _device_tensor(b, torch.float64)
    count = len(n)
    dev = A.device
    N = _device_tensor(np.asarray(n, dtype=np.int32))
    K = _device_tensor(np.asarray(kind, dtype=np.uint8))
    OA = _device_tensor(np.asarray(offsets_a, dtype=np.int64))
    OB = _device_tensor(np.asarray(offsets_b, dtype=np.int64))
    X = torch.zeros_like(B)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    iters = torch.zeros(count, dtype=torch.int32, device=dev)
    berr = torch.zeros(count, dtype=torch.float64, device=dev)
    lib = _lib.load()
    cfg = _lib.BatchCfg()
    lib.nmx_batch_cfg_default(ctypes.byref(cfg))
    cfg.tol = float(tol)
    cfg.max_iter = int(max_iter)
    if teams_per_cu is None:
        from numax_amd import autotune
        teams_per_cu = autotune.cached("batched", A.device.index or 0)
    if teams_per_cu is not None:
        for i, v in enumerate(teams_per_cu):
            cfg.teams_per_cu[i] = int(v)
    _lib.check(lib.nmx_batched_gesv(_ctx(A).handle, count, _ptr(N), _ptr(OA), _ptr(OB), _ptr(K), _ptr(A), A.numel(),
                                    _ptr(B), B.numel(), _ptr(X), _ptr(info), _ptr(iters), _ptr(berr),
                                    ctypes.byref(cfg)), "batched_solve")
    return BatchResult(X, info, iters, berr, np.asarray(offsets_b), np.asarray(n))


class BatchedSolveGraph:
    """One ``nmx_batched_gesv`` captured as a HIP graph for a fixed batch shape
    (counts, sizes, offsets): ``replay()`` re-solves whatever values ``a`` and
    ``b`` hold at that moment into ``result`` with one launch.  Built by
    :func:`batched_solve_graph`; the device buffers stay owned here."""

    def __init__(self, handle, keep, result: BatchResult, a: torch.Tensor, b: torch.Tensor):
        self._handle = handle
        self._keep = keep
        self.result = result
        self.a = a
        self.b = b

    def replay(self) -> BatchResult:
        _lib.check(_lib.load().nmx_graph_launch(self._handle), "graph replay")
        return self.result

    def close(self) -> None:
        if self._handle:
            _lib.load().nmx_graph_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batched_solve_graph(a, b, n, kind, offsets_a, offsets_b, tol: float = 1e-15, max_iter: int = 10,
                        teams_per_cu=None) -> BatchedSolveGraph:
    """:func:`batched_solve` run once eagerly (which also sizes the workspace),
    then captured into a graph on the same stream for replays; the eager
    result is in ``graph.result`` until the first replay overwrites it."""
    A = _device_tensor(a, torch.float64).clone()
    B = _device_tensor(b, torch.float64).clone()
    res = batched_solve(A, B, n, kind, offsets_a, offsets_b, tol=tol, max_iter=max_iter, teams_per_cu=teams_per_cu)
    count = len(n)
    N = _device_tensor(np.asarray(n, dtype=np.int32))
    K = _device_tensor(np.asarray(kind, dtype=np.uint8))
    OA = _device_tensor(np.asarray(offsets_a, dtype=np.int64))
    OB = _device_tensor(np.asarray(offsets_b, dtype=np.int64))
    lib = _lib.load()
    cfg = _lib.BatchCfg()
    lib.nmx_batch_cfg_default(ctypes.byref(cfg))
    cfg.tol = float(tol)
    cfg.max_iter = int(max_iter)
    if teams_per_cu is not None:
        for i, v in enumerate(teams_per_cu):
            cfg.teams_per_cu[i] = int(v)
    ctx = _ctx(A)
    torch.cuda.synchronize(A.device)
    _lib.check(lib.nmx_graph_begin(ctx.handle), "graph begin")
    st = lib.nmx_batched_gesv(ctx.handle, count, _ptr(N), _ptr(OA), _ptr(OB), _ptr(K), _ptr(A), A.numel(), _ptr(B),
                              B.numel(), _ptr(res.x), _ptr(res.info), _ptr(res.iters), _ptr(res.berr),
                              ctypes.byref(cfg))
    handle = ctypes.c_void_p()
    st_end = lib.nmx_graph_end(ctx.handle, ctypes.byref(handle))
    _lib.check(st, "batched_solve (captured)")
    _lib.check(st_end, "graph end")
    return BatchedSolveGraph(handle, (N, K, OA, OB, cfg), res, A, B)


# ---- batched small-matrix eigensolver and tall-skinny least squares --------------
def eigh_batched(a, n, offsets_a, offsets_w, sweeps: int = 12, vectors: bool = True):
    """numax's register-tier ``eigh`` (numax/linalg/_array/eigen.mojo:316-392)
    over a batch of symmetric float32 matrices of order 1..64 packed back to
    back (``offsets_a`` / ``offsets_w``: element offsets of each matrix and of
    its eigenvalues).  Returns ``(w, v)``: unsorted eigenvalues and the
    eigenvectors as columns (``v`` None unless ``vectors``)."""
    A = _device_tensor(a, torch.float32)
    count = len(n)
    N = _device_tensor(np.asarray(n, dtype=np.int32))
    OA = _device_tensor(np.asarray(offsets_a, dtype=np.int64))
    OW = _device_tensor(np.asarray(offsets_w, dtype=np.int64))
    w_len = int(np.max(np.asarray(offsets_w) + np.asarray(n))) if count else 1
    W = torch.zeros(max(w_len, 1), dtype=torch.float32, device=A.device)
    V = torch.zeros_like(A) if vectors else None
    _lib.check(_lib.load().nmx_batched_syevj_f32(_ctx(A).handle, count, _ptr(N), _ptr(OA), _ptr(OW), _ptr(A),
                                                 A.numel(), _ptr(W), W.numel(), _ptr(V) if vectors else None,
                                                 int(sweeps)), "eigh")
    return W, V


def lstsq_batched(a, b) -> torch.Tensor:
    """numax's ``lstsq`` (numax/linalg/qr.mojo:746-821, method "qr") for a
    batch of float32 problems: ``a`` (count, m, n) with 1 <= n <= 32 and m >= n,
    ``b`` (count, m); returns x (count, n)."""
    A = _device_tensor(a, torch.float32)
    B = _device_tensor(b, torch.float32)
    if A.dim() != 3 or B.dim() != 2 or B.shape[0] != A.shape[0] or B.shape[1] != A.shape[1]:
        raise NumaxError(f"lstsq: shapes {tuple(A.shape)} and {tuple(B.shape)} do not match (count, m, n) / (count, m)")
    count, m, n = A.shape
    X = torch.zeros((count, n), dtype=torch.float32, device=A.device)
    _lib.check(_lib.load().nmx_batched_gels_f32(_ctx(A).handle, count, m, n, _ptr(A), _ptr(B), _ptr(X)), "lstsq")
    return X


# ---- register-tier operations over batches of small matrices ---------------

def _small_batch(a, n, offsets_a):
    A = _device_tensor(a, torch.float64)
    count = len(n)
    N = _device_tensor(np.asarray(n, dtype=np.int32))
    OA = _device_tensor(np.asarray(offsets_a, dtype=np.int64))
    return A, count, N, OA


def batched_lu_nopiv(a, n, offsets_a) -> torch.Tensor:
    """numax's Array-tier ``lu`` (_array/lu.mojo:25-55) over a batch of float64
    matrices of order 1..64 packed back to back (``offsets_a``: element offset
    of each): unpivoted Doolittle, packed L\\U at the same offsets."""
    A, count, N, OA = _small_batch(a, n, offsets_a)
    out = A.clone()
    _lib.check(_lib.load().nmx_batched_lu_nopiv_f64(_ctx(A).handle, count, _ptr(N), _ptr(OA), _ptr(A), A.numel(),
                                                     _ptr(out)), "lu")
    return out


def batched_solve_nopiv(a, b, n, offsets_a, offsets_b) -> torch.Tensor:
    """numax's Array-tier ``solve`` (_array/basic.mojo:20-31): unpivoted LU and
    two substitutions per problem; ``x_p`` at ``offsets_b[p]`` of the result."""
    A, count, N, OA = _small_batch(a, n, offsets_a)
    B = _device_tensor(b, torch.float64)
    OB = _device_tensor(np.asarray(offsets_b, dtype=np.int64))
    X = torch.zeros_like(B)
    _lib.check(_lib.load().nmx_batched_solve_nopiv_f64(_ctx(A).handle, count, _ptr(N), _ptr(OA), _ptr(OB), _ptr(A),
                                                        A.numel(), _ptr(B), B.numel(), _ptr(X)), "solve")
    return X


def batched_inverse(a, n, offsets_a) -> torch.Tensor:
    """numax's Array-tier ``inverse`` (_array/basic.mojo:34-60), unpivoted,
    each inverse at its matrix's offset."""
    A, count, N, OA = _small_batch(a, n, offsets_a)
    out = torch.zeros_like(A)
    _lib.check(_lib.load().nmx_batched_inverse_nopiv_f64(_ctx(A).handle, count, _ptr(N), _ptr(OA), _ptr(A),
                                                          A.numel(), _ptr(out)), "inverse")
    return out


def batched_det(a, n, offsets_a, pivoted: bool = False) -> torch.Tensor:
    """numax's Array-tier ``det`` (_array/lu.mojo:219-233, unpivoted) or, with
    ``pivoted``, ``lu_factor(a).det()`` (:112-123), one value per matrix."""
    A, count, N, OA = _small_batch(a, n, offsets_a)
    out = torch.zeros(max(count, 1), dtype=torch.float64, device=A.device)
    _lib.check(_lib.load().nmx_batched_det_f64(_ctx(A).handle, count, _ptr(N), _ptr(OA), _ptr(A), A.numel(),
                                                int(bool(pivoted)), _ptr(out)), "det")
    return out[:count]


def batched_slogdet(a, n, offsets_a):
    """numax's Array-tier ``slogdet`` (_array/lu.mojo:236-258, through
    ``PivotedLU.slogdet``, :125-162): ``(sign, logabsdet)`` per matrix, sign 0
    and -inf for an exactly singular factor."""
    A, count, N, OA = _small_batch(a, n, offsets_a)
    sign = torch.zeros(max(count, 1), dtype=torch.float64, device=A.device)
    logabs = torch.zeros_like(sign)
    _lib.check(_lib.load().nmx_batched_slogdet_f64(_ctx(A).handle, count, _ptr(N), _ptr(OA), _ptr(A), A.numel(),
                                                    _ptr(sign), _ptr(logabs)), "slogdet")
    return sign[:count], logabs[:count]


def batched_slogdet_cholesky(lower, n, offsets_l) -> torch.Tensor:
    """numax's ``slogdet_cholesky`` (_array/cholesky.mojo:95-116) of a batch of
    Cholesky factors: ``2 sum ln L_ii`` per factor."""
    L, count, N, OL = _small_batch(lower, n, offsets_l)
    out = torch.zeros(max(count, 1), dtype=torch.float64, device=L.device)
    _lib.check(_lib.load().nmx_batched_slogdet_cholesky_f64(_ctx(L).handle, count, _ptr(N), _ptr(OL), _ptr(L),
                                                             L.numel(), _ptr(out)), "slogdet_cholesky")
    return out[:count]


MAP_BLOCKS_KINDS = {"lu": 0, "cholesky": 1, "nopivot": 2}


def map_blocks_solve(xs, n: int, kind: str = "lu") -> torch.Tensor:
    """A map_blocks batched solve (numax/core/functional.mojo:691-834) over its
    structure-of-arrays layout: ``xs`` is ``(n*n + n, batch)`` -- rows
    ``0 .. n*n-1`` hold every problem's row-major ``A``, rows ``n*n ..`` its
    ``b`` -- and the result is ``(n, batch)``.  ``kind`` picks the step body:
    ``"lu"`` (lu_factor + PivotedLU.solve), ``"cholesky"`` (cholesky +
    cholesky_solve) or ``"nopivot"`` (Array ``solve``).  float32 or float64."""
    X = _device_tensor(xs)
    if kind not in MAP_BLOCKS_KINDS:
        raise NumaxError(f"map_blocks: kind {kind!r} is not one of {sorted(MAP_BLOCKS_KINDS)}")
    if X.dim() != 2 or X.shape[0] != n * n + n:
        raise NumaxError(f"map_blocks: xs must be (n*n + n, batch) = ({n * n + n}, batch), got {tuple(X.shape)}")
    batch = int(X.shape[1])
    Y = torch.zeros((n, batch), dtype=X.dtype, device=X.device)
    _lib.check(_fn("map_blocks_solve", X.dtype)(_ctx(X).handle, MAP_BLOCKS_KINDS[kind], n, batch, _ptr(X), _ptr(Y)),
               "map_blocks")
    return Y
