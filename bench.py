"""Benchmark: dense LU solve on one MI355X at n = 16384.

Metric: solves/s + TFLOP/s at FP64 backward error, mixed-precision LU-IR vs
plain FP64 LU, n = 16384.

A step is one solve of A x = b from scratch: copy A into the work matrix
(numax's lu_factor copies its input too), factor, solve, and -- for the
mixed-precision path -- refine to FP64 backward error.  Inputs are resident
in HBM before timing starts.  The workload matrix is A = U diag(sigma) V^T
with U, V Haar-orthogonal (QR of seeded N(0,1) matrices, sign-fixed) and
sigma geometric from 1 to 1/kappa.

Printed: one JSON line (rank 0).  `value` is the headline path's solves/s;
the plain FP64 LU is reported beside it, with TFLOP/s counted as
2n^3/3 + 2 n^2 nrhs for both so they compare.  `roofline` is the dominant
kernel's achieved rate from per-launch HIP events on the library's stream;
`cpu_baseline` is the CPU oracle (numax's float64 host algorithm, restated)
timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm_GBs": 8000.0, "fp64_TFs": 78.6, "fp32_TFs": 157.3, "fp16_TFs": 2500.0, "bf16_TFs": 2500.0}
# what bounds each kernel class of the library's profiler
ROOF = {"gemm": ("mfma", "fp64_TFs"), "mp_gemm": ("hbm", None), "getf2_leaf": ("hbm", None),
        "laswp": ("hbm", None), "trsm_base": ("hbm", None), "residual": ("hbm", None), "mp_trsv": ("hbm", None),
        "mp_panel": ("hbm", None), "potrf_diag": ("hbm", None), "other": ("hbm", None)}


# the library's profiler classes are timed per class; the kernels each class
# holds on the paths the roofline objects name
CLASS_KERNELS = {("mp-ir", "getf2_leaf"): "tslu_tournament_kernel (tournament pivot selection of a 32-column leaf)",
                 ("plain-fp64", "gemm"): "gemm_kernel<double> (trailing and in-panel updates)"}


# HBM bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, both factors from a
# 1 GiB copy kernel) of the kernel classes a roofline object names, measured by
# rocprofv3 --pmc on numax_amd/lib/nmx_profile at this benchmark's sizes:
# n = 16384 for the LU paths, the 20,000-system batch for the batched leg.
# Keyed by (kernel class, n or batch count); another size reports null.
MEASURED_TRAFFIC = {
    # mixed-precision path: the tournament (leaf class), averaged over the factorization's leaves
    ("getf2_leaf", 16384): (2.16 + 0.01) * 2 ** 20,
    ("mp_gemm", 16384): (249.53 + 170.69) * 2 ** 20,  # next-panel and look-ahead parts, averaged
    # plain FP64 path: the 128-tile trailing updates and the 64-tile in-panel updates, per class launch
    ("gemm", 16384): 108.68 * 2 ** 20,
    ("batched_solve", 20000): (34144.7 + 7892.0) * 2 ** 20,  # the eight class launches of one call
    ("c2_lu_gemm", 8192): 28.04 * 2 ** 20,  # per class launch, FP64 LU + 16-RHS solve
    ("batched_eigh", 65536): (405.22 + 408.13 + 103.09 + 104.12) * 2 ** 20,  # both class launches
    ("batched_lstsq", 4096): (669.27 + 0.50) * 2 ** 20,
}


def profile_step(ctx, step) -> dict:
    """Per-kernel-class HIP-event times of one step on the library's stream."""
    ctx.set_profiling(True)
    step()
    prof = ctx.profile()
    ctx.set_profiling(False)
    log("profile: " + json.dumps({k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in prof.items()}))
    return prof


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def c3_matrix(n: int, kappa: float, torch, seed: int = 3):
    from numax_amd import random as R
    return R.device_svd_matrix(n, kappa, seed=seed)


def backward_error(a, x, b) -> float:
    r = b - a @ x
    if r.dim() == 1:
        r, x2 = r[:, None], x[:, None]
    else:
        x2 = x
    anorm = a.abs().sum(1).max()
    eta = r.abs().max(0).values / (anorm * x2.abs().max(0).values)
    return float(eta.max().item())


def cpu_baseline(n_sample: int, threads: int) -> dict:
    """The oracle (numax's float64 host LU: block 32, base 16, threaded
    trsm / GEMM) on an n_sample LU solve, scaled to n = 16384 by flops."""
    import numpy as np

    import oracle
    from numax_amd import random as R
    oracle.set_threads(threads)
    a = R.lu_matrix(5, n_sample)
    b = R.normal(5, (n_sample,))
    t0 = time.perf_counter()
    lu, piv, _, _ = oracle.getrf(a)
    x = oracle.getrs(lu, piv, b)
    dt = time.perf_counter() - t0
    flops = 2 * n_sample ** 3 / 3 + 2 * n_sample ** 2
    rate = flops / dt
    full = 2 * 16384 ** 3 / 3 + 2 * 16384 ** 2
    return {"value": rate / full, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"oracle LU factor+solve, n={n_sample}, f64, {dt:.2f} s -> {rate / 1e9:.1f} GFLOP/s, "
                      f"scaled to n=16384 by flop count",
            "gflops": rate / 1e9, "finite": bool(np.isfinite(x).all())}


def batched_leg(args, torch) -> dict:
    """Times the uneven batch (one nmx_batched_gesv per step) and reports its
    HBM roofline: algorithmic bytes = sum(8 n^2 + 16 n) (A and b read once,
    x written once, FP64)."""
    import ctypes

    import numpy as np

    from numax_amd import _lib
    from numax_amd import linalg as nl
    from numax_amd import random as R
    a, b, n, kind, offa, offb = R.device_uneven_batch(args.batch_count, 4)
    lib = _lib.load()
    ctx = nl._ctx(a)
    N = torch.as_tensor(n, device="cuda")
    K = torch.as_tensor(kind, device="cuda")
    OA = torch.as_tensor(offa, device="cuda")
    OB = torch.as_tensor(offb, device="cuda")
    X = torch.zeros_like(b)
    info = torch.zeros(len(n), dtype=torch.int32, device="cuda")
    iters = torch.zeros_like(info)
    berr = torch.zeros(len(n), dtype=torch.float64, device="cuda")
    cfg = _lib.BatchCfg()
    lib.nmx_batch_cfg_default(ctypes.byref(cfg))
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def solve(c):
        _lib.check(lib.nmx_batched_gesv(ctx.handle, len(n), P(N), P(OA), P(OB), P(K), P(a), a.numel(), P(b),
                                        b.numel(), P(X), P(info), P(iters), P(berr), ctypes.byref(c)),
                   "nmx_batched_gesv")

    teams = None
    if not args.no_autotune:
        # tuned on this batch, before the timed steps (the size classes share
        # the chip, so a per-class synthetic sweep does not carry over)
        from numax_amd import autotune
        teams = autotune.tune_batched_on(solve)
        for i, v in enumerate(teams):
            cfg.teams_per_cu[i] = v
        log(f"autotuned teams_per_cu {teams}")

    def step():
        solve(cfg)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    prof = profile_step(ctx, step)
    nn = n.astype(np.float64)
    bytes_alg = float((8 * nn ** 2 + 16 * nn).sum())
    flops = float(np.where(kind == 0, 2 * nn ** 3 / 3, nn ** 3 / 3).sum())
    be = berr.cpu().numpy()
    inf = info.cpu().numpy()
    it = iters.cpu().numpy()
    rec = prof.get("batched_solve", {"ms": dt * 1e3, "launches": 1})
    kern_s = rec["ms"] * 1e-3
    out = {"workload": f"{len(n)} uneven systems, n ~ U{{8..256}}, half LU (U[-1,1] + 0.1 n I), "
                       f"half Cholesky (G G^T / n + I), FP32 factor + FP64 refinement to {cfg.tol:g}",
           "ms_per_step": dt * 1e3, "solves_per_s": len(n) / dt, "gflops": flops / dt / 1e9,
           "roofline": {"bound": "hbm", "achieved": bytes_alg / kern_s / 1e9, "peak": PEAK["hbm_GBs"],
                        "unit": "GB/s", "frac": bytes_alg / kern_s / 1e9 / PEAK["hbm_GBs"],
                        "traffic": MEASURED_TRAFFIC.get(("batched_solve", len(n))),
                        "bytes_per_step": bytes_alg, "kernel_ms": rec["ms"]},
           "teams_per_cu": [int(cfg.teams_per_cu[i]) for i in range(3)], "autotuned": teams is not None,
           "max_backward_error": float(np.nanmax(be)), "max_iterations": int(it.max()),
           "failed": int((inf != 0).sum()), "kernel_profile": prof}
    if not args.no_cpu:
        import oracle
        threads, vis = host_threads()
        oracle.set_threads(threads)
        sa, sb, sn, skind, soa, sob = R.device_uneven_batch(args.batch_cpu_sample, 4)
        sa, sb = sa.cpu().numpy(), sb.cpu().numpy()
        t0 = time.perf_counter()
        oracle.array_gesv_batch(sa, sb, sn, skind, soa, sob)
        dts = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(sn) / dts, "unit": "systems/s", "cores": threads, "cores_visible": vis,
                               "kind": "port",
                               "sample": f"oracle register-tier lu_factor + PivotedLU.solve / cholesky + "
                                         f"cholesky_solve (FP64, one system per thread), {len(sn)} systems of the "
                                         f"same distribution, {dts:.2f} s"}
    log("batched: " + json.dumps({k: v for k, v in out.items() if k != "kernel_profile"}))
    return out


def host_threads() -> tuple[int, int]:
    """(threads used, cores visible): the CPU baselines use every core this
    process may run on, at most 16 (one GPU's share of an eight-GPU node's host cores)."""
    vis = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return min(vis, 16), vis


def timed_steps(step, steps, warmup, torch):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def class_roofline(prof, name, bound, flops=None, nbytes=None, peak_key=None, traffic=None) -> dict:
    """Roofline object of one kernel class from the library's per-launch HIP
    events: `flops` / `nbytes` are the algorithmic work of the whole step
    (None: the profiler's own per-launch counts)."""
    rec = prof[name]
    sec = rec["ms"] * 1e-3
    if bound == "mfma" or bound == "valu":
        work = rec["flops"] if flops is None else flops
        ach = work / sec / 1e12
        peak = PEAK[peak_key]
        unit = "TFLOP/s"
    else:
        work = rec["bytes"] if nbytes is None else nbytes
        ach = work / sec / 1e9
        peak = PEAK["hbm_GBs"]
        unit = "GB/s"
    return {"kernel": name, "bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
            "traffic": traffic, "launches_per_step": rec["launches"], "kernel_ms": rec["ms"],
            "avg_launch_ms": rec["ms"] / max(rec["launches"], 1)}


def c2_leg(args, torch) -> dict:
    """Plain FP64 GPU LU (A ~ N(0,1)) and Cholesky (A = G G^T / n + I),
    n = 8192, 16 right-hand sides ~ N(0,1), block 256: factor + solve per
    step, the FP64 MFMA trailing update's roofline, the oracle beside it."""
    import ctypes

    import numpy as np

    from numax_amd import _lib
    from numax_amd import linalg as nl
    from numax_amd import random as R
    n, k = args.c2_n, 16
    lib = _lib.load()
    a_lu = R.device_normal(21, (n, n))
    g = R.device_normal(22, (n, n))
    a_ch = g @ g.T / n + torch.eye(n, dtype=torch.float64, device="cuda")
    del g
    bm = R.device_normal(23, (n, k))
    work = torch.empty_like(a_lu)
    x = torch.empty_like(bm)
    piv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx = nl._ctx(work)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def lu_step():
        work.copy_(a_lu)
        _lib.check(lib.nmx_getrf_f64(ctx.handle, n, P(work), n, P(piv), P(info), 256, 32), "getrf")
        x.copy_(bm)
        _lib.check(lib.nmx_getrs_f64(ctx.handle, n, k, P(work), n, P(piv), P(x), k), "getrs")

    def chol_step():
        work.copy_(a_ch)
        _lib.check(lib.nmx_potrf_f64(ctx.handle, n, P(work), n, P(info), 256), "potrf")
        x.copy_(bm)
        _lib.check(lib.nmx_potrs_f64(ctx.handle, n, k, P(work), n, P(x), k), "potrs")

    out = {}
    threads, vis = host_threads()
    for name, step, a, fl, cpu_fn in (("lu", lu_step, a_lu, 2 * n ** 3 / 3 + 2 * n * n * k, "lu"),
                                      ("cholesky", chol_step, a_ch, n ** 3 / 3 + 2 * n * n * k, "chol")):
        dt = timed_steps(step, args.steps, 1, torch)
        eta = backward_error(a, x, bm)
        prof = profile_step(ctx, step)
        roof = class_roofline(prof, "gemm", "mfma", peak_key="fp64_TFs",
                              traffic=MEASURED_TRAFFIC.get(("c2_lu_gemm", n)) if name == "lu" else None)
        roof["share_of_step"] = prof["gemm"]["ms"] / (dt * 1e3)
        cpu = None
        if not args.no_cpu:
            cpu = c2_cpu_baseline(cpu_fn, args.c2_cpu_sample, n, k, threads, vis)
        out[name] = {"workload": f"FP64 {'LU with partial pivoting' if name == 'lu' else 'Cholesky'} factor + "
                                 f"solve, n={n}, {k} RHS, block 256",
                     "ms_per_step": dt * 1e3, "solves_per_s": k / dt, "tflops": fl / dt / 1e12,
                     "backward_error": eta, "roofline": roof, "cpu_baseline": cpu, "kernel_profile": prof}
        log(f"c2 {name}: {dt * 1e3:.2f} ms, {fl / dt / 1e12:.2f} TFLOP/s, eta {eta:.2e}, gemm "
            f"{roof['achieved']:.1f} TFLOP/s")
    del a_lu, a_ch, work
    torch.cuda.empty_cache()
    return out


def c2_cpu_baseline(kind: str, n_sample: int, n: int, k: int, threads: int, vis: int) -> dict:
    """The oracle (numax's host LU, block 32 / base 16, or Cholesky, block
    64 / tile 128, with its threaded trsm / GEMM) on an n_sample problem,
    scaled to n by flop count."""
    import numpy as np

    import oracle
    from numax_amd import random as R
    oracle.set_threads(threads)
    b = R.normal(7, (n_sample, k))
    if kind == "lu":
        a = R.normal(8, (n_sample, n_sample))
        t0 = time.perf_counter()
        lu, piv, _, _ = oracle.getrf(a)
        oracle.getrs(lu, piv, b)
        dt = time.perf_counter() - t0
        fl_s, fl = 2 * n_sample ** 3 / 3 + 2 * n_sample ** 2 * k, 2 * n ** 3 / 3 + 2 * n * n * k
    else:
        a = R.spd_matrix(9, n_sample)
        t0 = time.perf_counter()
        l, _ = oracle.potrf(a)
        oracle.potrs(l, b)
        dt = time.perf_counter() - t0
        fl_s, fl = n_sample ** 3 / 3 + 2 * n_sample ** 2 * k, n ** 3 / 3 + 2 * n * n * k
    rate = fl_s / dt
    return {"value": rate / fl * k, "unit": "solves/s", "cores": threads, "cores_visible": vis, "kind": "port",
            "sample": f"oracle {kind} factor + solve, n={n_sample}, {k} RHS, f64: {dt:.2f} s -> "
                      f"{rate / 1e9:.1f} GFLOP/s, scaled to n={n} by flop count"}


def c5_legs(args, torch) -> dict:
    """Batched Jacobi eigh (65,536 matrices, n in {24, 32, 48, 64}) and
    batched tall-skinny least squares (4096 problems 1024 x 32), FP32."""
    import ctypes

    import numpy as np

    import oracle
    from numax_amd import _lib
    from numax_amd import linalg as nl
    from numax_amd import random as R
    lib = _lib.load()
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    threads, vis = host_threads()
    out = {}
    # ---- eigh
    a, n, oa, ow = R.eigh_batch(args.eigh_count, 5)
    A = torch.as_tensor(a, device="cuda")
    N = torch.as_tensor(n, device="cuda")
    OA = torch.as_tensor(oa, device="cuda")
    OW = torch.as_tensor(ow, device="cuda")
    W = torch.zeros(int((ow + n).max()), dtype=torch.float32, device="cuda")
    V = torch.zeros_like(A)
    ctx = nl._ctx(A)

    def eigh_step():
        _lib.check(lib.nmx_batched_syevj_f32(ctx.handle, len(n), P(N), P(OA), P(OW), P(A), A.numel(), P(W), W.numel(),
                                             P(V), 12), "eigh")

    dt = timed_steps(eigh_step, args.steps, 1, torch)
    prof = profile_step(ctx, eigh_step)
    nn = n.astype(np.float64)
    flops = float((12 * nn * (nn - 1) / 2 * (18 * nn + 15)).sum())  # numax's rotation count x flops
    roof = class_roofline(prof, "batched_eigh", "valu", flops=flops, peak_key="fp32_TFs",
                          traffic=MEASURED_TRAFFIC.get(("batched_eigh", len(n))))
    cpu = None
    if not args.no_cpu:
        oracle.set_threads(threads)
        sample = np.concatenate([np.flatnonzero(n == s)[:args.eigh_cpu_sample // 4] for s in (24, 32, 48, 64)])
        sa = np.concatenate([a[oa[p]:oa[p] + n[p] * n[p]] for p in sample]).astype(np.float64)
        sn = n[sample]
        soa = np.zeros(len(sample), dtype=np.int64)
        soa[1:] = np.cumsum(sn.astype(np.int64) ** 2)[:-1]
        sow = np.zeros(len(sample), dtype=np.int64)
        sow[1:] = np.cumsum(sn)[:-1]
        t0 = time.perf_counter()
        oracle.array_eigh_batch(sa, sn, soa, sow)
        dts = time.perf_counter() - t0
        cpu = {"value": len(sample) / dts, "unit": "matrices/s", "cores": threads, "cores_visible": vis,
               "kind": "port", "sample": f"oracle register-tier eigh (12 cyclic Jacobi sweeps, FP64), "
                                         f"{len(sample)} of the matrices (equal shares of each n), {dts:.2f} s"}
    out["eigh"] = {"workload": f"{len(n)} symmetric eigenproblems, n in {{24, 32, 48, 64}} in equal shares, half "
                               "finite-difference Hamiltonians, half (G + G^T)/2; 12 Jacobi sweeps, FP32, vectors",
                   "ms_per_step": dt * 1e3, "matrices_per_s": len(n) / dt, "tflops": flops / dt / 1e12,
                   "roofline": roof, "cpu_baseline": cpu, "kernel_profile": prof}
    log(f"c5 eigh: {dt * 1e3:.2f} ms, {len(n) / dt:.3g} matrices/s, {roof['achieved']:.2f} TFLOP/s")
    del A, V
    # ---- lstsq
    m, k, cnt = 1024, 32, args.lstsq_count
    a, b = R.lstsq_batch(cnt, m, k, 6)
    A = torch.as_tensor(a, device="cuda")
    B = torch.as_tensor(b, device="cuda")
    X = torch.zeros((cnt, k), dtype=torch.float32, device="cuda")

    def lstsq_step():
        _lib.check(lib.nmx_batched_gels_f32(ctx.handle, cnt, m, k, P(A), P(B), P(X)), "gels")

    dt = timed_steps(lstsq_step, args.steps, 1, torch)
    prof = profile_step(ctx, lstsq_step)
    nbytes = 4.0 * (m * k + m + k) * cnt
    roof = class_roofline(prof, "batched_lstsq", "hbm", nbytes=nbytes, traffic=MEASURED_TRAFFIC.get(("batched_lstsq", cnt)))
    cpu = None
    if not args.no_cpu:
        s = args.lstsq_cpu_sample
        t0 = time.perf_counter()
        oracle.array_lstsq_batch(a[:s], b[:s])
        dts = time.perf_counter() - t0
        cpu = {"value": s / dts, "unit": "problems/s", "cores": threads, "cores_visible": vis, "kind": "port",
               "sample": f"oracle register-tier lstsq (Householder, FP64), {s} of the problems, {dts:.2f} s"}
    out["lstsq"] = {"workload": f"{cnt} least-squares problems, A {m} x {k} ~ N(0,1), b ~ N(0,1), FP32",
                    "ms_per_step": dt * 1e3, "problems_per_s": cnt / dt, "roofline": roof, "cpu_baseline": cpu,
                    "kernel_profile": prof}
    log(f"c5 lstsq: {dt * 1e3:.3f} ms, {cnt / dt:.3g} problems/s, {roof['achieved']:.0f} GB/s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e6)
    ap.add_argument("--precision", default="fp16s", choices=["fp16", "bf16", "fp16s", "fp32"])
    ap.add_argument("--tol", type=float, default=1e-15)
    ap.add_argument("--all-kappa", action="store_true", help="also time kappa = 1e2, 1e4, 1e6")
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the library's default launch shapes instead of tuning them first")
    ap.add_argument("--block-rhs", type=int, default=64,
                    help="also time one factor + refine of this many right-hand sides against plain FP64 (0: skip)")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-count", type=int, default=20000, help="systems in the uneven batch (0: skip it)")
    ap.add_argument("--pivot", default="tournament", choices=["tournament", "partial"],
                    help="how the mixed-precision factor pivots")
    ap.add_argument("--c2-n", type=int, default=8192, help="order of the plain FP64 LU / Cholesky leg (0: skip)")
    ap.add_argument("--c2-cpu-sample", type=int, default=4096)
    ap.add_argument("--eigh-count", type=int, default=65536)
    ap.add_argument("--eigh-cpu-sample", type=int, default=2048)
    ap.add_argument("--lstsq-count", type=int, default=4096)
    ap.add_argument("--lstsq-cpu-sample", type=int, default=512)
    ap.add_argument("--no-c5", action="store_true", help="skip the batched eigh / lstsq legs")
    ap.add_argument("--batch-cpu-sample", type=int, default=2000)
    ap.add_argument("--batched-only", action="store_true", help="time only the uneven batch (prints its object)")
    ap.add_argument("--matrix", default="haar", choices=["haar", "normal"],
                    help="haar: U diag(sigma) V^T (the workload); normal: N(0,1) entries (no QR in the setup, for "
                         "counter-collection runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the factor-and-solve path runs on a single MI355X "
              "(how it would shard across GPUs is not implemented); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from numax_amd import linalg as nl
    from numax_amd import random as R
    if args.batched_only:
        print(json.dumps(batched_leg(args, torch)), flush=True)
        return 0
    n, nrhs = args.n, args.nrhs
    dev = torch.device("cuda:0")
    log(f"generating A (n={n}, kappa={args.kappa:g}) on the GPU")
    t0 = time.perf_counter()
    a = c3_matrix(n, args.kappa, torch) if args.matrix == "haar" else R.device_normal(3, (n, n))
    b = R.device_normal(4, (n, nrhs)) if nrhs > 1 else R.device_normal(4, (n,))
    torch.cuda.synchronize()
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s")

    work = torch.empty_like(a)
    piv = torch.zeros(n, dtype=torch.int32, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    x = torch.empty_like(b)
    from numax_amd import _lib
    import ctypes
    lib = _lib.load()
    ctx = nl._ctx(work)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def plain_step():
        work.copy_(a)
        _lib.check(lib.nmx_getrf_f64(ctx.handle, n, P(work), n, P(piv), P(info), args.block, 32), "getrf")
        x.copy_(b)
        _lib.check(lib.nmx_getrs_f64(ctx.handle, n, nrhs, P(work), n, P(piv), P(x), max(nrhs, 1)), "getrs")

    def timed(step, steps, warmup, label):
        for i in range(warmup):
            step()
            torch.cuda.synchronize()
            log(f"{label} warmup {i + 1}/{warmup} done")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        log(f"{label}: {dt * 1e3:.2f} ms/step")
        return dt

    flops = 2 * n ** 3 / 3 + 2 * n ** 2 * nrhs
    dt_plain = timed(plain_step, args.steps, args.warmup, "plain fp64 LU")
    ctx.check_device()
    eta_plain = backward_error(a, x, b)
    log(f"plain fp64 backward error {eta_plain:.3e}")
    prof_plain = profile_step(ctx, plain_step)

    # ---- mixed-precision LU-IR on the same matrix, then at the other kappas
    cfg = _lib.IrCfg()
    lib.nmx_ir_cfg_default(ctypes.byref(cfg))
    cfg.precision = nl.PRECISIONS[args.precision]
    cfg.pivot = nl.PIVOTS[args.pivot]
    cfg.block = args.block
    if not args.no_autotune:
        lib_st = lib.nmx_ir_autotune(ctx.handle, n, ctypes.byref(cfg))
        _lib.check(lib_st, "nmx_ir_autotune")
        log(f"autotuned mixed-precision panel width {cfg.block}")
    mp_block = int(cfg.block)
    cfg.tol = args.tol
    rep = _lib.IrReport()
    xm = torch.empty(n, dtype=torch.float64, device=dev)
    bvec = b if nrhs == 1 else b[:, 0].contiguous()

    def mp_step():
        st = lib.nmx_gesv_mpir(ctx.handle, n, P(a), n, P(bvec), P(xm), ctypes.byref(cfg), ctypes.byref(rep))
        if st not in (0, 5):
            _lib.check(st, "nmx_gesv_mpir")

    results = {}
    kappas = [args.kappa] + [k for k in (1e2, 1e4, 1e6) if k != args.kappa and args.all_kappa]
    for kappa in kappas:
        if kappa != args.kappa:
            del a
            torch.cuda.empty_cache()
            a = c3_matrix(n, kappa, torch)
        dt_mp = timed(mp_step, args.steps, args.warmup, f"mp-ir {args.precision} kappa={kappa:g}")
        eta_mp = backward_error(a, xm, bvec)
        results[f"{kappa:g}"] = {"ms_per_step": dt_mp * 1e3, "solves_per_s": 1.0 / dt_mp,
                                 "effective_tflops": (2 * n ** 3 / 3 + 2 * n ** 2) / dt_mp / 1e12,
                                 "iterations": rep.iterations, "converged": bool(rep.converged),
                                 "backward_error": eta_mp, "factor_ms": rep.factor_ms, "refine_ms": rep.refine_ms,
                                 "history": [rep.history[i] for i in range(min(rep.iterations + 1, 32))]}
        log(f"mp-ir kappa={kappa:g}: {json.dumps(results[f'{kappa:g}'])}")
    head = results[f"{args.kappa:g}"]
    # per-kernel profile of one MP solve at the headline kappa's matrix
    if len(kappas) > 1:
        a = c3_matrix(n, args.kappa, torch)
    prof_mp = profile_step(ctx, mp_step)

    block_rhs = None
    if args.block_rhs > 1:
        k = args.block_rhs
        Bk = R.device_normal(5, (n, k))
        Xk = torch.empty_like(Bk)
        Xp = torch.empty_like(Bk)
        reps = (_lib.IrReport * k)()

        def mp_block_step():
            st = lib.nmx_gesv_mpir_nrhs(ctx.handle, n, k, P(a), n, P(Bk), k, P(Xk), k, ctypes.byref(cfg), reps)
            if st not in (0, 5):
                _lib.check(st, "nmx_gesv_mpir_nrhs")

        def plain_block_step():
            work.copy_(a)
            _lib.check(lib.nmx_getrf_f64(ctx.handle, n, P(work), n, P(piv), P(info), args.block, 32), "getrf")
            Xp.copy_(Bk)
            _lib.check(lib.nmx_getrs_f64(ctx.handle, n, k, P(work), n, P(piv), P(Xp), k), "getrs")

        bsteps = max(1, min(args.steps, 3))
        profile_step(ctx, mp_block_step)
        dt_pk = timed(plain_block_step, bsteps, 1, f"plain-fp64 nrhs={k}")
        dt_mk = timed(mp_block_step, bsteps, 1, f"mp-ir nrhs={k}")
        block_rhs = {"nrhs": k, "kappa": args.kappa,
                     "mp_ms_per_step": dt_mk * 1e3, "mp_solves_per_s": k / dt_mk,
                     "plain_ms_per_step": dt_pk * 1e3, "plain_solves_per_s": k / dt_pk,
                     "speedup_vs_plain_fp64": dt_pk / dt_mk,
                     "mp_backward_error_max": backward_error(a, Xk, Bk),
                     "plain_backward_error_max": backward_error(a, Xp, Bk),
                     "mp_iterations_max": max(reps[j].iterations for j in range(k)),
                     "all_converged": all(bool(reps[j].converged) for j in range(k))}
        log("block rhs: " + json.dumps(block_rhs))

    def roofline_of(prof, label):
        name, rec = max(prof.items(), key=lambda kv: kv[1]["ms"])
        bound = ROOF.get(name, ("hbm", None))
        sec = rec["ms"] * 1e-3
        if bound[0] == "mfma":
            ach = rec["flops"] / sec / 1e12
            peak = PEAK[bound[1]]
            unit = "TFLOP/s"
        else:
            ach = rec["bytes"] / sec / 1e9
            peak = PEAK["hbm_GBs"]
            unit = "GB/s"
        return {"kernel": name, "kernels": CLASS_KERNELS.get((label, name), name), "path": label, "bound": bound[0],
                "achieved": ach, "peak": peak, "unit": unit,
                "frac": ach / peak, "traffic": MEASURED_TRAFFIC.get((name, n)), "launches_per_step": rec["launches"],
                "avg_launch_ms": rec["ms"] / max(rec["launches"], 1), "share_of_step": None}

    roof = roofline_of(prof_mp, "mp-ir")
    roof["share_of_step"] = prof_mp[roof["kernel"]]["ms"] / head["ms_per_step"]
    roof_plain = roofline_of(prof_plain, "plain-fp64")
    roof_plain["share_of_step"] = prof_plain[roof_plain["kernel"]]["ms"] / (dt_plain * 1e3)

    batched = batched_leg(args, torch) if args.batch_count > 0 else None
    del a, work
    torch.cuda.empty_cache()
    c2 = c2_leg(args, torch) if args.c2_n > 0 else None
    c5 = None if args.no_c5 else c5_legs(args, torch)

    cpu = None
    if not args.no_cpu:
        threads, vis = host_threads()
        log(f"cpu baseline: oracle LU n={args.cpu_sample} on {threads} threads ({vis} visible)")
        cpu = cpu_baseline(args.cpu_sample, threads)
        cpu["cores_visible"] = vis

    value = 1.0 / (head["ms_per_step"] * 1e-3)
    out = {
        "metric": "solves/s + TFLOP/s at FP64 backward error: MP-IR vs plain FP64 LU, n=16384",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"mixed: {args.precision} MFMA update / f32 panel / f64 residual",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision LU-IR solve, n={n}, 1 RHS, A = U diag(sigma) V^T (Haar U, V), "
                               f"kappa={args.kappa:g}, {args.pivot} pivoting, {args.precision} trailing update, "
                               f"block {mp_block}, refined to FP64 backward error <= {args.tol:g}",
                   "n": n, "nrhs": 1, "kappa": args.kappa, "precision": args.precision, "pivot": args.pivot,
                   "block": mp_block,
                   "autotuned": not args.no_autotune},
        "tflops": head["effective_tflops"],
        "backward_error": head["backward_error"],
        "ir_iterations": head["iterations"],
        "speedup_vs_plain_fp64": dt_plain * 1e3 / head["ms_per_step"],
        "plain_fp64": {"ms_per_step": dt_plain * 1e3, "solves_per_s": nrhs / dt_plain,
                       "tflops": flops / dt_plain / 1e12, "backward_error": eta_plain, "nrhs": nrhs,
                       "roofline": roof_plain, "kernel_profile": prof_plain},
        "mp_by_kappa": results,
        "mp_block_rhs": block_rhs,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_uneven": batched,
        "c2_plain_fp64_n8192": c2,
        "c5_batched_eigh_lstsq": c5,
        "kernel_profile": prof_mp,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
