"""GPU: mixed-precision LU with iterative refinement (no reference
counterpart).  Pinned by the FP64 oracle's backward error: every result must
reach the target normwise backward error ||b - A x|| / (||A|| ||x||) <= 1e-15
and agree with the oracle's FP64 solution to what the condition number
allows."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from numax_amd import linalg as nl
from numax_amd import random as R

pytestmark = pytest.mark.gpu


def haar(seed, n):
    q, r = sl.qr(R.normal(seed, (n, n)))
    return q * np.sign(np.diag(r))


def svd_matrix(n, kappa, seed=3):
    u, v = haar(seed, n), haar(seed + 100, n)
    sigma = kappa ** (-np.arange(n) / (n - 1))
    return (u * sigma) @ v.T


def fp64_eta(a, b):
    lu, piv, _, _ = oracle.getrf(a)
    xo = oracle.getrs(lu, piv, b)
    return oracle.backward_error(a, xo, b), xo


@pytest.mark.parametrize("precision", ["fp16s", "fp16", "bf16", "fp32"])
def test_mpir_reaches_fp64_backward_error(precision):
    n = 1024
    a = svd_matrix(n, 1e2)
    b = R.normal(4, (n,))
    x, rep = nl.solve_mpir(a, b, precision=precision, max_iter=50)
    x = x.cpu().numpy()
    eta = oracle.backward_error(a, x, b)
    eta_o, xo = fp64_eta(a, b)
    assert rep["converged"] and eta <= max(1e-15, eta_o), (rep, eta, eta_o)
    assert np.max(np.abs(x - xo)) / np.max(np.abs(xo)) < 1e2 * 1e-14


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_mpir_condition_numbers(kappa):
    n = 2048
    a = svd_matrix(n, kappa, seed=7)
    b = R.normal(8, (n,))
    x, rep = nl.solve_mpir(a, b, precision="fp16s")
    eta = oracle.backward_error(a, x.cpu().numpy(), b)
    eta_o, _ = fp64_eta(a, b)
    assert rep["converged"] and eta <= max(1e-15, eta_o), (kappa, rep, eta, eta_o)


@pytest.mark.parametrize("n", [5, 64, 257, 1001, 1030])
def test_mpir_odd_sizes(n):
    """n not a multiple of 4 (butterfly tail), smaller than a block, ragged
    last block."""
    a = R.normal(20 + n, (n, n)) + n ** 0.5 * np.eye(n)
    b = R.normal(30 + n, (n,))
    x, rep = nl.solve_mpir(a, b, precision="fp16s")
    eta = oracle.backward_error(a, x.cpu().numpy(), b)
    assert rep["converged"] and eta <= 1e-15, (n, rep, eta)


@pytest.mark.parametrize("precision", ["fp16s", "fp32"])
def test_mpir_butterfly_mode_small(precision):
    """The opt-in unpivoted mode (random butterflies) on a diagonally strong
    matrix, where no-pivot elimination is stable."""
    n = 700
    a = R.normal(40, (n, n)) + 4 * n ** 0.5 * np.eye(n)
    b = R.normal(41, (n,))
    x, rep = nl.solve_mpir(a, b, precision=precision, pivot="butterfly")
    assert rep["converged"] and oracle.backward_error(a, x.cpu().numpy(), b) <= 1e-15


def test_mpir_matrix_rhs_and_errors():
    n = 300
    a = svd_matrix(n, 1e3, seed=11)
    bm = R.normal(12, (n, 3))
    x, reps = nl.solve_mpir(a, bm)
    assert x.shape == (n, 3) and all(r["converged"] for r in reps)
    with pytest.raises(nl.NumaxError, match="entries"):
        nl.solve_mpir(a, np.ones(n - 1))
    with pytest.raises(nl.NumaxError, match="precision"):
        nl.solve_mpir(a, bm[:, 0], precision="fp8")


def test_gmres_ir_converges_where_classical_refinement_fails():
    """Plain FP16 factor at kappa = 1e4 (kappa u_f ~ 5): classical steps do
    not contract, GMRES-IR (FP64 GMRES preconditioned by the factor) reaches
    FP64 backward error."""
    n = 1024
    a = svd_matrix(n, 1e4, seed=11)
    b = R.normal(12, (n,))
    _, rep_c = nl.solve_mpir(a, b, precision="fp16", solver="classical", max_iter=20, raise_on_failure=False)
    assert not rep_c["converged"]
    x, rep = nl.solve_mpir(a, b, precision="fp16", solver="gmres", max_iter=20)
    eta = oracle.backward_error(a, x.cpu().numpy(), b)
    eta_o, xo = fp64_eta(a, b)
    assert rep["converged"] and eta <= max(1e-15, 2 * eta_o), (rep, eta, eta_o)
    assert rep["gmres_iterations"] > 0
    assert np.max(np.abs(x.cpu().numpy() - xo)) / np.max(np.abs(xo)) < 1e-10


@pytest.mark.parametrize("kappa", [1e4, 1e6])
def test_gmres_ir_split_fp16(kappa):
    n = 1024
    a = svd_matrix(n, kappa, seed=11)
    b = R.normal(12, (n,))
    x, rep = nl.solve_mpir(a, b, precision="fp16s", solver="gmres", max_iter=20)
    eta = oracle.backward_error(a, x.cpu().numpy(), b)
    eta_o, _ = fp64_eta(a, b)
    assert rep["converged"] and eta <= max(1e-15, 2 * eta_o), (rep, eta, eta_o)
    assert rep["gmres_iterations"] > 0


def test_auto_switches_to_gmres_when_classical_stagnates():
    n = 1024
    a = svd_matrix(n, 1e4, seed=11)
    b = R.normal(12, (n,))
    x, rep = nl.solve_mpir(a, b, precision="fp16", solver="auto", max_iter=20)
    eta = oracle.backward_error(a, x.cpu().numpy(), b)
    assert rep["converged"] and eta <= 1e-15, rep
    assert rep["gmres_iterations"] > 0
    # where classical steps contract, auto stays classical
    _, rep2 = nl.solve_mpir(a, b, precision="fp16s", solver="auto")
    assert rep2["converged"] and rep2["gmres_iterations"] == 0


@pytest.mark.parametrize("n", [7, 257, 1030])
def test_gmres_ir_odd_sizes(n):
    a = svd_matrix(n, 1e3, seed=n)
    b = R.normal(n + 1, (n,))
    x, rep = nl.solve_mpir(a, b, precision="fp16", solver="gmres", gmres_restart=8, max_iter=30)
    eta = oracle.backward_error(a, x.cpu().numpy(), b)
    assert rep["converged"] and eta <= 1e-15, (rep, eta)


@pytest.mark.parametrize("nrhs,kappa,pivot", [(3, 1e4, "partial"), (64, 1e6, "partial"), (5, 1e2, "butterfly"),
                                              (20, 1e4, "tournament"), (100, 1e2, "tournament")])
def test_mpir_block_of_right_hand_sides(nrhs, kappa, pivot):
    """One factorization, every column refined to FP64 backward error; each
    column agrees with its single-right-hand-side solve to what kappa allows.
    The block solves take up to 64 right-hand sides per workgroup in groups
    of 16: 3, 5, 20 (a partial second group), 64 and 100 (two workgroups per
    block, the second partial) cover every shape."""
    n = 512 if pivot == "butterfly" else 1001
    a = svd_matrix(n, kappa, seed=5)
    b = R.normal(6, (n, nrhs))
    x, reps = nl.solve_mpir(a, b, precision="fp16s", pivot=pivot)
    x = x.cpu().numpy()
    assert len(reps) == nrhs
    for j in (0, nrhs // 2, nrhs - 1):
        eta = oracle.backward_error(a, x[:, j], b[:, j])
        assert reps[j]["converged"] and eta <= 1e-15, (j, reps[j], eta)
        xj, _ = nl.solve_mpir(a, b[:, j], precision="fp16s", pivot=pivot)
        xj = xj.cpu().numpy()
        assert np.max(np.abs(x[:, j] - xj)) / np.max(np.abs(xj)) < 10 * kappa * 1e-15
    assert max(r["backward_error"] for r in reps) <= 1e-15


def test_autotuners_pick_from_their_candidates(tmp_path, monkeypatch):
    from numax_amd import autotune
    monkeypatch.setenv("NUMAX_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    block = autotune.tune_mpir(2048)
    assert block in (128, 256, 512)
    assert autotune.cached("mpir", 0, n=2048, precision="fp16s") == block
    teams = autotune.tune_batched(samples=300)
    assert len(teams) == 3 and all(t in (1, 2, 3, 4, 6, 8) for t in teams)
    assert autotune.cached("batched", 0) == teams
    # the tuned shapes are the ones the solvers now use, and still solve
    a = svd_matrix(2048, 1e4, seed=1)
    b = R.normal(2, (2048,))
    x, rep = nl.solve_mpir(a, b)
    assert rep["converged"] and oracle.backward_error(a, x.cpu().numpy(), b) <= 1e-15


@pytest.mark.parametrize("pivot,kappa", [("tournament", 1e4), ("butterfly", 1e2)])
def test_mpir_graph_replay_matches_eager_bit_for_bit(pivot, kappa):
    """nmx_gesv_mpir with classical refinement captured whole (device-side
    stop test, max_iter + 1 steps in the graph): every replay gives the eager
    solve's x bit for bit and its report, with other eager solves on the same
    context between replays, and a new b written into the captured buffer is
    solved as an eager call solves it."""
    import torch
    n = 1024
    a = svd_matrix(n, kappa, seed=21)
    b = R.normal(22, (n,))
    g = nl.solve_mpir_graph(a, b, max_iter=30, block=256, pivot=pivot)
    x_e = g.x.clone()
    rep_e = g.eager_report
    assert rep_e["converged"] and 0 < rep_e["iterations"] < 30
    x_ref, rep_ref = nl.solve_mpir(a, b, solver="classical", block=256, pivot=pivot)
    assert torch.equal(x_ref, x_e) and rep_ref["iterations"] == rep_e["iterations"]
    for k in range(3):
        g.x.zero_()
        x, rep = g.replay()
        assert torch.equal(x, x_e), k
        assert rep["converged"] and rep["iterations"] == rep_e["iterations"], (k, rep, rep_e)
        assert rep["backward_error"] == rep_e["backward_error"]
        assert rep["history"] == rep_e["history"]
        nl.solve_mpir(a, 2.0 * b, solver="classical", block=256, pivot=pivot)  # eager work between replays
    b2 = R.normal(23, (n,))
    g.b.copy_(torch.as_tensor(b2, device=g.b.device))
    x, rep = g.replay()
    x2, rep2 = nl.solve_mpir(a, b2, solver="classical", block=256, pivot=pivot)
    assert torch.equal(x, x2) and rep["iterations"] == rep2["iterations"]
    assert oracle.backward_error(a, x.cpu().numpy(), b2) <= 1e-15
    g.close()


@pytest.mark.parametrize("solver,pivot", [(2, 2), (1, 1)])
def test_mpir_capture_refuses_gmres_and_partial_pivoting(solver, pivot):
    """GMRES-IR decides on the host and the partial-pivoting panel takes
    host-side epochs: neither is captured."""
    import ctypes

    import torch
    from numax_amd import _lib
    n = 256
    A = torch.as_tensor(svd_matrix(n, 1e2, seed=24), device="cuda")
    b = torch.as_tensor(R.normal(25, (n,)), device="cuda")
    x = torch.empty_like(b)
    lib = _lib.load()
    cfg = _lib.IrCfg()
    lib.nmx_ir_cfg_default(ctypes.byref(cfg))
    cfg.solver = solver
    cfg.pivot = pivot
    ctx = nl._ctx(A)
    nl.solve_mpir(A, b, pivot="partial" if pivot == 1 else "tournament")  # buffers sized
    torch.cuda.synchronize()
    _lib.check(lib.nmx_graph_begin(ctx.handle), "begin")
    st = lib.nmx_gesv_mpir(ctx.handle, n, nl._ptr(A), n, nl._ptr(b), nl._ptr(x), ctypes.byref(cfg), None)
    h = ctypes.c_void_p()
    _lib.check(lib.nmx_graph_end(ctx.handle, ctypes.byref(h)), "end")
    lib.nmx_graph_destroy(h)
    assert st == _lib.NMX_ERR_UNSUPPORTED
