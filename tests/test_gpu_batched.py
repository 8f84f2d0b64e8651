"""GPU parity of the batched uneven-size solver (nmx_batched_gesv) against the
oracle's register-tier algorithms (numax/linalg/_array/lu.mojo:165-216,
numax/linalg/_array/cholesky.mojo:26-101) in float64.

Bar: the GPU factors in FP32 and refines in FP64 until the backward error is
<= 1e-15 (the default target), so every problem must report <= 1e-15 and
measure <= 2e-15 with the oracle's residual, and its solution must agree with the FP64
oracle solve to within cond(A) * 1e-15 (asserted as 1e-12 relative here on
well-conditioned batches).  info / skip semantics are checked exactly.
"""
import numpy as np
import pytest

import oracle
from numax_amd import linalg as nl
from tests.golden import kat

pytestmark = pytest.mark.gpu


def _systems(sizes, seed):
    """Half LU (A ~ U[-1,1] + 0.1 n I), half Cholesky (A = G G^T / n + I)."""
    rng = np.random.default_rng(seed)
    mats, rhss, kinds = [], [], []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            a = rng.uniform(-1.0, 1.0, (n, n)) + 0.1 * n * np.eye(n)
            kinds.append(0)
        else:
            g = rng.standard_normal((n, n))
            a = g @ g.T / n + np.eye(n)
            kinds.append(1)
        mats.append(a)
        rhss.append(rng.standard_normal(n))
    return mats, rhss, np.array(kinds, dtype=np.uint8)


def _oracle_x(a, b, kind):
    if kind == 0:
        lu, perm, _ = oracle.array_lu_factor(a)
        return oracle.array_lu_solve(lu, perm, b)
    return oracle.array_cholesky_solve(oracle.array_cholesky(a), b)


def _run(mats, rhss, kinds, **kw):
    a, b, n, oa, ob = nl.pack_systems(mats, rhss)
    res = nl.batched_solve(a, b, n, kinds, oa, ob, **kw)
    return res, res.info.cpu().numpy(), res.iters.cpu().numpy(), res.berr.cpu().numpy()


def _check(mats, rhss, kinds, res, info, berr):
    x_all = res.x.cpu().numpy()
    for p, (a, b, k) in enumerate(zip(mats, rhss, kinds)):
        o = int(res._offb[p])
        x = x_all[o:o + len(b)]
        assert info[p] == 0, f"problem {p} (n={len(b)}, kind={k}) info {info[p]}"
        assert berr[p] <= 1e-15, f"problem {p}: backward error {berr[p]}"  # the refinement target
        assert oracle.backward_error(a, x, b) <= 2e-15, f"problem {p}: oracle-measured backward error"
        xo = _oracle_x(a, b, k)
        assert np.max(np.abs(x - xo)) <= 1e-12 * max(1.0, np.max(np.abs(xo))), f"problem {p} (n={len(b)}) x"


@pytest.mark.parametrize("sizes", [
    [1, 2, 3, 7, 8, 16, 31, 32, 33, 63, 64],          # one wavefront per problem
    [65, 66, 100, 127, 128],                          # four wavefronts
    [129, 130, 200, 255, 256],                        # sixteen wavefronts
])
def test_batched_size_classes_match_oracle(sizes):
    mats, rhss, kinds = _systems(sizes + sizes[::-1], seed=len(sizes) * 7 + sizes[0])
    res, info, iters, berr = _run(mats, rhss, kinds)
    _check(mats, rhss, kinds, res, info, berr)


def test_batched_mixed_uneven_batch():
    rng = np.random.default_rng(2024)
    sizes = [int(s) for s in rng.integers(8, 257, size=300)]
    mats, rhss, kinds = _systems(sizes, seed=11)
    res, info, iters, berr = _run(mats, rhss, kinds)
    _check(mats, rhss, kinds, res, info, berr)
    assert iters.max() <= 10


def test_batched_map_blocks_spd_batch():
    # tests/core/test_map_blocks.mojo:36-66: 1003 SPD 4x4 systems
    a, b = kat.map_blocks_systems(1003, 4)
    mats, rhss = list(a), list(b)
    kinds = np.ones(len(mats), dtype=np.uint8)
    res, info, iters, berr = _run(mats, rhss, kinds)
    _check(mats, rhss, kinds, res, info, berr)


def test_batched_known_answers():
    mats = [kat.EXCHANGE, kat.TEXTBOOK_A, kat.NONSYM4, kat.PLU_A, kat.SPD5]
    rhss = [kat.EXCHANGE_B, kat.TEXTBOOK_B, kat.NONSYM4_B, np.arange(1.0, 5.0), np.ones(5)]
    kinds = np.array([0, 0, 0, 0, 1], dtype=np.uint8)
    res, info, iters, berr = _run(mats, rhss, kinds)
    assert np.allclose(res.solution(0).cpu().numpy(), kat.EXCHANGE_X, atol=1e-15)
    assert np.allclose(res.solution(1).cpu().numpy(), kat.TEXTBOOK_X, atol=1e-14)
    _check(mats, rhss, kinds, res, info, berr)


def test_batched_info_and_skips():
    mats = [np.zeros((3, 3)), kat.NOT_PD3, kat.LATE_NOT_PD, np.eye(2), np.eye(3)]
    rhss = [np.ones(3), np.ones(3), np.ones(4), np.ones(2), np.ones(3)]
    kinds = np.array([0, 1, 1, 2, 0], dtype=np.uint8)  # kind 2 is not a solver
    a, b, n, oa, ob = nl.pack_systems(mats, rhss)
    n = n.copy()
    n[4] = 300  # past the 256 the kernels take
    res = nl.batched_solve(a, b, n, kinds, oa, ob)
    info = res.info.cpu().numpy()
    assert info[0] == 1                       # exactly-zero first pivot, reported not raised
    assert info[1] == kat.NOT_PD3_INFO        # 1-based, as the panel's info (test_panel.mojo:146-161)
    assert info[2] == kat.LATE_NOT_PD_INDEX + 1  # the 0-based index cholesky's message names, + 1
    assert info[3] == -1 and info[4] == -1    # skipped
    assert np.isnan(res.berr.cpu().numpy()[3])


def test_batched_rejects_out_of_range_offsets():
    mats, rhss, kinds = _systems([5, 6], seed=3)
    a, b, n, oa, ob = nl.pack_systems(mats, rhss)
    oa = oa.copy()
    oa[1] = len(a) - 10  # 6 x 6 would run past the end of A
    res = nl.batched_solve(a, b, n, kinds, oa, ob)
    info = res.info.cpu().numpy()
    assert info[0] == 0 and info[1] == -1


def test_batched_graph_replays_match_eager():
    """One nmx_batched_gesv captured into a HIP graph: every replay (with
    eager library work and synchronizations between them, and new values in
    the inputs) matches an eager solve of the same inputs bit for bit."""
    import torch
    sizes = [5, 17, 64, 65, 100, 130, 200, 256, 33, 8]
    mats, rhss, kinds = _systems(sizes, seed=21)
    a, b, n, oa, ob = nl.pack_systems(mats, rhss)
    g = nl.batched_solve_graph(a, b, n, kinds, oa, ob)
    rng = np.random.default_rng(5)
    for rep in range(3):
        # new values of the same shapes (diagonally dominant / SPD kept)
        scale = 1.0 + 0.25 * rep
        a2 = a * scale
        b2 = b + rng.standard_normal(b.shape)
        g.a.copy_(torch.from_numpy(a2))
        g.b.copy_(torch.from_numpy(b2))
        res = g.replay()
        torch.cuda.synchronize()
        x_graph = res.x.cpu().numpy().copy()
        info_graph = res.info.cpu().numpy().copy()
        # eager work between replays: an unrelated batch, then the same inputs eagerly
        om, orhs, okinds = _systems([9, 40], seed=rep)
        oa_, ob_, on_, ooa, oob = nl.pack_systems(om, orhs)
        other = nl.batched_solve(oa_, ob_, on_, okinds, ooa, oob)
        torch.cuda.synchronize()
        assert (other.info.cpu().numpy() == 0).all()
        eager = nl.batched_solve(a2, b2, n, kinds, oa, ob)
        np.testing.assert_array_equal(info_graph, eager.info.cpu().numpy())
        np.testing.assert_array_equal(x_graph, eager.x.cpu().numpy())
        assert (info_graph == 0).all()
    g.close()


def test_batched_graph_survives_workspace_growth():
    """A live graph keeps the workspace it was captured with: calls that grow
    the context's workspace between replays (a larger batch, an LU solve with
    many right-hand sides) must not free it under the graph."""
    import torch
    sizes = [40, 150, 256, 9]
    mats, rhss, kinds = _systems(sizes, seed=31)
    a, b, n, oa, ob = nl.pack_systems(mats, rhss)
    g = nl.batched_solve_graph(a, b, n, kinds, oa, ob)
    rng = np.random.default_rng(8)
    for rep in range(3):
        # grow the workspace past anything used so far, each time by more
        big = [int(s) for s in rng.integers(200, 257, size=60 * (rep + 1))]
        bm, br, bk = _systems(big, seed=100 + rep)
        ba, bb, bn, boa, bob = nl.pack_systems(bm, br)
        other = nl.batched_solve(ba, bb, bn, bk, boa, bob)
        lu = nl.lu_factor(rng.standard_normal((600, 600)), block=256, base=32)
        lu.solve(rng.standard_normal((600, 300 * (rep + 1))))
        torch.cuda.synchronize()
        assert (other.info.cpu().numpy() == 0).all()
        b2 = b + rng.standard_normal(b.shape)
        g.b.copy_(torch.from_numpy(b2))
        res = g.replay()
        torch.cuda.synchronize()
        eager = nl.batched_solve(a, b2, n, kinds, oa, ob)
        np.testing.assert_array_equal(res.x.cpu().numpy(), eager.x.cpu().numpy())
        assert (res.info.cpu().numpy() == 0).all()
    g.close()


def test_batched_two_streams_in_a_row():
    """Two asynchronous calls under different torch streams share the
    context's workspace: the second is ordered after the first."""
    import torch
    m1, r1, k1 = _systems([256, 200, 180, 256] * 8, seed=41)
    m2, r2, k2 = _systems([250, 60, 256, 7] * 8, seed=42)
    p1 = nl.pack_systems(m1, r1)
    p2 = nl.pack_systems(m2, r2)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s1):
        res1 = nl.batched_solve(p1[0], p1[1], p1[2], k1, p1[3], p1[4])
    with torch.cuda.stream(s2):
        res2 = nl.batched_solve(p2[0], p2[1], p2[2], k2, p2[3], p2[4])
    torch.cuda.synchronize()
    _check(m1, r1, k1, res1, res1.info.cpu().numpy(), res1.berr.cpu().numpy())
    _check(m2, r2, k2, res2, res2.info.cpu().numpy(), res2.berr.cpu().numpy())


def test_batched_problems_sharing_a_matrix():
    """Problems that share one A (several right-hand sides) are factored in
    separate scratch; with a pool too small for both, the later one is
    skipped with info -1."""
    rng = np.random.default_rng(12)
    n0 = 150
    a0 = rng.uniform(-1.0, 1.0, (n0, n0)) + 0.1 * n0 * np.eye(n0)
    bs = [rng.standard_normal(n0) for _ in range(3)]
    # the A buffer is padded so the scratch pool (as many elements as A) fits three copies
    abuf = np.zeros(3 * n0 * n0)
    abuf[:n0 * n0] = a0.ravel()
    bbuf = np.concatenate(bs)
    n = np.full(3, n0, dtype=np.int32)
    oa = np.zeros(3, dtype=np.int64)
    ob = np.arange(3, dtype=np.int64) * n0
    kinds = np.zeros(3, dtype=np.uint8)
    res = nl.batched_solve(abuf, bbuf, n, kinds, oa, ob)
    info = res.info.cpu().numpy()
    assert (info == 0).all()
    for p in range(3):
        x = res.solution(p).cpu().numpy()
        assert oracle.backward_error(a0, x, bs[p]) <= 2e-15
        xo = _oracle_x(a0, bs[p], 0)
        assert np.max(np.abs(x - xo)) <= 1e-12 * max(1.0, np.max(np.abs(xo)))
    # a pool of 2 n0^2 holds two copies: the third problem is skipped
    res2 = nl.batched_solve(abuf[:2 * n0 * n0], bbuf, n, kinds, oa, ob)
    info2 = res2.info.cpu().numpy()
    assert info2[0] == 0 and info2[1] == 0 and info2[2] == -1
    np.testing.assert_array_equal(res2.x.cpu().numpy()[:2 * n0], res.x.cpu().numpy()[:2 * n0])


def test_batched_tuned_on_the_batch_still_solves():
    """autotune.tune_batched_on: coordinate search of teams_per_cu on the
    caller's own batch; every candidate run solves the batch, and the chosen
    shape gives the same x as the defaults (problems are independent of the
    launch shape)."""
    import ctypes

    from numax_amd import _lib, autotune
    rng = np.random.default_rng(77)
    sizes = [int(s) for s in rng.integers(8, 257, size=120)]
    mats, rhss, kinds = _systems(sizes, seed=78)
    a, b, n, oa, ob = nl.pack_systems(mats, rhss)
    calls = []

    def run(cfg):
        calls.append([int(cfg.teams_per_cu[i]) for i in range(3)])
        nl.batched_solve(a, b, n, kinds, oa, ob, teams_per_cu=calls[-1])

    teams = autotune.tune_batched_on(run, reps=1)
    assert len(teams) == 3 and all(t in (1, 2, 3, 4, 6, 8) for t in teams)
    assert len(calls) == 2 * (1 + 3 * 5)
    cfg = _lib.BatchCfg()
    _lib.load().nmx_batch_cfg_default(ctypes.byref(cfg))
    ref, _, _, _ = _run(mats, rhss, kinds, teams_per_cu=[int(cfg.teams_per_cu[i]) for i in range(3)])
    res, info, iters, berr = _run(mats, rhss, kinds, teams_per_cu=teams)
    _check(mats, rhss, kinds, res, info, berr)
    np.testing.assert_array_equal(res.x.cpu().numpy(), ref.x.cpu().numpy())
