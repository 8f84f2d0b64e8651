"""nmx_batched_gesv_noalloc (the allocation-free variant of the batched
solve, scratch in a static device array) against nmx_batched_gesv: the same
kernels on the same inputs, so the solutions, info, iteration counts and
backward errors must be identical bit for bit -- on two calls in a row, the
second with a smaller batch (the array keeps the first call's contents), and
a call past the array's 64 MiB is refused without touching the outputs."""
import ctypes

import numpy as np
import pytest

from numax_amd import _lib
from numax_amd import linalg as nl

pytestmark = pytest.mark.gpu


def _systems(sizes, seed):
    rng = np.random.default_rng(seed)
    mats, rhss = [], []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            mats.append(rng.uniform(-1.0, 1.0, (n, n)) + 0.1 * n * np.eye(n))
        else:
            g = rng.standard_normal((n, n))
            mats.append(g @ g.T / n + np.eye(n))
        rhss.append(rng.standard_normal(n))
    kinds = (np.arange(len(sizes)) % 2).astype(np.uint8)
    return mats, rhss, kinds


def _call(fn_name, a, b, n, kinds, oa, ob):
    import torch
    A = torch.as_tensor(a, device="cuda")
    B = torch.as_tensor(b, device="cuda")
    N = torch.as_tensor(n, device="cuda")
    K = torch.as_tensor(kinds, device="cuda")
    OA = torch.as_tensor(oa, device="cuda")
    OB = torch.as_tensor(ob, device="cuda")
    X = torch.full_like(B, -7.0)
    info = torch.full((len(n),), -9, dtype=torch.int32, device="cuda")
    iters = torch.zeros_like(info)
    berr = torch.zeros(len(n), dtype=torch.float64, device="cuda")
    lib = _lib.load()
    cfg = _lib.BatchCfg()
    lib.nmx_batch_cfg_default(ctypes.byref(cfg))
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    st = getattr(lib, fn_name)(nl._ctx(A).handle, len(n), P(N), P(OA), P(OB), P(K), P(A), A.numel(), P(B),
                               B.numel(), P(X), P(info), P(iters), P(berr), ctypes.byref(cfg))
    torch.cuda.synchronize()
    return st, X.cpu().numpy(), info.cpu().numpy(), iters.cpu().numpy(), berr.cpu().numpy()


def test_noalloc_matches_the_allocating_entry_point_twice():
    rng = np.random.default_rng(77)
    first = [int(s) for s in rng.integers(1, 257, size=160)]
    second = [int(s) for s in rng.integers(1, 129, size=40)]  # smaller: the array keeps the first call's bytes
    for sizes, seed in ((first, 1), (second, 2)):
        mats, rhss, kinds = _systems(sizes, seed)
        a, b, n, oa, ob = nl.pack_systems(mats, rhss)
        st_ref, x_ref, info_ref, it_ref, be_ref = _call("nmx_batched_gesv", a, b, n, kinds, oa, ob)
        st, x, info, it, be = _call("nmx_batched_gesv_noalloc", a, b, n, kinds, oa, ob)
        assert st_ref == 0 and st == 0
        assert (info_ref == 0).all()
        np.testing.assert_array_equal(info, info_ref)
        np.testing.assert_array_equal(it, it_ref)
        np.testing.assert_array_equal(x, x_ref)
        np.testing.assert_array_equal(be, be_ref)
        assert be.max() <= 1e-15


def test_noalloc_refuses_past_its_array():
    # 4 a_elems bytes of working copies alone pass 64 MiB: refused, outputs untouched
    n0 = 256
    count = 260
    a = np.zeros(count * n0 * n0)
    b = np.zeros(count * n0)
    n = np.full(count, n0, dtype=np.int32)
    oa = np.arange(count, dtype=np.int64) * n0 * n0
    ob = np.arange(count, dtype=np.int64) * n0
    kinds = np.zeros(count, dtype=np.uint8)
    st, x, info, it, be = _call("nmx_batched_gesv_noalloc", a, b, n, kinds, oa, ob)
    assert st == 3  # NMX_ERR_UNSUPPORTED
    assert (info == -9).all() and (x == -7.0).all()
