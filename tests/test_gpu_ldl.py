"""numax's ldl on the GPU (nmx_sytrf + nmx_ldl_unpack) against the oracle's
restatement of sytf2 / SciPy's unpacking and the reference's known answers.
The kernel rounds every product and difference as the restatement does, so
float64 factors, pivots and permutations must be equal bit for bit; float32
is compared with the float64 oracle at its precision."""
import numpy as np
import pytest

import oracle
from numax_amd import linalg as nl
from tests.golden import kat

pytestmark = pytest.mark.gpu


def test_ldl_known_answers():
    """tests/linalg/test_ldl.mojo:49-111 (SciPy's (lu, d, perm))."""
    for a, lu_w, d_w, p_w in ((kat.LDL_A, kat.LDL_LU, kat.LDL_D, kat.LDL_PERM),
                              (kat.LDL_Z, kat.LDL_Z_LU, kat.LDL_Z_D, kat.LDL_Z_PERM)):
        lu, d, perm = (t.cpu().numpy() for t in nl.ldl(a))
        np.testing.assert_allclose(lu, lu_w, atol=1e-14)
        np.testing.assert_allclose(d, d_w, atol=1e-14)
        np.testing.assert_array_equal(perm, p_w)
        np.testing.assert_allclose(lu @ d @ lu.T, a, atol=1e-14)


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (9, 3), (64, 4), (257, 5), (600, 6)])
def test_ldl_bit_exact_against_the_oracle(n, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n))
    a = (g + g.T) / 2
    f = nl.ldl(a)
    fo, ipo, info_o = oracle.sytf2_lower(a)
    np.testing.assert_array_equal(f.ipiv.cpu().numpy(), ipo)
    got = np.tril(f.factored.cpu().numpy())
    np.testing.assert_array_equal(got, np.tril(fo))
    lu_o, d_o, p_o = oracle.ldl(a)
    np.testing.assert_array_equal(f.lu.cpu().numpy(), lu_o)
    np.testing.assert_array_equal(f.d.cpu().numpy(), d_o)
    np.testing.assert_array_equal(f.perm.cpu().numpy(), p_o)
    assert f.info == info_o == 0
    lu, d = lu_o, d_o
    np.testing.assert_allclose(lu @ d @ lu.T, a, rtol=0, atol=1e-11 * n)


def test_ldl_float32_and_only_the_lower_triangle_is_read():
    import torch
    rng = np.random.default_rng(7)
    n = 120
    g = rng.standard_normal((n, n))
    a = (g + g.T) / 2
    junk = a + np.triu(rng.standard_normal((n, n)) * 1e3, 1)  # the strict upper triangle is ignored
    f64 = nl.ldl(junk)
    np.testing.assert_array_equal(f64.lu.cpu().numpy(), oracle.ldl(a)[0])
    f32 = nl.ldl(torch.as_tensor(a, dtype=torch.float32))
    lu, d, perm = (t.cpu().numpy().astype(np.float64) for t in f32)
    np.testing.assert_allclose(lu @ d @ lu.T, a, rtol=0, atol=5e-4)


def test_ldl_singular_reports_info():
    a = np.zeros((4, 4))
    a[1:, 1:] = [[2.0, 1.0, 0.0], [1.0, 3.0, 1.0], [0.0, 1.0, 4.0]]
    f = nl.ldl(a)
    assert f.info == 1
    _, ipo, info = oracle.sytf2_lower(a)
    assert info == 1
    # after the zero pivot the trailing entries are NaN; no 2 x 2 block may
    # start at the last column (it would store one entry past ipiv)
    np.testing.assert_array_equal(f.ipiv.cpu().numpy(), ipo)
    assert ipo[-1] > 0 or ipo[-2] == ipo[-1] < 0
