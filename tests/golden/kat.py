"""Known-answer data copied as numbers from the reference's own tests (the
matrices and expected factors, not the test code)."""
import numpy as np

# tests/linalg/test_lu_plu.mojo:15-109 -- SciPy's scipy.linalg.lu of a 4x4
# whose first pivot is zero (A = P L U).
PLU_A = np.array([[0.0, 2.0, 1.0, 3.0], [4.0, 1.0, -2.0, 1.0], [2.0, 5.0, 3.0, -1.0], [1.0, -3.0, 2.0, 6.0]])
PLU_P = np.array([[0.0, 0.0, 0.0, 1.0], [1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0]])
PLU_L = np.array([[1.0, 0.0, 0.0, 0.0], [0.5, 1.0, 0.0, 0.0], [0.25, -0.7222222222222222, 1.0, 0.0],
                  [0.0, 0.4444444444444444, -0.14432989690721645, 1.0]])
PLU_U = np.array([[4.0, 1.0, -2.0, 1.0], [0.0, 4.5, 4.0, -1.5], [0.0, 0.0, 5.388888888888889, 4.666666666666667],
                  [0.0, 0.0, 0.0, 4.34020618556701]])

# tests/linalg/test_panel.mojo:146-161 -- not positive definite, info == 2.
NOT_PD3 = np.array([[4.0, 2.0, 2.0], [2.0, 1.0, 0.0], [2.0, 0.0, 9.0]])
NOT_PD3_INFO = 2

# tests/linalg/test_panel.mojo:258-274 -- first pivot is row 2.
PIVOT3 = np.array([[0.0, 1.0, 2.0], [4.0, 5.0, 6.0], [7.0, 8.0, 10.0]])
PIVOT3_FIRST = 2

# tests/linalg/test_tensor_linalg.mojo:816-832 -- the exchange matrix.
EXCHANGE = np.array([[0.0, 1.0], [1.0, 0.0]])
EXCHANGE_B = np.array([3.0, 5.0])
EXCHANGE_X = np.array([5.0, 3.0])
EXCHANGE_DET = -1.0

# tests/linalg/test_linalg.mojo:134-139 -- the textbook 3x3, x = (2, 3, -1).
TEXTBOOK_A = np.array([[2.0, 1.0, -1.0], [-3.0, -1.0, 2.0], [-2.0, 1.0, 2.0]])
TEXTBOOK_B = np.array([8.0, -11.0, -3.0])
TEXTBOOK_X = np.array([2.0, 3.0, -1.0])


def sin48() -> np.ndarray:
    """tests_gpu/linalg/test_lu_plu_gpu.mojo:17-24: a[i] = sin(i*1.37 +
    (i%7)*0.71) in float32, row-major 48x48."""
    i = np.arange(48 * 48)
    x = i.astype(np.float32) * np.float32(1.37) + (i % 7).astype(np.float32) * np.float32(0.71)
    return np.sin(x).astype(np.float32).reshape(48, 48)

# tests/linalg/test_tensor_linalg.mojo:504-530 -- SPD 5x5, diagonally dominant.
SPD5 = np.array([[9.0, 3.0, 1.0, 2.0, 0.0], [3.0, 10.0, 2.0, 1.0, 1.0], [1.0, 2.0, 8.0, 3.0, 2.0],
                 [2.0, 1.0, 3.0, 11.0, 4.0], [0.0, 1.0, 2.0, 4.0, 12.0]])

# tests/linalg/test_tensor_linalg.mojo:690-720 -- PD leading 2x2, fails at
# index 3 (the second block at block=2).
LATE_NOT_PD = np.array([[4.0, 1.0, 0.0, 0.0], [1.0, 4.0, 0.0, 0.0], [0.0, 0.0, 1.0, 3.0], [0.0, 0.0, 3.0, 1.0]])
LATE_NOT_PD_INDEX = 3

# tests/linalg/test_tensor_linalg.mojo:723-745 -- nonsymmetric 4x4, rhs 1..4.
NONSYM4 = np.array([[4.0, 1.0, 2.0, 0.5], [1.0, 5.0, 0.0, 2.0], [2.0, 0.0, 6.0, 1.0], [0.5, 2.0, 1.0, 7.0]])
NONSYM4_B = np.array([1.0, 2.0, 3.0, 4.0])

# tests/linalg/test_tensor_linalg.mojo (_tall_6x3) -- full column rank.
TALL_6x3 = np.array([[1.0, -2.0, 3.0], [4.0, 5.0, -1.0], [-3.0, 2.0, 7.0], [0.5, -4.0, 1.5], [6.0, 1.0, -2.5],
                     [2.0, 3.5, 4.0]])

# tests/linalg/test_linalg.mojo:65-89 -- spd3 (det 67 at :171-173).
SPD3_DET = 67.0


def spd_n(n: int) -> np.ndarray:
    """tests/linalg/test_tensor_linalg.mojo:486-502 (and bench_linalg.mojo's
    _spd_entry): n on the diagonal, 1/(1+|i-j|) off it."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = 1.0 / (1.0 + np.abs(i - j))
    a[np.arange(n), np.arange(n)] = float(n)
    return a


def dynamic_lu_entry(n: int) -> np.ndarray:
    """tests/linalg/test_dynamic_lu.mojo:22-25."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = ((i * 7 + j * 3) % 5) * 0.25 - 0.5
    a[np.arange(n), np.arange(n)] = n + 1.0
    return a


def map_blocks_systems(batch: int = 1003, n: int = 4):
    """tests/core/test_map_blocks.mojo:36-66: `batch` SPD systems A = M M^T +
    4 I from a multiplicative hash, b_r = 1 + r/4.  Returns (A[batch,n,n],
    b[batch,n])."""
    p = np.arange(batch)[:, None]
    jj = np.arange(n * n)[None, :]
    h = ((p * n * n + jj) * 2654435761) % 100003
    m = (h / 100003.0 - 0.5).reshape(batch, n, n)
    a = np.einsum("pik,pjk->pij", m, m) + 4.0 * np.eye(n)
    b = np.tile(1.0 + 0.25 * np.arange(n), (batch, 1))
    return a, b

# tests/linalg/test_slogdet.mojo:36-48 -- a sign-negative 3x3, and a
# singular matrix (row 1 = 2 row 0): sign 0, logabsdet -inf.  The test's
# docstring calls the first determinant -30; its assertions compare with
# det() instead, and the cofactor expansion 2(3(-2) - 2) + (-2) gives -18.
SLOGDET_M3 = np.array([[2.0, -1.0, 0.0], [1.0, 3.0, 1.0], [0.0, 2.0, -2.0]])
SLOGDET_M3_DET = -18.0
SLOGDET_SINGULAR = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [7.0, 8.0, 10.0]])

# tests/linalg/test_linalg.mojo:65-82, :170-173 -- spd3 (det 67 by cofactors)
# and general3 (leading minors nonzero, so the unpivoted LU is valid).
SPD3 = np.array([[4.0, 2.0, 1.0], [2.0, 5.0, 3.0], [1.0, 3.0, 6.0]])
GENERAL3 = np.array([[2.0, 1.0, -1.0], [-3.0, -1.0, 2.0], [-2.0, 1.0, 2.0]])

# tests/linalg/test_ldl.mojo:25-111 -- SciPy's scipy.linalg.ldl of an
# indefinite 4x4 (one 2 x 2 pivot block, interchanges) and of a 3x3 with an
# all-zero diagonal.
LDL_A = np.array([[1.0, 2.0, 3.0, 4.0], [2.0, -1.0, 0.5, 2.0], [3.0, 0.5, 0.2, -1.0], [4.0, 2.0, -1.0, 0.1]])
LDL_LU = np.array([[1.0, 0.0, 0.0, 0.0], [0.49056603773584906, 0.37735849056603776, -0.3249656121045393, 1.0],
                   [-0.27044025157232704, 0.8176100628930818, 1.0, 0.0], [0.0, 1.0, 0.0, 0.0]])
LDL_D = np.array([[1.0, 4.0, 0.0, 0.0], [4.0, 0.1, 0.0, 0.0], [0.0, 0.0, 1.828930817610063, 0.0],
                  [0.0, 0.0, 0.0, -2.9289889958734525]])
LDL_PERM = np.array([0, 3, 2, 1])
LDL_Z = np.array([[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [2.0, 3.0, 0.0]])
LDL_Z_LU = np.array([[1.0, 0.0, 0.0], [1.5, 0.5, 1.0], [0.0, 1.0, 0.0]])
LDL_Z_D = np.array([[0.0, 2.0, 0.0], [2.0, 0.0, 0.0], [0.0, 0.0, -3.0]])
LDL_Z_PERM = np.array([0, 2, 1])
