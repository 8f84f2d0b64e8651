"""Writes tests/golden/goldens.npz: LAPACK results (SciPy 1.15 / NumPy 2.2,
run in the build container) on the seeded inputs the oracle tests use.

    python tests/golden/make_goldens.py
"""
import os
import sys

import numpy as np
import scipy.linalg as sl

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from numax_amd import random as R  # noqa: E402
from tests.golden import kat  # noqa: E402


def hamiltonian(n: int, omega: float) -> np.ndarray:
    """Finite-difference H on [-4, 4]: tridiagonal kinetic 1/h^2 on the
    diagonal, -1/(2h^2) off it, plus omega^2 x^2 / 2 (quantum_well.mojo:66-80)."""
    h = 8.0 / (n - 1)
    x = -4.0 + h * np.arange(n)
    a = np.diag(1.0 / h**2 + 0.5 * omega**2 * x**2)
    a += np.diag(np.full(n - 1, -0.5 / h**2), 1) + np.diag(np.full(n - 1, -0.5 / h**2), -1)
    return a


def main():
    out = {}
    a = R.lu_matrix(1, 512)
    b = R.normal(1, (512,))
    lu, piv = sl.lu_factor(a)
    out["c1_piv"] = piv.astype(np.int32)
    out["c1_x"] = sl.lu_solve((lu, piv), b)
    for n in (4, 17, 64, 130):
        a = R.normal(1000 + n, (n, n))
        lu, piv = sl.lu_factor(a)
        out[f"lu{n}_lu"] = lu
        out[f"lu{n}_piv"] = piv.astype(np.int32)
    out["plu_piv"] = sl.lu_factor(kat.PLU_A)[1].astype(np.int32)
    for n in (5, 40, 150):
        a = R.spd_matrix(2000 + n, n)
        out[f"chol{n}"] = sl.cholesky(a, lower=True)
    out["spd5_chol"] = sl.cholesky(kat.SPD5, lower=True)
    a = R.normal(50, (50, 20))
    out["qr50x20_r"] = sl.qr(a, mode="r")[0][:20]
    bq = R.normal(51, (50,))
    out["qr50x20_x"] = np.linalg.lstsq(a, bq, rcond=None)[0]
    out["tall6x3_r"] = sl.qr(kat.TALL_6x3, mode="r")[0][:3]
    out["tall6x3_x"] = np.linalg.lstsq(kat.TALL_6x3, np.arange(1.0, 7.0), rcond=None)[0]
    out["ham24_w"] = np.linalg.eigvalsh(hamiltonian(24, 1.0))
    g = R.normal(60, (32, 32))
    out["sym32_w"] = np.linalg.eigvalsh((g + g.T) / 2)
    a = R.normal(61, (1024, 32)).astype(np.float32).astype(np.float64)
    bb = R.normal(62, (1024,)).astype(np.float32).astype(np.float64)
    out["ts1024x32_x"] = np.linalg.lstsq(a, bb, rcond=None)[0]
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "goldens.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
