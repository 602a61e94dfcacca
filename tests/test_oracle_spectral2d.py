"""Pin the 2-D spectral double-slab oracle (CPU only): exactness on a harmonic polynomial,
spectral convergence to exp(x) sin(y), an exact plane wave of the Helmholtz operator,
the Laplace row-sum invariant, and a planted sign error those checks catch."""
import numpy as np
import pytest

from oracle import spectral2d as os2


class PDO2:
    def __init__(self, c=None, c1=None, c2=None):
        one = lambda P: np.ones(P.shape[0])  # noqa: E731
        self.c11 = self.c22 = one
        self.c1, self.c2, self.c = c1, c2, c
        self.c12 = None


def solve(res):
    from oracle.hps import stot_dense
    return np.linalg.solve(stot_dense(res["S_rk_list"]), np.concatenate(res["rhs_list"]))


def poly(P):
    return P[:, 0] ** 2 - P[:, 1] ** 2 + 3 * P[:, 0] * P[:, 1] + P[:, 0]


def test_exact_on_harmonic_polynomial():
    res = os2.oms_stot(PDO2(), 5, 6, 12, bc=poly)
    assert np.max(np.abs(solve(res) - poly(np.concatenate(res["Ic_pts"])))) < 1e-12


def test_spectral_convergence():
    f = lambda P: np.exp(P[:, 0]) * np.sin(P[:, 1])  # noqa: E731
    errs = []
    for ordy in (6, 10, 14):
        res = os2.oms_stot(PDO2(), 4, 8, ordy, bc=f)
        errs.append(np.max(np.abs(solve(res) - f(np.concatenate(res["Ic_pts"])))))
    assert errs[0] > 1e-9 and errs[1] < 1e-4 * errs[0] and errs[2] < 1e-12, errs


def wave(P):
    k = 3.0 / np.sqrt(2.0)
    return np.sin(k * P[:, 0] + k * P[:, 1])


def test_helmholtz_plane_wave_and_planted_sign():
    """-Lap u - 9 u = 0 holds for the plane wave: exact up to the discretization; the
    same oracle with the sign of the c term flipped (planted) misses it by far."""
    pdo = PDO2(c=lambda P: -9.0 * np.ones(P.shape[0]))
    res = os2.oms_stot(pdo, 4, 12, 20, bc=wave)
    pts = np.concatenate(res["Ic_pts"])
    assert np.max(np.abs(solve(res) - wave(pts))) < 1e-9
    bad = os2.oms_stot(pdo, 4, 12, 20, bc=wave, _plant="c_sign")
    assert np.max(np.abs(solve(bad) - wave(pts))) > 1e-3


def test_laplace_row_sum_invariant():
    res = os2.oms_stot(PDO2(), 5, 6, 10, bc=lambda P: np.ones(P.shape[0]))
    for blocks, R in zip(res["S_rk_list"], res["R_list"]):
        lhs = np.ones(res["nc"]) + sum(B @ np.ones(B.shape[1]) for B in blocks)
        assert np.allclose(lhs + R @ np.ones(R.shape[1]), 0.0, atol=1e-10)
