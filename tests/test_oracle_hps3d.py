"""Pin the 3-D HPS / oms oracle (CPU only): exactness on a harmonic polynomial, exact
domain decomposition against a single global solve, spectral convergence to a smooth
harmonic function, the Laplace row-sum invariant, and a planted flux-sign error that
those checks catch."""
import numpy as np
import pytest

from oracle import hps3d as o3


class PDO3:
    def __init__(self, c=None, c2=None):
        one = lambda P: np.ones(P.shape[0])  # noqa: E731
        self.c11 = self.c22 = self.c33 = one
        self.c1, self.c2, self.c3, self.c = None, c2, None, c


LAP = PDO3()
H3 = (1 / 3, 1 / 2, 1 / 2)


def poly(P):
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    return x ** 2 + y ** 2 - 2 * z ** 2 + x * y * z + 3 * x - y


def smooth(P):
    return np.exp(np.sqrt(2) * np.pi * P[:, 0]) * np.sin(np.pi * P[:, 1]) * np.sin(np.pi * P[:, 2])


def solve(res):
    S = o3.stot_dense(res["S_rk_list"])
    return np.linalg.solve(S, np.concatenate(res["rhs_list"]))


@pytest.mark.parametrize("q", [5, 6])
def test_exact_on_harmonic_polynomial(q):
    res = o3.oms_stot(LAP, 3, H3, q, bc=poly)
    u = solve(res)
    assert np.max(np.abs(u - poly(np.concatenate(res["Ic_pts"])))) < 1e-11


@pytest.mark.parametrize("pdo", ["laplace", "helmholtz", "convection"])
def test_dd_equals_global_solve(pdo):
    P = {"laplace": LAP,
         "helmholtz": PDO3(c=lambda X: -16.0 * (1 + 0.3 * X[:, 0])),
         "convection": PDO3(c2=lambda X: 2.0 + X[:, 2])}[pdo]
    N, q = 4, 5
    h = (1 / 4, 1 / 2, 1 / 2)
    res = o3.oms_stot(P, N, h, q, bc=smooth)
    u = solve(res)
    ug, pts = o3.global_solution_on_planes(P, N, h, q, smooth)
    assert np.allclose(np.concatenate(res["Ic_pts"]), pts, atol=1e-15)
    assert np.linalg.norm(u - ug) / np.linalg.norm(ug) < 1e-11


def test_spectral_convergence():
    errs = []
    for q in (5, 7, 9):
        res = o3.oms_stot(LAP, 3, H3, q, bc=smooth)
        pts = np.concatenate(res["Ic_pts"])
        errs.append(np.max(np.abs(solve(res) - smooth(pts))) / np.max(np.abs(smooth(pts))))
    assert errs[0] > 5 * errs[1] and errs[1] > 100 * errs[2] and errs[2] < 1e-6, errs


def test_laplace_row_sum_invariant():
    res = o3.oms_stot(LAP, 4, (1 / 4, 1 / 2, 1 / 2), 5, bc=lambda P: np.ones(P.shape[0]))
    for blocks, R in zip(res["S_rk_list"], res["R_list"]):
        lhs = np.ones(res["nc"])
        for B in blocks:
            lhs = lhs + B @ np.ones(B.shape[1])
        assert np.allclose(lhs + R @ np.ones(R.shape[1]), 0.0, atol=1e-10)


def test_planted_flux_sign_error_is_caught():
    """Flipping one face's outward normal inside the oracle (on the CPU) breaks the
    exactness and domain-decomposition checks above: they are not vacuous."""
    res = o3.oms_stot(LAP, 3, H3, 5, bc=poly, _plant="flux_sign")
    pts = np.concatenate(res["Ic_pts"])
    assert np.max(np.abs(solve(res) - poly(pts))) > 1e-3


def test_periodic_dd_equals_global_solve():
    """x-periodic cube (geometry/geom_3D/squareTorus.py dSlabs): the periodic S-system
    equals the global HPS solve with the faces x = 0 and x = 1 identified."""
    from oracle.hps import stot_dense_cyclic
    u = lambda P: np.sin(2 * np.pi * P[:, 0]) * np.sin(np.pi * P[:, 2]) * np.exp(  # noqa: E731
        np.sqrt(5) * np.pi * (P[:, 1] - 1))
    N, q, h = 4, 5, (1 / 4, 1 / 2, 1 / 2)
    res = o3.oms_stot(LAP, N, h, q, bc=u, periodic=True)
    x = np.linalg.solve(stot_dense_cyclic(res["S_rk_list"]), np.concatenate(res["rhs_list"]))
    ug, pts = o3.global_solution_on_planes(LAP, N, h, q, u, periodic=True)
    assert np.allclose(np.concatenate(res["Ic_pts"]), pts, atol=1e-15)
    assert np.linalg.norm(x - ug) / np.linalg.norm(ug) < 1e-11
