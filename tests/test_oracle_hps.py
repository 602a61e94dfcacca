"""Pin the HPS / oms oracle (CPU only): closed forms, exact solutions, exact domain
decomposition against a single global solve, and the Laplace row-sum invariant."""
import numpy as np
import pytest

from oracle import blocktri as ob
from oracle import hps as oh


class PDO2d:
    def __init__(self, c11, c22, c1=None, c2=None, c=None):
        self.c11, self.c22, self.c1, self.c2, self.c = c11, c22, c1, c2, c
        self.c12 = None


def ones(P):
    return np.ones(P.shape[0])


LAPLACE = PDO2d(ones, ones)


def helmholtz(kappa=20.0):
    def c(P):
        r2 = (P[:, 0] - 0.5) ** 2 + (P[:, 1] - 0.5) ** 2
        return -kappa ** 2 * (1.0 - 0.5 * np.exp(-r2 / 0.05))
    return PDO2d(ones, ones, c=c)


@pytest.mark.parametrize("q", [4, 8, 12, 18])
def test_cheb_exact_on_polynomials(q):
    t, D = oh.unit_cheb(q)
    for k in range(q):
        assert np.allclose(D @ t ** k, k * t ** (k - 1) if k else 0 * t, atol=1e-9 * q * q)
    x, Dx = oh.cheb_diff(q - 1)
    assert np.allclose(Dx @ np.ones(q), 0, atol=1e-12)


def solve_stot(res):
    S = res["S_rk_list"]
    nb = len(S)
    m = res["nc"]
    rhs = np.concatenate(res["rhs_list"])
    Sd = oh.stot_dense(S)
    return np.linalg.solve(Sd, rhs), Sd


def exp_sin(P):
    return np.exp(P[:, 0]) * np.sin(P[:, 1])


def test_laplace_exact_solution_small():
    """Config-1 data u = exp(x) sin(y) on a reduced grid: spectral accuracy."""
    res = oh.oms_stot(LAPLACE, 4, (1 / 8, 1 / 4), 10, bc=exp_sin)
    u, _ = solve_stot(res)
    pts = np.concatenate(res["Ic_pts"])
    assert np.max(np.abs(u - exp_sin(pts))) < 1e-9


@pytest.mark.parametrize("pdo_name", ["laplace", "helmholtz"])
def test_dd_equals_global_solve(pdo_name):
    """The S-system is an exact domain decomposition of the global HPS system."""
    pdo = LAPLACE if pdo_name == "laplace" else helmholtz(8.0)
    N, leaf_h, q = 4, (1 / 8, 1 / 4), 8

    def bc(P):
        return np.cos(2 * P[:, 0] + 1) * np.sinh(P[:, 1]) + P[:, 0] * P[:, 1] ** 2

    res = oh.oms_stot(pdo, N, leaf_h, q, bc=bc)
    u, _ = solve_stot(res)
    ug, pts = oh.global_solution_on_lines(pdo, N, leaf_h, q, bc)
    assert np.allclose(np.concatenate(res["Ic_pts"]), pts, atol=1e-15)
    assert np.linalg.norm(u - ug) / np.linalg.norm(ug) < 1e-11


def test_laplace_row_sum_invariant():
    """Constant data is reproduced exactly: 1 + sum_J S_J 1 + R 1 = ... -> u = 1, i.e.
    (I + S_l + S_r) 1 = -(rhs of f = 1) row-wise (multislab/oms.py:184-185,292)."""
    res = oh.oms_stot(LAPLACE, 4, (1 / 8, 1 / 4), 8, bc=lambda P: np.ones(P.shape[0]))
    for k, (blocks, R) in enumerate(zip(res["S_rk_list"], res["R_list"])):
        lhs = np.ones(res["nc"])
        for B in blocks:
            lhs = lhs + B @ np.ones(B.shape[1])
        assert np.allclose(lhs + R @ np.ones(R.shape[1]), 0.0, atol=1e-10)


def test_planted_flux_sign_error_is_caught():
    """Flipping the sign of one leaf's outward normal breaks DD = global; the check is
    not vacuous (a mistake planted in the oracle, on the CPU)."""
    N, leaf_h, q = 4, (1 / 8, 1 / 4), 6
    res = oh.oms_stot(LAPLACE, N, leaf_h, q, bc=exp_sin)
    u, _ = solve_stot(res)
    # Corrupt one S block as a sign error in its flux rows would.
    res["S_rk_list"][1][0] = -res["S_rk_list"][1][0]
    ubad, _ = solve_stot(res)
    pts = np.concatenate(res["Ic_pts"])
    assert np.max(np.abs(u - exp_sin(pts))) < 1e-6
    assert np.max(np.abs(ubad - exp_sin(pts))) > 1e-3


def test_redblack_on_stot():
    res = oh.oms_stot(LAPLACE, 6, (1 / 12, 1 / 4), 6, bc=exp_sin)
    S = res["S_rk_list"]
    m = res["nc"]
    nb = len(S)
    pairs = [(np.zeros((m, m)), S[0][0])] + [(s[0], s[1]) for s in S[1:-1]] + \
            [(S[-1][0], np.zeros((m, m)))]
    rhs = np.concatenate(res["rhs_list"])
    x = ob.redblack_direct(pairs, [np.eye(m)] * nb, rhs)
    u, _ = solve_stot(res)
    assert np.linalg.norm(x - u) / np.linalg.norm(u) < 1e-12


def periodic_harmonic(P):
    return (np.sin(2 * np.pi * P[:, 0]) * np.exp(2 * np.pi * (P[:, 1] - 1))
            + np.cos(4 * np.pi * P[:, 0]) * np.cosh(4 * np.pi * (P[:, 1] - 0.5)) / np.cosh(2 * np.pi))


@pytest.mark.parametrize("pdo_name", ["laplace", "helmholtz"])
def test_periodic_dd_equals_global_solve(pdo_name):
    """x-periodic square (squareTorus-style dSlabs): the periodic S-system is an exact
    domain decomposition of the global HPS system with the lines x = 0 and x = 1
    identified (the coefficients must be periodic in x too)."""
    pdo = LAPLACE if pdo_name == "laplace" else PDO2d(
        ones, ones, c=lambda P: -64.0 * (1 + 0.3 * np.sin(2 * np.pi * P[:, 0])))
    N, leaf_h, q = 4, (1 / 8, 1 / 4), 8
    res = oh.oms_stot(pdo, N, leaf_h, q, bc=periodic_harmonic, periodic=True)
    assert all(len(b) == 2 for b in res["S_rk_list"])
    u = np.linalg.solve(oh.stot_dense_cyclic(res["S_rk_list"]), np.concatenate(res["rhs_list"]))
    ug, pts = oh.global_solution_on_lines(pdo, N, leaf_h, q, periodic_harmonic, periodic=True)
    assert np.allclose(np.concatenate(res["Ic_pts"]), pts, atol=1e-15)
    assert np.linalg.norm(u - ug) / np.linalg.norm(ug) < 1e-11


def test_periodic_spectral_accuracy():
    errs = []
    for q in (8, 12):
        res = oh.oms_stot(LAPLACE, 4, (1 / 8, 1 / 4), q, bc=periodic_harmonic, periodic=True)
        u = np.linalg.solve(oh.stot_dense_cyclic(res["S_rk_list"]), np.concatenate(res["rhs_list"]))
        errs.append(np.max(np.abs(u - periodic_harmonic(np.concatenate(res["Ic_pts"])))))
    assert errs[1] < 1e-3 * errs[0] and errs[1] < 1e-8, errs
