"""Pins of the 3-D spectral slab oracle (oracle/spectral3d.py) on the CPU."""
import numpy as np

from oracle import blocktri as ob
from oracle import spectral3d as o3


def _solve(SiM, SiP, rhs):
    nb = SiM.shape[0]
    S = [(SiM[i], SiP[i]) for i in range(nb)]
    return ob.thomas_direct(S, [np.eye(SiM.shape[1])] * nb, rhs)


def test_cheb_d2_exact_on_polynomials():
    x, D2 = o3.cheb_points_d2(7, -0.3, 0.5)
    for k in range(7):
        d2 = k * (k - 1) * x ** (k - 2) if k >= 2 else 0 * x
        assert np.allclose(D2 @ x ** k, d2, atol=1e-9 * max(1, k ** 2))


def _harmonic_quadratic(P):
    return P[:, 0] ** 2 - P[:, 1] ** 2 + 3 * P[:, 2] * P[:, 0]


def test_harmonic_quadratic_is_exact():
    SiM, SiP, rhs, pts = o3.oms_spectral3d(3, 5, 6, 7, _harmonic_quadratic)
    u = _solve(SiM, SiP, rhs)[:, 0]
    assert np.abs(u - _harmonic_quadratic(pts)).max() < 1e-10


def test_exp_sin_spectral_accuracy():
    def g(P):
        return np.exp(np.sqrt(2) * np.pi * P[:, 0]) * np.sin(np.pi * P[:, 1]) * np.sin(np.pi * P[:, 2])
    SiM, SiP, rhs, pts = o3.oms_spectral3d(4, 9, 12, 12, g)
    u = _solve(SiM, SiP, rhs)[:, 0]
    assert np.abs(u - g(pts)).max() / np.abs(g(pts)).max() < 1e-6


def test_planted_sign_error_is_caught():
    SiM, SiP, rhs, pts = o3.oms_spectral3d(3, 5, 6, 7, _harmonic_quadratic)
    u = _solve(-SiM, SiP, rhs)[:, 0]
    assert np.abs(u - _harmonic_quadratic(pts)).max() > 1e-3
