"""Pin the 2-D SOMS oracle (CPU only): exactness on a harmonic cubic, exact polynomial
interpolation of the gluing map, spectral convergence for Helmholtz plane waves, the
unknown layout, and a planted sign error that the exactness check catches."""
import numpy as np
import pytest

from oracle import soms2d as so


def harmonic(P):
    x, y = P[:, 0], P[:, 1]
    return x ** 3 - 3 * x * y ** 2 + 2 * x * y - y + 0.5


def plane(P, kh, th=0.4):
    return np.cos(kh * (np.cos(th) * P[:, 0] + np.sin(th) * P[:, 1]))


@pytest.mark.parametrize("px,py,nbx,nby", [(8, 8, 3, 2), (10, 6, 2, 4), (6, 10, 4, 3)])
def test_exact_on_harmonic_cubic(px, py, nbx, nby):
    S, XY, Ii, Ib = so.soms_system(px, py, nbx, nby, Lx=1.5, Ly=1.0, kh=0.0)
    u = harmonic(XY)
    assert np.max(np.abs(so.soms_solve(S, Ii, Ib, u[Ib]) - u[Ii])) < 1e-10


def test_layout():
    S, XY, Ii, Ib = so.soms_system(8, 6, 3, 2, Lx=1.5, Ly=1.0)
    assert len(np.unique(np.round(XY, 12), axis=0)) == XY.shape[0]
    hx, hy = 0.5, 0.5
    on_v = np.abs(XY[:, 0] / hx - np.round(XY[:, 0] / hx)) < 1e-12
    on_h = np.abs(XY[:, 1] / hy - np.round(XY[:, 1] / hy)) < 1e-12
    assert np.all(on_v ^ on_h)                     # every unknown on exactly one edge line
    # unknowns: (nbx + 1) nby vertical edges of py - 1 nodes, nbx (nby + 1) horizontal of px - 1
    assert XY.shape[0] == 4 * 2 * 5 + 3 * 3 * 7
    assert np.allclose(S[np.ix_(Ib, Ib)], np.eye(len(Ib)))


def test_gluing_interpolates_polynomials():
    t, _ = so.cheb_grid(8, 0.5)
    tj, _ = so.cheb_grid(12, 1.0)
    C = so.glue(t, tj, 7)
    deg = len(tj) - 3
    rng = np.random.default_rng(1)
    c = rng.standard_normal(deg + 1)
    f = lambda s: np.polynomial.polynomial.polyval(s, c)  # noqa: E731
    src = np.concatenate([t[1:-1], 0.5 + t[1:-1]])
    e = 5
    v = np.concatenate([np.zeros(e), f(src), np.zeros(2 * len(t) - 4), np.zeros(e)])
    out = C @ v
    assert np.max(np.abs(out[e:e + len(tj) - 2] - f(tj[1:-1]))) < 1e-10


def test_helmholtz_plane_wave_converges():
    errs = []
    for p in (8, 12, 16):
        S, XY, Ii, Ib = so.soms_system(p, p, 3, 3, kh=12.0)
        u = plane(XY, 12.0)
        errs.append(np.max(np.abs(so.soms_solve(S, Ii, Ib, u[Ib]) - u[Ii])))
    assert errs[0] > 10 * errs[1] > 100 * errs[2] and errs[2] < 1e-7, errs


def test_planted_sign_error_is_caught():
    S, XY, Ii, Ib = so.soms_system(8, 8, 3, 2, Lx=1.5, kh=0.0, _plant="flip_x")
    u = harmonic(XY)
    assert np.max(np.abs(so.soms_solve(S, Ii, Ib, u[Ib]) - u[Ii])) > 1e-2
