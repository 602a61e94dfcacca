"""GPU: solverWrapper's condensed skeleton system (solver/solver.py:100-144 attributes
XX, Ii, Ib, Aib, Abi, Abb, solver_ii; pde_solver.solve_dir) on 2-D and 3-D HPS double
slabs.

* solver_ii (A_ii^{-1}, modes N and T) applied to A_ib's columns equals the S maps of the
  pair solve (1e-12), and its transpose is the adjoint of its apply.
* Laplace with a harmonic polynomial u: the centre-line values -A_ii^{-1} A_ib u_b
  (solve_dir) equal u exactly, and A_bi u_c + A_bb u_b equals the exact outward normal
  derivative on the double slab's boundary (the DtN map of the double slab).
* A body load: solve_dir(u_b, f) = A_ii^{-1}(f - A_ib u_b), checked by applying A_ii.
* 3-D: solver_ii and A_ib against the 3-D S maps, and the harmonic-polynomial trace."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    b = b.cpu().numpy() if hasattr(b, "cpu") else np.asarray(b)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def T(gpu):
    import torch
    from sslablu_amd import _lib
    _lib.load()
    return torch


class Lap2:
    c11 = c22 = staticmethod(lambda P: np.ones(P.shape[0]))
    c1 = c2 = c = None


class Lap3:
    c11 = c22 = c33 = staticmethod(lambda P: np.ones(P.shape[0]))
    c1 = c2 = c3 = c = None


def u2(P):
    x, y = P[:, 0], P[:, 1]
    return x ** 3 - 3 * x * y ** 2 + 2 * x * y - y + 0.5


def grad_u2(P):
    x, y = P[:, 0], P[:, 1]
    return np.stack([3 * x ** 2 - 3 * y ** 2 + 2 * y, -6 * x * y + 2 * x - 1], axis=1)


def test_2d_skeleton_system(T):
    from sslablu_amd.solver.solver import solverOptions, solverWrapper
    geom = np.array([[0.25, 0.0], [0.75, 1.0]])
    W = solverWrapper(solverOptions("hps", [8, 8], [1 / 8, 1 / 8])).construct(geom, Lap2())
    m, nbd = len(W.Ii), len(W.Ib)
    rng = np.random.default_rng(0)
    V = rng.standard_normal((nbd, 3))
    Sii = W.solver_ii
    for which, cols in (("l", slice(0, m)), ("r", slice(m, 2 * m)), ("gb", slice(2 * m, nbd))):
        Vj = T.tensor(V[cols], device="cuda")
        assert rel(Sii @ (W.Aib[:, cols] @ Vj), W.smap(which) @ Vj) < 1e-12, which
    x, y = rng.standard_normal((m, 2)), rng.standard_normal((m, 2))
    assert abs(np.sum((Sii @ x) * y) - np.sum(x * (Sii.T @ y))) < 1e-12 * np.abs(Sii @ x).sum()
    assert rel(Sii.rmatvec(y[:, 0]), np.linalg.solve(W.Aii.cpu().numpy().T, y[:, 0])) < 1e-12
    ub, uc = u2(W.XXb), u2(W.XXi)
    assert np.max(np.abs(W.solve_dir(ub) - uc)) < 1e-10
    # boundary rows: outward normal derivative of u on the double slab
    flux = (W.Abi @ T.tensor(uc, device="cuda") + W.Abb @ T.tensor(ub, device="cuda")).cpu().numpy()
    P = W.XXb
    n = np.zeros_like(P)
    n[np.abs(P[:, 0] - 0.25) < 1e-12, 0] = -1
    n[np.abs(P[:, 0] - 0.75) < 1e-12, 0] = 1
    n[np.abs(P[:, 1]) < 1e-12, 1] = -1
    n[np.abs(P[:, 1] - 1) < 1e-12, 1] = 1
    assert np.all(np.abs(n).sum(1) == 1)
    dn = (grad_u2(P) * n).sum(1)
    assert np.max(np.abs(flux - dn)) < 1e-8 * np.max(np.abs(dn))
    # body load on the skeleton
    f = rng.standard_normal(m)
    ui = W.solve_dir(ub, f)
    Aii = W.Aii.cpu().numpy()
    assert rel(Aii @ ui + W.Aib.cpu().numpy() @ ub, f) < 1e-11


def test_3d_skeleton_system(T):
    from sslablu_amd.solver.solver import solverOptions, solverWrapper
    geom = np.array([[0.0, 0.0, 0.0], [0.5, 1.0, 1.0]])
    W = solverWrapper(solverOptions("hpsalt", [6, 6, 6], [1 / 8, 1 / 4, 1 / 4])).construct(
        geom, Lap3())
    m, nbd = len(W.Ii), len(W.Ib)
    rng = np.random.default_rng(1)
    V = T.tensor(rng.standard_normal((nbd, 2)), device="cuda")
    Sii, Aib = W.solver_ii, W.Aib
    assert Aib.shape == (m, nbd)
    for which, cols in (("l", slice(0, m)), ("r", slice(m, 2 * m)), ("gb", slice(2 * m, nbd))):
        assert rel(Sii @ (Aib[:, cols] @ V[cols]), W.smap(which).matmat(V[cols])) < 1e-12, which
    x, y = rng.standard_normal((m, 2)), rng.standard_normal((m, 2))
    assert abs(np.sum((Sii @ x) * y) - np.sum(x * (Sii.T @ y))) < 1e-12 * np.abs(Sii @ x).sum()

    def u3(P):
        return P[:, 0] ** 2 + P[:, 1] ** 2 - 2 * P[:, 2] ** 2 + P[:, 0] * P[:, 1] * P[:, 2]

    assert np.max(np.abs(W.solve_dir(u3(W.XXb)) - u3(W.XXi))) < 1e-10
