"""Pin the block-tridiagonal oracle (CPU only).

Every solver is checked against numpy.linalg.solve of the assembled dense system --
the closed form the reference's own test uses (direct_solve/test_RB.py:51-60) --
including its nine (nSlabs, m) known-answer cases at its tolerance 1e-8.
"""
import numpy as np
import pytest

from oracle import blocktri as ob

KAT_CASES = [(8, 3), (2, 1), (2, 5), (4, 3), (8, 1), (8, 5), (8, 10), (16, 3), (32, 3)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nSlabs,m", KAT_CASES)
def test_redblack_kat(nSlabs, m):
    S, T, dense, rhs = ob.make_block_tridiagonal(nSlabs, m, seed=42)
    x_ref = np.linalg.solve(dense, rhs)
    x = ob.redblack_solve(ob.redblack_factor(S, T), rhs)
    assert rel(x, x_ref) < 1e-8
    assert rel(x, x_ref) < 1e-13  # the system is strongly diagonally dominant


@pytest.mark.parametrize("nSlabs,m", KAT_CASES)
def test_thomas_kat(nSlabs, m):
    S, T, dense, rhs = ob.make_block_tridiagonal(nSlabs, m, seed=42)
    x_ref = np.linalg.solve(dense, rhs)
    assert rel(ob.thomas_direct(S, T, rhs), x_ref) < 1e-13


def test_generator_matches_assembly():
    S, T, dense, _ = ob.make_block_tridiagonal(4, 3, seed=42)
    assert np.array_equal(ob.assemble_dense(S, T), dense)


def _random_system(nb, m, seed, cyclic=False, scale=0.3):
    rng = np.random.default_rng(seed)
    T = [np.eye(m) * 3 + 0.2 * rng.standard_normal((m, m)) for _ in range(nb)]
    S = []
    for i in range(nb):
        left = scale * rng.standard_normal((m, m)) / np.sqrt(m)
        right = scale * rng.standard_normal((m, m)) / np.sqrt(m)
        if not cyclic:
            if i == 0:
                left = np.zeros((m, m))
            if i == nb - 1:
                right = np.zeros((m, m))
        S.append((left, right))
    return S, T


@pytest.mark.parametrize("nb", [2, 3, 5, 7, 12])
def test_redblack_padding_nonpow2(nb):
    m = 4
    S, T = _random_system(nb, m, seed=nb)
    rhs = np.random.default_rng(1).standard_normal((nb * m, 3))
    x_ref = np.linalg.solve(ob.assemble_dense(S, T), rhs)
    assert rel(ob.redblack_direct(S, T, rhs), x_ref) < 1e-12


@pytest.mark.parametrize("nb", [4, 8])
def test_cyclic_thomas_and_redblack(nb):
    m = 5
    S, T = _random_system(nb, m, seed=10 + nb, cyclic=True)
    rhs = np.random.default_rng(2).standard_normal(nb * m)
    A = ob.assemble_dense(S, T, cyclic=True)
    x_ref = np.linalg.solve(A, rhs)
    assert rel(ob.thomas_direct(S, T, rhs, cyclic=True), x_ref) < 1e-12
    assert rel(ob.redblack_direct(S, T, rhs, cyclic=True), x_ref) < 1e-12


def test_identity_thomas():
    nb, m = 6, 4
    S, _ = _random_system(nb, m, seed=3)
    T = [np.eye(m)] * nb
    rhs = np.random.default_rng(3).standard_normal((nb * m, 2))
    idx = [range(k * m, (k + 1) * m) for k in range(nb)]
    x = ob.identity_thomas_solve(ob.identity_thomas_factor(S), rhs, idx)
    assert rel(x, np.linalg.solve(ob.assemble_dense(S, T), rhs)) < 1e-12


def test_planted_mistake_is_caught():
    """A sign error in one coupling term (a plausible restatement mistake) must show
    up against the dense closed form: the check is not vacuous."""
    S, T, dense, rhs = ob.make_block_tridiagonal(8, 3, seed=42)
    S_bad = [(s[0], -s[1] if i == 3 else s[1]) for i, s in enumerate(S)]
    x_bad = ob.redblack_solve(ob.redblack_factor(S_bad, T), rhs)
    assert rel(x_bad, np.linalg.solve(dense, rhs)) > 1e-6
