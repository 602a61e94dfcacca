"""Pin the strong-admissibility HBS oracle (CPU only) and the cluster-tree adjacency lists
it consumes: exact recovery of operators that the strong format holds exactly (line,
rectangle 'partial' / 'full', quad trees), the closed form of the near-field error tiles,
far-field annihilation, planted mistakes that those checks catch, and slabTree's
level_adj_list against brute-force geometry of its boxes."""
import numpy as np
import pytest

from oracle import hbs_strong as hs
from sslablu_amd.matAssembly.HBS.slabTree import slabTree


def grid(nx, ny):
    g = np.stack(np.meshgrid((np.arange(nx) + 0.5) / nx, (np.arange(ny) + 0.5) / ny,
                             indexing="ij"), -1)
    return g.reshape(-1, 2)


def trees():
    return {
        "line": slabTree((np.arange(128) + 0.5) / 128, False, 8),
        "rect_partial": slabTree(grid(16, 8), False, 8, adjacency="partial"),
        "rect_full": slabTree(grid(16, 16), False, 16, adjacency="full"),
        "quad": slabTree(grid(16, 16), True, 16),
    }


def bounds(tree, lvl):
    return [tree.get_node(g).bounds for g in tree.get_boxes_level(lvl)]


@pytest.mark.parametrize("name", ["line", "rect_partial", "rect_full", "quad"])
def test_adjacency_matches_brute_force_geometry(name):
    t = trees()[name]
    edge_only = name in ("rect_partial", "quad")
    ref = hs.brute_adjacency([bounds(t, l) for l in range(t.nlevels)], edge_only)
    assert [[set(a) for a in lv] for lv in t.level_adj_list] == ref
    # symmetric, self excluded
    for lv in t.level_adj_list:
        for i, a in enumerate(lv):
            assert i not in a and all(i in lv[j] for j in a)


def test_partial_and_full_differ_by_the_corner_neighbours():
    P = grid(16, 16)
    tp = slabTree(P, False, 16, adjacency="partial")
    tf = slabTree(P, False, 16, adjacency="full")
    L = tp.nlevels - 1
    assert max(len(a) for a in tp.level_adj_list[L]) == 4
    assert max(len(a) for a in tf.level_adj_list[L]) == 8
    assert all(a <= b for a, b in zip(tp.level_adj_list[L], tf.level_adj_list[L]))


def _case(name, rk=4, seed=0):
    t = trees()[name]
    adj = t.level_adj_list
    fac = 4 if name == "quad" else 2
    b = len(t.perm_leaf) // t.nleaves
    A = hs.exact_strong_operator(adj, b, rk, fac, seed=seed)
    s = hs.sample_count(adj, b, rk, fac=fac)
    rng = np.random.default_rng(seed + 1)
    N = A.shape[0]
    Om, Psi = rng.standard_normal((N, s)), rng.standard_normal((N, s))
    return A, adj, fac, b, Om, Psi


@pytest.mark.parametrize("basis", ["svd", "pqr"])
@pytest.mark.parametrize("name", ["line", "rect_partial", "rect_full", "quad"])
def test_exact_recovery(name, basis):
    A, adj, fac, b, Om, Psi = _case(name)
    H = hs.construct(Om, Psi, A @ Om, A.T @ Psi, adj, 4, fac, basis=basis)
    assert len(H["levels"]) >= 2                     # several compressed levels
    X = np.random.default_rng(3).standard_normal((A.shape[0], 3))
    nA = np.linalg.norm(A)
    assert np.linalg.norm(hs.apply(H, X) - A @ X) / np.linalg.norm(A @ X) < 1e-10
    assert np.linalg.norm(hs.apply(H, X, "T") - A.T @ X) / np.linalg.norm(A.T @ X) < 1e-10
    assert np.linalg.norm(hs.dense(H, A.shape[0]) - A) / nA < 1e-10


def test_error_tiles_closed_form_and_far_field_annihilation():
    A, adj, fac, b, Om, Psi = _case("rect_full")
    H = hs.construct(Om, Psi, A @ Om, A.T @ Psi, adj, 4, fac)
    U, V, E, nf = H["levels"][0]
    nb = len(U)
    blk = lambda i, j: A[i * b:(i + 1) * b, j * b:(j + 1) * b]  # noqa: E731
    for i in range(nb):
        for j in range(nb):
            Aij = blk(i, j)
            if j in nf[i]:
                ref = Aij - U[i] @ (U[i].T @ Aij @ V[j]) @ V[j].T
                assert np.linalg.norm(E[(i, j)] - ref) <= 1e-10 * np.linalg.norm(Aij)
            else:
                assert np.linalg.norm(Aij - U[i] @ (U[i].T @ Aij)) <= 1e-10 * np.linalg.norm(A)
                assert np.linalg.norm(Aij - (Aij @ V[j]) @ V[j].T) <= 1e-10 * np.linalg.norm(A)
        assert np.allclose(U[i].T @ U[i], np.eye(U[i].shape[1]), atol=1e-12)


@pytest.mark.parametrize("plant", ["drop_term2", "no_projection"])
def test_planted_mistakes_are_caught(plant):
    """Each planted change (inside the oracle, on the CPU) breaks the exact recovery that
    test_exact_recovery checks: the pins are not vacuous."""
    A, adj, fac, b, Om, Psi = _case("rect_partial")
    H = hs.construct(Om, Psi, A @ Om, A.T @ Psi, adj, 4, fac, _plant=plant)
    assert np.linalg.norm(hs.dense(H, A.shape[0]) - A) / np.linalg.norm(A) > 1e-3


def test_weak_admissibility_is_the_special_case_without_neighbours():
    """With empty neighbour lists every off-diagonal block is far field: the oracle's
    strong construction then reproduces an exactly weak-HBS operator (diagonal near
    field only) to ~1e-12 -- the weak format is the k = 1 case."""
    L, nb, b, rk = 4, 16, 12, 3
    adj = [[set() for _ in range(2 ** l)] for l in range(L + 1)]
    A = hs.exact_strong_operator(adj, b, rk, 2, seed=5)
    s = hs.sample_count(adj, b, rk)
    rng = np.random.default_rng(6)
    Om, Psi = rng.standard_normal((nb * b, s)), rng.standard_normal((nb * b, s))
    H = hs.construct(Om, Psi, A @ Om, A.T @ Psi, adj, rk, 2)
    assert np.linalg.norm(hs.dense(H, nb * b) - A) / np.linalg.norm(A) < 1e-10
