"""CPU: host-side geometry of the 3-D HPS slab assembly -- merge schedules, leaf order,
merge index maps, the leaf flux rows and the node sets -- checked against brute force,
closed forms and the 3-D oracle's nodes (oracle/hps3d.py)."""
import numpy as np
import pytest

from oracle import hps3d as o3
from sslablu_amd import hps3d as h3


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 2, 2), (2, 4, 2), (1, 16, 16), (4, 1, 2)])
def test_leaf_positions_cover_slab_and_pair_neighbours(shape):
    sched = h3.merge_schedule(*shape)
    pos = h3.leaf_positions(sched)
    assert len(pos) == np.prod(shape)
    assert sorted(map(tuple, pos)) == sorted(np.ndindex(*shape))
    # every merge level pairs consecutive boxes that touch along its axis
    boxes = [np.array([p]) for p in pos]
    for axis in sched:
        new = []
        for a, b in zip(boxes[0::2], boxes[1::2]):
            assert a[:, axis].max() + 1 == b[:, axis].min()
            others = [d for d in range(3) if d != axis]
            assert np.array_equal(np.unique(a[:, others], axis=0), np.unique(b[:, others], axis=0))
            new.append(np.concatenate([a, b]))
        boxes = new
    assert len(boxes) == 1


@pytest.mark.parametrize("q", [4, 5, 8])
@pytest.mark.parametrize("shape,axis", [((1, 1, 1), 1), ((1, 2, 1), 2), ((2, 2, 4), 0),
                                        ((1, 4, 2), 1)])
def test_merge_maps_by_brute_force(q, shape, axis):
    ca, cb, na, nbm, new = h3.merge_maps3(shape, axis, q)
    Q = q - 1
    off = [0, 0, 0]
    off[axis] = shape[axis] * Q
    ka = h3.box_nodes(shape, q)
    kb = h3.box_nodes(shape, q, off)
    kn = h3.box_nodes(new, q)
    assert np.array_equal(ka[ca], kb[cb])                   # the shared face, same order
    assert np.all(ka[ca][:, axis] == shape[axis] * Q)
    for i, k in enumerate(kn):
        src = ka[na[i]] if na[i] >= 0 else kb[nbm[i]]
        assert np.array_equal(src, k)
    # every non-shared node of alpha and beta is a node of the merged box exactly once
    assert sorted(na[na >= 0].tolist() + ca.tolist()) == list(range(len(ka)))
    assert sorted(nbm[nbm >= 0].tolist() + cb.tolist()) == list(range(len(kb)))
    p = q - 2
    assert len(kn) == 2 * p * p * (new[0] * new[1] + new[1] * new[2] + new[0] * new[2])


def _leaf_grid_values(f, q, h):
    t = h3.cheb_unit(q)[0]
    A, B, C = np.meshgrid(t * h[0], t * h[1], t * h[2], indexing="ij")
    return A, B, C, f(A, B, C)


@pytest.mark.parametrize("q", [5, 8])
def test_leaf_flux_rows_exact_on_polynomials(q):
    """Nb u_b + Ni u_i is the outward normal derivative at every face node for a
    polynomial of degree <= q - 1 per variable."""
    h = (0.3, 0.5, 0.25)
    c = h3.leaf_constants3(q, *h)
    p = q - 2
    f = lambda x, y, z: x ** 3 * y - 2 * y ** 2 * z + x * z ** 2 + 1.5 * z ** 3
    fx = lambda x, y, z: 3 * x ** 2 * y + z ** 2
    fy = lambda x, y, z: x ** 3 - 4 * y * z
    fz = lambda x, y, z: -2 * y ** 2 + 2 * x * z + 4.5 * z ** 2
    A, B, C, U = _leaf_grid_values(f, q, h)
    K = h3.box_nodes((1, 1, 1), q)
    ub = U[K[:, 0], K[:, 1], K[:, 2]]
    ui = U[1:-1, 1:-1, 1:-1].ravel()
    flux = c["Nb"] @ ub + c["Ni"] @ ui
    X, Y, Z = A[K[:, 0], K[:, 1], K[:, 2]], B[K[:, 0], K[:, 1], K[:, 2]], C[K[:, 0], K[:, 1], K[:, 2]]
    want = np.empty(len(K))
    p2 = p * p
    for face in range(6):
        d, hi = divmod(face, 2)
        s = slice(face * p2, (face + 1) * p2)
        g = (fx, fy, fz)[d](X[s], Y[s], Z[s])
        want[s] = g if hi else -g
    assert np.allclose(flux, want, atol=1e-9 * np.abs(want).max())


def test_grid_nodes_match_oracle():
    """Interface and global-boundary nodes of SlabGrid3D are the oracle's Ic / Igb point
    sets (the S-system unknowns in the same (y, z) order)."""
    N, q = 3, 5
    g = h3.SlabGrid3D(N, 1, 2, 2, q)
    res = o3.oms_stot(o3_laplace(), N, (1 / 3, 1 / 2, 1 / 2), q)
    assert np.allclose(g.interface_points(), np.concatenate(res["Ic_pts"]), atol=1e-15)
    gb = g.gb_points()
    assert np.all(o3.gb_unit_cube(gb))
    allgb = np.unique(np.round(np.concatenate(res["XXgb_list"]), 13), axis=0)
    assert np.array_equal(np.unique(np.round(gb, 13), axis=0), allgb)
    assert len(gb) == len(np.unique(np.round(gb, 13), axis=0))        # no duplicates


def o3_laplace():
    class P:
        c11 = c22 = c33 = staticmethod(lambda P: np.ones(P.shape[0]))
        c1 = c2 = c3 = c = None
    return P()
