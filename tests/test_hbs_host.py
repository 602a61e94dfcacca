"""CPU: host-side bookkeeping of the HBS engine that needs no GPU -- the leaf-permutation
cache (identity orders need no device copy and no gathers) and the event-based stage
clocks of HBSMAT (pending intervals fold into the total when it is read or reset)."""
import numpy as np


def test_identity_permutation_needs_no_device_copy():
    from sslablu_amd import hbs
    hbs._PERMS.clear()
    assert hbs.device_perm(None, "cuda") is None
    assert hbs.device_perm(np.arange(9216), "cuda") is None
    assert hbs.device_perm(list(range(5)), "cuda:0") is None
    # cached by content and device: a second identity of the same length is a hit
    n = len(hbs._PERMS)
    assert hbs.device_perm(np.arange(9216, dtype=np.int32), "cuda") is None
    assert len(hbs._PERMS) == n


class _Ev:
    """Stands in for a timing event pair: elapsed_time in milliseconds."""

    def __init__(self, t):
        self.t = t
        self.waited = False

    def synchronize(self):
        self.waited = True

    def elapsed_time(self, other):
        return other.t - self.t


def test_stage_clock_folds_pending_intervals():
    from sslablu_amd.matAssembly.HBS.HBSnew import _StageClock
    c = _StageClock()
    e1, e2, e3, e4 = _Ev(0.0), _Ev(250.0), _Ev(1000.0), _Ev(1500.0)
    c.open += [(e1, e2), (e3, e4)]
    assert abs(c.value() - 0.75) < 1e-12 and e2.waited and e4.waited
    assert c.open == [] and abs(c.value() - 0.75) < 1e-12
    c.open.append((_Ev(0.0), _Ev(100.0)))
    c.set(2.0)  # assignment drops what is pending
    assert c.value() == 2.0


def test_fp64_kernels_refuse_host_and_float32_operands():
    """The batched FP64 primitives take raw pointers: a float32 or host tensor must be
    refused on the host (TypeError) before any kernel could read past it."""
    import pytest
    import torch
    from sslablu_amd import hbs
    a = torch.zeros(2, 4, 4, dtype=torch.float64)
    with pytest.raises(TypeError):
        hbs.bgemm(a, a)
    with pytest.raises(TypeError):
        hbs.bgemm(a.float(), a.float())
    with pytest.raises(TypeError):
        hbs.gdesc(a, [(a, a, 1.0)])
    with pytest.raises(TypeError):
        hbs.bqr(a)
    with pytest.raises(TypeError):
        hbs.btri_inv(a)
    with pytest.raises(TypeError):
        hbs.DeviceHBS([a], [a, a[:1]], [a])
    with pytest.raises(TypeError):
        hbs.DeviceHBS.from_sketches(a[0], a[0], a[0], a[0], 2, 2, 1)
    with pytest.raises(TypeError):
        hbs.bgemm(np.zeros((2, 4, 4)), a)


def _tree_pts(nx, ny, x0=0.0, y0=0.0, w=1.0, h=1.0, z=None):
    xs = x0 + (np.arange(nx) + 0.5) * w / nx
    ys = y0 + (np.arange(ny) + 0.5) * h / ny
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    cols = [X.ravel(), Y.ravel()]
    if z is not None:  # a face embedded in 3-D (constant first coordinate)
        cols = [np.full(X.size, z)] + cols
    return np.stack(cols, 1)


def _brute_boxes(P, quad, leaf):
    """Leaf boxes by brute force: split the bounding rectangle at midpoints (binary:
    alternate y then x; quad: both) until a box holds <= leaf points; returns the list of
    point-index sets in the tree's leaf order (heap numbering, children left to right)."""
    out = []

    def rec(inds, lo, hi, level, idx):
        if len(inds) <= leaf:
            out.append((idx, sorted(inds.tolist())))
            return
        mid = (lo + hi) / 2
        if quad:
            sx, sy = P[inds, 0] <= mid[0], P[inds, 1] <= mid[1]
            for q, msk, l2, h2 in ((0, sx & sy, lo, mid), (1, ~sx & sy, [mid[0], lo[1]], [hi[0], mid[1]]),
                                   (2, sx & ~sy, [lo[0], mid[1]], [mid[0], hi[1]]), (3, ~sx & ~sy, mid, hi)):
                rec(inds[msk], np.array(l2, float), np.array(h2, float), level + 1, 4 * idx + 1 + q)
        else:
            ax = 1 if level % 2 == 0 else 0
            below = P[inds, ax] <= mid[ax]
            h1, l2 = hi.copy(), lo.copy()
            h1[ax] = l2[ax] = mid[ax]
            rec(inds[below], lo, h1, level + 1, 2 * idx + 1)
            rec(inds[~below], l2, hi, level + 1, 2 * idx + 2)

    rec(np.arange(len(P)), P.min(0), P.max(0), 0, 0)
    return [s for _, s in sorted(out)]


def test_slabtree_rect_and_quad_leaves_by_brute_force():
    """perm_leaf of the rectangle (binary) and quad trees equals a brute-force midpoint
    partition of the same points, leaf by leaf, for a planar face and for the same face
    embedded in 3-D at constant x (the interfaces of 3-D slabs)."""
    from sslablu_amd.matAssembly.HBS.slabTree import slabTree
    for nx, ny, leaf in ((8, 8, 4), (16, 8, 8), (12, 12, 9)):
        P2 = _tree_pts(nx, ny)
        for quad in (False, True):
            if quad and nx != ny:
                continue
            for P in (P2, _tree_pts(nx, ny, z=0.5)):
                t = slabTree(P, quad=quad, min_leaf_size=leaf)
                boxes = _brute_boxes(P2, quad, leaf)
                got = [sorted(t.get_box_inds(b).tolist()) for b in t.get_leaves()]
                assert got == boxes, (nx, ny, leaf, quad)
                assert sorted(t.perm_leaf.tolist()) == list(range(len(P)))
                assert t.perm_leaf.tolist() == [i for b in boxes for i in b]
                assert t.nleaves == len(boxes) and not t.is_line_input


def test_slabtree_line_detection():
    from sslablu_amd.matAssembly.HBS.slabTree import slabTree
    x = np.linspace(0, 1, 33)
    line = np.stack([x, 2 * x + 1], 1)
    t = slabTree(line, min_leaf_size=4)
    assert t.is_line_input and t.nleaves == 16
    import pytest
    with pytest.raises(ValueError):
        slabTree(line, quad=True)
    assert not slabTree(_tree_pts(4, 4), min_leaf_size=4).is_line_input
