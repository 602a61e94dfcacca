"""3-D Chebyshev HPS slab assembly on the GPU: leaves -> slab DtN maps -> pair factors.

Geometry: the box [x0, x0+W] x [y0, y0+Hy] x [z0, z0+Hz] (default the unit cube) cut
into N single slabs of width H = W / N along x; every single slab is tiled by
nx x ny x nz leaves of q^3 Chebyshev points (p = q - 2 face nodes per leaf side, edges
and corners excluded).  The interface between single slabs i and i+1 carries
m = ny nz p^2 nodes -- the centre plane Ic of double slab i (multislab/oms.py:38-52,
the double slabs of geometry/geom_3D/cube.py:32-46), ordered lexicographically by
(y, z).

The reference's 3-D drivers discretize each double slab with HPS leaves of the absent
hpsalt submodule and build the S maps as sparse-solve linear operators, which the
rkHBS assembler then sketches (solver/solver.py:120-144, multislab/oms.py:99-129,
matAssembly/matAssembler.py:103-133, test/3D/cubeTest.py:114-126).  Here every single
slab is discretized once, on the GPU:

1. leaf DtN maps for all leaves in one batch (sslab_hps3_leaf_dtn);
2. binary merge trees over each slab's leaves (sslab_hps_merge, batched over all
   slabs), alternating the axes so that boxes stay near cubic -> one DtN map T_s per
   single slab, boundary ordered [x-lo | x-hi | y-lo | y-hi | z-lo | z-hi];
3. per interface i the pair factor: the explicit inverse of T^i_RR + T^{i+1}_LL (batched
   block-elimination inverse on MFMA GEMMs, checked, pivoted LU as its fallback), and the
   coupling blocks T^i_RL, T^{i+1}_LR and the side-face blocks, gathered once.

The S maps are then operators, as in the reference (never formed densely unless asked):
    S_l(i) v = (T^i_RR + T^{i+1}_LL)^{-1} T^i_RL v,    S_r(i) v = (...)^{-1} T^{i+1}_LR v,
applied as GEMMs with those inverses (transposes through the transposed operand), and the
S-system right-hand side for boundary data F is -(...)^{-1} (sum of the side-face and
end-face terms).  Global boundary data are ordered [side faces of slab 0 | ... | side
faces of slab N-1 | left end face | right end face], each slab's side faces as
[y-lo | y-hi | z-lo | z-hi], so double slab i reads one contiguous row range.

All arithmetic runs in libsslab_hip.so; this module prepares constants, index maps and
coefficient samples (user callables evaluated on the host).
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, sizes2, stream_handle
from .hps import cheb_unit

F64 = torch.float64
I32 = torch.int32

_AXES = "xyz"


def _is_pow2(k):
    return k >= 1 and (k & (k - 1)) == 0


def merge_schedule(nx, ny, nz):
    """Axes of the successive binary merges of a slab of nx x ny x nz leaves: x first,
    then whichever of y, z has the fewer merged leaves so far (y on ties)."""
    sched = []
    b = [1, 1, 1]
    n = [nx, ny, nz]
    while b[0] < n[0]:
        sched.append(0)
        b[0] *= 2
    while b[1] < n[1] or b[2] < n[2]:
        d = 1 if (b[1] < n[1] and (b[1] <= b[2] or b[2] >= n[2])) else 2
        sched.append(d)
        b[d] *= 2
    return sched


def leaf_positions(sched):
    """(2^L, 3) leaf positions in merge order: bit t of the leaf index steps along the
    axis of merge t, so every merge pairs consecutive boxes."""
    L = len(sched)
    k = np.arange(1 << L)
    pos = np.zeros((1 << L, 3), dtype=np.int64)
    seen = [0, 0, 0]
    for t, d in enumerate(sched):
        pos[:, d] += ((k >> t) & 1) << seen[d]
        seen[d] += 1
    return pos


def box_nodes(shape, q, origin=(0, 0, 0)):
    """Integer lattice keys (n, 3) of the boundary nodes of a box of shape[0] x shape[1]
    x shape[2] leaves, in the canonical order: faces [x-lo, x-hi, y-lo, y-hi, z-lo,
    z-hi], each in lexicographic order of its two tangential coordinates.  A leaf spans
    q - 1 lattice units; face nodes have tangential local indices 1..q-2."""
    Q = q - 1
    o = np.asarray(origin, dtype=np.int64)

    def fc(nl):
        return (np.arange(nl)[:, None] * Q + np.arange(1, q - 1)[None, :]).ravel()

    out = []
    for d in range(3):
        t0, t1 = [e for e in range(3) if e != d]
        A, B = np.meshgrid(fc(shape[t0]), fc(shape[t1]), indexing="ij")
        for end in (0, shape[d] * Q):
            K = np.empty((A.size, 3), dtype=np.int64)
            K[:, d] = end
            K[:, t0] = A.ravel()
            K[:, t1] = B.ravel()
            out.append(K + o)
    return np.concatenate(out)


def _face_ranges(shape, p):
    """[start, stop) of each of the six faces in box_nodes order."""
    sizes = []
    for d in range(3):
        t0, t1 = [e for e in range(3) if e != d]
        f = shape[t0] * shape[t1] * p * p
        sizes += [f, f]
    off = np.concatenate([[0], np.cumsum(sizes)])
    return [(int(off[k]), int(off[k + 1])) for k in range(6)]


def _encode(K, big=1 << 20):
    return (K[:, 0] * big + K[:, 1]) * big + K[:, 2]


def merge_maps3(shape, axis, q):
    """Index maps of merging two boxes of `shape` leaves along `axis` (alpha below
    beta): (ca, cb, na, nb, new_shape) in the layout of sslab_hps_merge."""
    Q = q - 1
    off = [0, 0, 0]
    off[axis] = shape[axis] * Q
    ka = box_nodes(shape, q)
    kb = box_nodes(shape, q, off)
    new = list(shape)
    new[axis] *= 2
    kn = box_nodes(new, q)
    pos_a = {int(c): i for i, c in enumerate(_encode(ka))}
    pos_b = {int(c): i for i, c in enumerate(_encode(kb))}
    p = q - 2
    s, e = _face_ranges(shape, p)[2 * axis + 1]        # alpha's upper face along axis
    ca = np.arange(s, e)
    cb = np.array([pos_b[int(c)] for c in _encode(ka[s:e])])
    en = _encode(kn)
    na = np.array([pos_a.get(int(c), -1) for c in en])
    nbm = np.array([pos_b.get(int(c), -1) for c in en])
    assert np.all((na >= 0) != (nbm >= 0)), "merged node must come from exactly one box"
    return ca, cb, na, nbm, tuple(new)


def leaf_constants3(q, hx, hy, hz):
    """Derivative matrices of one leaf and the outward-normal flux rows on its faces
    (Nb: 6p^2 x 6p^2 on the face nodes, Ni: 6p^2 x p^3 on the interior nodes), in the
    face order of box_nodes((1, 1, 1), q) and the interior order ((a-1) p + (b-1)) p +
    (c-1)."""
    t, D = cheb_unit(q)
    p = q - 2
    h = (hx, hy, hz)
    D1 = [D / h[d] for d in range(3)]
    p2 = p * p
    Nb = np.zeros((6 * p2, 6 * p2))
    Ni = np.zeros((6 * p2, p2 * p))
    ii = np.arange(p)
    U, V = np.meshgrid(ii, ii, indexing="ij")
    U, V = U.ravel(), V.ravel()
    for d in range(3):
        t0, t1 = [e for e in range(3) if e != d]
        for side, end, sgn in ((0, 0, -1.0), (1, q - 1, 1.0)):
            rows = (2 * d + side) * p2 + np.arange(p2)
            for s in range(q):
                coef = sgn * D1[d][end, s]
                if s == 0 or s == q - 1:            # the face nodes at either end
                    col = (2 * d + (0 if s == 0 else 1)) * p2 + np.arange(p2)
                    Nb[rows, col] += coef
                else:                               # interior node (s along axis d)
                    g = np.zeros((p2, 3), dtype=np.int64)
                    g[:, d] = s - 1
                    g[:, t0] = U
                    g[:, t1] = V
                    col = (g[:, 0] * p + g[:, 1]) * p + g[:, 2]
                    Ni[rows, col] += coef
    return dict(t=t, Dx=D1[0], Dxx=D1[0] @ D1[0], Dy=D1[1], Dyy=D1[1] @ D1[1], Dz=D1[2],
                Dzz=D1[2] @ D1[2], Nb=Nb, Ni=Ni)


class SlabGrid3D:
    """Leaf tiling of the box origin + [0, W] x [0, Hy] x [0, Hz] by N single slabs
    (along x) of nx x ny x nz leaves with q^3 points each."""

    def __init__(self, N, nx, ny, nz, q, origin=(0.0, 0.0, 0.0), size=(1.0, 1.0, 1.0),
                 periodic=False):
        if N < 2 or (periodic and N < 3):
            raise ValueError("need at least two single slabs (one interface), three when "
                             "periodic")
        if not (_is_pow2(nx) and _is_pow2(ny) and _is_pow2(nz)):
            raise NotImplementedError("leaves per slab along each axis must be powers of two")
        if q < 3:
            raise ValueError("q >= 3 Chebyshev points per leaf side")
        self.N, self.nx, self.ny, self.nz, self.q = N, nx, ny, nz, q
        self.p = q - 2
        self.origin = np.asarray(origin, dtype=np.float64)
        self.size = np.asarray(size, dtype=np.float64)
        self.H = self.size[0] / N
        self.hx, self.hy, self.hz = self.H / nx, self.size[1] / ny, self.size[2] / nz
        self.m = ny * nz * self.p ** 2
        # periodic in x: N interfaces, interface k on x0 + k H between single slabs
        # (k-1) mod N and k (geometry/geom_3D/squareTorus.py:268-277)
        self.periodic = bool(periodic)
        self.nb = N if periodic else N - 1
        self.sched = merge_schedule(nx, ny, nz)
        self.pos = leaf_positions(self.sched)        # (leaves per slab, 3)
        self.leaves_per_slab = nx * ny * nz
        self.nleaves = N * self.leaves_per_slab
        self.t = cheb_unit(q)[0]
        shape = (nx, ny, nz)
        self.faces = _face_ranges(shape, self.p)
        self.n_s = self.faces[-1][1]
        self.n_side = self.n_s - 2 * self.m           # y/z faces of one single slab

    def leaf_origins(self):
        s = np.repeat(np.arange(self.N), self.leaves_per_slab)
        P = np.tile(self.pos, (self.N, 1))
        x = self.origin[0] + s * self.H + P[:, 0] * self.hx
        y = self.origin[1] + P[:, 1] * self.hy
        z = self.origin[2] + P[:, 2] * self.hz
        return np.stack([x, y, z], 1)

    def interior_points(self):
        """(nleaves * p^3, 3) leaf-interior collocation nodes, leaf-major, row
        ((a-1) p + (b-1)) p + (c-1) inside a leaf."""
        O = self.leaf_origins()
        ti = self.t[1:-1]
        A, B, C = np.meshgrid(ti, ti, ti, indexing="ij")
        loc = np.stack([A.ravel() * self.hx, B.ravel() * self.hy, C.ravel() * self.hz], 1)
        return (O[:, None, :] + loc[None, :, :]).reshape(-1, 3)

    def _coords(self, K, slab):
        """Points of lattice keys K of single slab `slab`."""
        Q = self.q - 1
        cell, a = np.divmod(K, Q)
        h = np.array([self.hx, self.hy, self.hz])
        base = self.origin + np.array([slab * self.H, 0.0, 0.0])
        return base + (cell + self.t[a]) * h

    def slab_boundary_points(self, slab):
        return self._coords(box_nodes((self.nx, self.ny, self.nz), self.q), slab)

    def gb_points(self):
        """Global boundary nodes: [sides of slab 0 | ... | sides of slab N-1 | x = x0
        face | x = x0 + W face], each slab's sides [y-lo | y-hi | z-lo | z-hi] (periodic:
        the sides only)."""
        m = self.m
        out = [self.slab_boundary_points(s)[2 * m:] for s in range(self.N)]
        if not self.periodic:
            out.append(self.slab_boundary_points(0)[:m])
            out.append(self.slab_boundary_points(self.N - 1)[m:2 * m])
        return np.concatenate(out)

    def interface_points(self):
        """(nb m, 3) interface nodes in (y, z) order: interface i on x = x0 + (i+1) H
        (periodic: on x0 + i H)."""
        m = self.m
        if self.periodic:
            return np.concatenate([self.slab_boundary_points(i)[:m] for i in range(self.nb)])
        return np.concatenate([self.slab_boundary_points(i)[m:2 * m] for i in range(self.nb)])


def _coef(fn, P):
    if fn is None:
        return None
    v = fn(P)
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().numpy()
    v = np.asarray(v, dtype=np.float64)
    if v.ndim == 0:
        v = np.full(P.shape[0], float(v))
    return v.reshape(-1)


class HPSSlabOperator3D:
    """Device-resident oms S-system of one PDE on one SlabGrid3D, in factored form.

    After build(): per interface i (batched over nb), the LU factors of
    Tcc_i = T^i_RR + T^{i+1}_LL (`lu`, `piv`), the coupling blocks Bl_i = T^i_RL and
    Br_i = T^{i+1}_LR (`Bl`, `Br`, (nb, m, m)) and the side-face blocks
    G_i = [T^i_R,side | T^{i+1}_L,side] (`G`, (nb, m, 2 n_side)).  S maps apply as
    S_l(i) = Tcc_i^{-1} Bl_i, S_r(i) = Tcc_i^{-1} Br_i (smap / smaps_apply), dense
    copies on request (dense_blocks)."""

    def __init__(self, pdo, grid: SlabGrid3D, device=None, slab_chunk=None):
        _lib.require_gpu()
        _lib.load()
        self.pdo = pdo
        self.g = grid
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.slab_chunk = slab_chunk or grid.N
        self.info_bad = 0
        self._infos = []

    def _dev(self, a, dtype=F64):
        return torch.as_tensor(np.ascontiguousarray(a)).to(device=self.device, dtype=dtype)

    # -- 1. leaves -------------------------------------------------------------------
    def leaf_dtn(self, slabs, stream=None):
        """DtN maps (n, 6p^2, 6p^2) of every leaf of the single slabs `slabs` (a range)."""
        g = self.g
        q, p = g.q, g.p
        c = leaf_constants3(q, g.hx, g.hy, g.hz)
        lps = g.leaves_per_slab
        P = g.interior_points().reshape(g.N, lps * p ** 3, 3)[slabs.start:slabs.stop]
        P = P.reshape(-1, 3)
        coefs = []
        for name in ("c11", "c22", "c33", "c1", "c2", "c3", "c"):
            v = _coef(getattr(self.pdo, name, None), P)
            coefs.append(self._dev(v) if v is not None else None)
        consts = {k: self._dev(c[k]) for k in ("Dx", "Dxx", "Dy", "Dyy", "Dz", "Dzz", "Nb", "Ni")}
        nl = len(slabs) * lps
        nf = 6 * p * p
        T = torch.empty(nl, nf, nf, dtype=F64, device=self.device)
        wd, wi = sizes2("sslab_hps3_leaf_sizes", nl, q)
        work = torch.empty(wd, dtype=F64, device=self.device)
        iwork = torch.empty(wi, dtype=I32, device=self.device)
        info = torch.zeros(nl, dtype=I32, device=self.device)
        cp = [ptr(x) if x is not None else None for x in coefs]
        call("sslab_hps3_leaf_dtn", nl, q, *[ptr(consts[k]) for k in
                                              ("Dx", "Dxx", "Dy", "Dyy", "Dz", "Dzz", "Nb", "Ni")],
             *cp, ptr(T), None, ptr(work), ptr(iwork), ptr(info), stream_handle(stream))
        self._infos.append(info)
        return T

    # -- 2. merges -------------------------------------------------------------------
    def merge_tree(self, T, stream=None):
        """Merge each slab's leaves (consecutive in T) into one DtN map per slab."""
        g = self.g
        shape = (1, 1, 1)
        boxes = T
        for axis in g.sched:
            ca, cb, na, nbm, new = merge_maps3(shape, axis, g.q)
            n_old = boxes.shape[1]
            n_new = len(na)
            nmerge = boxes.shape[0] // 2
            c = len(ca)
            maps = [self._dev(x, I32) for x in (ca, cb, na, nbm)]
            out = torch.empty(nmerge, n_new, n_new, dtype=F64, device=self.device)
            wd, wi = sizes2("sslab_hps_merge_sizes", nmerge, n_new, c)
            work = torch.empty(wd, dtype=F64, device=self.device)
            iwork = torch.empty(wi, dtype=I32, device=self.device)
            info = torch.zeros(nmerge, dtype=I32, device=self.device)
            s_old = n_old * n_old
            call("sslab_hps_merge", nmerge, n_old, n_old, n_new, c, ptr(boxes), 2 * s_old,
                 ptr(boxes) + 8 * s_old, 2 * s_old, *[ptr(x) for x in maps], ptr(out),
                 n_new * n_new, ptr(work), ptr(iwork), ptr(info), stream_handle(stream))
            self._infos.append(info)
            del work, iwork
            boxes = out
            shape = new
        return boxes

    def slab_dtn(self, stream=None):
        """(N, n_s, n_s) DtN map of every single slab, `slab_chunk` slabs per pass."""
        g = self.g
        Ts = torch.empty(g.N, g.n_s, g.n_s, dtype=F64, device=self.device)
        for s0 in range(0, g.N, self.slab_chunk):
            sl = range(s0, min(g.N, s0 + self.slab_chunk))
            T = self.leaf_dtn(sl, stream)
            Ts[sl.start:sl.stop] = self.merge_tree(T, stream)
            del T
        return Ts

    # -- 3. pair factors -------------------------------------------------------------
    def pair_factor(self, Ts, stream=None):
        g = self.g
        if g.periodic:
            # interface 0 joins the last single slab to the first: factor it from a copy of
            # those two DtN maps, the others as in the chain, interface 0 first
            reg, wrap = self._pair_factor(Ts, g.N - 1, stream), None
            Tw = torch.stack([Ts[g.N - 1], Ts[0]])
            wrap = self._pair_factor(Tw, 1, stream)
            del Tw
            self.tinv, self.Bl, self.Br, self.G = (torch.cat([w, r]) for w, r in zip(wrap, reg))
            return
        self.tinv, self.Bl, self.Br, self.G = self._pair_factor(Ts, g.nb, stream)

    def _pair_factor(self, Ts, nb, stream=None):
        g = self.g
        m, n_s, ns = g.m, g.n_s, g.n_side
        dev = self.device
        s = stream_handle(stream)
        L = np.arange(m)
        R = np.arange(m, 2 * m)
        side = np.arange(2 * m, n_s)
        neg_m = -np.ones(m, dtype=np.int64)
        dR, dL, dside = (self._dev(x, I32) for x in (R, L, side))
        sT = n_s * n_s
        lu = torch.empty(nb, m, m, dtype=F64, device=dev)
        # Tcc_i = T^i[R, R] + T^{i+1}[L, L]
        call("sslab_gather2", nb, m, m, ptr(Ts), n_s, sT, ptr(Ts) + 8 * sT, n_s, sT,
             ptr(dR), ptr(dR), ptr(dL), ptr(dL), ptr(lu), m, m * m, s)
        Bl = torch.empty(nb, m, m, dtype=F64, device=dev)
        Br = torch.empty(nb, m, m, dtype=F64, device=dev)
        call("sslab_gather2", nb, m, m, ptr(Ts), n_s, sT, None, 0, 0, ptr(dR), ptr(dL), None,
             None, ptr(Bl), m, m * m, s)
        call("sslab_gather2", nb, m, m, ptr(Ts) + 8 * sT, n_s, sT, None, 0, 0, ptr(dL), ptr(dR),
             None, None, ptr(Br), m, m * m, s)
        # G_i = [T^i[R, side] | T^{i+1}[L, side]]
        G = torch.empty(nb, m, 2 * ns, dtype=F64, device=dev)
        call("sslab_gather2", nb, m, ns, ptr(Ts), n_s, sT, None, 0, 0, ptr(dR), ptr(dside),
             None, None, ptr(G), 2 * ns, m * 2 * ns, s)
        call("sslab_gather2", nb, m, ns, ptr(Ts) + 8 * sT, n_s, sT, None, 0, 0, ptr(dL),
             ptr(dside), None, None, ptr(G) + 8 * ns, 2 * ns, m * 2 * ns, s)
        # (T^i_RR + T^{i+1}_LL)^{-1}: every later solve with it is then one GEMM (the LU's
        # one-workgroup panels over m rows were most of the 3-D assembly time)
        tinv = torch.empty(nb, m, m, dtype=F64, device=dev)
        wd, wi = sizes2("sslab_inv_sizes", m, nb)
        work = torch.empty(max(wd, 1), dtype=F64, device=dev)
        iwork = torch.empty(max(wi, 1), dtype=I32, device=dev)
        info = torch.zeros(nb, dtype=I32, device=dev)
        call("sslab_dinv_batched", m, ptr(lu), m, m * m, ptr(tinv), m, m * m, nb, ptr(work),
             ptr(iwork), ptr(info), 0.0, None, s)
        self._infos.append(info)
        del neg_m, work, iwork, lu
        return tinv, Bl, Br, G

    def build(self, stream=None):
        """Run the whole assembly; returns self."""
        with torch.cuda.device(self.device):
            Ts = self.slab_dtn(stream)
            self.pair_factor(Ts, stream)
            del Ts
            self.info_bad = int(sum(int((i != 0).sum()) for i in self._infos))
            self._infos = []
        if self.info_bad:
            warnings.warn(f"{self.info_bad} singular local solve(s) in the 3-D HPS assembly "
                          "(a leaf or slab Dirichlet problem at resonance?)", RuntimeWarning)
        return self

    # -- applies ---------------------------------------------------------------------
    def _solve(self, B, trans=False, stream=None):
        """Tcc_i^{-1} B_i (or Tcc_i^{-T} B_i) for B (nb, m, r) contiguous; new tensor
        (one batched GEMM with the inverses)."""
        from .hbs import bgemm
        return bgemm(self.tinv, B, ta=trans)

    def smaps_apply(self, side, X, transpose=False, stream=None, fp32=False):
        """S_l(i) X_i (side 'l') or S_r(i) X_i ('r') for every interface at once; X
        (nb, m, r).  transpose: S^T X = B^T Tcc^{-T} X.  fp32: the coupling-block product
        on the FP32 MFMA GEMM (B and its operand rounded to float, FP32 accumulation);
        the LU solves stay FP64."""
        from .hbs import bgemm
        B = self.Bl if side == "l" else self.Br
        X = X.to(self.device, F64).contiguous()
        prod = self._bgemm32 if fp32 else (lambda A, Y, ta: bgemm(A, Y, ta=ta))
        if transpose:
            return prod(B, self._solve(X, True, stream), True)
        return self._solve(prod(B, X, False), False, stream)

    def _bgemm32(self, B, X, ta):
        """op(B_i) X_i in FP32 for (nb, m, m) B and (nb, m, r) X; returns float64."""
        nb, m, _ = B.shape
        r = X.shape[2]
        s = stream_handle()
        f32 = torch.float32

        def cast(t):
            y = torch.empty(t.shape, dtype=f32, device=self.device)
            call("sslab_cast_f64_to_f32", t.numel(), ptr(t), ptr(y), s)
            return y

        B32, X32 = cast(B.contiguous()), cast(X)
        Y32 = torch.empty(nb, m, r, dtype=f32, device=self.device)
        call("sslab_sgemm_strided_batched", int(ta), 0, m, r, m, 1.0, ptr(B32), m, m * m,
             ptr(X32), r, m * r, 0.0, ptr(Y32), r, m * r, nb, s)
        del B32, X32
        return Y32.to(F64)

    def smap(self, i, side):
        """The S map of interface i, side 'l' (from the left neighbour's interface) or 'r',
        as a device linear operator (matmat / rmatmat on (m, r) device tensors)."""
        return PairSMap(self, i, side)

    def dense_blocks(self, stream=None):
        """(SiM, SiP): dense (nb, m, m) S blocks (SiM_i = S_l(i), SiP_i = S_r(i)); the
        end blocks SiM_0 and SiP_{nb-1} couple to the end faces' boundary data."""
        return (self._solve(self.Bl.contiguous(), False, stream),
                self._solve(self.Br.contiguous(), False, stream))

    def rhs_from_boundary(self, F: torch.Tensor, out=None, stream=None):
        """S-system right-hand sides rhs_i = -(A_ii^{-1} A_ib[:, Igb] f)[Ic] for r columns
        of global boundary data F (ngb x r, order of SlabGrid3D.gb_points())."""
        from .hbs import bgemm, gdesc, ggemm
        g = self.g
        m, nb, ns, N = g.m, g.nb, g.n_side, g.N
        if F.dim() == 1:
            F = F[:, None]
        F = F.to(device=self.device, dtype=F64).contiguous()
        r = F.shape[1]
        ngb = N * ns + (0 if g.periodic else 2 * m)
        if F.shape[0] != ngb:
            raise ValueError(f"boundary data needs {ngb} rows, got {F.shape[0]}")
        W = torch.empty(nb, m, r, dtype=F64, device=self.device)
        if g.periodic:
            # interface k reads the side faces of slabs (k-1) mod N and k: gathered, since
            # the wrapping pair's rows are not one window
            k = np.arange(N)
            rows = np.concatenate([((k - 1) % N)[:, None] * ns + np.arange(ns),
                                   k[:, None] * ns + np.arange(ns)], axis=1)
            Fw = F.index_select(0, torch.as_tensor(rows.ravel(), device=self.device))
            ggemm(gdesc(W, [(self.G, Fw.view(nb, 2 * ns, r), 1.0)]), stream=stream)
        else:
            # W_i = G_i F[i ns : (i+2) ns]  (consecutive double slabs overlap by one slab)
            Fw = F.as_strided((nb, 2 * ns, r), (ns * r, r, 1))
            ggemm(gdesc(W, [(self.G, Fw, 1.0)]), stream=stream)
            left = F[N * ns:N * ns + m].reshape(1, m, r)
            right = F[N * ns + m:].reshape(1, m, r)
            bgemm(self.Bl[0:1], left, out=W[0:1], beta=1.0)
            bgemm(self.Br[nb - 1:nb], right, out=W[nb - 1:nb], beta=1.0)
        X = self._solve(W, False, stream)
        X.neg_()
        X = X.reshape(nb * m, r)
        if out is not None:
            out.copy_(X)
            return out
        return X

    def S_rk_list(self):
        """[[S_r], [S_l, S_r], ..., [S_l]] as device operators (multislab/oms.py:239-244);
        periodic: [S_l, S_r] for every interface."""
        nb = self.g.nb
        if self.g.periodic:
            return [[self.smap(i, "l"), self.smap(i, "r")] for i in range(nb)]
        out = []
        for i in range(nb):
            if i == 0:
                out.append([self.smap(0, "r")])
            elif i == nb - 1:
                out.append([self.smap(i, "l")])
            else:
                out.append([self.smap(i, "l"), self.smap(i, "r")])
        return out

    def tridiag_apply(self):
        """X -> S X for the whole S-system (N_tot x r device tensors): X_i + S_l(i)
        X_{i-1} + S_r(i) X_{i+1}, through the factors (the exact operator the reference's
        GMRES applies, multislab/oms.py:267-304)."""
        nb, m = self.g.nb, self.g.m

        periodic = self.g.periodic

        def apply(X):
            r = X.shape[1]
            X3 = X.reshape(nb, m, r)
            Zl = torch.zeros(nb, m, r, dtype=F64, device=self.device)
            Zr = torch.zeros(nb, m, r, dtype=F64, device=self.device)
            Zl[1:] = X3[:-1]
            Zr[:-1] = X3[1:]
            if periodic:
                Zl[0] = X3[nb - 1]
                Zr[nb - 1] = X3[0]
            from .hbs import bgemm
            W = bgemm(self.Bl, Zl)
            bgemm(self.Br, Zr, out=W, beta=1.0)
            return (X3 + self._solve(W)).reshape(nb * m, r)
        return apply

    @property
    def nbytes(self):
        return 8 * sum(t.numel() for t in (self.tinv, self.Bl, self.Br, self.G))


class PairSMap:
    """One S block of a 3-D slab operator as a device linear operator (the reference's
    S maps are LinearOperators of sparse solves, multislab/oms.py:99-129)."""

    device_op = True
    dtype = np.dtype(np.float64)

    def __init__(self, op: HPSSlabOperator3D, i: int, side: str):
        if side not in ("l", "r"):
            raise ValueError("side is 'l' or 'r'")
        if not 0 <= i < op.g.nb:
            raise IndexError(f"interface {i} out of range (the operator has {op.g.nb})")
        self.op, self.i, self.side = op, i, side
        m = op.g.m
        self.shape = (m, m)
        self.tree = self.quad = None

    def _run(self, X, transpose):
        op, i = self.op, self.i
        was_np = not isinstance(X, torch.Tensor)
        Xt = torch.as_tensor(np.asarray(X, dtype=np.float64)) if was_np else X
        Xt = Xt.to(op.device, F64)
        one_d = Xt.dim() == 1
        X2 = Xt[:, None] if one_d else Xt
        from .hbs import bgemm
        B = (op.Bl if self.side == "l" else op.Br)[i:i + 1]
        sub = _SubFactor(op, i)
        if transpose:
            Y = bgemm(B, sub.solve(X2.contiguous()[None], True), ta=True)[0]
        else:
            Y = sub.solve(bgemm(B, X2.contiguous()[None]), False)[0]
        Y = Y[:, 0] if one_d else Y
        return Y.cpu().numpy() if was_np else Y

    def matmat(self, X):
        return self._run(X, False)

    def rmatmat(self, X):
        return self._run(X, True)

    matvec = matmat
    rmatvec = rmatmat

    def __matmul__(self, X):
        return self._run(X, False)

    @property
    def T(self):
        """The transposed map (the reference's LinearOperator .T)."""
        outer = self

        class _Tr:
            device_op = True
            dtype = np.dtype(np.float64)
            shape = (outer.shape[1], outer.shape[0])
            tree = quad = None

            def matmat(self, X):
                return outer._run(X, True)

            def rmatmat(self, X):
                return outer._run(X, False)

            matvec = matmat
            rmatvec = rmatmat
            __matmul__ = matmat

            @property
            def T(self):
                return outer
        return _Tr()


class SideMap:
    """The boundary-data map of one 3-D double slab (interface 0 of its operator):
    columns [side faces of the first single slab | side faces of the second] ->
    (A_ii^{-1} A_ib)[Ic, side columns] = Tcc^{-1} G, a device linear operator."""

    device_op = True
    dtype = np.dtype(np.float64)

    def __init__(self, op: HPSSlabOperator3D, i: int = 0):
        if not 0 <= i < op.g.nb:
            raise IndexError(f"interface {i} out of range (the operator has {op.g.nb})")
        self.op, self.i = op, i
        self.shape = (op.g.m, op.G.shape[2])

    def matmat(self, X):
        from .hbs import bgemm
        op, i = self.op, self.i
        was_np = not isinstance(X, torch.Tensor)
        Xt = torch.as_tensor(np.asarray(X, dtype=np.float64)) if was_np else X
        Xt = Xt.to(op.device, F64)
        one_d = Xt.dim() == 1
        X2 = (Xt[:, None] if one_d else Xt).contiguous()
        Y = _SubFactor(op, i).solve(bgemm(op.G[i:i + 1], X2[None]), False)[0]
        Y = Y[:, 0] if one_d else Y
        return Y.cpu().numpy() if was_np else Y

    matvec = matmat


class _SubFactor:
    """The pair factor of one interface (a batch-of-one view of its inverse)."""

    def __init__(self, op, i):
        self.op, self.i = op, i

    def solve(self, B, trans):
        from .hbs import bgemm
        return bgemm(self.op.tinv[self.i:self.i + 1], B.contiguous(), ta=trans)


def compress_s_blocks(op: HPSSlabOperator3D, tree, rk: int, samples: int, seed: int = 0,
                      sketch_fp32: bool = False, strong: bool = False):
    """HBS compressions of every S block of `op` (matAssembly/matAssembler.py:103-133,
    the rkHBS assembler of multislab/oms.py:160-205) with the sketches of all
    interfaces formed together: Y = S Omega and Z = S^T Psi for the nb left blocks in
    one batched product and one batched (transposed) LU solve, likewise the right
    blocks; every block shares the seeded test matrices of its side.  Returns
    (H_l, H_r): lists of nb HBStorch.HBSMAT (H_l[0] and H_r[nb-1] are the end blocks,
    which couple to the end faces' boundary data, not to an interface).  sketch_fp32:
    the coupling-block products of the sketches on the FP32 MFMA GEMM (the LU solves,
    QRs and interpolative bases stay FP64).  strong: strong-admissibility compressions
    (the HBStorch_strong mirror; `tree` a slabTree over one interface's points, with its
    adjacency lists)."""
    from .matAssembly.HBS import HBStorch, HBStorch_strong
    g = op.g
    nb, m = g.nb, g.m
    dev = op.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    out = []
    for side in ("l", "r"):
        Om = torch.randn(m, samples, generator=gen, device=dev, dtype=F64)
        Psi = torch.randn(m, samples, generator=gen, device=dev, dtype=F64)
        Y = op.smaps_apply(side, Om.expand(nb, m, samples), fp32=sketch_fp32)
        Z = op.smaps_apply(side, Psi.expand(nb, m, samples), transpose=True, fp32=sketch_fp32)
        hs = []
        for i in range(nb):
            if strong:
                h = HBStorch_strong.HBSMAT(device=dev, tree=tree)
            else:
                h = HBStorch.HBSMAT(op.smap(i, side), dev, tree, quad=False)
            h.construct(rk, Om, Psi, Y[i], Z[i])
            hs.append(h)
        del Y, Z
        out.append(hs)
    return out[0], out[1]


def hbs_S_rk_list(H_l, H_r):
    """S_rk_list ([[S_r], [S_l, S_r], ..., [S_l]]) of compressed blocks."""
    nb = len(H_l)
    return [[H_r[0]]] + [[H_l[i], H_r[i]] for i in range(1, nb - 1)] + [[H_l[nb - 1]]]
