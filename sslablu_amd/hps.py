"""2-D Chebyshev HPS slab assembly on the GPU: leaves -> slab DtN maps -> S-system.

Geometry: [0,1]^2 cut into N single slabs of width H = 1/N (the double slabs of
geometry/geom_2D/square.py:33-47 are neighbouring pairs of them); every single slab is
tiled by nx x ny leaves of q x q Chebyshev points (p = q - 2 edge nodes per leaf side,
corners excluded).  The interface between single slabs i and i+1 carries m = ny*p
nodes -- the centre line Ic of double slab i (multislab/oms.py:38-52).

The reference discretizes each double slab separately and builds its S maps by sparse
solves (multislab/oms.py:99-129,160-205).  Here every single slab is discretized once:

1. leaf DtN maps for all leaves in one batch (sslab_hps_leaf_dtn);
2. binary merge trees, first across (nx) then up (ny), batched over all slabs
   (sslab_hps_merge) -> one DtN map T_s per single slab, boundary ordered
   [left, right, bottom, top];
3. one batched pair solve per double slab (sslab_hps_pair_solve):
   Y_i = (T^i_RR + T^{i+1}_LL)^{-1} [T^i_RL | T^{i+1}_LR | T^i_R,bot | T^{i+1}_L,bot |
   T^i_R,top | T^{i+1}_L,top], which holds S_l = Y[:, :m], S_r = Y[:, m:2m] and the
   boundary-data maps (the reference's A_ii^{-1} A_ib[:, J] restricted to Ic).

Global boundary data are ordered [bottom (N*nx*p, by x) | top (N*nx*p) | left (m) |
right (m)], so each double slab reads contiguous row ranges.

All arithmetic runs in libsslab_hip.so; this module only prepares constants, index
maps and coefficient samples (user callables evaluated on the host).
"""
from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, sizes2, stream_handle

F64 = torch.float64
I32 = torch.int32

# read-only device constants shared by all operators, keyed by what they depend on
_DEV_CONST: dict = {}


def cheb_unit(q: int):
    """q ascending Chebyshev points on [0, 1] and the first-derivative matrix on them
    (the Chebyshev collocation matrix of Trefethen, Spectral Methods in MATLAB, 2000)."""
    n = q - 1
    x = np.cos(np.pi * np.arange(q) / n)
    c = np.ones(q)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(q)
    D = np.outer(c, 1.0 / c) / (x[:, None] - x[None, :] + np.eye(q))
    D -= np.diag(D.sum(axis=1))
    return (1.0 - x) / 2.0, -2.0 * D


@functools.lru_cache(maxsize=32)
def leaf_constants(q: int, hx: float, hy: float):
    """Derivative matrices of one leaf and its outward-normal flux rows
    (Nb: 4p x 4p on the edge nodes, Ni: 4p x p^2 on the interior nodes)."""
    t, D = cheb_unit(q)
    p = q - 2
    Dx, Dy = D / hx, D / hy
    Dxx, Dyy = Dx @ Dx, Dy @ Dy
    Nb = np.zeros((4 * p, 4 * p))
    Ni = np.zeros((4 * p, p * p))
    for k in range(p):
        b = k + 1
        for e, a0, sgn in ((k, 0, -1.0), (p + k, q - 1, 1.0)):      # left, right edges
            Nb[e, k] += sgn * Dx[a0, 0]
            Nb[e, p + k] += sgn * Dx[a0, q - 1]
            for a2 in range(1, q - 1):
                Ni[e, (a2 - 1) * p + (b - 1)] += sgn * Dx[a0, a2]
        a = k + 1
        for e, b0, sgn in ((2 * p + k, 0, -1.0), (3 * p + k, q - 1, 1.0)):  # bottom, top
            Nb[e, 2 * p + k] += sgn * Dy[b0, 0]
            Nb[e, 3 * p + k] += sgn * Dy[b0, q - 1]
            for b2 in range(1, q - 1):
                Ni[e, (a - 1) * p + (b2 - 1)] += sgn * Dy[b0, b2]
    out = dict(t=t, Dx=Dx, Dxx=Dxx, Dy=Dy, Dyy=Dyy, Nb=Nb, Ni=Ni)
    for a in out.values():
        a.setflags(write=False)  # cached: shared by every caller
    return out


def _box_layout(nxb, nyb, p):
    """Offsets of the [left, right, bottom, top] boundary segments of a box of
    nxb x nyb leaves."""
    L = nyb * p
    B = nxb * p
    return dict(left=0, right=L, bottom=2 * L, top=2 * L + B, n=2 * (L + B), L=L, B=B)


@functools.lru_cache(maxsize=256)
def merge_maps(kind, nxb, nyb, p):
    """Index maps of merging two nxb x nyb boxes across (kind 'h': alpha left of beta)
    or up (kind 'v': alpha below beta).  Returns (ca, cb, na, nb, n_new, (nxb', nyb'))."""
    s = _box_layout(nxb, nyb, p)
    ar = np.arange
    none = lambda k: -np.ones(k, dtype=np.int64)
    if kind == "h":
        ca = s["right"] + ar(s["L"])
        cb = s["left"] + ar(s["L"])
        na = np.concatenate([s["left"] + ar(s["L"]), none(s["L"]),
                             s["bottom"] + ar(s["B"]), none(s["B"]),
                             s["top"] + ar(s["B"]), none(s["B"])])
        nbm = np.concatenate([none(s["L"]), s["right"] + ar(s["L"]),
                              none(s["B"]), s["bottom"] + ar(s["B"]),
                              none(s["B"]), s["top"] + ar(s["B"])])
        new = (2 * nxb, nyb)
    else:
        ca = s["top"] + ar(s["B"])
        cb = s["bottom"] + ar(s["B"])
        na = np.concatenate([s["left"] + ar(s["L"]), none(s["L"]),
                             s["right"] + ar(s["L"]), none(s["L"]),
                             s["bottom"] + ar(s["B"]), none(s["B"])])
        nbm = np.concatenate([none(s["L"]), s["left"] + ar(s["L"]),
                              none(s["L"]), s["right"] + ar(s["L"]),
                              none(s["B"]), s["top"] + ar(s["B"])])
        new = (nxb, 2 * nyb)
    for a in (ca, cb, na, nbm):
        a.setflags(write=False)  # cached: shared by every caller
    return ca, cb, na, nbm, len(na), new


@functools.lru_cache(maxsize=64)
def pair_maps(nx, ny, p):
    """Index maps of the double-slab pair solve: shared edge and the column layout
    [alpha.left | beta.right | alpha.bottom | beta.bottom | alpha.top | beta.top]."""
    s = _box_layout(nx, ny, p)
    ar = np.arange
    none = lambda k: -np.ones(k, dtype=np.int64)
    m, B = s["L"], s["B"]
    ca = s["right"] + ar(m)
    cb = s["left"] + ar(m)
    cola = np.concatenate([s["left"] + ar(m), none(m), s["bottom"] + ar(B), none(B),
                           s["top"] + ar(B), none(B)])
    colb = np.concatenate([none(m), s["right"] + ar(m), none(B), s["bottom"] + ar(B),
                           none(B), s["top"] + ar(B)])
    for a in (ca, cb, cola, colb):
        a.setflags(write=False)  # cached: shared by every caller
    return ca, cb, cola, colb


def _is_pow2(k):
    return k >= 1 and (k & (k - 1)) == 0


@functools.lru_cache(maxsize=16)
def _interior_points(N, nx, ny, q, x0, width, y0, height):
    g = SlabGrid2D(N, nx, ny, q, x0, width, y0, height)
    lx, ly = g.leaf_origins()
    ti = g.t[1:-1]
    ga, gb = np.meshgrid(ti, ti, indexing="ij")
    P = np.empty((lx.shape[0] * ti.shape[0] ** 2, 2))
    P[:, 0].reshape(lx.shape[0], -1)[:] = lx[:, None] + ga.ravel()[None, :] * g.hx
    P[:, 1].reshape(lx.shape[0], -1)[:] = ly[:, None] + gb.ravel()[None, :] * g.hy
    P.setflags(write=False)
    return P


# GEMM tile policy of the HPS assembly: the inverses' many small batched products (e.g.
# 1024 x 64^3, 31 x 256^3) fill the chip better with more, smaller workgroups (a 4096-
# workgroup target, 64 x 32 split tiles) than under the library default; measured on the
# C2 assembly 3.5 -> 3.2 ms for the leaves and 1.3 -> 1.15 ms for the merges.  (As the
# library-wide default it slowed the FP32 solves of the mixed path.)
def _assembly_policy():
    from .tuning import Policy
    return Policy(4096, 512, 64, 256, 3, 4, 0)


class SlabGrid2D:
    """Leaf tiling of the rectangle [x0, x0+width] x [y0, y0+height] (default the unit
    square) by N single slabs of nx x ny leaves with q points."""

    def __init__(self, N: int, nx: int, ny: int, q: int, x0: float = 0.0,
                 width: float = 1.0, y0: float = 0.0, height: float = 1.0,
                 periodic: bool = False):
        if N < 2 or (periodic and N < 3):
            raise ValueError("need at least two single slabs (one interface), three when "
                             "periodic")
        if not (_is_pow2(nx) and _is_pow2(ny)):
            raise NotImplementedError("leaves per slab across / up must be powers of two")
        if q < 3:
            raise ValueError("q >= 3 Chebyshev points per leaf side")
        self.N, self.nx, self.ny, self.q = N, nx, ny, q
        self.p = q - 2
        self.x0, self.y0, self.width, self.height = x0, y0, width, height
        self.H = width / N
        self.hx = self.H / nx
        self.hy = height / ny
        self.m = ny * self.p
        # periodic in x (x0 and x0 + width identified): N interfaces, interface k at
        # x0 + k H between single slabs k-1 (mod N) and k, the ordering of the reference's
        # periodic dSlabs (geometry/geom_3D/squareTorus.py:268-277)
        self.periodic = bool(periodic)
        self.nb = N if periodic else N - 1
        self.nleaves = N * nx * ny
        self.t = cheb_unit(q)[0]

    # leaf L = (s*ny + j)*nx + ix
    def leaf_origins(self):
        s, j, ix = np.meshgrid(np.arange(self.N), np.arange(self.ny), np.arange(self.nx),
                               indexing="ij")
        x0 = self.x0 + s.ravel() * self.H + ix.ravel() * self.hx
        y0 = self.y0 + j.ravel() * self.hy
        return x0, y0

    def interior_points(self):
        """(nleaves*p^2, 2) interior collocation nodes, row r = (a-1)*p + (b-1); a
        read-only array shared by every grid of the same tiling (built once)."""
        return _interior_points(self.N, self.nx, self.ny, self.q, self.x0, self.width,
                                self.y0, self.height)

    def edge_coords(self, n_leaves, h):
        ti = self.t[1:-1]
        return (np.arange(n_leaves)[:, None] + ti[None, :]).ravel() * h

    def gb_points(self):
        """Global boundary nodes in the order [bottom | top | left | right] (periodic:
        [bottom | top])."""
        xs = self.x0 + self.edge_coords(self.N * self.nx, self.hx)
        ys = self.y0 + self.edge_coords(self.ny, self.hy)
        bottom = np.stack([xs, np.full_like(xs, self.y0)], 1)
        top = np.stack([xs, np.full_like(xs, self.y0 + self.height)], 1)
        if self.periodic:
            return np.concatenate([bottom, top])
        left = np.stack([np.full_like(ys, self.x0), ys], 1)
        right = np.stack([np.full_like(ys, self.x0 + self.width), ys], 1)
        return np.concatenate([bottom, top, left, right])

    def interface_points(self):
        """(nb*m, 2) interface nodes, bottom to top: interface i at x = x0 + (i+1) H
        (periodic: at x0 + i H)."""
        ys = self.y0 + self.edge_coords(self.ny, self.hy)
        sh = 0 if self.periodic else 1
        pts = [np.stack([np.full_like(ys, self.x0 + (i + sh) * self.H), ys], 1)
               for i in range(self.nb)]
        return np.concatenate(pts)


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


def _child_maps(ca, cb, na, nbm, n_old, dev):
    """Gather maps of the downward pass through one merge level: child alpha's boundary
    node k is parent boundary node inv_na[k] or shared-edge node inv_ca[k] (the other
    index -1); likewise beta."""
    def inv(idx):
        out = -np.ones(n_old, dtype=np.int64)
        for i, k in enumerate(idx):
            if k >= 0:
                out[k] = i
        return out
    ia, iac, ib, ibc = inv(na), inv(ca), inv(nbm), inv(cb)
    assert np.all((ia >= 0) != (iac >= 0)) and np.all((ib >= 0) != (ibc >= 0))
    return [dev(x, I32) for x in (ia, iac, ib, ibc)]


class HPSSlabOperator:
    """Device-resident S-system of the oms method for one PDE on one SlabGrid2D.

    After build(): Y (nb, m, 2m + 4*nx*p) float64 on the GPU (S maps and boundary-data
    maps of every double slab), SiM / SiP (nb, m, m) contiguous copies of the coupling
    blocks, and rhs_from_boundary() to map boundary data to S-system right-hand sides.
    """

    def __init__(self, pdo, grid: SlabGrid2D, device=None, keep_leaf_ops=False,
                 keep_solution_ops=False):
        _lib.require_gpu()
        _lib.load()
        self.pdo = pdo
        self.g = grid
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        # the leaf and merge solution operators are what solve_dir needs
        self.keep_solution_ops = keep_solution_ops
        self.keep_leaf_ops = keep_leaf_ops or keep_solution_ops
        self._down = []
        self.info_bad = 0

    # -------------------------------------------------------------------------
    def _dev(self, a, dtype=F64):
        a = np.ascontiguousarray(a)
        if not a.flags.writeable:  # torch wraps only writable arrays
            a = a.copy()
        return torch.as_tensor(a).to(device=self.device, dtype=dtype)

    def _const(self, key, make, dtype=F64):
        """Device copy of a read-only constant (leaf matrices, index maps) that depends on
        `key` only, shared by every operator on the device: later builds skip the host
        construction and the synchronous upload."""
        k = (key, str(self.device), dtype)
        t = _DEV_CONST.get(k)
        if t is None:
            if len(_DEV_CONST) >= 512:
                _DEV_CONST.clear()
            t = self._dev(make(), dtype)
            _DEV_CONST[k] = t
        return t

    def leaf_dtn(self, stream=None):
        g = self.g
        q, p = g.q, g.p
        P = g.interior_points()
        pdo = self.pdo
        vals = [_coef(getattr(pdo, name, None), P) for name in ("c11", "c22", "c1", "c2", "c")]
        # the coefficient samples go up in one copy from pinned memory
        have = [i for i, v in enumerate(vals) if v is not None]
        coefs = [None] * len(vals)
        if have:
            host = torch.empty((len(have), P.shape[0]), dtype=F64, pin_memory=True)
            for r, i in enumerate(have):
                host[r].numpy()[:] = vals[i]
            dev = host.to(self.device, non_blocking=True)
            for r, i in enumerate(have):
                coefs[i] = dev[r]
        ckey = ("leaf2", q, float(g.hx), float(g.hy))
        consts = {k: self._const(ckey + (k,), lambda k=k: leaf_constants(q, g.hx, g.hy)[k])
                  for k in ("Dx", "Dxx", "Dy", "Dyy", "Nb", "Ni")}
        nl = g.nleaves
        T = torch.empty(nl, 4 * p, 4 * p, dtype=F64, device=self.device)
        X = torch.empty(nl, p * p, 4 * p, dtype=F64, device=self.device) if self.keep_leaf_ops else None
        wd, wi = sizes2("sslab_hps_leaf_sizes", nl, q)
        work = torch.empty(wd, dtype=F64, device=self.device)
        iwork = torch.empty(wi, dtype=I32, device=self.device)
        info = torch.zeros(nl, dtype=I32, device=self.device)
        cp = [ptr(x) if x is not None else None for x in coefs]
        call("sslab_hps_leaf_dtn", nl, q, ptr(consts["Dx"]), ptr(consts["Dxx"]),
             ptr(consts["Dy"]), ptr(consts["Dyy"]), ptr(consts["Nb"]), ptr(consts["Ni"]),
             cp[0], cp[1], cp[2], cp[3], cp[4], ptr(T), ptr(X) if X is not None else None,
             ptr(work), ptr(iwork), ptr(info), stream_handle(stream))
        self._infos = [info]
        self.leaf_X = X
        return T

    def merge_tree(self, T, stream=None):
        g = self.g
        p = g.p
        self._down = []
        nxb, nyb = 1, 1
        boxes = T
        while nxb < g.nx or nyb < g.ny:
            kind = "h" if nxb < g.nx else "v"
            ca, cb, na, nbm, n_new, new = merge_maps(kind, nxb, nyb, p)
            n_old = boxes.shape[1]
            nmerge = boxes.shape[0] // 2
            c = len(ca)
            maps = [self._const(("merge2", kind, nxb, nyb, p, j), lambda x=x: x, I32)
                    for j, x in enumerate((ca, cb, na, nbm))]
            out = torch.empty(nmerge, n_new, n_new, dtype=F64, device=self.device)
            wd, wi = sizes2("sslab_hps_merge_sizes", nmerge, n_new, c)
            work = torch.empty(wd, dtype=F64, device=self.device)
            iwork = torch.empty(wi, dtype=I32, device=self.device)
            info = torch.zeros(nmerge, dtype=I32, device=self.device)
            s_old = n_old * n_old
            if self.keep_solution_ops:
                Xs = torch.empty(nmerge, c, n_new, dtype=F64, device=self.device)
                call("sslab_hps_merge_ex", nmerge, n_old, n_old, n_new, c, ptr(boxes),
                     2 * s_old, ptr(boxes) + 8 * s_old, 2 * s_old, *[ptr(x) for x in maps],
                     ptr(out), n_new * n_new, ptr(work), ptr(iwork), ptr(info), ptr(Xs),
                     c * n_new, stream_handle(stream))
                self._down.append(dict(X=Xs, n_old=n_old, n_new=n_new, c=c,
                                       maps=_child_maps(ca, cb, na, nbm, n_old, self._dev)))
            else:
                call("sslab_hps_merge", nmerge, n_old, n_old, n_new, c, ptr(boxes),
                     2 * s_old, ptr(boxes) + 8 * s_old, 2 * s_old, *[ptr(x) for x in maps],
                     ptr(out), n_new * n_new, ptr(work), ptr(iwork), ptr(info),
                     stream_handle(stream))
            self._infos.append(info)
            boxes = out
            nxb, nyb = new
        return boxes

    def pair_solve(self, Ts, stream=None):
        g = self.g
        if g.periodic:  # interface 0 joins the last slab to the first; 1.. as usual
            Yr = self._pair_solve(Ts, g.N - 1, stream)
            Yw = self._pair_solve(torch.stack([Ts[g.N - 1], Ts[0]]), 1, stream)
            return torch.cat([Yw, Yr])
        return self._pair_solve(Ts, g.nb, stream)

    def _pair_solve(self, Ts, nb, stream=None):
        g = self.g
        p, m = g.p, g.m
        ca, cb, cola, colb = pair_maps(g.nx, g.ny, p)
        ncols = len(cola)
        n_s = Ts.shape[1]
        maps = [self._const(("pair2", g.nx, g.ny, p, j), lambda x=x: x, I32)
                for j, x in enumerate((ca, cb, cola, colb))]
        Y = torch.empty(nb, m, ncols, dtype=F64, device=self.device)
        wd, wi = sizes2("sslab_hps_merge_sizes", nb, ncols, m)
        work = torch.empty(wd, dtype=F64, device=self.device)
        iwork = torch.empty(wi, dtype=I32, device=self.device)
        info = torch.zeros(nb, dtype=I32, device=self.device)
        call("sslab_hps_pair_solve", nb, n_s, m, ncols, ptr(Ts), n_s * n_s,
             *[ptr(x) for x in maps], ptr(Y), ptr(work), ptr(iwork), ptr(info),
             stream_handle(stream))
        self._infos.append(info)
        return Y

    def build(self, stream=None):
        """Run the whole assembly; returns self.  Its GEMM launches (the batched inverses'
        recursion over 1024 leaves and the merge / pair systems) run under ASSEMBLY_POLICY
        unless the caller has set a tuned policy for the key "hps_assembly"."""
        from . import tuning
        pol = tuning.lookup("hps_assembly") or ASSEMBLY_POLICY
        with torch.cuda.device(self.device), tuning.policy(pol):
            T = self.leaf_dtn(stream)
            Ts = self.merge_tree(T, stream)
            del T
            self.Ts = Ts
            self.Y = self.pair_solve(Ts, stream)
            m = self.g.m
            self.SiM = self.Y[:, :, 0:m].contiguous()
            self.SiP = self.Y[:, :, m:2 * m].contiguous()
            self.info_bad = int(sum(int((i != 0).sum()) for i in self._infos))
        if self.info_bad:
            import warnings
            warnings.warn(f"{self.info_bad} singular local solve(s) in the HPS assembly "
                          "(a leaf or slab Dirichlet problem at resonance?)", RuntimeWarning)
        return self

    # -------------------------------------------------------------------------
    def S_rk_list(self):
        """[[S_r], [S_l, S_r], ..., [S_l]] device views (multislab/oms.py:239-244);
        periodic: [S_l, S_r] for every interface."""
        nb = self.g.nb
        if self.g.periodic:
            return [[self.SiM[i], self.SiP[i]] for i in range(nb)]
        out = []
        for i in range(nb):
            if i == 0:
                out.append([self.SiP[0]])
            elif i == nb - 1:
                out.append([self.SiM[i]])
            else:
                out.append([self.SiM[i], self.SiP[i]])
        return out

    def rhs_from_boundary(self, F: torch.Tensor, out: torch.Tensor | None = None, stream=None):
        """S-system right-hand sides rhs_i = -(A_ii^{-1} A_ib[:, Igb] f)[Ic] for r columns
        of global boundary data F (N_gb x r, order of SlabGrid2D.gb_points())."""
        g = self.g
        m, nb, nxp = g.m, g.nb, g.nx * g.p
        if F.dim() == 1:
            F = F[:, None]
        F = F.to(device=self.device, dtype=F64).contiguous()
        r = F.shape[1]
        if g.periodic:
            return self._rhs_periodic(F, out)
        ngb = 2 * g.N * nxp + 2 * m
        if F.shape[0] != ngb:
            raise ValueError(f"boundary data needs {ngb} rows, got {F.shape[0]}")
        if out is None:
            out = torch.empty(nb * m, r, dtype=F64, device=self.device)
        ncols = self.Y.shape[2]
        call("sslab_oms_rhs", nb, m, nxp, ncols, ptr(self.Y), ptr(F), F.stride(0), r,
             g.N, ptr(out), out.stride(0), stream_handle(stream))
        return out

    def _rhs_periodic(self, F, out):
        """rhs of the x-periodic S-system: interface k reads the bottom / top windows of
        single slabs (k-1) mod N and k, gathered into one batch (the wrapping pair has
        no contiguous window), then one batched product with the maps."""
        from .hbs import bgemm
        g = self.g
        m, N, nxp = g.m, g.N, g.nx * g.p
        if F.shape[0] != 2 * N * nxp:
            raise ValueError(f"boundary data needs {2 * N * nxp} rows, got {F.shape[0]}")
        k = np.arange(N)
        a, b = (k - 1) % N, k
        w = np.arange(nxp)
        rows = np.concatenate([(a[:, None] * nxp + w), (b[:, None] * nxp + w),
                               (N * nxp + a[:, None] * nxp + w), (N * nxp + b[:, None] * nxp + w)],
                              axis=1)                                   # (N, 4 nxp)
        idx = torch.as_tensor(rows.ravel(), device=self.device)
        Fp = F.index_select(0, idx).view(N, 4 * nxp, F.shape[1])
        X = bgemm(self.Y[:, :, 2 * m:], Fp, alpha=-1.0).reshape(N * m, -1)
        if out is not None:
            out.copy_(X)
            return out
        return X

    # -------------------------------------------------------------------------
    def solve_dir(self, x_iface: torch.Tensor, F: torch.Tensor, ff_body=None, stream=None):
        """Solution at every leaf's interior Chebyshev nodes (rows in the order of
        SlabGrid2D.interior_points()) from the interface solution x_iface (nb*m x r, as
        the S-system solve returns it) and the global boundary data F (boundary-node
        order of SlabGrid2D.gb_points()): the reference's solve_dir
        (solver/pde_solver.py:169-201, u_i = A_ii^{-1}(ff_body - A_ix u_b)) per slab, by
        the HPS downward pass -- shared-edge values -X u at every merge (top down), then
        -X_leaf u_b at the leaves; batched GEMMs and gathers only.  Needs
        keep_solution_ops=True at construction.  Body loads (ff_body) are not supported:
        the leaf and merge operators here are those of the homogeneous equation."""
        if ff_body is not None:
            raise NotImplementedError("solve_dir: body loads are not supported")
        if self.g.periodic:
            raise NotImplementedError("solve_dir: periodic slabs are not provided")
        if not self.keep_solution_ops or self.leaf_X is None or (
                len(self._down) == 0 and (self.g.nx > 1 or self.g.ny > 1)):
            raise ValueError("solve_dir: build the operator with keep_solution_ops=True")
        from .hbs import bgemm
        g = self.g
        m, nb, p, N = g.m, g.nb, g.p, g.N
        nxp = g.nx * p
        one_d = x_iface.dim() == 1
        X = x_iface.reshape(x_iface.shape[0], -1).to(self.device, F64).contiguous()
        if F.dim() == 1:
            F = F[:, None]
        F = F.to(self.device, F64).contiguous()
        r = X.shape[1]
        ngb = 2 * N * nxp + 2 * m
        if X.shape[0] != nb * m or F.shape != (ngb, r):
            raise ValueError(f"solve_dir: x_iface ({nb * m}, r) and F ({ngb}, r) expected")
        dev = self.device
        ident = torch.arange(r, dtype=I32, device=dev)
        n_s = 2 * m + 2 * nxp
        U = torch.empty(N, n_s, r, dtype=F64, device=dev)
        s = stream_handle(stream)
        i = np.arange(n_s)
        bot = (i >= 2 * m) & (i < 2 * m + nxp)
        top = i >= 2 * m + nxp
        sref = np.where(bot, i - 2 * m, np.where(top, N * nxp + i - 2 * m - nxp, -1))

        keep = []  # index arrays stay referenced until every gather is enqueued

        def gather(out, batch, s0, rid0, s1, rid1, s1_stride):
            d0, d1 = self._dev(rid0, I32), self._dev(rid1, I32)
            keep.extend((d0, d1))
            call("sslab_gather2", batch, n_s, r, s0, r, m * r, s1, r, s1_stride,
                 ptr(d0), ptr(ident), ptr(d1), ptr(ident), out, r, n_s * r, s)

        # slab boundary [left | right | bottom | top]: interfaces from X, the rest from F
        if N > 2:  # slabs 1 .. N-2: left = interface s-1, right = interface s
            gather(ptr(U) + 8 * n_s * r, N - 2, ptr(X), np.where(i < 2 * m, i, -1),
                   ptr(F) + 8 * nxp * r, np.where(bot | top, sref, -1), nxp * r)
        left_f = np.where(i < m, 2 * N * nxp + i, -1)
        gather(ptr(U), 1, ptr(X) - 8 * m * r, np.where((i >= m) & (i < 2 * m), i, -1),
               ptr(F), np.where(bot | top, sref, left_f), 0)
        right_f = np.where((i >= m) & (i < 2 * m), 2 * N * nxp + i, -1)
        last = N - 1
        sref_last = np.where(bot, sref + last * nxp, np.where(top, sref + last * nxp, -1))
        gather(ptr(U) + 8 * last * n_s * r, 1, ptr(X) + 8 * (last - 1) * m * r,
               np.where(i < m, i, -1), ptr(F),
               np.where(bot | top, sref_last, right_f), 0)
        # downward pass through the merge levels, top down
        for lvl in reversed(self._down):
            Xs, n_old, c = lvl["X"], lvl["n_old"], lvl["c"]
            nmerge = Xs.shape[0]
            Uc = bgemm(Xs, U, alpha=-1.0)                     # shared-edge values
            Uch = torch.empty(2 * nmerge, n_old, r, dtype=F64, device=dev)
            ia, iac, ib, ibc = lvl["maps"]
            for half, (r0, r1) in enumerate(((ia, iac), (ib, ibc))):
                call("sslab_gather2", nmerge, n_old, r, ptr(U), r, U.stride(0), ptr(Uc), r,
                     Uc.stride(0), ptr(r0), ptr(ident), ptr(r1), ptr(ident),
                     ptr(Uch) + 8 * half * n_old * r, r, 2 * n_old * r, s)
            U = Uch
        out = bgemm(self.leaf_X, U, alpha=-1.0)               # leaf interiors
        out = out.reshape(-1, r)
        return out[:, 0] if one_d else out

    def fold(self, fac, stream=None) -> "BoundaryFold":
        """The forward sweep of `fac` (red-black, T = None, built from this operator's
        SiM / SiP) folded into boundary maps: see BoundaryFold."""
        return BoundaryFold(self, fac, stream)

    def rb0_maps(self, stream=None):
        """Boundary maps that fold the first red-black elimination level of the S-system
        (identity diagonal blocks) into the right-hand side: for even i,
            v_i - SiM_i v_{i-1} - SiP_i v_{i+1}
              = Ycb_i Fbot[(i-1) nxp : (i+3) nxp] + Yct_i Ftop[...] (+ boundary terms),
        since v_k = -(Y_k F)[Ic] reads bottom/top windows k, k+1 (window w = rows
        [w nxp, (w+1) nxp)).  Built once per operator with the library's GEMM; cached."""
        if getattr(self, "_rb0", None) is not None:
            return self._rb0
        from .hbs import bgemm
        g = self.g
        if g.periodic:
            raise NotImplementedError("rb0_maps: periodic operators are solved unfolded")
        m, nb, nxp = g.m, g.nb, g.nx * g.p
        if nb < 2:
            raise ValueError("rb0_maps: needs at least two interfaces")
        Y = self.Y
        nE, nO = (nb + 1) // 2, nb // 2
        w = 2 * nxp
        out = []
        with torch.cuda.device(self.device), torch.cuda.stream(
                stream if stream is not None else torch.cuda.current_stream()):
            for c0 in (2 * m, 2 * m + w):  # bottom, then top maps
                Yc = torch.zeros(nE, m, 2 * w, dtype=F64, device=self.device)
                Yc[:, :, nxp:nxp + w] = -Y[0::2, :, c0:c0 + w]          # own double slab
                if nE > 1:                                               # left neighbour
                    bgemm(Y[2::2, :, 0:m], Y[1:nb - 1:2, :, c0:c0 + w], out=Yc[1:, :, 0:w],
                          beta=1.0)
                bgemm(Y[0:2 * nO:2, :, m:2 * m], Y[1::2, :, c0:c0 + w],  # right neighbour
                      out=Yc[:nO, :, w:2 * w], beta=1.0)
                out.append(Yc)
            Ycl = (-Y[0, :, 0:m]).contiguous()
            if (nb - 1) % 2 == 0:
                Ycr = (-Y[nb - 1, :, m:2 * m]).contiguous()
            else:
                Ycr = bgemm(Y[nb - 2:nb - 1, :, m:2 * m], Y[nb - 1:nb, :, m:2 * m])[0]
        self._rb0 = (out[0], out[1], Ycl, Ycr)
        return self._rb0

    def rhs_rb0_from_boundary(self, F: torch.Tensor, V: torch.Tensor | None = None,
                              X: torch.Tensor | None = None, stream=None):
        """rhs_from_boundary with the first red-black level applied; returns (V, X), the
        entry state of BlockTriFactor.solve_(V, X, level0_reduced=True) for the factor
        built with T = None: V's even block rows hold rhs_i - SiM_i rhs_{i-1} -
        SiP_i rhs_{i+1}, X's odd block rows rhs_j and X's even block rows zero (one zero
        kernel and one grouped GEMM launch)."""
        g = self.g
        m, nb, nxp = g.m, g.nb, g.nx * g.p
        if F.dim() == 1:
            F = F[:, None]
        F = F.to(device=self.device, dtype=F64).contiguous()
        r = F.shape[1]
        ngb = 2 * g.N * nxp + 2 * m
        if F.shape[0] != ngb:
            raise ValueError(f"boundary data needs {ngb} rows, got {F.shape[0]}")
        Ycb, Yct, Ycl, Ycr = self.rb0_maps(stream)
        if V is None:
            V = torch.empty(nb * m, r, dtype=F64, device=self.device)
        if X is None:
            X = torch.empty(nb * m, r, dtype=F64, device=self.device)
        for B in (V, X):
            if B.shape != (nb * m, r) or B.dtype != F64 or (r > 0 and B.stride(1) != 1):
                raise ValueError("rhs_rb0_from_boundary: V and X must be float64 (nb*m, r) "
                                 "with unit column stride")
        call("sslab_oms_rhs_rb0", nb, m, nxp, self.Y.shape[2], ptr(self.Y), ptr(Ycb),
             ptr(Yct), ptr(Ycl), ptr(Ycr), ptr(F), max(F.stride(0), r), r, g.N, ptr(V),
             max(V.stride(0), r), ptr(X), max(X.stride(0), r), stream_handle(stream))
        return V, X


class BoundaryFold:
    """Solve of an oms S-system straight from global boundary data, with the red-black
    forward sweep folded into boundary maps (sslab_oms_fold_*).

    Everything before the backward sweep is linear in the boundary data F: the sweep's
    entry state (y_j = T_j^{-1} v_j at every level's odd block rows, x at the coarsest
    level's rows) is M F.  M is built once here -- the rhs maps (multislab/oms.py:180-186)
    and the forward sweep applied to F = I -- and its blocks keep only the boundary
    windows they read.  solve_ is then one grouped GEMM launch and the backward sweep."""

    def __init__(self, op: HPSSlabOperator, fac, stream=None):
        if fac.method != "rb" or not fac.ident0 or fac.cyclic or fac.factor is None:
            raise ValueError("BoundaryFold: needs the FP64 red-black factor built with T = None")
        if op.g.periodic:
            raise ValueError("BoundaryFold: periodic operators are solved unfolded")
        g = op.g
        m, nb, nxp = g.m, g.nb, g.nx * g.p
        if fac.nb != nb or fac.m != m:
            raise ValueError("BoundaryFold: factor and operator sizes differ")
        self.op, self.fac = op, fac
        self.nF = 2 * g.N * nxp + 2 * m
        self.handle = None
        s = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.device(op.device), torch.cuda.stream(s):
            if nb >= 2:
                # Boundary columns reordered [left | (bottom_s, top_s) per slab s | right], so
                # block row i of the rhs maps reads the one interval of slabs i and i+1 (plus
                # the left / right column block at the ends); the windowed forward sweep
                # keeps every product on the columns its operands can touch.
                u = 2 * nxp
                nsl = g.N
                bot = np.arange(nsl * nxp).reshape(nsl, nxp)
                top = nsl * nxp + bot
                left = 2 * nsl * nxp + np.arange(m)
                perm = np.concatenate([left, np.concatenate([bot, top], axis=1).ravel(), m + left])
                lo = (ctypes.c_int * nb)(*[0 if i == 0 else m + u * i for i in range(nb)])
                hi = (ctypes.c_int * nb)(*[self.nF if i == nb - 1 else m + u * (i + 2)
                                           for i in range(nb)])
                pd = torch.as_tensor(perm, device=op.device)
                R = self._permuted_rhs_maps(op)
                Mp = torch.empty_like(R)
                call("sslab_rb_forward_windowed", nb, m, self.nF, ptr(fac.factor), ptr(R),
                     R.stride(0), ptr(Mp), Mp.stride(0), ctypes.addressof(lo),
                     ctypes.addressof(hi), stream_handle(s))
                del R
                inv = torch.empty_like(pd)
                inv[pd] = torch.arange(self.nF, device=op.device)
                M = Mp.index_select(1, inv)
                del Mp
            else:
                R = op.rhs_from_boundary(torch.eye(self.nF, dtype=F64, device=op.device),
                                         stream=s)
                M = torch.empty_like(R)
                fac.solve_(R, M, stream=s, phase="forward")
                del R
            h = ctypes.c_void_p()
            call("sslab_oms_fold_create", nb, m, self.nF, ptr(M), M.stride(0),
                 ctypes.addressof(h), stream_handle(s))
            s.synchronize()  # the compact copies have read M
            del M
        self.handle = h.value

    def _permuted_rhs_maps(self, op):
        """The rhs maps of the column-permuted identity, rhs_from_boundary(I[:, perm]):
        block row i holds -Y_i's boundary-data columns at its window [m + u i, m + u (i+2))
        (slab i's bottom and top, then slab i+1's), block 0 also -Y_0[:, :m] at the left
        columns [0, m) and block nb-1 -Y_{nb-1}[:, m:2m] at the right ones -- the values the
        rhs GEMM computes for this identity, placed by one gather instead."""
        g = op.g
        m, nb, nxp = g.m, g.nb, g.nx * g.p
        u, nF = 2 * nxp, self.nF
        R = torch.zeros(nb * m, nF, dtype=F64, device=op.device)
        o = np.arange(nxp)
        ycols = 2 * m + np.concatenate([o, 2 * nxp + o, nxp + o, 3 * nxp + o])
        Yb = op.Y.index_select(2, torch.as_tensor(ycols, device=op.device))
        W = R.as_strided((nb, m, 2 * u), (m * nF + u, nF, 1), m)
        W.copy_(Yb).neg_()
        R[:m, :m] = -op.Y[0, :, :m]
        R[(nb - 1) * m:, nF - m:] = -op.Y[nb - 1, :, m:2 * m]
        return R

    @property
    def nbytes(self) -> int:
        return int(_lib.load().sslab_oms_fold_bytes(self.handle)) if self.handle else 0

    @property
    def groups(self) -> int:
        return int(_lib.load().sslab_oms_fold_groups(self.handle)) if self.handle else 0

    def tune_key(self, r: int) -> str:
        return f"oms_fold:{self.fac.nb}:{self.fac.m}:{self.nF}:{r}"

    def autotune(self, r: int, reps: int = 10, cands=None):
        """Pick the GEMM tile policy of the boundary-map launch alone (its shapes differ
        from the backward sweep's) by timing every candidate; later apply_ / solve_ calls
        of width r use it.  Returns (policy, microseconds per call)."""
        from . import tuning
        nbm = self.fac.nb * self.fac.m
        F = torch.randn(self.nF, r, dtype=F64, device=self.op.device)
        X = torch.empty(nbm, r, dtype=F64, device=self.op.device)
        key = self.tune_key(r)
        with tuning.policy(tuning.DEFAULT):
            tuning.lookup(key)
            saved = tuning._cache.pop(key, None)
            try:
                return tuning.autotune(lambda: self.apply_(F, X), key, cands=cands, reps=reps)
            except Exception:
                if saved is not None:
                    tuning._cache[key] = saved
                raise

    def apply_(self, F: torch.Tensor, X: torch.Tensor, stream=None):
        """X = M F: the backward sweep's entry state for boundary data F (nF x r), under the
        policy tuned for this width if there is one (else the caller's)."""
        from . import tuning
        r = F.shape[1]
        nbm = self.fac.nb * self.fac.m
        if F.shape[0] != self.nF or F.dtype != F64 or X.shape != (nbm, r) or X.dtype != F64:
            raise ValueError(f"BoundaryFold: F must be float64 ({self.nF}, r) and X ({nbm}, r)")
        if r > 0 and (F.stride(1) != 1 or X.stride(1) != 1):
            raise ValueError("BoundaryFold: unit column stride required")
        with tuning.policy(tuning.lookup(self.tune_key(r))):
            call("sslab_oms_fold_apply", self.handle, ptr(F), max(F.stride(0), r), r, ptr(X),
                 max(X.stride(0), r), stream_handle(stream))
        return X

    def solve_(self, F: torch.Tensor, X: torch.Tensor, stream=None):
        """X = S^{-1} rhs(F): interface values for boundary data F (nF x r)."""
        self.apply_(F, X, stream)
        self.fac.solve_(None, X, stream=stream, phase="backward")
        return X

    def solve(self, F: torch.Tensor, stream=None) -> torch.Tensor:
        if F.dim() == 1:
            F = F[:, None]
        F = F.to(device=self.op.device, dtype=F64).contiguous()
        X = torch.empty(self.fac.nb * self.fac.m, F.shape[1], dtype=F64, device=self.op.device)
        return self.solve_(F, X, stream)

    # -- mixed precision ------------------------------------------------------------------
    def enable_mixed(self, stream=None):
        """Keep an FP32 copy of the maps for solve_mixed_ (the factor needs its FP32 copy
        first: BlockTriFactor.enable_mixed with the S-system's SiM / SiP)."""
        if getattr(self.fac, "factor32", None) is None or getattr(self.fac, "_M64", None) is None:
            raise ValueError("BoundaryFold.enable_mixed: call the factor's enable_mixed first")
        call("sslab_oms_fold_enable_f32", self.handle, stream_handle(stream))
        return self

    def mixed_work(self, r: int):
        """Workspace of solve_mixed_ for r columns: (FP64 residual, FP32 panels, FP64 rhs)."""
        N = self.fac.nb * self.fac.m
        dev = self.op.device
        return (torch.empty(N * r, dtype=F64, device=dev),
                torch.empty((2 * N + self.nF) * r, dtype=torch.float32, device=dev),
                torch.empty(N, r, dtype=F64, device=dev))

    def solve_mixed_(self, F: torch.Tensor, X: torch.Tensor, iters: int = 1, work=None,
                     stream=None):
        """X = S^{-1} rhs(F) in mixed precision: the folded solve in FP32 (FP32 maps and
        backward sweep), then `iters` refinement steps with FP64 residuals and FP32
        red-black corrections (sslab_oms_fold_solve_mixed).  F: contiguous float64 (nF, r)."""
        r = F.shape[1]
        nbm = self.fac.nb * self.fac.m
        if (F.shape[0] != self.nF or F.dtype != F64 or not F.is_contiguous()
                or X.shape != (nbm, r) or X.dtype != F64 or not X.is_contiguous()):
            raise ValueError(f"solve_mixed_: contiguous float64 F ({self.nF}, r) and X ({nbm}, r)")
        if work is None:
            work = self.mixed_work(r)
        w64, w32, B = work
        self.op.rhs_from_boundary(F, out=B, stream=stream)
        fac = self.fac
        call("sslab_oms_fold_solve_mixed", self.handle, ptr(fac.factor32), ptr(fac._M64),
             ptr(fac._P64), ptr(F), r, r, ptr(B), ptr(X), ptr(w64), ptr(w32), int(iters),
             stream_handle(stream))
        return X

    def __del__(self):
        h, self.handle = getattr(self, "handle", None), None
        if h and _lib is not None:
            try:
                _lib.load().sslab_oms_fold_destroy(h)
            except Exception:
                pass


ASSEMBLY_POLICY = _assembly_policy()
