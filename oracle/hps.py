"""Oracle: 2-D Chebyshev HPS slab discretization and the oms S-system (numpy/scipy FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference's HPS discretization (the `hpsalt` / `hps` solvers behind
solver/solver.py:99-144) is an absent submodule, so this restates the published
Chebyshev HPS collocation (Gillman & Martinsson; Martinsson, "Fast Direct Solvers for
Elliptic PDEs", ch. 25) with the reference's conventions:

* a leaf is a q x q tensor Chebyshev grid (q = p + 2, the "p_disc = p + 2" of
  test/3D/cubeTest.py:115-117); its p**2 interior nodes carry the PDE collocation and
  its 4p edge nodes (corners excluded) are shared with the neighbours, where the
  equation is continuity of the normal derivative;
* the operator follows pdo/pdo.py:19-39:
  A[u] = -c11 u_xx - c22 u_yy + c1 u_x + c2 u_y + c u  (c12 = 0);
* a double slab is exposed exactly like solverWrapper.construct (solver/solver.py:79-163):
  XX, Ii, Ib, Aib and A_ii (sparse), so the oms maps of multislab/oms.py:38-52, 99-129,
  180-186 apply verbatim:  S_J = (A_ii^{-1} A_ib[:, J])[Ic],
  rhs = -(A_ii^{-1} A_ib[:, Igb] f)[Ic];
* slabs come from dSlabs (geometry/geom_2D/square.py:33-47).

Parity pins (tests/test_oracle_hps.py):
  * the Chebyshev differentiation matrix is exact on polynomials of degree <= q-1;
  * the oms S-system solution equals the solution of the single global HPS system on the
    interface lines (exact domain decomposition, ~1e-12);
  * spectral accuracy against the harmonic u = exp(x) sin(y) (config 1);
  * rows of I + sum_J S_J + R_gb sum to zero for Laplace with constant data.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


# ---------------------------------------------------------------------------
# Chebyshev collocation

def cheb_diff(n: int):
    """Points x_k = cos(pi k / n), k = 0..n, and the differentiation matrix on them
    (Trefethen, Spectral Methods in MATLAB, 2000, program cheb; the reference uses the
    same textbook matrix, solver/spectral/spectralSolver.py:132-151)."""
    x = np.cos(np.pi * np.arange(n + 1) / n)
    c = np.ones(n + 1)
    c[0] = c[-1] = 2.0
    c *= (-1.0) ** np.arange(n + 1)
    dX = x[:, None] - x[None, :]
    D = np.outer(c, 1.0 / c) / (dX + np.eye(n + 1))
    D -= np.diag(D.sum(axis=1))
    return x, D


def unit_cheb(q: int):
    """q ascending Chebyshev points on [0, 1] and d/dt on them."""
    x, D = cheb_diff(q - 1)
    t = (1.0 - x) / 2.0          # ascending from 0 to 1
    return t, -2.0 * D           # dt = -dx/2


@dataclass
class LeafRef:
    q: int
    t: np.ndarray      # (q,) nodes on [0,1]
    D: np.ndarray      # (q,q) d/dt
    D2: np.ndarray     # (q,q) d2/dt2

    @staticmethod
    def make(q):
        t, D = unit_cheb(q)
        return LeafRef(q, t, D, D @ D)


def eval_coef(f, X):
    if f is None:
        return None
    v = f(X)
    v = np.asarray(v, dtype=np.float64)
    if v.ndim == 0:
        v = np.full(X.shape[0], float(v))
    return v.reshape(-1)


# ---------------------------------------------------------------------------
# Sparse HPS system on a rectangle of leaves

class HPSRect:
    """All leaf-interior and leaf-edge nodes of the rectangle [x0, x0+nxl*hx] x
    [y0, y0+nyl*hy] tiled by nxl x nyl leaves of q x q Chebyshev points.

    Node numbering: edge nodes first -- vertical lines k = 0..nxl (each nyl*p nodes,
    bottom to top), then horizontal lines j = 0..nyl (each nxl*p nodes, left to right)
    -- then leaf interiors (leaf-major, x-major inside a leaf).
    Equations: PDE at leaf interiors, flux continuity at internal edges."""

    def __init__(self, pdo, x0, y0, hx, hy, nxl, nyl, q, periodic_x=False):
        self.pdo, self.x0, self.y0, self.hx, self.hy = pdo, x0, y0, hx, hy
        self.nxl, self.nyl, self.q = nxl, nyl, q
        self.periodic_x = periodic_x
        p = q - 2
        self.p = p
        self.ref = LeafRef.make(q)
        t = self.ref.t
        # periodic in x: the vertical line k = nxl is line 0 (one set of nodes)
        nv = (nxl + (0 if periodic_x else 1)) * nyl * p
        nh = (nyl + 1) * nxl * p
        self.nv, self.nh = nv, nh
        self.nedge = nv + nh
        self.nint = nxl * nyl * p * p
        self.n = self.nedge + self.nint
        XX = np.zeros((self.n, 2))
        for k in range(nxl + (0 if periodic_x else 1)):
            for j in range(nyl):
                s = self.vidx(k, j)
                XX[s:s + p, 0] = x0 + k * hx
                XX[s:s + p, 1] = y0 + (j + t[1:-1]) * hy
        for j in range(nyl + 1):
            for i in range(nxl):
                s = self.hidx(j, i)
                XX[s:s + p, 0] = x0 + (i + t[1:-1]) * hx
                XX[s:s + p, 1] = y0 + j * hy
        for i in range(nxl):
            for j in range(nyl):
                s = self.iidx(i, j)
                gx, gy = np.meshgrid(t[1:-1], t[1:-1], indexing="ij")
                XX[s:s + p * p, 0] = x0 + (i + gx.ravel()) * hx
                XX[s:s + p * p, 1] = y0 + (j + gy.ravel()) * hy
        self.XX = XX

    # index helpers -------------------------------------------------------------
    def vidx(self, k, j):
        if self.periodic_x:
            k %= self.nxl
        return (k * self.nyl + j) * self.p

    def hidx(self, j, i):
        return self.nv + (j * self.nxl + i) * self.p

    def iidx(self, i, j):
        return self.nedge + (i * self.nyl + j) * self.p * self.p

    def leaf_local_map(self, i, j):
        """Global index of every q x q grid node of leaf (i, j), -1 at the corners."""
        q, p = self.q, self.p
        g = -np.ones((q, q), dtype=np.int64)
        g[1:-1, 1:-1] = self.iidx(i, j) + np.arange(p * p).reshape(p, p)
        g[0, 1:-1] = self.vidx(i, j) + np.arange(p)
        g[q - 1, 1:-1] = self.vidx(i + 1, j) + np.arange(p)
        g[1:-1, 0] = self.hidx(j, i) + np.arange(p)
        g[1:-1, q - 1] = self.hidx(j + 1, i) + np.arange(p)
        return g

    # assembly --------------------------------------------------------------------
    def assemble(self):
        """Sparse square matrix with one row per node: PDE rows at leaf interiors,
        flux-continuity rows at internal edges, identity rows at the outer edges."""
        q, p, pdo = self.q, self.p, self.pdo
        D, D2 = self.ref.D, self.ref.D2
        Dx, Dxx = D / self.hx, D2 / self.hx ** 2
        Dy, Dyy = D / self.hy, D2 / self.hy ** 2
        rows, cols, vals = [], [], []
        flux = {}  # edge node -> list of (col, val) contributions (outward normal derivative)
        for i in range(self.nxl):
            for j in range(self.nyl):
                g = self.leaf_local_map(i, j)
                Xi = self.XX[g[1:-1, 1:-1].ravel()]
                c11 = eval_coef(pdo.c11, Xi)
                c22 = eval_coef(pdo.c22, Xi)
                c1 = eval_coef(getattr(pdo, "c1", None), Xi)
                c2 = eval_coef(getattr(pdo, "c2", None), Xi)
                c0 = eval_coef(getattr(pdo, "c", None), Xi)
                r = 0
                for a in range(1, q - 1):
                    for b in range(1, q - 1):
                        row = g[a, b]
                        acc = {}
                        for a2 in range(q):   # x-direction line through (a, b)
                            v = -c11[r] * Dxx[a, a2]
                            if c1 is not None:
                                v += c1[r] * Dx[a, a2]
                            acc[g[a2, b]] = acc.get(g[a2, b], 0.0) + v
                        for b2 in range(q):   # y-direction line
                            v = -c22[r] * Dyy[b, b2]
                            if c2 is not None:
                                v += c2[r] * Dy[b, b2]
                            acc[g[a, b2]] = acc.get(g[a, b2], 0.0) + v
                        if c0 is not None:
                            acc[row] = acc.get(row, 0.0) + c0[r]
                        for cidx, v in acc.items():
                            rows.append(row); cols.append(cidx); vals.append(v)
                        r += 1
                # outward normal derivative at the leaf's edge nodes
                for b in range(1, q - 1):
                    for side, a0, sgn in ((0, 0, -1.0), (1, q - 1, 1.0)):
                        node = g[a0, b]
                        lst = flux.setdefault(node, [])
                        for a2 in range(q):
                            lst.append((g[a2, b], sgn * Dx[a0, a2]))
                for a in range(1, q - 1):
                    for b0, sgn in ((0, -1.0), (q - 1, 1.0)):
                        node = g[a, b0]
                        lst = flux.setdefault(node, [])
                        for b2 in range(q):
                            lst.append((g[a, b2], sgn * Dy[b0, b2]))
        outer = self.outer_edge_nodes()
        is_outer = np.zeros(self.n, dtype=bool)
        is_outer[outer] = True
        for node, lst in flux.items():
            if is_outer[node]:
                continue
            for cidx, v in lst:
                rows.append(node); cols.append(cidx); vals.append(v)
        for node in outer:
            rows.append(node); cols.append(node); vals.append(1.0)
        A = sp.csr_matrix((vals, (rows, cols)), shape=(self.n, self.n))
        A.sum_duplicates()
        return A

    def outer_edge_nodes(self):
        p, nyl, nxl = self.p, self.nyl, self.nxl
        if self.periodic_x:  # only the bottom and top lines bound the cylinder
            return np.concatenate([self.hidx(0, 0) + np.arange(nxl * p),
                                   self.hidx(nyl, 0) + np.arange(nxl * p)])
        left = self.vidx(0, 0) + np.arange(nyl * p)
        right = self.vidx(nxl, 0) + np.arange(nyl * p)
        bottom = self.hidx(0, 0) + np.arange(nxl * p)
        top = self.hidx(nyl, 0) + np.arange(nxl * p)
        return np.concatenate([left, right, bottom, top])


class SlabSolver:
    """solverWrapper.construct-style view of one double slab (solver/solver.py:79-163):
    XX, Ii, Ib, Aib, Abi, Abb and solver_ii (a sparse LU of A_ii)."""

    def __init__(self, pdo, bounds, leaf_h, q):
        (xl, yl), (xr, yr) = np.asarray(bounds, dtype=np.float64)
        hx, hy = leaf_h
        nxl = int(round((xr - xl) / hx))
        nyl = int(round((yr - yl) / hy))
        self.rect = HPSRect(pdo, xl, yl, hx, hy, nxl, nyl, q)
        A = self.rect.assemble()
        outer = self.rect.outer_edge_nodes()
        mask = np.ones(self.rect.n, dtype=bool)
        mask[outer] = False
        self.Ib = outer
        self.Ii = np.where(mask)[0]
        self.XX = self.rect.XX
        self.Aii = A[self.Ii][:, self.Ii].tocsc()
        self.Aib = A[self.Ii][:, self.Ib].tocsc()
        self.lu = spla.splu(self.Aii)
        self.solver_ii = spla.LinearOperator(self.Aii.shape, matvec=self.lu.solve,
                                             matmat=self.lu.solve,
                                             rmatvec=lambda v: self.lu.solve(v, trans="T"),
                                             rmatmat=lambda v: self.lu.solve(v, trans="T"))


def gb_unit_square(P):
    """Global-boundary predicate of [0,1]^2 (geometry/geom_2D/square.py:6-7)."""
    return ((np.abs(P[:, 0]) < 1e-14) | (np.abs(P[:, 0] - 1) < 1e-14)
            | (np.abs(P[:, 1]) < 1e-14) | (np.abs(P[:, 1] - 1) < 1e-14))


def dSlabs_periodic(N):
    """Double slabs of [0,1]^2 periodic in x and their connectivity, the convention of
    geometry/geom_3D/squareTorus.py:268-277: double slab n is centred on x = n H
    (n = 0..N-1, the first straddling x = 0 = 1), connected to (n-1) % N and (n+1) % N."""
    H = 1.0 / N
    slabs = [[[n * H - H, 0.0], [n * H + H, 1.0]] for n in range(N)]
    return slabs, [[(n - 1) % N, (n + 1) % N] for n in range(N)], H


def gb_bottom_top(P):
    """Global boundary of the x-periodic unit square: y = 0 and y = 1."""
    return (np.abs(P[:, 1]) < 1e-14) | (np.abs(P[:, 1] - 1) < 1e-14)


def dSlabs(N):
    """Double slabs of [0,1]^2 and their connectivity (geometry/geom_2D/square.py:33-47)."""
    H = 1.0 / N
    slabs, conn = [], []
    for n in range(N - 1):
        c = (n + 1) * H
        slabs.append([[c - H, 0.0], [c + H, 1.0]])
        if n == 0:
            conn.append([-1, 1])
        elif n == N - 2:
            conn.append([n - 1, -1])
        else:
            conn.append([n - 1, n + 1])
    return slabs, conn, H


def slab_index_sets(solver, bounds, gb=gb_unit_square):
    """Il, Ir, Ic, Igb, XXi, XXb of multislab/oms.py:38-52 (tolerance 1e-14)."""
    XX = solver.XX
    XXb = XX[solver.Ib]
    XXi = XX[solver.Ii]
    xl = bounds[0][0]
    xr = bounds[1][0]
    xc = (xl + xr) / 2.0
    Il = np.where((np.abs(XXb[:, 0] - xl) < 1e-14) & ~gb(XXb))[0]
    Ir = np.where((np.abs(XXb[:, 0] - xr) < 1e-14) & ~gb(XXb))[0]
    Ic = np.where(np.abs(XXi[:, 0] - xc) < 1e-14)[0]
    Igb = np.where(gb(XXb))[0]
    return Il, Ir, Ic, Igb, XXi, XXb


def oms_stot(pdo, N, leaf_h, q, bc=None, gb=gb_unit_square, periodic=False):
    """S_rk_list, rhs maps and rhs of multislab/oms.py:131-264 with the dense assembler
    (matAssembly/matAssembler.py:93-98: M = linOp @ I):
    per double slab S_l = (A_ii^{-1} A_ib[:, Il])[Ic], S_r likewise,
    R = (A_ii^{-1} A_ib[:, Igb])[Ic], rhs = -R bc(XXb[Igb]).
    periodic: the x-periodic square (dSlabs_periodic; every S_rk_list entry [S_l, S_r]).
    Returns dict(S_rk_list, R_list, XXgb_list, Ic_pts, rhs_list, nc)."""
    slabs, conn, H = dSlabs_periodic(N) if periodic else dSlabs(N)
    if periodic:
        gb = gb_bottom_top
    S_rk_list, R_list, XXgb_list, Ic_pts, rhs_list = [], [], [], [], []
    for k, bnd in enumerate(slabs):
        sol = SlabSolver(pdo, bnd, leaf_h, q)
        Il, Ir, Ic, Igb, XXi, XXb = slab_index_sets(sol, bnd, gb)
        def smap(J):
            Y = sol.lu.solve(sol.Aib[:, J].toarray())
            return Y[Ic]
        S_l = smap(Il) if len(Il) else None
        S_r = smap(Ir) if len(Ir) else None
        R = smap(Igb)
        if conn[k][0] < 0:
            S_rk_list.append([S_r])
        elif conn[k][1] < 0:
            S_rk_list.append([S_l])
        else:
            S_rk_list.append([S_l, S_r])
        R_list.append(R)
        XXgb_list.append(XXb[Igb])
        Ic_pts.append(XXi[Ic])
        if bc is not None:
            rhs_list.append(-R @ bc(XXb[Igb]))
    return dict(S_rk_list=S_rk_list, R_list=R_list, XXgb_list=XXgb_list, Ic_pts=Ic_pts,
                rhs_list=rhs_list, nc=len(Ic_pts[0]), conn=conn, H=H)


def stot_dense_cyclic(S_rk_list):
    """Dense periodic S = I + S_l (block column i-1 mod n) + S_r (block column i+1 mod n)."""
    nb = len(S_rk_list)
    m = S_rk_list[0][0].shape[0]
    S = np.eye(nb * m)
    for i, (Sl, Sr) in enumerate(S_rk_list):
        jl, jr = (i - 1) % nb, (i + 1) % nb
        S[i * m:(i + 1) * m, jl * m:(jl + 1) * m] += Sl
        S[i * m:(i + 1) * m, jr * m:(jr + 1) * m] += Sr
    return S


def stot_dense(S_rk_list):
    """Dense S = I + off-diagonal blocks (multislab/oms.py:267-304 with contiguous dofs)."""
    nb = len(S_rk_list)
    m = S_rk_list[0][0].shape[0]
    S = np.eye(nb * m)
    for i, blocks in enumerate(S_rk_list):
        if i == 0:
            S[0:m, m:2 * m] += blocks[0]
        elif i == nb - 1:
            S[i * m:(i + 1) * m, (i - 1) * m:i * m] += blocks[0]
        else:
            S[i * m:(i + 1) * m, (i - 1) * m:i * m] += blocks[0]
            S[i * m:(i + 1) * m, (i + 1) * m:(i + 2) * m] += blocks[1]
    return S


def global_solution_on_lines(pdo, N, leaf_h, q, bc, periodic=False):
    """Brute force: the single HPS system on all of [0,1]^2 (periodic in x when
    `periodic`) with Dirichlet data bc, solved by sparse LU; returns (values, points) on
    the interface lines x = k/N, k = 1..N-1 (periodic: k = 0..N-1), ordered like the
    S-system unknowns."""
    hx, hy = leaf_h
    nxl = int(round(1.0 / hx))
    nyl = int(round(1.0 / hy))
    rect = HPSRect(pdo, 0.0, 0.0, hx, hy, nxl, nyl, q, periodic_x=periodic)
    A = rect.assemble().tocsc()
    outer = rect.outer_edge_nodes()
    rhs = np.zeros(rect.n)
    rhs[outer] = bc(rect.XX[outer])
    u = spla.spsolve(A, rhs)
    per = int(round((1.0 / N) / hx))
    vals, pts = [], []
    for k in range(0 if periodic else 1, N):
        s = rect.vidx(k * per, 0)
        sl = slice(s, s + nyl * rect.p)
        vals.append(u[sl])
        pts.append(rect.XX[sl])
    return np.concatenate(vals), np.concatenate(pts)


def global_solution(pdo, N, leaf_h, q, bc):
    """Brute force: the single HPS system on all of [0,1]^2 with Dirichlet data bc, solved
    by sparse LU; returns (values, points) at every node.  Pins interior reconstruction
    (solver/pde_solver.py:169-201 solve_dir per slab): with the exact interface values,
    each slab's interior solution is the global one restricted (exact domain
    decomposition)."""
    hx, hy = leaf_h
    rect = HPSRect(pdo, 0.0, 0.0, hx, hy, int(round(1.0 / hx)), int(round(1.0 / hy)), q)
    A = rect.assemble().tocsc()
    outer = rect.outer_edge_nodes()
    rhs = np.zeros(rect.n)
    rhs[outer] = bc(rect.XX[outer])
    return spla.spsolve(A, rhs), rect.XX
