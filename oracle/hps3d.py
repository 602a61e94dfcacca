"""Oracle: 3-D Chebyshev HPS slab discretization and the oms S-system (numpy/scipy FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference's 3-D drivers discretize each double slab with HPS leaves from the absent
`hpsalt` submodule (solver/solver.py:120-144, `Domain_Driver(...).build("reduced_cpu",
"MUMPS")`, driven by test/3D/cubeTest.py:114-126 with p_disc = p + 2 points per leaf
side).  This restates the published 3-D Chebyshev HPS collocation (Hao & Martinsson,
"A direct solver for elliptic PDEs in three dimensions based on hierarchical merging of
Poincare-Steklov operators", 2016 -- the paper pdo/pdo.py:50-52 cites) with the
reference's conventions:

* a leaf is a q x q x q tensor Chebyshev grid (q = p_disc); its p^3 interior nodes
  (p = q - 2) carry the PDE collocation, its 6 p^2 face nodes (edges and corners
  excluded) are shared with the neighbours, where the equation is continuity of the
  normal derivative;
* the operator follows pdo/pdo.py:41-65 (PDO3d) without mixed derivatives:
  A[u] = -c11 u_xx - c22 u_yy - c33 u_zz + c1 u_x + c2 u_y + c3 u_z + c u;
* a double slab is exposed like solverWrapper.construct (solver/solver.py:79-163):
  XX, Ii, Ib, A_ii, A_ib, so the oms maps of multislab/oms.py:38-52, 99-129, 180-186
  apply verbatim: S_J = (A_ii^{-1} A_ib[:, J])[Ic], rhs = -(A_ii^{-1} A_ib[:, Igb] f)[Ic];
* slabs come from dSlabs (geometry/geom_3D/cube.py:32-46).

Interface nodes (Ic, Il, Ir) are ordered lexicographically by (y, z).

Parity pins (tests/test_oracle_hps3d.py): exactness on a harmonic polynomial (every
leaf reproduces polynomials of degree <= q - 1 per variable exactly), exact domain
decomposition (the oms S-system solution equals one global HPS solve restricted to the
interface planes, ~1e-12), spectral convergence to exp(sqrt(2) pi x) sin(pi y)
sin(pi z), the Laplace row-sum invariant, a planted flux-sign error that the checks
catch, and agreement with the 2-D oracle's Chebyshev matrices (shared `unit_cheb`).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .hps import eval_coef, unit_cheb


def gb_unit_cube(P):
    """Global-boundary predicate of [0,1]^3 (geometry/geom_3D/cube.py:7-8)."""
    P = np.asarray(P)
    return np.any((np.abs(P) < 1e-14) | (np.abs(P - 1.0) < 1e-14), axis=1)


def dSlabs(N):
    """Double slabs of [0,1]^3 and their connectivity (geometry/geom_3D/cube.py:32-46)."""
    H = 1.0 / N
    slabs, conn = [], []
    for n in range(N - 1):
        c = (n + 1) * H
        slabs.append([[c - H, 0.0, 0.0], [c + H, 1.0, 1.0]])
        conn.append([-1 if n == 0 else n - 1, -1 if n == N - 2 else n + 1])
    return slabs, conn, H


class HPSBox3D:
    """All leaf-interior and leaf-face nodes of the box origin + [0, nl * h] tiled by
    nl[0] x nl[1] x nl[2] leaves of q^3 Chebyshev points.

    Nodes are identified by integer lattice keys (i (q-1) + a, j (q-1) + b, l (q-1) + c)
    for leaf (i, j, l) and local grid index (a, b, c); a key with two or more of its
    local indices on a leaf boundary (an edge or corner) is not a node.  Equations: PDE
    at leaf interiors, flux continuity at internal faces, identity at the outer faces."""

    def __init__(self, pdo, origin, h, nl, q, periodic_x=False):
        self.pdo = pdo
        self.periodic_x = periodic_x
        self.origin = np.asarray(origin, dtype=np.float64)
        self.h = np.asarray(h, dtype=np.float64)
        self.nl = tuple(int(v) for v in nl)
        self.q = q
        self.p = q - 2
        self.t, D = unit_cheb(q)
        self.D, self.D2 = D, D @ D
        Q = q - 1
        loc = np.stack(np.meshgrid(np.arange(q), np.arange(q), np.arange(q), indexing="ij"),
                       -1).reshape(-1, 3)
        on_bnd = (loc == 0) | (loc == Q)
        self.loc = loc
        self.loc_ok = on_bnd.sum(1) <= 1
        keys = []
        for i in range(self.nl[0]):
            for j in range(self.nl[1]):
                for l in range(self.nl[2]):
                    keys.append(self._wrap(loc[self.loc_ok] + np.array([i, j, l]) * Q))
        keys = np.unique(np.concatenate(keys), axis=0)
        self.keys = keys
        self.n = len(keys)
        self._dims = np.array(self.nl) * Q + 1
        self._code = {int(c): k for k, c in enumerate(self._encode(keys))}
        self.XX = self.coords(keys)

    def _wrap(self, K):
        """Periodic in x: the face x = nl[0] (q-1) is the face x = 0 (one set of nodes)."""
        if self.periodic_x:
            K = K.copy()
            K[:, 0] %= self.nl[0] * (self.q - 1)
        return K

    def _encode(self, K):
        d = self._dims
        return (K[:, 0] * d[1] + K[:, 1]) * d[2] + K[:, 2]

    def index(self, K):
        return np.array([self._code.get(int(c), -1) for c in self._encode(K)])

    def coords(self, K):
        Q = self.q - 1
        cell, a = np.divmod(K, Q)
        return self.origin + (cell + self.t[a]) * self.h

    def leaf_map(self, i, j, l):
        """Node index of every local grid point (q^3, C order over (a, b, c)); -1 at
        edges and corners."""
        g = -np.ones(len(self.loc), dtype=np.int64)
        K = self._wrap(self.loc[self.loc_ok] + np.array([i, j, l]) * (self.q - 1))
        g[self.loc_ok] = self.index(K)
        return g.reshape(self.q, self.q, self.q)

    def outer_nodes(self):
        Q = self.q - 1
        K = self.keys
        lim = np.array(self.nl) * Q
        d0 = 1 if self.periodic_x else 0
        return np.where(np.any((K[:, d0:] == 0) | (K[:, d0:] == lim[d0:]), axis=1))[0]

    def assemble(self, _plant=None):
        """Sparse square matrix with one row per node (see the class docstring).
        _plant: None, or 'flux_sign' to flip the sign of the x-hi face flux of every
        leaf -- a deliberately wrong oracle for the tests that prove the checks bite."""
        q, p, pdo = self.q, self.p, self.pdo
        h = self.h
        D1 = [self.D / h[d] for d in range(3)]
        D2 = [self.D2 / h[d] ** 2 for d in range(3)]
        names2 = ("c11", "c22", "c33")
        names1 = ("c1", "c2", "c3")
        rows, cols, vals = [], [], []
        flux_r, flux_c, flux_v = [], [], []
        ii = np.arange(1, q - 1)
        I3 = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
        for i in range(self.nl[0]):
            for j in range(self.nl[1]):
                for l in range(self.nl[2]):
                    g = self.leaf_map(i, j, l)
                    row = g[I3[:, 0], I3[:, 1], I3[:, 2]]
                    Xi = self.XX[row]
                    for d in range(3):
                        k2 = eval_coef(getattr(pdo, names2[d], None), Xi)
                        k1 = eval_coef(getattr(pdo, names1[d], None), Xi)
                        for s in range(q):
                            idx = I3.copy()
                            idx[:, d] = s
                            v = np.zeros(len(row))
                            if k2 is not None:
                                v -= k2 * D2[d][I3[:, d], s]
                            if k1 is not None:
                                v += k1 * D1[d][I3[:, d], s]
                            rows.append(row)
                            cols.append(g[idx[:, 0], idx[:, 1], idx[:, 2]])
                            vals.append(v)
                    k0 = eval_coef(getattr(pdo, "c", None), Xi)
                    if k0 is not None:
                        rows.append(row); cols.append(row); vals.append(k0)
                    # outward normal derivative at the six faces' nodes
                    F2 = np.stack(np.meshgrid(ii, ii, indexing="ij"), -1).reshape(-1, 2)
                    for d in range(3):
                        other = [e for e in range(3) if e != d]
                        for end, sgn in ((0, -1.0), (q - 1, 1.0)):
                            if _plant == "flux_sign" and d == 0 and end == q - 1:
                                sgn = -sgn
                            base = np.zeros((len(F2), 3), dtype=np.int64)
                            base[:, other[0]] = F2[:, 0]
                            base[:, other[1]] = F2[:, 1]
                            base[:, d] = end
                            node = g[base[:, 0], base[:, 1], base[:, 2]]
                            for s in range(q):
                                idx = base.copy()
                                idx[:, d] = s
                                flux_r.append(node)
                                flux_c.append(g[idx[:, 0], idx[:, 1], idx[:, 2]])
                                flux_v.append(np.full(len(node), sgn * D1[d][end, s]))
        outer = self.outer_nodes()
        is_outer = np.zeros(self.n, dtype=bool)
        is_outer[outer] = True
        fr, fc, fv = (np.concatenate(x) for x in (flux_r, flux_c, flux_v))
        keep = ~is_outer[fr]
        rows += [fr[keep], outer]
        cols += [fc[keep], outer]
        vals += [fv[keep], np.ones(len(outer))]
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(self.n, self.n))
        A.sum_duplicates()
        return A


def _yz_order(P):
    return np.lexsort((P[:, 2], P[:, 1]))


class SlabSolver3D:
    """solverWrapper.construct-style view of one 3-D double slab: XX, Ii, Ib, A_ii, A_ib
    and a sparse LU of A_ii (solver/solver.py:79-163)."""

    def __init__(self, pdo, bounds, leaf_h, q, _plant=None):
        lo, hi = (np.asarray(b, dtype=np.float64) for b in bounds)
        h = np.asarray(leaf_h, dtype=np.float64)
        nl = np.rint((hi - lo) / h).astype(int)
        self.box = HPSBox3D(pdo, lo, h, nl, q)
        A = self.box.assemble(_plant)
        outer = self.box.outer_nodes()
        mask = np.ones(self.box.n, dtype=bool)
        mask[outer] = False
        self.Ib = outer
        self.Ii = np.where(mask)[0]
        self.XX = self.box.XX
        self.Aii = A[self.Ii][:, self.Ii].tocsc()
        self.Aib = A[self.Ii][:, self.Ib].tocsc()
        self.lu = spla.splu(self.Aii)


def slab_index_sets(solver, bounds, gb=gb_unit_cube):
    """Il, Ir, Ic, Igb, XXi, XXb of multislab/oms.py:38-52 (tolerance 1e-14); Il, Ir
    and Ic in lexicographic (y, z) order."""
    XX = solver.XX
    XXb = XX[solver.Ib]
    XXi = XX[solver.Ii]
    xl, xr = bounds[0][0], bounds[1][0]
    xc = (xl + xr) / 2.0
    gbb = gb(XXb)
    Il = np.where((np.abs(XXb[:, 0] - xl) < 1e-14) & ~gbb)[0]
    Ir = np.where((np.abs(XXb[:, 0] - xr) < 1e-14) & ~gbb)[0]
    Ic = np.where(np.abs(XXi[:, 0] - xc) < 1e-14)[0]
    Il, Ir, Ic = Il[_yz_order(XXb[Il])], Ir[_yz_order(XXb[Ir])], Ic[_yz_order(XXi[Ic])]
    Igb = np.where(gbb)[0]
    return Il, Ir, Ic, Igb, XXi, XXb


def dSlabs_periodic(N):
    """Double slabs of [0,1]^3 periodic in x (geometry/geom_3D/squareTorus.py:268-277):
    double slab n centred on x = n H, connected to (n-1) % N and (n+1) % N."""
    H = 1.0 / N
    slabs = [[[n * H - H, 0.0, 0.0], [n * H + H, 1.0, 1.0]] for n in range(N)]
    return slabs, [[(n - 1) % N, (n + 1) % N] for n in range(N)], H


def gb_sides(P):
    """Global boundary of the x-periodic cube: the faces y, z = 0, 1
    (geometry/geom_3D/squareTorus.py:43-48)."""
    P = np.asarray(P)
    return np.any((np.abs(P[:, 1:]) < 1e-14) | (np.abs(P[:, 1:] - 1.0) < 1e-14), axis=1)


def oms_stot(pdo, N, leaf_h, q, bc=None, gb=gb_unit_cube, _plant=None, periodic=False):
    """S_rk_list, rhs maps and rhs of multislab/oms.py:131-264 for cube.dSlabs(N) with
    the dense assembler (matAssembly/matAssembler.py:93-98): per double slab
    S_l = (A_ii^{-1} A_ib[:, Il])[Ic], S_r likewise, R = (A_ii^{-1} A_ib[:, Igb])[Ic],
    rhs = -R bc(XXb[Igb]).  Returns dict(S_rk_list, R_list, XXgb_list, Ic_pts, rhs_list,
    nc, conn, H).  periodic: the x-periodic cube (dSlabs_periodic, every entry
    [S_l, S_r])."""
    slabs, conn, H = dSlabs_periodic(N) if periodic else dSlabs(N)
    if periodic:
        gb = gb_sides
    S_rk_list, R_list, XXgb_list, Ic_pts, rhs_list = [], [], [], [], []
    for k, bnd in enumerate(slabs):
        sol = SlabSolver3D(pdo, bnd, leaf_h, q, _plant)
        Il, Ir, Ic, Igb, XXi, XXb = slab_index_sets(sol, bnd, gb)

        def smap(J):
            Y = sol.lu.solve(sol.Aib[:, J].toarray())
            return Y[Ic]

        blocks = []
        if conn[k][0] >= 0:
            blocks.append(smap(Il))
        if conn[k][1] >= 0:
            blocks.append(smap(Ir))
        S_rk_list.append(blocks)
        R = smap(Igb)
        R_list.append(R)
        XXgb_list.append(XXb[Igb])
        Ic_pts.append(XXi[Ic])
        if bc is not None:
            rhs_list.append(-R @ bc(XXb[Igb]))
    return dict(S_rk_list=S_rk_list, R_list=R_list, XXgb_list=XXgb_list, Ic_pts=Ic_pts,
                rhs_list=rhs_list, nc=len(Ic_pts[0]), conn=conn, H=H)


def global_solution_on_planes(pdo, N, leaf_h, q, bc, periodic=False):
    """Brute force: the single HPS system on all of [0,1]^3 with Dirichlet data bc,
    solved by sparse LU; returns (values, points) on the interface planes x = k/N,
    k = 1..N-1, each in (y, z) order (the S-system unknowns)."""
    h = np.asarray(leaf_h, dtype=np.float64)
    nl = np.rint(1.0 / h).astype(int)
    box = HPSBox3D(pdo, np.zeros(3), h, nl, q, periodic_x=periodic)
    A = box.assemble().tocsc()
    outer = box.outer_nodes()
    rhs = np.zeros(box.n)
    rhs[outer] = bc(box.XX[outer])
    u = spla.spsolve(A, rhs)
    vals, pts = [], []
    for k in range(0 if periodic else 1, N):
        sel = np.where(np.abs(box.XX[:, 0] - k / N) < 1e-14)[0]
        sel = sel[_yz_order(box.XX[sel])]
        vals.append(u[sel])
        pts.append(box.XX[sel])
    return np.concatenate(vals), np.concatenate(pts)


def stot_dense(S_rk_list):
    """Dense S = I + off-diagonal blocks (multislab/oms.py:267-304, contiguous dofs)."""
    from .hps import stot_dense as _sd
    return _sd(S_rk_list)
