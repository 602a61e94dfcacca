"""3-D oms S-system on the Chebyshev tensor-grid (spectral) slab discretization, on the
MI355X.

The unit cube is cut into N single slabs along x (the double slabs of
geometry/geom_3D/cube.py:32-46); every double slab [c-H, c+H] x [0,1]^2 is one
nx x ny x nz Chebyshev grid, as the reference's spectral discretization
(solver/spectral/spectralSolver.py:132-195, 226-251: L = -C11 Dxx(x)I(x)I - C22 I(x)Dyy(x)I
- C33 I(x)I(x)Dzz + C on x-major, y, z node order).  nx is odd, so the centre plane
x = c is a grid plane: the interface carries m = (ny-2)(nz-2) nodes, ordered by y then z
(the order of multislab/oms.py:38-52 on this grid).

For constant coefficients the interior operator is a Kronecker sum, so it diagonalizes
with the eigenvectors of the 1-D interior second-derivative blocks (real eigenvalues,
well-conditioned eigenvectors for Chebyshev Dirichlet problems):

    A_ii = (Vx (x) Vy (x) Vz) diag(-c11 lx - c22 ly - c33 lz + c) (Vx (x) Vy (x) Vz)^{-1}.

The S maps of multislab/oms.py:99-129 then act mode by mode in (y, z):
    S_left = (Vy (x) Vz) diag(s_l) (Vy (x) Vz)^{-1},
    s_l(mu, nu) = e_c^T (-c11 Dxx_ii + sigma_{mu nu})^{-1} (-c11 Dxx[int, 0]),
and likewise S_right with the last x column.  The dense S blocks are formed by one
MFMA GEMM; the boundary-data right-hand sides (oms.py:180-186) by mode transforms along
each axis (strided-batched GEMMs).  With uniform slabs and constant coefficients the two
S blocks are the same for every interface.

The face tree for HBS compression bisects the (ny-2) x (nz-2) interface grid by index,
alternating y and z, so every leaf has the same size (the reduced HBS form needs it;
coordinate midpoints would not, on Chebyshev-spaced points).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_handle
from .hps import cheb_unit

F64 = torch.float64


def cheb_d2(q: int, a: float, b: float):
    """q Chebyshev points ascending on [a, b] and the second-derivative matrix."""
    t, D = cheb_unit(q)
    L = b - a
    D = D / L
    return a + L * t, D @ D


def _bgemm(A, B, strideA, batch, C=None, beta=0.0):
    """C_b = A_b B_b for row-major float64 device matrices (A of (M, K), B of (batch, K, N)
    contiguous); strideA = 0 shares one A."""
    Mdim, K = A.shape[-2], A.shape[-1]
    N = B.shape[-1]
    if C is None:
        C = torch.empty(batch, Mdim, N, dtype=F64, device=B.device)
    call("sslab_dgemm_strided_batched", 0, 0, Mdim, N, K, 1.0, ptr(A), K, strideA, ptr(B), N,
         K * N, beta, ptr(C), N, Mdim * N, batch, stream_handle())
    return C


def _apply_axis(M, X, axis):
    """Contract the small matrix M (n_out x n_in) with axis `axis` of X (a float64 device
    tensor), as batched GEMMs over the other axes."""
    Xp = X.movedim(axis, 0)
    shape = Xp.shape
    lead = Xp.reshape(shape[0], -1).contiguous()
    out = _bgemm(M, lead[None], 0, 1)[0]
    return out.reshape((M.shape[0],) + tuple(shape[1:])).movedim(0, axis)


def face_tree(ny_i: int, nz_i: int, min_leaf: int):
    """Leaf-ordered permutation of the ny_i x nz_i face grid (point (a, b) has index
    a*nz_i + b) by alternating index bisection, y first; every split halves one dimension
    of every box, until boxes have <= min_leaf points.  Returns (perm, Nb, L)."""
    boxes = [((0, ny_i), (0, nz_i))]
    axis = 0
    levels = 1
    while (boxes[0][0][1] - boxes[0][0][0]) * (boxes[0][1][1] - boxes[0][1][0]) > min_leaf:
        ext = [b[axis][1] - b[axis][0] for b in boxes]
        if any(e % 2 for e in ext):
            other = 1 - axis
            if any((b[other][1] - b[other][0]) % 2 for b in boxes):
                raise ValueError("face grid does not bisect into equal leaves")
            axis = other
        nxt = []
        for b in boxes:
            lo, hi = b[axis]
            mid = (lo + hi) // 2
            for half in ((lo, mid), (mid, hi)):
                nb_ = list(b)
                nb_[axis] = half
                nxt.append(tuple(nb_))
        boxes = nxt
        levels += 1
        axis = 1 - axis
    perm = np.concatenate([(np.arange(*b[0])[:, None] * nz_i + np.arange(*b[1])[None, :]).ravel()
                           for b in boxes])
    return perm, len(boxes), levels


class FaceTree:
    """The face tree in the form the HBS mirrors read (perm_leaf, nleaves, nlevels)."""

    def __init__(self, ny_i, nz_i, min_leaf):
        self.perm_leaf, self.nleaves, self.nlevels = face_tree(ny_i, nz_i, min_leaf)

    def uniform(self):
        return True


class SpectralSlabs3D:
    """S-system of a constant-coefficient PDO on the unit cube, N single slabs along x,
    each double slab one nx x ny x nz Chebyshev grid (nx odd)."""

    def __init__(self, N, nx=9, ny=98, nz=98, c11=1.0, c22=1.0, c33=1.0, c=0.0, device=None):
        _lib.require_gpu()
        _lib.load()
        if N < 2:
            raise ValueError("need at least two single slabs (one interface)")
        if nx < 3 or nx % 2 == 0:
            raise ValueError("nx must be odd (the centre plane is a grid plane) and >= 3")
        if ny < 3 or nz < 3:
            raise ValueError("ny, nz >= 3")
        self.N, self.nx, self.ny, self.nz = N, nx, ny, nz
        self.coef = (float(c11), float(c22), float(c33), float(c))
        self.H = 1.0 / N
        self.nb = N - 1
        self.m = (ny - 2) * (nz - 2)
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())

    # -- 1-D spectra -------------------------------------------------------------------
    def _spectra(self):
        H = self.H
        self.xl, D2x = cheb_d2(self.nx, -H, H)
        self.y, D2y = cheb_d2(self.ny, 0.0, 1.0)
        self.z, D2z = cheb_d2(self.nz, 0.0, 1.0)
        out = {}
        for name, D2 in (("x", D2x), ("y", D2y), ("z", D2z)):
            A = D2[1:-1, 1:-1]
            lam, V = np.linalg.eig(A)
            if np.abs(lam.imag).max() > 1e-8 * np.abs(lam).max():
                raise RuntimeError("complex spectrum of the Chebyshev second derivative")
            lam, V = lam.real, V.real
            out[name] = (lam, V, np.linalg.inv(V), D2)
        return out

    def build(self):
        """Dense S blocks (S_left, S_right, m x m float64 on the device)."""
        c11, c22, c33, c = self.coef
        sp = self._spectra()
        self._sp = sp
        lx, Vx, Wx, D2x = sp["x"]
        ly, Vy, Wy, D2y = sp["y"]
        lz, Vz, Wz, D2z = sp["z"]
        ic = (self.nx - 1) // 2 - 1  # centre plane among the interior x nodes
        sigma = (-c22 * ly[:, None] - c33 * lz[None, :] + c).ravel()  # (y, z) mode shifts
        den = -c11 * lx[:, None] + sigma[None, :]                      # (x mode, yz mode)
        if np.abs(den).min() == 0.0:
            raise ZeroDivisionError("singular slab problem (resonant constant coefficient)")
        s = {}
        for side, col in (("left", 0), ("right", self.nx - 1)):
            w = Wx @ (-c11 * D2x[1:-1, col])                               # x-modes of the data
            s[side] = (Vx[ic, :] * w) @ (1.0 / den)                       # (m,)
        dev = self.device
        P = torch.kron(torch.tensor(Vy, device=dev), torch.tensor(Vz, device=dev))
        Pinv = torch.kron(torch.tensor(Wy, device=dev), torch.tensor(Wz, device=dev))
        m = self.m
        self.S = {}
        for side in ("left", "right"):
            Ps = P * torch.tensor(s[side], device=dev)[None, :]
            self.S[side] = _bgemm(Ps, Pinv[None], 0, 1)[0]
        self.S_left, self.S_right = self.S["left"], self.S["right"]
        del P, Pinv
        self._dev_sp = {k: tuple(torch.tensor(a, device=dev) for a in v[:3])
                        for k, v in sp.items()}
        self._den_full = None
        return self

    # -- geometry -----------------------------------------------------------------------
    def interface_points(self):
        yi, zi = self.y[1:-1], self.z[1:-1]
        Y, Z = np.meshgrid(yi, zi, indexing="ij")
        pts = []
        for i in range(self.nb):
            X = np.full(Y.size, (i + 1) * self.H)
            pts.append(np.stack([X, Y.ravel(), Z.ravel()], axis=1))
        return np.concatenate(pts)

    def S_rk_list(self):
        """[[S_r], [S_l, S_r], ..., [S_l]] (multislab/oms.py:239-244)."""
        nb = self.nb
        return [[self.S_right]] + [[self.S_left, self.S_right] for _ in range(nb - 2)] + \
            ([[self.S_left]] if nb > 1 else [])

    def blocks(self):
        """(SiM, SiP) as (nb, m, m) device views of the two S blocks (no copies)."""
        nb, m = self.nb, self.m
        return self.S_left.expand(nb, m, m), self.S_right.expand(nb, m, m)

    # -- right-hand sides ---------------------------------------------------------------
    def rhs_from_function(self, g) -> torch.Tensor:
        """rhs_i = -(A_ii^{-1} A_ib[:, Igb] f)[Ic] for Dirichlet data f = g(P) on the cube
        boundary, g: (n, 3) points -> (n,) or (n, r) values.  Returns (nb*m, r)."""
        c11, c22, c33, c = self.coef
        dev = self.device
        lx, Vx, Wx, D2x = self._sp["x"]
        ly, Vy, Wy, D2y = self._sp["y"]
        lz, Vz, Wz, D2z = self._sp["z"]
        nxi, nyi, nzi = self.nx - 2, self.ny - 2, self.nz - 2
        xi, yi, zi = self.xl[1:-1], self.y[1:-1], self.z[1:-1]
        nb = self.nb

        def ev(P):
            v = np.asarray(g(P), dtype=np.float64)
            return v.reshape(P.shape[0], -1)

        # faces y = 0 / 1 at (x_int, z_int), z = 0 / 1 at (x_int, y_int), per double slab
        Xs = np.stack([(i + 1) * self.H + xi for i in range(nb)])           # (nb, nxi)
        def face(fix, val):
            if fix == "y":
                A_, B_ = np.meshgrid(np.arange(nxi), np.arange(nzi), indexing="ij")
                P = np.stack([Xs[:, A_.ravel()].ravel(), np.full(nb * A_.size, val),
                              np.tile(zi[B_.ravel()], nb)], axis=1)
                return ev(P).reshape(nb, nxi, nzi, -1)
            A_, B_ = np.meshgrid(np.arange(nxi), np.arange(nyi), indexing="ij")
            P = np.stack([Xs[:, A_.ravel()].ravel(), np.tile(yi[B_.ravel()], nb),
                          np.full(nb * A_.size, val)], axis=1)
            return ev(P).reshape(nb, nxi, nyi, -1)
        Gy0, Gy1, Gz0, Gz1 = face("y", 0.0), face("y", 1.0), face("z", 0.0), face("z", 1.0)
        r = Gy0.shape[-1]
        Yg, Zg = np.meshgrid(yi, zi, indexing="ij")
        Gx0 = ev(np.stack([np.zeros(Yg.size), Yg.ravel(), Zg.ravel()], 1)).reshape(nyi, nzi, r)
        Gx1 = ev(np.stack([np.ones(Yg.size), Yg.ravel(), Zg.ravel()], 1)).reshape(nyi, nzi, r)
        # b = -A_ib f on the interior grid (nb, nxi, nyi, nzi, r), built on the host (small)
        b = np.zeros((nb, nxi, nyi, nzi, r))
        dy0, dy1 = D2y[1:-1, 0], D2y[1:-1, -1]
        dz0, dz1 = D2z[1:-1, 0], D2z[1:-1, -1]
        b += c22 * (dy0[None, None, :, None, None] * Gy0[:, :, None, :, :] +
                    dy1[None, None, :, None, None] * Gy1[:, :, None, :, :])
        b += c33 * (dz0[None, None, None, :, None] * Gz0[:, :, :, None, :] +
                    dz1[None, None, None, :, None] * Gz1[:, :, :, None, :])
        b[0] += c11 * D2x[1:-1, 0][:, None, None, None] * Gx0[None]
        b[nb - 1] += c11 * D2x[1:-1, -1][:, None, None, None] * Gx1[None]
        # u = A_ii^{-1} b at the centre plane, by the mode transforms on the device
        lxd, Vxd, Wxd = self._dev_sp["x"]
        lyd, Vyd, Wyd = self._dev_sp["y"]
        lzd, Vzd, Wzd = self._dev_sp["z"]
        B = torch.tensor(b, device=dev)
        for ax, W in ((1, Wxd), (2, Wyd), (3, Wzd)):
            B = _apply_axis(W.contiguous(), B, ax)
        den = (-c11 * lxd[:, None, None] - c22 * lyd[None, :, None] - c33 * lzd[None, None, :]
               + c)
        B = B / den[None, :, :, :, None]
        ic = (self.nx - 1) // 2 - 1
        B = _apply_axis(Vxd[ic:ic + 1].contiguous(), B, 1)                   # (nb, 1, ...)
        for ax, V in ((2, Vyd), (3, Vzd)):
            B = _apply_axis(V.contiguous(), B, ax)
        return B.reshape(nb * self.m, r).contiguous()
