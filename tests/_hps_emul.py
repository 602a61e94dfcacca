"""numpy emulation of the device HPS routines, used only by CPU tests to check the
host-side logic of sslablu_amd.hps (leaf constants, merge / pair index maps, boundary
ordering) against the sparse oracle without a GPU.  Not a fallback: the product never
imports this file."""
import numpy as np

from sslablu_amd import hps


def leaf_dtn(grid, pdo):
    q, p = grid.q, grid.p
    c = hps.leaf_constants(q, grid.hx, grid.hy)
    P = grid.interior_points()
    co = {n: hps._coef(getattr(pdo, n, None), P) for n in ("c11", "c22", "c1", "c2", "c")}
    nl = grid.nleaves
    T = np.zeros((nl, 4 * p, 4 * p))
    p2 = p * p
    for L in range(nl):
        Aii = np.zeros((p2, p2))
        Aib = np.zeros((p2, 4 * p))
        for r in range(p2):
            a, b = r // p + 1, r % p + 1
            k = {n: (v[L * p2 + r] if v is not None else 0.0) for n, v in co.items()}
            for col in range(p2 + 4 * p):
                if col < p2:
                    a2, b2 = col // p + 1, col % p + 1
                else:
                    e = col - p2
                    side, t = e // p, e % p + 1
                    a2, b2 = [(0, t), (q - 1, t), (t, 0), (t, q - 1)][side]
                v = 0.0
                if b2 == b:
                    v += -k["c11"] * c["Dxx"][a, a2] + k["c1"] * c["Dx"][a, a2]
                if a2 == a:
                    v += -k["c22"] * c["Dyy"][b, b2] + k["c2"] * c["Dy"][b, b2]
                if a2 == a and b2 == b:
                    v += k["c"]
                if col < p2:
                    Aii[r, col] = v
                else:
                    Aib[r, col - p2] = v
        X = np.linalg.solve(Aii, Aib)
        T[L] = c["Nb"] - c["Ni"] @ X
    return T


def gather(Ta, Tb, rid0, cid0, rid1, cid1):
    out = np.zeros((len(rid0), len(cid0)))
    for i in range(len(rid0)):
        for j in range(len(cid0)):
            v = 0.0
            if rid0[i] >= 0 and cid0[j] >= 0:
                v += Ta[rid0[i], cid0[j]]
            if rid1[i] >= 0 and cid1[j] >= 0:
                v += Tb[rid1[i], cid1[j]]
            out[i, j] = v
    return out


def merge_tree(grid, T):
    nxb, nyb = 1, 1
    boxes = T
    while nxb < grid.nx or nyb < grid.ny:
        kind = "h" if nxb < grid.nx else "v"
        ca, cb, na, nbm, n_new, new = hps.merge_maps(kind, nxb, nyb, grid.p)
        out = []
        for b in range(boxes.shape[0] // 2):
            A, B = boxes[2 * b], boxes[2 * b + 1]
            Tcc = gather(A, B, ca, ca, cb, cb)
            R = gather(A, B, ca, na, cb, nbm)
            E = gather(A, B, na, ca, nbm, cb)
            Tn = gather(A, B, na, na, nbm, nbm)
            out.append(Tn - E @ np.linalg.solve(Tcc, R))
        boxes = np.stack(out)
        nxb, nyb = new
    return boxes


def pair_solve(grid, Ts):
    ca, cb, cola, colb = hps.pair_maps(grid.nx, grid.ny, grid.p)
    Y = []
    for i in range(grid.nb):
        A, B = Ts[i], Ts[i + 1]
        Tcc = gather(A, B, ca, ca, cb, cb)
        R = gather(A, B, ca, cola, cb, colb)
        Y.append(np.linalg.solve(Tcc, R))
    return np.stack(Y)


def oms_rhs(grid, Y, F):
    m, nb, nxp, N = grid.m, grid.nb, grid.nx * grid.p, grid.N
    out = np.zeros((nb * m, F.shape[1]))
    bot, top = F[:N * nxp], F[N * nxp:2 * N * nxp]
    left, right = F[2 * N * nxp:2 * N * nxp + m], F[2 * N * nxp + m:]
    for i in range(nb):
        v = Y[i, :, 2 * m:2 * m + 2 * nxp] @ bot[i * nxp:(i + 2) * nxp]
        v += Y[i, :, 2 * m + 2 * nxp:] @ top[i * nxp:(i + 2) * nxp]
        if i == 0:
            v += Y[i, :, :m] @ left
        if i == nb - 1:
            v += Y[i, :, m:2 * m] @ right
        out[i * m:(i + 1) * m] = -v
    return out
