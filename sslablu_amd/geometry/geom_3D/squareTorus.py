"""The periodic slab box of the reference's square-torus geometry.

Mirror of geometry/geom_3D/squareTorus.py (SimonDirckx/SslabLU): bnds, gb_np / gb (:43-48,
217-223: the faces y = 0, 1 and z = 0, 1 -- the box is periodic in x), box_geom, bounds,
the torus parameterization z1, z2, z3 (:19-23, 115-135) and dSlabs(N) (:268-277): N
double slabs, double slab n centred on x = n H (the first straddles x = 0 = 1), with
periodic connectivity [(n-1) % N, (n+1) % N].

param_geom (:238-259), the change of variables that turns a PDO on the torus into one on
this box, produces mixed second derivatives, which the GPU HPS leaves do not discretize;
it raises NotImplementedError.  PDOs without mixed terms on the x-periodic box are
assembled (sslablu_amd.hps / hps3d with periodic grids).
"""
import numpy as np

bnds = [[0, 0, 0], [1, 1, 1]]
const_theta = 1 / (2 * np.pi)


def gb_np(p):
    p = np.asarray(p)
    return ((np.abs(p[:, 1] - bnds[0][1]) < 1e-14) | (np.abs(p[:, 1] - bnds[1][1]) < 1e-14)
            | (np.abs(p[:, 2] - bnds[0][2]) < 1e-14) | (np.abs(p[:, 2] - bnds[1][2]) < 1e-14))


def gb(p, jax_avail=False, torch_avail=False):
    if hasattr(p, "detach"):
        p = p.detach().cpu().numpy()
    return gb_np(p)


def z1(p, jax_avail=False, torch_avail=False):
    p = np.asarray(p)
    return (1 + (p[:, 1] - .5)) * np.cos((p[:, 0] - .5) / const_theta)


def z2(p, jax_avail=False, torch_avail=False):
    p = np.asarray(p)
    return (1 + (p[:, 1] - .5)) * np.sin((p[:, 0] - .5) / const_theta)


def z3(p, jax_avail=False, torch_avail=False):
    return np.asarray(p)[:, 2] - .5


def box_geom(jax_avail=False, torch_avail=True):
    return np.array(bnds)


def bounds():
    return bnds


def param_geom(jax_avail=False, torch_avail=False, hpsalt=False):
    raise NotImplementedError("the torus change of variables yields mixed second derivatives, "
                              "which the GPU HPS leaves do not discretize")


def dSlabs(N):
    H = (bnds[1][0] - bnds[0][0]) / N
    slabs, connectivity = [], []
    for n in range(N):
        c = bnds[0][0] + n * H
        slabs.append([[c - H, bnds[0][1], bnds[0][2]], [c + H, bnds[1][1], bnds[1][2]]])
        connectivity.append([(n - 1) % N, (n + 1) % N])
    return slabs, connectivity, H
