"""Partial differential operator coefficient containers.

Mirror of pdo/pdo.py (SimonDirckx/SslabLU): const (:4-16), PDO2d (:19-39), PDO3d
(:41-65), get_known_greens (:67-102).  The 2-D operator is
A[u] = -c11 u_xx - c22 u_yy - 2 c12 u_xy + c1 u_x + c2 u_y + c u; coefficients are
callables of an (N, d) point array (numpy or torch).  The GPU HPS assembly supports
c12 = None (no mixed derivative), like the Chebyshev leaves it implements.
"""
import numpy as np
from scipy.special import j0


def const(c=1):
    def const_func(xxloc):
        assert xxloc.shape[1] == 2 or xxloc.shape[1] == 3
        return c * np.ones(xxloc.shape[0],)
    return const_func


class PDO2d:
    def __init__(self, c11, c22, c12=None, c1=None, c2=None, c=None):
        self.c11, self.c22 = c11, c22
        self.c12 = c12
        self.c1, self.c2 = c1, c2
        self.c = c


class PDO3d:
    def __init__(self, c11, c22, c33, c12=None, c13=None, c23=None, c1=None, c2=None, c3=None,
                 c=None):
        self.c11, self.c22, self.c33 = c11, c22, c33
        self.c12, self.c13, self.c23 = c12, c13, c23
        self.c1, self.c2, self.c3 = c1, c2, c3
        self.c = c


def get_known_greens(xx, kh, center=None):
    """Free-space Green's function centred outside the domain: log r / J0(kh r) in 2-D,
    1/r / sin(kh r)/r in 3-D (pdo.py:67-102)."""
    xx = np.asarray(xx, dtype=np.float64)
    ndim = xx.shape[-1]
    if center is None:
        center = np.ones(ndim) * 10
    rr = np.sqrt(((xx - np.asarray(center)[None, :]) ** 2).sum(axis=1))
    if ndim == 2:
        u = np.log(rr) if kh == 0 else j0(kh * rr)
    else:
        u = 1 / rr if kh == 0 else np.sin(kh * rr) / rr
    return u[:, None]
