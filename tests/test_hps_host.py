"""Host logic of the GPU HPS assembly (CPU only): leaf constants, merge / pair index
maps and the boundary-data ordering, run through a numpy emulation of the device
routines and compared with the sparse oracle (oracle/hps.py)."""
import numpy as np
import pytest

from oracle import hps as oh
from sslablu_amd import hps
from tests import _hps_emul as em
from tests.test_oracle_hps import LAPLACE, helmholtz


def test_leaf_constants_match_oracle_cheb():
    t, D = hps.cheb_unit(9)
    t2, D2 = oh.unit_cheb(9)
    assert np.allclose(t, t2) and np.allclose(D, D2)


def test_leaf_dtn_harmonic_polynomial():
    """A leaf DtN applied to a harmonic polynomial's edge values gives its exact
    outward normal derivative (degree <= q-1 is resolved exactly)."""
    grid = hps.SlabGrid2D(2, 1, 2, 7)
    T = em.leaf_dtn(grid, LAPLACE)
    p, q = grid.p, grid.q
    x0, y0 = grid.leaf_origins()
    t = grid.t[1:-1]
    u = lambda x, y: x ** 3 - 3 * x * y ** 2 + 2 * x * y     # harmonic
    ux = lambda x, y: 3 * x ** 2 - 3 * y ** 2 + 2 * y
    uy = lambda x, y: -6 * x * y + 2 * x
    L = 1
    X0, Y0, hx, hy = x0[L], y0[L], grid.hx, grid.hy
    xs, ys = X0 + t * hx, Y0 + t * hy
    ub = np.concatenate([u(X0, ys), u(X0 + hx, ys), u(xs, Y0), u(xs, Y0 + hy)])
    flux = np.concatenate([-ux(X0, ys), ux(X0 + hx, ys), -uy(xs, Y0), uy(xs, Y0 + hy)])
    assert np.allclose(T[L] @ ub, flux, atol=1e-9)


@pytest.mark.parametrize("N,nx,ny,q,pdo", [(3, 1, 2, 6, "laplace"), (4, 2, 2, 5, "helmholtz"),
                                           (3, 2, 4, 5, "laplace")])
def test_emulated_assembly_matches_oracle(N, nx, ny, q, pdo):
    pd = LAPLACE if pdo == "laplace" else helmholtz(6.0)
    grid = hps.SlabGrid2D(N, nx, ny, q)
    Y = em.pair_solve(grid, em.merge_tree(grid, em.leaf_dtn(grid, pd)))
    ref = oh.oms_stot(pd, N, (grid.hx, grid.hy), q)
    m = grid.m
    for i, blocks in enumerate(ref["S_rk_list"]):
        if i == 0:
            assert np.allclose(Y[i, :, m:2 * m], blocks[0], atol=1e-11)
        elif i == grid.nb - 1:
            assert np.allclose(Y[i, :, :m], blocks[0], atol=1e-11)
        else:
            assert np.allclose(Y[i, :, :m], blocks[0], atol=1e-11)
            assert np.allclose(Y[i, :, m:2 * m], blocks[1], atol=1e-11)
    # boundary data -> rhs agrees with the oracle's rhs = -R bc(XXb[Igb])
    bc = lambda P: np.cos(P[:, 0] + 2 * P[:, 1]) + P[:, 0] ** 2
    F = bc(grid.gb_points())[:, None]
    rhs = em.oms_rhs(grid, Y, F)[:, 0]
    ref_rhs = np.concatenate([-R @ bc(X) for R, X in zip(ref["R_list"], ref["XXgb_list"])])
    assert np.allclose(rhs, ref_rhs, atol=1e-11)
    assert np.allclose(grid.interface_points(), np.concatenate(ref["Ic_pts"]), atol=1e-15)
