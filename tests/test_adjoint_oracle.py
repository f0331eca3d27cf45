"""Pins the FP64 host adjoint (oracle/adjoint.py): the dot-product identity
<A x, y> = <x, A^T y> against the C forward kernels, <W x, y> = <x, W^T y>, and the
adjoint gradient of the receiver misfit against central finite differences."""
import numpy as np
import pytest

from bbwadg_amd import media
import oracle as orc
import adjoint as adjm
from cases import Problem


@pytest.mark.parametrize("N", [2, 3, 5])
def test_dot_product_identity(N):
    P = Problem(N, 2, seed=N)
    oo = orc.OracleOps(P.ops, np.float64)
    adj = adjm.AdjointOps(P.ops, P.dm)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((P.K, 4, P.ops.Np))
    y = rng.standard_normal((P.K, 4, P.ops.Np))
    Ax = adjm.forward_rhs(oo, P.K, P.dm, x)
    ATy = adj.AT(y)
    lhs, rhs = np.sum(Ax * y), np.sum(x * ATy)
    assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0)
    Wx = adjm.W(P.ops, P.c2q, x)
    WTy = adjm.WT(P.ops, P.c2q, y)
    assert abs(np.sum(Wx * y) - np.sum(x * WTy)) < 1e-10 * abs(np.sum(Wx * y))


def small_inverse_problem(N=2, n=2, nsteps=12, seed=0):
    P = Problem(N, n, seed=seed)
    oo = orc.OracleOps(P.ops, np.float64)
    adj = adjm.AdjointOps(P.ops, P.dm)
    elems, coef = media.gaussian_source(P.mesh, P.ops, (0.1, -0.05, 0.0), a=6.0)
    src = np.zeros((P.K, P.ops.Np)); src[elems] = coef
    relems, B = media.receivers(P.mesh, P.ops, media.surface_receivers(3, 5))
    dt = P.dt_cfl()
    Q0 = np.zeros((P.K, 4, P.ops.Np))
    c2_true = P.c2q
    _, d, _ = adjm.forward(oo, P.ops, P.dm, c2_true, Q0, dt, nsteps, relems, B, src, 10.0, 50.0)
    c2 = np.ones_like(c2_true)
    return P, oo, adj, src, relems, B, dt, Q0, d, c2


def test_gradient_matches_central_differences():
    P, oo, adj, src, relems, B, dt, Q0, d, c2 = small_inverse_problem()
    nsteps = d.shape[0]
    g, J = adjm.gradient(oo, P.ops, P.dm, adj, c2, Q0, dt, nsteps, relems, B, d, src, 10.0, 50.0)
    assert J > 0
    rng = np.random.default_rng(11)
    f = lambda c: adjm.misfit(adjm.forward(oo, P.ops, P.dm, c, Q0, dt, nsteps, relems, B, src, 10.0, 50.0)[1], d)
    for _ in range(3):
        v = rng.standard_normal(c2.shape)
        ad = float(np.sum(g * v))
        # central differences converge as eps^2 to the adjoint value: O(1e-5) at 1e-4 and
        # O(1e-7) at 1e-5 relative on this problem
        for eps, tol in ((1e-4, 5e-5), (1e-5, 2e-6)):
            fd = (f(c2 + eps * v) - f(c2 - eps * v)) / (2 * eps)
            assert abs(fd - ad) < tol * abs(ad), (eps, fd, ad)
