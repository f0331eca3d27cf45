"""Pins the FP64 host adjoint of the elastic scheme (oracle/adjoint_elas.py): the dot-product
identity <A x, y> = <x, A^T y> against the C forward kernels (orc_volume_elas +
orc_surface_elas) with unequal penalties, <W x, y> = <x, W^T y>, and the adjoint gradient of the
receiver misfit in 1/rho, lambda and mu against central finite differences.  A planted swap of
the two penalties in the transpose, or of lambda and mu in the gradient, breaks the checks."""
import numpy as np
import pytest

from bbwadg_amd import media
import oracle as orc
import adjoint_elas as ae
from cases import Problem

TAU = (0.8, 1.3)


def media_at(P, seed=3):
    lam, mu, rho = media.lame_fourier(seed)
    return (1.0 / rho(P.xq, P.yq, P.zq), lam(P.xq, P.yq, P.zq), mu(P.xq, P.yq, P.zq))


def dot_gap(P, adj, oo, tau, seed=7):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((P.K, 9, P.ops.Np))
    y = rng.standard_normal((P.K, 9, P.ops.Np))
    lhs = np.sum(ae.forward_rhs(oo, P.K, P.dm, x, *tau) * y)
    rhs = np.sum(x * adj.AT(y))
    return abs(lhs - rhs) / abs(lhs)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_dot_product_identity(N):
    P = Problem(N, 2, nfields=9, seed=N)
    oo = orc.OracleOps(P.ops, np.float64)
    adj = ae.ElasAdjointOps(P.ops, P.dm, *TAU)
    assert dot_gap(P, adj, oo, TAU) < 1e-12
    m = media_at(P)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((P.K, 9, P.ops.Np))
    y = rng.standard_normal((P.K, 9, P.ops.Np))
    a, b = np.sum(ae.W(P.ops, m, x) * y), np.sum(x * ae.WT(P.ops, m, y))
    assert abs(a - b) < 1e-12 * abs(a)
    # W is the oracle's rk_update_elas projection (FP64 C kernel with fa = 0, fb = 0, fdt = 1)
    res = np.zeros_like(x)
    orc.update_elas(oo, P.K, *m, 0.0, 0.0, 1.0, x, res, np.zeros_like(x))
    assert np.abs(res - ae.W(P.ops, m, x)).max() < 1e-12 * np.abs(res).max()


def test_planted_penalty_swap_breaks_the_identity():
    P = Problem(2, 2, nfields=9, seed=2)
    oo = orc.OracleOps(P.ops, np.float64)
    bad = ae.ElasAdjointOps(P.ops, P.dm, TAU[1], TAU[0])
    assert dot_gap(P, bad, oo, TAU) > 1e-3


def small_inverse_problem(N=2, n=2, nsteps=10, seed=0, rec_field=0):
    P = Problem(N, n, nfields=9, seed=seed)
    oo = orc.OracleOps(P.ops, np.float64)
    adj = ae.ElasAdjointOps(P.ops, P.dm, *TAU)
    elems, coef = media.gaussian_source(P.mesh, P.ops, (0.1, -0.05, 0.0), a=6.0)
    src = np.zeros((P.K, P.ops.Np)); src[elems] = coef
    srcw = np.array([0.2, 0, 0, 1, 1, 1, 0, 0.3, 0])
    relems, B = media.receivers(P.mesh, P.ops, media.surface_receivers(3, 5))
    dt = 0.5 * P.dt_cfl()
    Q0 = np.zeros((P.K, 9, P.ops.Np))
    m_true = media_at(P, seed=4)
    _, d, _ = ae.forward(oo, P.ops, P.dm, m_true, Q0, dt, nsteps, relems, B, rec_field, src, srcw,
                         10.0, 50.0, *TAU)
    m0 = tuple(np.full_like(a, float(np.mean(a))) for a in m_true)
    return P, oo, adj, src, srcw, relems, B, dt, Q0, d, m0, rec_field


@pytest.mark.parametrize("rec_field", [0, 5])
def test_gradient_matches_central_differences(rec_field):
    P, oo, adj, src, srcw, relems, B, dt, Q0, d, m0, rf = small_inverse_problem(rec_field=rec_field)
    nsteps = d.shape[0]
    g, J = ae.gradient(oo, P.ops, P.dm, adj, m0, Q0, dt, nsteps, relems, B, d, rf, src, srcw, 10.0, 50.0)
    assert J > 0

    def f(m):
        return ae.misfit(ae.forward(oo, P.ops, P.dm, m, Q0, dt, nsteps, relems, B, rf, src, srcw,
                                    10.0, 50.0, *TAU)[1], d)

    rng = np.random.default_rng(11)
    for which in range(3):
        v = rng.standard_normal(m0[0].shape)
        ad = float(np.sum(g[which] * v))
        assert abs(ad) > 0
        for eps, tol in ((1e-4, 5e-5), (1e-5, 2e-6)):
            mp = list(m0); mm = list(m0)
            mp[which] = m0[which] + eps * v
            mm[which] = m0[which] - eps * v
            fd = (f(tuple(mp)) - f(tuple(mm))) / (2 * eps)
            assert abs(fd - ad) < tol * abs(ad), (which, eps, fd, ad)
    # a gradient with lambda and mu swapped fails the same check
    v = rng.standard_normal(m0[0].shape)
    eps = 1e-5
    mp = list(m0); mm = list(m0)
    mp[1] = m0[1] + eps * v
    mm[1] = m0[1] - eps * v
    fd = (f(tuple(mp)) - f(tuple(mm))) / (2 * eps)
    assert abs(fd - float(np.sum(g[2] * v))) > 1e-2 * abs(fd)
