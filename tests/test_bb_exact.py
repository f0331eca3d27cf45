"""Pins the exact Bernstein-Bezier WADG oracle (oracle/bb_exact.py, the algorithm of
rk_update_BBWADG, cpp/okl/WaveKernelsBBWADG.okl:14-321) and through it the quadrature WADG
update: the c_j projection formula equals the L2 projection built by an exact cubature; the BB
product equals the pointwise product; for c^2 in P_1 (and for constant c^2) the quadrature WADG
update of the oracle equals the exact one, since its degree 2N+1 rule then integrates exactly."""
import numpy as np
import pytest

from bbwadg_amd import basis, media, mesh as M, operators as O
import bb_exact as X
import oracle as orc


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_projection_formula_is_the_l2_projection(N):
    r, s, t, w = basis.tet_cubature(3 * N)
    VN, V2 = basis.bern_tet(N, r, s, t), basis.bern_tet(2 * N, r, s, t)
    P = np.linalg.solve(VN.T @ (w[:, None] * VN), VN.T @ (w[:, None] * V2))
    assert np.abs(X.projection(N) - P).max() < 1e-11 * max(1.0, np.abs(P).max())
    # and it reproduces degree-N polynomials given in degree 2N
    c = np.random.default_rng(N).standard_normal(P.shape[0])
    assert np.allclose(X.projection(N) @ (X.elevation(N, 2 * N) @ c), c, atol=1e-11)


@pytest.mark.parametrize("N", [2, 4, 5])
def test_product_is_pointwise(N):
    rng = np.random.default_rng(10 + N)
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    c, p = rng.standard_normal((2, Np))
    pts = rng.uniform(-1, -0.1, (3, 20))
    pts = pts[:, pts.sum(axis=0) < -1.0]
    f = X.bb_product(N, c, p)[0]
    VN, V2 = basis.bern_tet(N, *pts), basis.bern_tet(2 * N, *pts)
    assert np.allclose(V2 @ f, (VN @ c) * (VN @ p), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("c2kind", ["linear", "constant"])
@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_quadrature_wadg_equals_exact_for_linear_c2(N, c2kind):
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(2)
    xq, yq, zq = media.cubature_points(m, ops)
    c2q = 1.0 + 0.3 * xq - 0.2 * yq + 0.25 * zq if c2kind == "linear" else np.full_like(xq, 1.7)
    rng = np.random.default_rng(N)
    rhs = rng.standard_normal((m.K, 4, ops.Np))
    res0 = rng.standard_normal(rhs.shape)
    Q0 = rng.standard_normal(rhs.shape)
    fa, fb, fdt = ops.rk4a[2], ops.rk4b[2], 1e-2
    oo = orc.OracleOps(ops, np.float64)
    Qq, rq = Q0.copy(), res0.copy()
    orc.update_wadg(oo, m.K, c2q, fa, fb, fdt, rhs.copy(), rq, Qq, nproj=1)
    Qe, re = X.update_bbwadg(ops, X.c2_bb(ops, c2q), fa, fb, fdt, rhs, res0.copy(), Q0.copy())
    assert np.abs(re - rq).max() < 1e-11 * np.abs(re).max()
    assert np.abs(Qe - Qq).max() < 1e-11 * np.abs(Qe).max()
    # a c^2 of degree 2 is not integrated exactly: the two updates then differ (the pins bite)
    c2quad = c2q + 0.5 * xq * yq
    rq2 = res0.copy()
    orc.update_wadg(oo, m.K, c2quad, fa, fb, fdt, rhs.copy(), rq2, Q0.copy(), nproj=1)
    _, re2 = X.update_bbwadg(ops, X.c2_bb(ops, c2quad), fa, fb, fdt, rhs, res0.copy(), Q0.copy())
    assert np.abs(re2 - rq2).max() > 1e-9
