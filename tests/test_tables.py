"""Pins the node/cubature tables (data from cpp/include/NodeData.h) by mathematics:
cubature exactness on every monomial up to the rule's degree, positive weights and
interior points, node counts on each face, node symmetry under the vertex permutations."""
import itertools

import numpy as np
import pytest

from bbwadg_amd import basis, tables
import ref_ops


def _tol(deg):
    # the degree 15..21 rules are tabulated to about 1e-11 relative (monomials of degree 19
    # sum 2e3 point values of either sign)
    return 1e-12 if deg <= 13 else 1e-9


@pytest.mark.parametrize("deg", [1, 3, 5, 7, 9, 11, 13, 15, 17, 19])
def test_tet_rule_exact(deg):
    r, s, t, w = basis.tet_cubature(deg)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                exact = ref_ops.monomial_integral_tet(a, b, c)
                got = np.sum(w * r ** a * s ** b * t ** c)
                assert abs(got - exact) < _tol(deg) * max(1.0, abs(exact)), (a, b, c)
    assert np.all(w > 0)
    lam = np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2])
    assert lam.min() > -1e-12  # points inside the reference tetrahedron


@pytest.mark.parametrize("deg", [1, 3, 5, 7, 9, 11, 13, 15, 17, 19])
def test_tri_rule_exact(deg):
    r, s, w = basis.tri_cubature(deg)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = ref_ops.monomial_integral_tri(a, b)
            assert abs(np.sum(w * r ** a * s ** b) - exact) < _tol(deg) * max(1.0, abs(exact))
    assert np.all(w > 0)


def test_rule_sizes_match_reference_counts():
    # p_Nq_N7 = 31, p_Nq_N9 = 57, p_Nq_N11 = 95; p_Nfq_N7 = 15, N9 = 19, N11 = 28
    assert [len(tables.tet_rule(d)["w"]) for d in (7, 9, 11)] == [31, 57, 95]
    assert [len(tables.tri_rule(d)["w"]) for d in (7, 9, 11)] == [15, 19, 28]


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_nodes(N):
    r, s, t = basis.nodes3d(N)
    Np, Nfp = basis.np_tet(N), basis.np_tri(N)
    assert len(r) == Np
    lam = np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2])
    assert lam.min() > -1e-12
    for m in range(4):
        assert np.sum(np.abs(lam[m]) < 1e-10) == Nfp
    # the node set is invariant under all 24 vertex permutations
    key = {tuple(np.round(lam[:, i], 10)) for i in range(Np)}
    for p in itertools.permutations(range(4)):
        key2 = {tuple(np.round(lam[list(p), i], 10)) for i in range(Np)}
        assert key2 == key
    # interpolation is well conditioned (Lebesgue-type bound on the BB interpolation matrix)
    VB = basis.bern_tet(N, r, s, t)
    assert np.linalg.cond(VB) < 1e4


@pytest.mark.parametrize("N", [8, 9])
def test_wadg_tensor_rule(N):
    """Orders with 2N+1 > 15 project on the (N+1)^3 tensor-product set (WaveOKL3d.cpp:435-457):
    its 1D parts are the Gauss-Jacobi(0,0) and (2,0) rules (exact to degree 2N+1 against
    weights 1 and (1-x)^2), and the projection still reproduces P_N: Pq Vq = I."""
    from bbwadg_amd import operators as O
    r, s, t, w = O.wadg_rule(N)
    assert len(w) == (N + 1) ** 3 and np.all(w > 0)
    a, ia = np.unique(np.round(r, 13), return_inverse=True)
    c, ic = np.unique(np.round(t, 13), return_inverse=True)
    assert len(a) == N + 1 and len(c) == N + 1

    def mono(m):  # int_{-1}^{1} x^m dx
        return (1 + (-1) ** m) / (m + 1)

    # marginals: sum over r = a_i of w is wa_i * sum(wa) * sum(wc) = wa_i * 2 * 8/3;
    # sum over t = c_k of w is wc_k * 2 * 2
    wa = np.bincount(ia.ravel(), w) / (2 * 8 / 3)
    wc = np.bincount(ic.ravel(), w) / 4
    for m in range(2 * N + 2):
        assert abs(np.sum(wa * a ** m) - mono(m)) < 1e-12
        exact = mono(m) - 2 * mono(m + 1) + mono(m + 2)  # int (1 - x)^2 x^m
        assert abs(np.sum(wc * c ** m) - exact) < 1e-12
    ops = O.build_ref_ops(N)
    assert np.abs(ops.Pq @ ops.Vq - np.eye(ops.Np)).max() < 1e-8


@pytest.mark.parametrize("N", [8, 9])
def test_wadg_collapsed_rule(N):
    """The collapsed-coordinate form of the tensor-product set: exact to degree 2N+1 on the
    tetrahedron, points inside, positive weights."""
    from bbwadg_amd import operators as O
    r, s, t, w = O.wadg_rule_collapsed(N)
    deg = 2 * N + 1
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                exact = ref_ops.monomial_integral_tet(a, b, c)
                got = np.sum(w * r ** a * s ** b * t ** c)
                assert abs(got - exact) < 1e-9 * max(1.0, abs(exact)), (a, b, c)
    assert np.all(w > 0)
    lam = np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2])
    assert lam.min() > 0


@pytest.mark.parametrize("N", [8, 9])
def test_wadg_tensor_rules_mass(N):
    """Why the collapsed form exists: the reference's un-collapsed set is no rule for the
    tetrahedron, so Vq^T W Vq is not the mass matrix (off by a factor ~1e8) and the WADG
    weighted mass M (Pq diag(c2) Vq)^-1 is not symmetric for variable c2: the scheme loses its
    energy bound (the oracle diverges on it, tests/test_oracle.py).  The collapsed form gives
    the mass matrix to rounding."""
    from bbwadg_amd import operators as O
    errs = {}
    for rule in ("reference", "collapsed"):
        ops = O.build_ref_ops(N, tp_rule=rule)
        Mq = ops.Vq.T @ (ops.wq[:, None] * ops.Vq)
        errs[rule] = np.abs(Mq - ops.MB).max() / np.abs(ops.MB).max()
        assert np.abs(ops.Pq @ ops.Vq - np.eye(ops.Np)).max() < 1e-7
    assert errs["reference"] > 1e3
    assert errs["collapsed"] < 1e-12
