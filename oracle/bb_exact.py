"""Exact Bernstein-Bezier WADG update, FP64 -- TEST INFRASTRUCTURE ONLY (a second oracle for the
WADG projection; nothing in the product imports it).

Restates the algorithm of ``rk_update_BBWADG`` (cpp/okl/WaveKernelsBBWADG.okl:14-321) with its
setup (cpp/src/WaveOKL3d.cpp:486-702):

1. c^2 is projected to degree N in BB coefficients, ``c2_bb = Pq c2q`` (:486);
2. the product of two degree-N BB polynomials is the degree-2N BB polynomial
   f_g = sum_{a + b = g} C(N,a) C(N,b) / C(2N,g) c_a p_b (``CNscale``, ``invC2Nscale``,
   ``subtet_ids``, :488-571; kernel lines 55-96);
3. f is projected back to degree N by P = sum_j c_j E_{N-j}^N (E_{N-j}^{2N})^T, with E the BB
   degree elevation operators and c_j solving the reference's (N+1)x(N+1) system L c = b
   (:655-690): the degree reductions 2N -> N (kernel lines 127-180) followed by the two sweeps
   of sum_j c_j E E^T (lines 199-300).  As committed, the kernel has the 2N -> N reduction
   disabled (``#if 0``) and so projects only the first Np product coefficients; this restates the
   complete algorithm, which is what the pins below check.

Pins (tests/test_bb_exact.py): P equals M_N^-1 M_{N,2N} built by an exact cubature; the
product equals the pointwise product at random points; for c^2 in P_1 the quadrature WADG
update of the oracle equals this exact update (its degree 2N+1 rule integrates c^2 p B exactly).
"""
from functools import lru_cache
from math import factorial

import numpy as np

from bbwadg_amd import basis


def _fr(a, b):
    """a! / b! for a >= b >= 0 (``factorial_ratio``)."""
    r = 1
    for x in range(b + 1, a + 1):
        r *= x
    return r


def _multinom(N, a):
    c = factorial(N)
    for x in a:
        c //= factorial(int(x))
    return c


@lru_cache(maxsize=None)
def elevation(a, b):
    """E (Np(b) x Np(a)): degree-a BB coefficients -> degree-b coefficients of the same polynomial,
    E[g, al] = C(a, al) C(b - a, g - al) / C(b, g) for g >= al componentwise."""
    ma, mb = basis.tet_multi_indices(a), basis.tet_multi_indices(b)
    E = np.zeros((len(mb), len(ma)))
    for i, g in enumerate(mb):
        for j, al in enumerate(ma):
            d = g - al
            if np.all(d >= 0):
                E[i, j] = _multinom(a, al) * _multinom(b - a, d) / _multinom(b, g)
    return E


@lru_cache(maxsize=None)
def projection_cj(N, M=None):
    """c_j of P_{N <- M} = sum_j c_j E_{N-j}^N (E_{N-j}^M)^T (reference: M = 2N, d = 3)."""
    M = 2 * N if M is None else M
    d = 3
    L = np.zeros((N + 1, N + 1))
    for j in range(N + 1):
        Nj = N - j
        for i in range(Nj + 1):
            L[i, j] = _fr(Nj, Nj - i) / _fr(Nj + i + d, Nj)
    b = np.array([_fr(M, M - i) / _fr(M + i + d, M) for i in range(N + 1)], dtype=np.float64)
    return np.linalg.solve(L, b)


@lru_cache(maxsize=None)
def projection(N, M=None):
    """P_{N <- M} (Np(N) x Np(M)) from the c_j formula."""
    M = 2 * N if M is None else M
    cj = projection_cj(N, M)
    P = np.zeros((len(basis.tet_multi_indices(N)), len(basis.tet_multi_indices(M))))
    for j in range(N + 1):
        P += cj[j] * elevation(N - j, N) @ elevation(N - j, M).T
    return P


@lru_cache(maxsize=None)
def product_tables(N):
    """(g, a, b, w) arrays: f_g += w c_a p_b with w = C(N,a) C(N,b) / C(2N,g), g the index of a+b in
    the degree-2N ordering."""
    mi = basis.tet_multi_indices(N)
    m2 = basis.tet_multi_indices(2 * N)
    index2 = {tuple(x): i for i, x in enumerate(m2)}
    G, A, B, W = [], [], [], []
    for a, al in enumerate(mi):
        for b, be in enumerate(mi):
            g = al + be
            G.append(index2[tuple(g)])
            A.append(a)
            B.append(b)
            W.append(_multinom(N, al) * _multinom(N, be) / _multinom(2 * N, g))
    return np.array(G), np.array(A), np.array(B), np.array(W, dtype=np.float64)


def bb_product(N, c, p):
    """Degree-2N BB coefficients of c * p; c, p (K, Np) -> (K, Np(2N))."""
    G, A, B, W = product_tables(N)
    c = np.atleast_2d(np.asarray(c, dtype=np.float64))
    p = np.atleast_2d(np.asarray(p, dtype=np.float64))
    S = np.zeros((len(basis.tet_multi_indices(2 * N)), len(G)))
    S[G, np.arange(len(G))] = W                       # scatter of the Np^2 pair products
    return (c[:, A] * p[:, B]) @ S.T


def c2_bb(ops, c2q):
    """Degree-N BB coefficients of c^2 per element, Pq c2q (:486); c2q (K, Nq)."""
    return np.asarray(c2q, dtype=np.float64) @ ops.Pq.T


def wadg_exact(ops, c2b, rhs_p):
    """P_N(c2 p) for every element: (K, Np) from c2b (K, Np) and rhs_p (K, Np)."""
    f = bb_product(ops.N, c2b, rhs_p)
    return f @ projection(ops.N).T


def update_bbwadg(ops, c2b, fa, fb, fdt, rhs, res, Q):
    """LSRK stage with the exact BB-WADG pressure update (velocity rows unweighted), in place:
    rhs_p <- P_N(c2 rhs_p); res = fa res + fdt rhs; Q += fb res."""
    r = np.array(rhs, dtype=np.float64, copy=True)
    r[:, 0] = wadg_exact(ops, c2b, r[:, 0])
    res *= fa
    res += fdt * r
    Q += fb * res
    return Q, res
