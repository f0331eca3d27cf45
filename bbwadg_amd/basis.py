"""Bernstein--Bezier bases on the reference tetrahedron and triangle (FP64, host).

Reference element: the simplex with vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1),
(-1,-1,1) in (r, s, t).  Barycentric coordinates follow the reference
(``cpp/src/Basis.cpp:541-546``):

    lam1 = -(1+r+s+t)/2,  lam2 = (1+r)/2,  lam3 = (1+s)/2,  lam4 = (1+t)/2.

Multi-index ordering (``cpp/src/Basis.cpp:548-553``): ``l`` outermost, then
``k``, then ``j``, with ``i = N-j-k-l``; the basis function is
``N!/(i!j!k!l!) lam1^i lam2^j lam3^k lam4^l``.  The triangle basis orders
``k`` outer, ``j`` inner, ``i = N-j-k`` (``cpp/src/Basis.cpp:517-537``).
"""
from functools import lru_cache
from math import factorial

import numpy as np

from . import tables


def np_tet(N):
    return (N + 1) * (N + 2) * (N + 3) // 6


def np_tri(N):
    return (N + 1) * (N + 2) // 2


@lru_cache(maxsize=None)
def tet_multi_indices(N):
    """(Np, 4) int array of (i, j, k, l) in the reference's BB ordering."""
    idx = []
    for l in range(N + 1):
        for k in range(N + 1 - l):
            for j in range(N + 1 - l - k):
                idx.append((N - j - k - l, j, k, l))
    return np.array(idx, dtype=np.int64)


@lru_cache(maxsize=None)
def tri_multi_indices(N):
    """(Nfp, 3) int array of (i, j, k) in the reference's triangle BB ordering."""
    idx = []
    for k in range(N + 1):
        for j in range(N + 1 - k):
            idx.append((N - j - k, j, k))
    return np.array(idx, dtype=np.int64)


def _multinomial(N, alpha):
    c = factorial(N)
    for a in alpha:
        c //= factorial(int(a))
    return float(c)


def tet_barycentric(r, s, t):
    r, s, t = (np.asarray(x, dtype=np.float64) for x in (r, s, t))
    return np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2])


def bern_tet(N, r, s, t):
    """BB Vandermonde V[q, alpha] = B_alpha(r_q, s_q, t_q) (``BernTet``)."""
    lam = tet_barycentric(r, s, t)
    mi = tet_multi_indices(N)
    V = np.empty((lam.shape[1], len(mi)))
    for c, a in enumerate(mi):
        V[:, c] = _multinomial(N, a) * np.prod([lam[m] ** a[m] for m in range(4)], axis=0)
    return V


def grad_bern_tet(N, r, s, t):
    """Derivatives of the BB basis w.r.t. each barycentric coordinate (``GradBernTet``).

    Returns [V1, V2, V3, V4] with Vm[q, alpha] = d B_alpha / d lam_m, treating the
    four barycentric coordinates as independent (``cpp/src/Basis.cpp:568-620``).
    """
    lam = tet_barycentric(r, s, t)
    mi = tet_multi_indices(N)
    out = [np.zeros((lam.shape[1], len(mi))) for _ in range(4)]
    for c, a in enumerate(mi):
        C = _multinomial(N, a)
        for m in range(4):
            if a[m] == 0:
                continue
            prod = np.ones(lam.shape[1])
            for mm in range(4):
                e = a[mm] - (1 if mm == m else 0)
                prod = prod * lam[mm] ** e
            out[m][:, c] = C * a[m] * prod
    return out


def bern_tri(N, r, s):
    """Triangle BB Vandermonde (``BernTri``), lam = (-(r+s)/2, (1+r)/2, (1+s)/2)."""
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    lam = np.stack([-(r + s) / 2, (1 + r) / 2, (1 + s) / 2])
    mi = tri_multi_indices(N)
    V = np.empty((lam.shape[1], len(mi)))
    for c, a in enumerate(mi):
        V[:, c] = _multinomial(N, a) * lam[0] ** a[0] * lam[1] ** a[1] * lam[2] ** a[2]
    return V


def nodes3d(N):
    """Warburton interpolation nodes of order N (``Nodes3D``)."""
    d = tables.nodes(N)
    return d["r"], d["s"], d["t"]


def tet_cubature(deg):
    d = tables.tet_rule(deg)
    return d["r"], d["s"], d["t"], d["w"]


def tri_cubature(deg):
    d = tables.tri_rule(deg)
    return d["r"], d["s"], d["w"]
