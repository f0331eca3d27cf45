"""Reference-element operators for the BB-WADG solver (FP64 host setup).

Restates the Bernstein block of ``StartUp3d`` (``cpp/src/StartUp3d.cpp:22-57``,
``:228-407``, ``:517-537``) and the WADG projection of ``InitWADG_subelem``
(``cpp/src/WaveOKL3d.cpp:419-483``):

* ``Fmask``: face-node lists (faces t=-1, s=-1, r+s+t=-1, r=-1);
* ``D_ids``/``D_vals``: the four barycentric derivative operators, 4 non-zeros
  per row, as fixed-width sparse rows (``get_sparse_ids``, relative tolerance
  1e-5, columns ascending; ``cpp/src/Basis.cpp:980-1007``).  The reference keeps
  one value array shared by all four operators (the values of the last one); the
  per-row value pattern (l, k, j, i) is the same for all four, so this is exact;
* the BB lift ``LIFTB = MB^-1 EmatB`` factored as ``EEL . blkdiag(L0, L0, L0, L0)``
  with ``EEL`` (Np x 4Nfp, Nfp+3 non-zeros per row) built from triangle degree
  elevation matrices scaled by ``(-1)^i C(N,i)/(1+i)`` and a row permutation per
  face, and ``L0 = LIFTB[:Nfp, :Nfp]`` (at most 7 non-zeros per row);
* ``Vq``/``Pq``: BB interpolation to the degree 2N+1 cubature and the L2
  projection back, ``Pq = (Vq^T W Vq)^-1 Vq^T W``;
* the Carpenter--Kennedy LSRK45 coefficients.
"""
from dataclasses import dataclass, field
from math import comb

import numpy as np

from . import basis

NODETOL = 1e-6  # cpp/include/Mesh.h:13
NFACES = 4


def get_sparse_ids(A, tol=1e-5):
    """Fixed-width sparse rows of A: (cols, vals), ascending columns, zero padded.

    Entries with |A_ij| > tol * max|A| are kept (``cpp/src/Basis.cpp:980-1007``).
    """
    A = np.asarray(A, dtype=np.float64)
    thresh = tol * np.abs(A).max()
    mask = np.abs(A) > thresh
    width = int(mask.sum(axis=1).max())
    cols = np.zeros((A.shape[0], width), dtype=np.int32)
    vals = np.zeros((A.shape[0], width))
    for i in range(A.shape[0]):
        nz = np.nonzero(mask[i])[0]
        cols[i, : len(nz)] = nz
        vals[i, : len(nz)] = A[i, nz]
    return cols, vals


def _fixed_width(cols, vals, width):
    """Pad or truncate sparse rows to ``width`` (the reference's fixed nnz)."""
    n = cols.shape[0]
    c = np.zeros((n, width), dtype=np.int32)
    v = np.zeros((n, width))
    w = min(width, cols.shape[1])
    c[:, :w] = cols[:, :w]
    v[:, :w] = vals[:, :w]
    return c, v, cols.shape[1]


def face_masks(r, s, t, N):
    """Fmask[f] = node ids on face f, in node order (``StartUp3d.cpp:41-57``)."""
    conds = [np.abs(1 + t), np.abs(1 + s), np.abs(1 + r + s + t), np.abs(1 + r)]
    Fm = [np.nonzero(c < NODETOL)[0] for c in conds]
    nfp = basis.np_tri(N)
    for f in Fm:
        if len(f) != nfp:
            raise RuntimeError("face node count mismatch")
    return np.stack(Fm).astype(np.int32)  # (4, Nfp)


def lsrk45_coefficients():
    """Carpenter--Kennedy (1994) 5-stage 4th-order 2N-storage RK (``StartUp3d.cpp:517-537``)."""
    a = np.array([0.0,
                  -567301805773.0 / 1357537059087.0,
                  -2404267990393.0 / 2016746695238.0,
                  -3550918686646.0 / 2091501179385.0,
                  -1275806237668.0 / 842570457699.0])
    b = np.array([1432997174477.0 / 9575080441755.0,
                  5161836677717.0 / 13612068292357.0,
                  1720146321549.0 / 2090206949498.0,
                  3134564353537.0 / 4481467310338.0,
                  2277821191437.0 / 14882151754819.0])
    c = np.array([0.0,
                  1432997174477.0 / 9575080441755.0,
                  2526269341429.0 / 6820363962896.0,
                  2006345519317.0 / 3224310063776.0,
                  2802321613138.0 / 2924317926251.0,
                  1.0])
    return a, b, c


@dataclass
class RefOps:
    """Reference-element operators for one polynomial order N (all FP64)."""
    N: int
    Np: int
    Nfp: int
    r: np.ndarray
    s: np.ndarray
    t: np.ndarray
    Fmask: np.ndarray          # (4, Nfp) int32
    VB: np.ndarray             # (Np, Np) BB -> nodal values
    invVB: np.ndarray
    DB: list                   # 4 dense (Np, Np) barycentric derivative matrices
    D_ids: np.ndarray          # (4 ops, Np, 4) int32: D_ids[m, i, j] = j-th column of row i of D_{m+1}
    D_vals: np.ndarray         # (Np, 4) values shared by the four operators
    LIFTB: np.ndarray          # (Np, 4Nfp)
    EEL: np.ndarray            # (Np, 4Nfp) dense
    EEL_ids: np.ndarray        # (Np, EEL_nnz) int32
    EEL_vals: np.ndarray       # (Np, EEL_nnz)
    EEL_nnz: int
    L0: np.ndarray             # (Nfp, Nfp) dense
    L0_ids: np.ndarray         # (Nfp, L0_nnz) int32
    L0_vals: np.ndarray
    L0_nnz: int
    cEL: np.ndarray            # (N+1,)
    rowperm: np.ndarray        # (Np, 3)
    MB: np.ndarray             # BB mass matrix on the reference element
    rq: np.ndarray
    sq: np.ndarray
    tq: np.ndarray
    wq: np.ndarray
    Vq: np.ndarray             # (Nq, Np) BB interpolation to cubature (WADG rule)
    Pq: np.ndarray             # (Np, Nq) L2 projection
    rk4a: np.ndarray = field(default_factory=lambda: lsrk45_coefficients()[0])
    rk4b: np.ndarray = field(default_factory=lambda: lsrk45_coefficients()[1])
    rk4c: np.ndarray = field(default_factory=lambda: lsrk45_coefficients()[2])
    sparse_widths: dict = field(default_factory=dict)

    @property
    def Nq(self):
        return len(self.wq)

    @property
    def NfpNfaces(self):
        return NFACES * self.Nfp


def build_ref_ops(N, tp_rule="reference"):
    """Build every reference-element operator for order N (``StartUp3d`` + WADG).

    ``tp_rule`` picks the WADG cubature for 2N+1 > 15 (N = 8, 9): "reference" is the
    reference's un-collapsed tensor-product set (``wadg_rule``), "collapsed" the same 1D rules
    mapped through the collapsed coordinates (``wadg_rule_collapsed``).  Orders N <= 7 use the
    tetrahedral rule of degree 2N+1 either way."""
    if N < 1:
        raise ValueError("N must be >= 1")
    Np, Nfp = basis.np_tet(N), basis.np_tri(N)
    r, s, t = basis.nodes3d(N)
    Fmask = face_masks(r, s, t, N)

    VB = basis.bern_tet(N, r, s, t)
    invVB = np.linalg.inv(VB)
    G = basis.grad_bern_tet(N, r, s, t)
    DB = [np.linalg.solve(VB, Gm) for Gm in G]

    widths = {}
    D_ids = np.zeros((4, Np, 4), dtype=np.int32)
    D_vals = None
    for m in range(4):
        cols, vals = get_sparse_ids(DB[m])
        c, v, w = _fixed_width(cols, vals, 4)
        widths[f"D{m + 1}"] = w
        D_ids[m] = c
        D_vals = v  # overwritten by each operator, as StartUp3d.cpp:255-259

    # --- BB lift: LIFTB = MB \ EmatB (StartUp3d.cpp:286-307)
    qdeg = min(21, 2 * N + 1)
    rqt, sqt, wqt = basis.tri_cubature(qdeg)
    Vface = basis.bern_tri(N, rqt, sqt)
    massFace = Vface.T @ (wqt[:, None] * Vface)
    EmatB = np.zeros((Np, NFACES * Nfp))
    for f in range(NFACES):
        for i in range(Nfp):
            EmatB[Fmask[f, i], f * Nfp:(f + 1) * Nfp] += massFace[i]
    rq, sq, tq, wq = basis.tet_cubature(qdeg)
    VBq = basis.bern_tet(N, rq, sq, tq)
    MB = VBq.T @ (wq[:, None] * VBq)
    LIFTB = np.linalg.solve(MB, EmatB)

    # --- EEL1: stacked cEL_i * E_{N-i}^{N T} (StartUp3d.cpp:312-327)
    cEL = np.array([(-1) ** i * comb(N, i) / (1 + i) for i in range(N + 1)])
    VBTri = basis.bern_tri(N, rqt, sqt)
    blocks = []
    for i in range(N + 1):
        VBi = basis.bern_tri(N - i, rqt, sqt)
        Ei = np.linalg.lstsq(VBTri, VBi, rcond=None)[0]  # (Nfp, NpElev): elevation N-i -> N
        blocks.append(cEL[i] * Ei.T)
    EEL1 = np.vstack(blocks)
    assert EEL1.shape == (Np, Nfp)

    # --- face permutations by matching LIFTB rows (StartUp3d.cpp:330-356)
    LIFTBf1 = LIFTB[:, :Nfp]
    maxVal = np.abs(LIFTB).max()
    rowperm = np.zeros((Np, NFACES - 1), dtype=np.int32)
    for f in range(1, NFACES):
        LIFTBf = LIFTB[:, f * Nfp:(f + 1) * Nfp]
        for i in range(Np):
            diff = np.abs(LIFTBf1 - LIFTBf[i][None, :]).max(axis=1)
            hits = np.nonzero(diff < NODETOL * maxVal)[0]
            if len(hits) == 0:
                raise RuntimeError(f"no lift row match for face {f}, row {i}")
            rowperm[i, f - 1] = hits[0]
    EEL = np.zeros((Np, NFACES * Nfp))
    EEL[:, :Nfp] = EEL1
    for f in range(1, NFACES):
        EEL[:, f * Nfp:(f + 1) * Nfp] = EEL1[rowperm[:, f - 1]]
    cols, vals = get_sparse_ids(EEL)
    EEL_nnz = Nfp + 3  # StartUp3d.cpp:375
    EEL_ids, EEL_vals, widths["EEL"] = _fixed_width(cols, vals, EEL_nnz)

    L0 = LIFTB[:Nfp, :Nfp].copy()
    cols, vals = get_sparse_ids(L0)
    L0_nnz = min(Nfp, 7)  # StartUp3d.cpp:395, WaveOKL3d.cpp:1492
    L0_ids, L0_vals, widths["L0"] = _fixed_width(cols, vals, L0_nnz)

    # --- WADG projection on the degree 2N+1 rule (WaveOKL3d.cpp:425-483)
    if N > 9:
        raise ValueError("orders above 9 are not supported")
    if tp_rule == "reference":
        rqw, sqw, tqw, wqw = wadg_rule(N)
    elif tp_rule == "collapsed":
        rqw, sqw, tqw, wqw = wadg_rule_collapsed(N)
    else:
        raise ValueError("tp_rule must be 'reference' or 'collapsed'")
    Vq = basis.bern_tet(N, rqw, sqw, tqw)
    if 2 * N + 1 <= 15:
        Mq = Vq.T @ (wqw[:, None] * Vq)
        Pq = np.linalg.solve(Mq, Vq.T * wqw[None, :])
    else:
        # same operator through a QR of W^1/2 Vq: the tensor-product points lie partly outside
        # the element, where Bernstein values reach 1e4 and Vq^T W Vq has condition ~1e13
        sw = np.sqrt(wqw)
        Qf, Rf = np.linalg.qr(Vq * sw[:, None])
        Pq = np.linalg.solve(Rf, Qf.T) * sw[None, :]

    return RefOps(N=N, Np=Np, Nfp=Nfp, r=r, s=s, t=t, Fmask=Fmask, VB=VB, invVB=invVB,
                  DB=DB, D_ids=D_ids, D_vals=D_vals, LIFTB=LIFTB, EEL=EEL,
                  EEL_ids=EEL_ids, EEL_vals=EEL_vals, EEL_nnz=EEL_nnz, L0=L0,
                  L0_ids=L0_ids, L0_vals=L0_vals, L0_nnz=L0_nnz, cEL=cEL, rowperm=rowperm,
                  MB=MB, rq=rqw, sq=sqw, tq=tqw, wq=wqw, Vq=Vq, Pq=Pq, sparse_widths=widths)


def wadg_rule(N):
    """The reference's WADG cubature (``InitWADG_subelem``, cpp/src/WaveOKL3d.cpp:425-458): the
    tetrahedral rule of degree 2N+1 when 2N+1 <= 15, else its (N+1)^3-point tensor-product set:
    r = a_i, s = a_j, t = c_k with a, wa the (N+1)-point Gauss-Jacobi(0,0) rule and c, wc the
    Gauss-Jacobi(2,0) rule, w = wa_i wa_j wc_k (k outer, i, j inner).  That set is used as the
    reference uses it, without a collapsed-coordinate map, so its points are not all inside the
    tetrahedron: Pq = (Vq^T W Vq)^-1 Vq^T W is then a weighted least-squares fit on those points
    (Pq Vq = I still holds)."""
    if 2 * N + 1 <= 15:
        return basis.tet_cubature(min(21, 2 * N + 1))
    from scipy.special import roots_jacobi
    a, wa = roots_jacobi(N + 1, 0.0, 0.0)
    c, wc = roots_jacobi(N + 1, 2.0, 0.0)
    K, I, J = np.meshgrid(np.arange(N + 1), np.arange(N + 1), np.arange(N + 1), indexing="ij")
    K, I, J = K.ravel(), I.ravel(), J.ravel()
    return a[I], a[J], c[K], wa[I] * wa[J] * wc[K]


def wadg_rule_collapsed(N):
    """A degree 2N+1 rule on the reference tetrahedron with (N+1)^3 points for 2N+1 > 15: the
    Gauss-Jacobi (0,0), (1,0) and (2,0) rules in the collapsed coordinates (a, b, c), mapped by
    t = c, s = (1+b)(1-c)/2 - 1, r = (1+a)(1-b)(1-c)/4 - 1 (Jacobian (1-b)(1-c)^2/8); the
    weights of the (1,0) and (2,0) rules absorb (1-b) and (1-c)^2, so w = wa wb wc / 8.  All
    points lie inside the element, so the projection is a true L2 projection there.

    The reference's set (``wadg_rule``) skips this map: its points fill the cube [-1,1]^3 and the
    WADG operator Pq diag(c2) Vq has eigenvalues ~1e7 once c2 varies, so explicit time
    stepping diverges for heterogeneous media at N = 8, 9; with constant c2 both give Pq Vq = I.
    For 2N+1 <= 15 the tetrahedral rule is returned, as by ``wadg_rule``."""
    if 2 * N + 1 <= 15:
        return basis.tet_cubature(min(21, 2 * N + 1))
    from scipy.special import roots_jacobi
    a, wa = roots_jacobi(N + 1, 0.0, 0.0)
    b, wb = roots_jacobi(N + 1, 1.0, 0.0)
    c, wc = roots_jacobi(N + 1, 2.0, 0.0)
    K, I, J = np.meshgrid(np.arange(N + 1), np.arange(N + 1), np.arange(N + 1), indexing="ij")
    A, B, C = a[I.ravel()], b[J.ravel()], c[K.ravel()]
    t = C
    s = 0.5 * (1 + B) * (1 - C) - 1
    r = 0.25 * (1 + A) * (1 - B) * (1 - C) - 1
    w = wa[I.ravel()] * wb[J.ravel()] * wc[K.ravel()] / 8.0
    return r, s, t, w


def slice_lift_ops(ops):
    """Tables of the slice-by-slice Bernstein lift (``rk_surface_bern_slice``,
    cpp/okl/WaveKernels.okl:652-939; set up in cpp/src/StartUp3d.cpp:409-505).

    The lift of face f's data phi_f is EEL (L0 phi_f): the face's block of EEL puts
    cEL_s (E_{N-s}^N)^T (L0 phi_f) on the nodes of slice s (the nodes whose barycentric index for
    the vertex opposite face f equals s), and (E_{N-s}^N)^T is s one-step degree reductions, so
    slice s follows from slice s - 1 with a 3-non-zero operator.  Returns

    * ``slice_ids`` (Np, 4) int32: row sk = (slice s, its k-th node), column f: the volume node.
      Within a slice the nodes are in ascending volume order, which is the triangle BB order of
      the face's nodes (``Fmask``) at degree N - s;
    * ``red_ids`` / ``red_vals`` (Np - Nfp, 3): the one-step reductions (E_{N-s}^{N-s+1})^T for
      s = 1..N stacked (rows of slice s), at most 3 non-zeros (zero-padded);
    * ``cEL`` (N + 1,).
    """
    N, Np = ops.N, ops.Np
    mi = basis.tet_multi_indices(N)
    opposite = [3, 2, 0, 1]   # faces t = -1, s = -1, r+s+t = -1, r = -1: lam4, lam3, lam1, lam2 = 0
    slice_ids = np.zeros((Np, 4), dtype=np.int32)
    for f in range(4):
        v = opposite[f]
        order = sorted(range(Np), key=lambda n: (mi[n][v], n))
        slice_ids[:, f] = order
    # one-step elevation E (degree d-1 -> d on the triangle) from the BB identity
    # B^{d-1}_a = sum_m (a_m + 1)/d B^d_{a + e_m}; its transpose has <= 3 non-zeros per row
    rows_ids, rows_vals = [], []
    for sl in range(1, N + 1):
        d = N - sl + 1
        hi = {tuple(x): n for n, x in enumerate(basis.tri_multi_indices(d))}
        for a in basis.tri_multi_indices(d - 1):
            ents = []
            for m in range(3):
                b = a.copy()
                b[m] += 1
                ents.append((hi[tuple(b)], (a[m] + 1) / d))
            ents.sort()
            rows_ids.append([e[0] for e in ents])
            rows_vals.append([e[1] for e in ents])
    red_ids = np.array(rows_ids, dtype=np.int32).reshape(-1, 3)
    red_vals = np.array(rows_vals, dtype=np.float64).reshape(-1, 3)
    return dict(slice_ids=slice_ids, red_ids=red_ids, red_vals=red_vals, cEL=ops.cEL.copy())


def slice_lift_apply(ops, sl, phi):
    """Host restatement of the slice-by-slice lift for one element: phi (4, Nfp) face data in
    Fmask order -> (Np,) = LIFTB @ phi.ravel() computed slice by slice (for checking the tables)."""
    N, Nfp = ops.N, ops.Nfp
    out = np.zeros(ops.Np)
    for f in range(4):
        cur = ops.L0 @ phi[f]
        sk, r0 = 0, 0
        for s_ in range(N + 1):
            nn = (N - s_ + 1) * (N - s_ + 2) // 2
            if s_ > 0:
                nxt = np.zeros(nn)
                for r in range(nn):
                    for c in range(3):
                        nxt[r] += sl["red_vals"][r0 + r, c] * cur[sl["red_ids"][r0 + r, c]]
                r0 += nn
                cur = nxt
            out[sl["slice_ids"][sk:sk + nn, f]] += sl["cEL"][s_] * cur
            sk += nn
    return out


def packed_device_ops(ops):
    """Pack operators in the device layouts of ``WaveInitOCCA3d`` (FP32/int32).

    * ``D_ids`` as 4 int4 arrays: ``D_idsj[i] = (D1[i][j], D2[i][j], D3[i][j], D4[i][j])``
      (``cpp/src/WaveOKL3d.cpp:1452-1480``); here shape (4 j, Np, 4 m);
    * ``Dvals4[i] = D_vals[i][0:4]`` (row-major float4);
    * ``EEL``, ``L0`` ids/values column-major: entry (i, j) at ``i + j*rows``
      (``cpp/src/WaveOKL3d.cpp:1485-1501``);
    * ``Vq`` column-major Nq x Np, ``Pq`` column-major Np x Nq
      (``cpp/okl/WaveKernelsWADG.okl:1812,1849``);
    * ``Fmask`` row-major (4, Nfp).
    """
    f32 = np.float32
    D_ids = np.ascontiguousarray(np.transpose(ops.D_ids, (2, 1, 0)), dtype=np.int32)  # [j][i][m]
    return {
        "Fmask": np.ascontiguousarray(ops.Fmask, dtype=np.int32),
        "D_ids": D_ids,
        "Dvals": np.ascontiguousarray(ops.D_vals, dtype=f32),
        "EEL_ids": np.ascontiguousarray(ops.EEL_ids.T, dtype=np.int32),
        "EEL_vals": np.ascontiguousarray(ops.EEL_vals.T, dtype=f32),
        "L0_ids": np.ascontiguousarray(ops.L0_ids.T, dtype=np.int32),
        "L0_vals": np.ascontiguousarray(ops.L0_vals.T, dtype=f32),
        "Vq": np.ascontiguousarray(ops.Vq.T, dtype=f32),   # column-major Nq x Np
        "Pq": np.ascontiguousarray(ops.Pq.T, dtype=f32),   # column-major Np x Nq
    }


def bbwadg_exact_ops(N):
    """Operators of the exact Bernstein-Bezier WADG update (``rk_update_BBWADG``), restating the
    BB branch of ``InitWADG_subelem`` (``cpp/src/WaveOKL3d.cpp:488-690``):

    * ``subtet_ids`` [Np][Np]: degree-2N index of the multi-index sum a + b (the BB product
      rule B^N_a B^N_b = C(N,a) C(N,b) / C(2N,a+b) B^2N_{a+b}); symmetric;
    * ``CNscale`` [Np] = C(N, a), ``invC2Nscale`` [N2p] = 1 / C(2N, g) (multinomials);
    * ``Ei_ids4``/``Ei_vals4`` [2N][N2p][4]: the one-step degree elevations E_i (degree i ->
      i + 1, i = 0 .. 2N-1) as fixed-width sparse rows (at most 4 non-zeros, value g_m / (i+1)
      at the column of g - e_m), rows beyond Np(i+1) zero; the reference obtains them as
      VB_{i+1} \\ VB_i at the degree-(i+1) nodes;
    * ``cj`` [N+1]: P_{N <- 2N} = sum_j c_j E_{N-j}^N (E_{N-j}^{2N})^T, from the reference's
      (N+1) x (N+1) system L c = b (d = 3).
    """
    from math import factorial

    def mult(M, a):
        c = factorial(M)
        for x in a:
            c //= factorial(int(x))
        return c

    def fr(a, b):
        r = 1
        for x in range(b + 1, a + 1):
            r *= x
        return r

    mi = basis.tet_multi_indices(N)
    Np = len(mi)
    m2 = basis.tet_multi_indices(2 * N)
    N2p = len(m2)
    idx2 = {tuple(x): i for i, x in enumerate(m2)}
    subtet = np.array([[idx2[tuple(a + b)] for b in mi] for a in mi], dtype=np.int32)
    CN = np.array([mult(N, a) for a in mi], dtype=np.float64)
    invC2N = np.array([1.0 / mult(2 * N, g) for g in m2], dtype=np.float64)
    Eids = np.zeros((2 * N, N2p, 4), dtype=np.int32)
    Evals = np.zeros((2 * N, N2p, 4), dtype=np.float64)
    for i in range(2 * N):
        lo = {tuple(x): n for n, x in enumerate(basis.tet_multi_indices(i))}
        for row, g in enumerate(basis.tet_multi_indices(i + 1)):
            ents = []
            for m in range(4):
                if g[m] > 0:
                    a = g.copy()
                    a[m] -= 1
                    ents.append((lo[tuple(a)], g[m] / (i + 1)))
            for w, (col, v) in enumerate(sorted(ents)):
                Eids[i, row, w], Evals[i, row, w] = col, v
    L = np.zeros((N + 1, N + 1))
    for j in range(N + 1):
        Nj = N - j
        for i in range(Nj + 1):
            L[i, j] = fr(Nj, Nj - i) / fr(Nj + i + 3, Nj)
    b = np.array([fr(2 * N, 2 * N - i) / fr(2 * N + i + 3, 2 * N) for i in range(N + 1)])
    cj = np.linalg.solve(L, b)
    return dict(subtet_ids=subtet, CNscale=CN, invC2Nscale=invC2N, Ei_ids4=Eids, Ei_vals4=Evals,
                cj=cj, Np=Np, N2p=N2p)
