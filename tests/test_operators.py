"""Pins the BB operators of bbwadg_amd.operators (restating StartUp3d.cpp:228-407 and
WaveOKL3d.cpp:479-483) against constructions that share none of their steps
(oracle/ref_ops.py): closed-form derivative and mass matrices, a face-quadrature lift,
and a nodal DG operator from a monomial basis."""
import numpy as np
import pytest

from bbwadg_amd import basis, mesh as M, operators as O
import oracle as orc
import ref_ops

NS = [1, 2, 3, 4, 5, 6, 7, 8, 9]


@pytest.fixture(scope="module", params=NS)
def ops(request):
    return O.build_ref_ops(request.param)


def test_derivatives_closed_form(ops):
    Dc = ref_ops.bb_deriv_closed(ops.N)
    for m in range(4):
        assert np.abs(ops.DB[m] - Dc[m]).max() < 1e-10
    # the sparse rows reproduce the dense operators (4 nnz, shared values)
    for m in range(4):
        D = np.zeros((ops.Np, ops.Np))
        for i in range(ops.Np):
            for j in range(4):
                D[i, ops.D_ids[m, i, j]] += ops.D_vals[i, j]
        assert np.abs(D - Dc[m]).max() < 1e-10
    # 1 = (sum lam)^N has coefficients all 1 and d/dlam_m of it is N (coefficients all N)
    one = np.ones(ops.Np)
    for m in range(4):
        assert np.abs(Dc[m] @ one - ops.N).max() < 1e-12
    # d/dr of r: r = 2 lam2 - 1 -> BB coefficients of lam2 are j/N
    lam2 = basis.tet_multi_indices(ops.N)[:, 1] / ops.N
    r_coef = 2 * lam2 - 1
    assert np.abs(0.5 * (Dc[1] - Dc[0]) @ r_coef - 1.0).max() < 1e-12
    assert np.abs(0.5 * (Dc[2] - Dc[0]) @ r_coef).max() < 1e-12


def test_mass_closed_form(ops):
    assert np.abs(ops.MB - ref_ops.bb_mass_closed(ops.N)).max() < 1e-13


def test_lift_factorisation(ops):
    Nfp = ops.Nfp
    B = np.zeros((4 * Nfp, 4 * Nfp))
    for f in range(4):
        B[f * Nfp:(f + 1) * Nfp, f * Nfp:(f + 1) * Nfp] = ops.L0
    # EEL . blkdiag(L0) = LIFTB and both equal the independent face-quadrature lift
    Lref = ref_ops.lift_face_quadrature(ops.N, ops.Fmask)
    scale = np.abs(Lref).max()
    tol = 1e-11 if ops.N <= 7 else 1e-10  # the BB mass matrix's condition grows ~4^N
    assert np.abs(ops.LIFTB - Lref).max() < tol * scale
    assert np.abs(ops.EEL @ B - Lref).max() < tol * scale
    # the fixed-width sparse forms hold every non-zero (nothing truncated)
    assert ops.sparse_widths["EEL"] <= ops.EEL_nnz
    assert ops.sparse_widths["L0"] <= ops.L0_nnz
    E = np.zeros_like(ops.EEL)
    for i in range(ops.Np):
        for j in range(ops.EEL_nnz):
            E[i, ops.EEL_ids[i, j]] += ops.EEL_vals[i, j]
    assert np.abs(E - ops.EEL).max() < 1e-5 * np.abs(ops.EEL).max()


def test_projection(ops):
    # orders 8, 9 project on the tensor-product set, whose outside points make Vq ~1e4
    tol = 1e-11 if ops.N <= 7 else 1e-7
    assert np.abs(ops.Pq @ ops.Vq - np.eye(ops.Np)).max() < tol
    # Pq is the L2 projection: Pq f reproduces degree-N polynomials sampled at cubature
    rng = np.random.default_rng(0)
    c = rng.standard_normal(ops.Np)
    assert np.abs(ops.Pq @ (ops.Vq @ c) - c).max() < tol


def test_slice_lift(ops):
    """The slice-by-slice lift tables (StartUp3d.cpp:409-505, WaveKernels.okl:652-939): each
    face's slices partition the nodes, slice s holds the nodes s steps from the face, the
    one-step reductions are the transposed degree elevations, and lifting slice by slice
    equals LIFTB (itself pinned to the face-quadrature lift above)."""
    N, Np, Nfp = ops.N, ops.Np, ops.Nfp
    sl = O.slice_lift_ops(ops)
    mi = basis.tet_multi_indices(N)
    opposite = [3, 2, 0, 1]
    for f in range(4):
        assert sorted(sl["slice_ids"][:, f]) == list(range(Np))
        # slice 0 is the face itself, in Fmask order
        assert list(sl["slice_ids"][:Nfp, f]) == list(ops.Fmask[f])
        sk = 0
        for s_ in range(N + 1):
            nn = (N - s_ + 1) * (N - s_ + 2) // 2
            assert np.all(mi[sl["slice_ids"][sk:sk + nn, f], opposite[f]] == s_)
            sk += nn
    # a row of E^T (degree d -> d - 1, multi-index a) holds (a_m + 1)/d, m = 0..2: sum 1 + 2/d
    r0 = 0
    for s_ in range(1, N + 1):
        d = N - s_ + 1
        nn = (d + 1) * d // 2
        rows = sl["red_vals"][r0:r0 + nn]
        assert np.allclose(rows.sum(axis=1), 1.0 + 2.0 / d)
        r0 += nn
    assert r0 == Np - Nfp
    rng = np.random.default_rng(N)
    phi = rng.standard_normal((4, Nfp))
    ref = ops.LIFTB @ phi.ravel()
    got = O.slice_lift_apply(ops, sl, phi)
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()


def test_lsrk_coefficients():
    a, b, c = O.lsrk45_coefficients()
    # consistency: sum of b-weights of the equivalent Butcher form = 1 -> the scheme
    # integrates y' = 1 exactly; check on y' = lam y to fourth order in dt
    for lam_dt in (1e-2, 5e-3):
        y, res = 1.0, 0.0
        for i in range(5):
            res = a[i] * res + lam_dt * y
            y = y + b[i] * res
        err = abs(y - np.exp(lam_dt))
        assert err < 2 * lam_dt ** 5


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_bb_rhs_equals_nodal_rhs(N):
    """Whole acoustic RHS: oracle BB kernels (volume + surface) == VB^-1 . nodal DG RHS."""
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(2)
    M.face_pair(m)
    vmapM, vmapP = M.build_maps(m, ops)
    dm = M.device_arrays(m, ops, 4, vmapP=vmapP)
    rng = np.random.default_rng(N)
    Q = rng.standard_normal((m.K, 4, ops.Np))
    oo = orc.OracleOps(ops, np.float64)
    rhs = np.zeros_like(Q)
    orc.volume(oo, m.K, dm.vgeo.astype(np.float64), Q, rhs)
    # FP64 geometry for the FP64 comparison
    g = M.geometric_factors(m)
    nx, ny, nz, sJ = M.normals(g)
    vgeo = np.stack([g[k] for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")], axis=1)
    fgeo = np.stack([sJ / g["J"][:, None], nx, ny, nz], axis=2)
    rhs = np.zeros_like(Q)
    orc.volume(oo, m.K, vgeo, Q, rhs)
    orc.surface(oo, m.K, fgeo, dm.vmapP, Q, rhs)
    Dr, Ds, Dt, LIFT = ref_ops.nodal_operators(N, ops.r, ops.s, ops.t, ops.Fmask)
    qn = np.einsum("ij,kfj->kfi", ops.VB, Q)
    rn = ref_ops.nodal_rhs(Dr, Ds, Dt, LIFT, ops.Fmask, vgeo, fgeo, vmapM, vmapP, qn)
    rbb = np.einsum("ij,kfj->kfi", ops.invVB, rn)
    assert np.abs(rbb - rhs).max() < 1e-8 * np.abs(rhs).max()
