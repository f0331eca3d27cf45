"""Independent FP64 constructions used to pin the operators -- TEST INFRASTRUCTURE ONLY.

None of these reuse the product's construction path (``bbwadg_amd.operators``
solves VB \\ GradBernTet, MB \\ EmatB with BernTri, and matches lift rows):

* ``bb_mass_closed``      BB mass matrix from the product rule
  B^N_a B^N_b = C(N,a) C(N,b) / C(2N,a+b) B^{2N}_{a+b} and int_T B^M = |T| 3! M!/(M+3)!;
* ``bb_deriv_closed``     barycentric derivative operators from the multi-index rule
  (D_m c)_a = sum_j a_j c_{a - e_j + e_m}  (derivative then degree elevation);
* ``lift_face_quadrature`` LIFTB = MB^-1 E with E[a, (f, n)] the face integrals of tet BB
  functions evaluated at face cubature points mapped onto each face;
* ``nodal_operators`` / ``nodal_rhs`` a nodal DG operator built from a monomial basis at
  the interpolation nodes (Lagrange basis by quadrature), i.e. the textbook nodal form of
  ``rk_volume``/``rk_surface`` (``cpp/okl/WaveKernels.okl:44-260``); BB and nodal RHS agree
  through the change of basis VB.
"""
from math import comb, factorial

import numpy as np

from bbwadg_amd import basis


def _multinom(N, a):
    c = factorial(N)
    for x in a:
        c //= factorial(int(x))
    return c


def bb_mass_closed(N):
    mi = basis.tet_multi_indices(N)
    Np = len(mi)
    vol = 4.0 / 3.0
    M = np.zeros((Np, Np))
    for i in range(Np):
        for j in range(Np):
            a, b = mi[i], mi[j]
            M[i, j] = _multinom(N, a) * _multinom(N, b) / _multinom(2 * N, a + b) \
                * vol * 6.0 * factorial(2 * N) / factorial(2 * N + 3)
    return M


def bb_deriv_closed(N):
    mi = basis.tet_multi_indices(N)
    index = {tuple(a): i for i, a in enumerate(mi)}
    Np = len(mi)
    D = [np.zeros((Np, Np)) for _ in range(4)]
    for m in range(4):
        for i, a in enumerate(mi):
            for j in range(4):
                if a[j] == 0:
                    continue
                b = list(a)
                b[j] -= 1
                b[m] += 1
                D[m][i, index[tuple(b)]] += a[j]
    return D


def face_points(deg):
    """Face cubature points mapped to faces 1..4 (``StartUp3d.cpp:119-146``), weights."""
    r, s, w = basis.tri_cubature(deg)
    one = np.ones_like(r)
    faces = [(r, s, -one), (r, -one, s), (-(1 + r + s), r, s), (-one, r, s)]
    return faces, w


def lift_face_quadrature(N, Fmask):
    Np, Nfp = basis.np_tet(N), basis.np_tri(N)
    faces, w = face_points(2 * N + 1)
    E = np.zeros((Np, 4 * Nfp))
    for f, (rf, sf, tf) in enumerate(faces):
        B = basis.bern_tet(N, rf, sf, tf)  # (nq, Np)
        E[:, f * Nfp:(f + 1) * Nfp] = B.T @ (w[:, None] * B[:, Fmask[f]])
    return np.linalg.solve(bb_mass_closed(N), E)


def _monomials(N, r, s, t):
    cols = []
    for a in range(N + 1):
        for b in range(N + 1 - a):
            for c in range(N + 1 - a - b):
                cols.append((a, b, c))
    V = np.stack([r ** a * s ** b * t ** c for a, b, c in cols], axis=1)
    Vr = np.stack([a * r ** max(a - 1, 0) * s ** b * t ** c for a, b, c in cols], axis=1)
    Vs = np.stack([b * r ** a * s ** max(b - 1, 0) * t ** c for a, b, c in cols], axis=1)
    Vt = np.stack([c * r ** a * s ** b * t ** max(c - 1, 0) for a, b, c in cols], axis=1)
    return V, Vr, Vs, Vt


def nodal_operators(N, r, s, t, Fmask):
    """Dr, Ds, Dt and LIFT (Np x 4Nfp) of the nodal basis at nodes (r, s, t)."""
    V, Vr, Vs, Vt = _monomials(N, r, s, t)
    Vi = np.linalg.inv(V)
    Dr, Ds, Dt = Vr @ Vi, Vs @ Vi, Vt @ Vi
    rq, sq, tq, wq = basis.tet_cubature(2 * N + 1)
    Lq = _monomials(N, rq, sq, tq)[0] @ Vi
    M = Lq.T @ (wq[:, None] * Lq)
    faces, w = face_points(2 * N + 1)
    Nfp = Fmask.shape[1]
    E = np.zeros((len(r), 4 * Nfp))
    for f, (rf, sf, tf) in enumerate(faces):
        Lf = _monomials(N, rf, sf, tf)[0] @ Vi
        E[:, f * Nfp:(f + 1) * Nfp] = Lf.T @ (w[:, None] * Lf[:, Fmask[f]])
    return Dr, Ds, Dt, np.linalg.solve(M, E)


def nodal_rhs(Dr, Ds, Dt, LIFT, Fmask, vgeo, fgeo, vmapM, vmapP, q):
    """Acoustic strong-form upwind DG RHS on nodal values q (K, 4, Np), FP64.

    vmapM/vmapP: (K, 4Nfp) node ids k*Np + n; boundary faces (vmapM == vmapP) use
    p+ = -p-, Un+ = Un- (``cpp/okl/WaveKernels.okl:562-565``).
    """
    K, _, Np = q.shape
    g = np.asarray(vgeo, dtype=np.float64)
    rx, ry, rz, sx, sy, sz, tx, ty, tz = (g[:, i:i + 1] for i in range(9))
    p, u, v, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]

    def grad(f):
        fr, fs, ft = f @ Dr.T, f @ Ds.T, f @ Dt.T
        return rx * fr + sx * fs + tx * ft, ry * fr + sy * fs + ty * ft, rz * fr + sz * fs + tz * ft

    px, py, pz = grad(p)
    divU = grad(u)[0] + grad(v)[1] + grad(w)[2]
    flat = q.transpose(1, 0, 2).reshape(4, -1)  # field, k*Np+n
    Nfp4 = vmapM.shape[1]
    Nfp = Nfp4 // 4
    fg = np.asarray(fgeo, dtype=np.float64)  # (K, 4, 4)
    F = np.repeat(fg, Nfp, axis=1)  # (K, 4Nfp, 4)
    Fscale, nx, ny, nz = F[..., 0], F[..., 1], F[..., 2], F[..., 3]
    qM = flat[:, vmapM]
    qP = flat[:, vmapP]
    pjump = qP[0] - qM[0]
    Unjump = (qP[1] - qM[1]) * nx + (qP[2] - qM[2]) * ny + (qP[3] - qM[3]) * nz
    bnd = vmapM == vmapP
    pjump = np.where(bnd, -2.0 * qM[0], pjump)
    Unjump = np.where(bnd, 0.0, Unjump)
    pflux = 0.5 * (pjump - Unjump) * Fscale
    Uflux = 0.5 * (Unjump - pjump) * Fscale
    out = np.empty_like(q)
    out[:, 0] = -divU + pflux @ LIFT.T
    out[:, 1] = -px + (Uflux * nx) @ LIFT.T
    out[:, 2] = -py + (Uflux * ny) @ LIFT.T
    out[:, 3] = -pz + (Uflux * nz) @ LIFT.T
    return out


def monomial_integral_tet(a, b, c):
    """Exact integral of r^a s^b t^c over the reference tet (vertices (-1,-1,-1), (1,-1,-1),
    (-1,1,-1), (-1,-1,1)), via x = (1+r)/2 etc. on the unit simplex and the Dirichlet
    integral int x^i y^j z^k = i! j! k! / (i+j+k+3)!."""
    total = 0.0
    for i in range(a + 1):
        for j in range(b + 1):
            for k in range(c + 1):
                coef = comb(a, i) * comb(b, j) * comb(c, k) * 2.0 ** (i + j + k) * (-1.0) ** (a - i + b - j + c - k)
                total += coef * factorial(i) * factorial(j) * factorial(k) / factorial(i + j + k + 3)
    return 8.0 * total


def monomial_integral_tri(a, b):
    total = 0.0
    for i in range(a + 1):
        for j in range(b + 1):
            coef = comb(a, i) * comb(b, j) * 2.0 ** (i + j) * (-1.0) ** (a - i + b - j)
            total += coef * factorial(i) * factorial(j) / factorial(i + j + 2)
    return 4.0 * total


def nodal_rhs_elastic(Dr, Ds, Dt, LIFT, vgeo, fgeo, vmapM, vmapP, q, tau_v=1.0, tau_s=1.0):
    """Elastic velocity-stress nodal DG RHS, FP64: the reference's nodal kernels
    rk_volume_elas / rk_surface_elas (``elas_cpp/okl/ElasKernelsWADG.okl:25-284``) on nodal
    values q (K, 9, Np): fields v1 v2 v3 sxx syy szz syz sxz sxy; boundary Q+ = 0."""
    K, _, Np = q.shape
    g = np.asarray(vgeo, dtype=np.float64)
    rx, ry, rz, sx, sy, sz, tx, ty, tz = (g[:, i:i + 1] for i in range(9))
    Qx, Qy, Qz = [], [], []
    for f in range(9):
        fr, fs, ft = q[:, f] @ Dr.T, q[:, f] @ Ds.T, q[:, f] @ Dt.T
        Qx.append(rx * fr + sx * fs + tx * ft)
        Qy.append(ry * fr + sy * fs + ty * ft)
        Qz.append(rz * fr + sz * fs + tz * ft)
    out = np.empty_like(q)
    out[:, 0] = Qx[3] + Qy[8] + Qz[7]
    out[:, 1] = Qx[8] + Qy[4] + Qz[6]
    out[:, 2] = Qx[7] + Qy[6] + Qz[5]
    out[:, 3] = Qx[0]
    out[:, 4] = Qy[1]
    out[:, 5] = Qz[2]
    out[:, 6] = Qy[2] + Qz[1]
    out[:, 7] = Qx[2] + Qz[0]
    out[:, 8] = Qx[1] + Qy[0]
    flat = q.transpose(1, 0, 2).reshape(9, -1)
    Nfp = vmapM.shape[1] // 4
    F = np.repeat(np.asarray(fgeo, dtype=np.float64), Nfp, axis=1)
    Fscale, nx, ny, nz = F[..., 0], F[..., 1], F[..., 2], F[..., 3]
    bnd = vmapM == vmapP
    dQ = np.where(bnd[None], 0.0, flat[:, vmapP]) - flat[:, vmapM]
    fc = [dQ[3] * nx + dQ[8] * ny + dQ[7] * nz,
          dQ[8] * nx + dQ[4] * ny + dQ[6] * nz,
          dQ[7] * nx + dQ[6] * ny + dQ[5] * nz,
          dQ[0] * nx, dQ[1] * ny, dQ[2] * nz,
          dQ[2] * ny + dQ[1] * nz, dQ[2] * nx + dQ[0] * nz, dQ[1] * nx + dQ[0] * ny]
    fl = [fc[0] + tau_v * (fc[3] * nx + fc[8] * ny + fc[7] * nz),
          fc[1] + tau_v * (fc[8] * nx + fc[4] * ny + fc[6] * nz),
          fc[2] + tau_v * (fc[7] * nx + fc[6] * ny + fc[5] * nz),
          fc[3] + tau_s * fc[0] * nx, fc[4] + tau_s * fc[1] * ny, fc[5] + tau_s * fc[2] * nz,
          fc[6] + tau_s * (fc[2] * ny + fc[1] * nz), fc[7] + tau_s * (fc[2] * nx + fc[0] * nz),
          fc[8] + tau_s * (fc[1] * nx + fc[0] * ny)]
    for f in range(9):
        out[:, f] += (0.5 * Fscale * fl[f]) @ LIFT.T
    return out
