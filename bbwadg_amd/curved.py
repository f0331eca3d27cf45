"""Host setup of the curvilinear WADG path (FP64): nodal operators at volume and face cubature,
the Gordon-Hall sphere map and the geometric factors at cubature points.

Restates ``InitQuadratureArrays`` (cpp/src/StartUp3d.cpp:60-160, :829-960), the operator part of
``InitWADG_curved`` (cpp/src/WaveOKL3d.cpp:164-417), ``GordonHallSphere`` (:929-1297) and
``checkCurvedGeo`` (:1299-1376).  The reference builds this path with the nodal basis (its
``main`` target, ``B=0``): the state holds nodal values, volume terms are the split ("skew")
form at the cubature points of a degree 2N+1 rule, face terms use a degree 2N+1 rule on every
face, and the WADG update divides by the Jacobian at the cubature points.

Nodal matrices are formed from the Bernstein basis: the nodal Lagrange basis is B(x) VB^-1, so
interpolation and derivative matrices at any points are (B or dB/dr)(points) VB^-1, and the
reference-element mass inverse (the reference's V V^T for its orthonormal modal basis) is
(Vq^T W Vq)^-1, exact for the degree 2N+1 rule.
"""
from dataclasses import dataclass

import numpy as np

from . import basis, mesh as M

SPHERE_TOL = 1e-6   # cpp/src/WaveOKL3d.cpp:960


@dataclass
class NodalOps:
    """Nodal operators of order N for the curvilinear path (all FP64)."""
    N: int
    Np: int
    Nfp: int
    Nq: int
    Nfq: int
    Fmask: np.ndarray      # (4, Nfp)
    Dr: np.ndarray         # (Np, Np) nodal derivatives at the nodes
    Ds: np.ndarray
    Dt: np.ndarray
    LIFT: np.ndarray       # (Np, 4 Nfp)
    Vq: np.ndarray         # (Nq, Np) nodal -> volume cubature
    Vrq: np.ndarray        # (Nq, Np) d/dr at volume cubature
    Vsq: np.ndarray
    Vtq: np.ndarray
    wq: np.ndarray         # (Nq,)
    rq: np.ndarray
    sq: np.ndarray
    tq: np.ndarray
    Pq: np.ndarray         # (Np, Nq) M^-1 Vq^T W
    Prq: np.ndarray        # (Np, Nq) M^-1 Vrq^T W
    Psq: np.ndarray
    Ptq: np.ndarray
    Vfq: np.ndarray        # (4 Nfq, Np) nodal -> face cubature, faces in order
    Vrf: np.ndarray        # (4 Nfq, Np) d/dr at face cubature
    Vsf: np.ndarray
    Vtf: np.ndarray
    wfq: np.ndarray        # (4 Nfq,)
    VfqFace: np.ndarray    # (Nfq, Nfp) face nodes (Fmask order) -> face cubature
    Pfq: np.ndarray        # (Np, 4 Nfq) M^-1 Vfq^T Wf
    invM: np.ndarray


def _nodal_at(N, invVB, r, s, t):
    """Nodal interpolation and r, s, t derivative matrices at the points (r, s, t)."""
    V = basis.bern_tet(N, r, s, t) @ invVB
    G = basis.grad_bern_tet(N, r, s, t)
    Vr = 0.5 * (G[1] - G[0]) @ invVB
    Vs = 0.5 * (G[2] - G[0]) @ invVB
    Vt = 0.5 * (G[3] - G[0]) @ invVB
    return V, Vr, Vs, Vt


def face_cubature(N):
    """The degree 2N+1 triangle rule placed on the four faces as ``StartUp3d`` does
    (cpp/src/StartUp3d.cpp:108-151): face 1 t = -1 (r, s) = (a, b); face 2 s = -1 (r, t) = (a, b);
    face 3 r + s + t = -1 (s, t) = (a, b); face 4 r = -1 (s, t) = (a, b)."""
    a, b, w = basis.tri_cubature(min(21, 2 * N + 1))
    m1 = -np.ones_like(a)
    r = np.concatenate([a, a, -(1 + a + b), m1])
    s = np.concatenate([b, m1, a, a])
    t = np.concatenate([m1, b, b, b])
    return r, s, t, np.tile(w, 4), len(w)


def build_nodal_ops(ops):
    """Nodal operators from the order-N reference operators ``ops`` (operators.RefOps)."""
    N, Np, Nfp = ops.N, ops.Np, ops.Nfp
    if 2 * N + 1 > 15:
        raise ValueError("the curvilinear path uses the degree 2N+1 tetrahedral rule (N <= 7)")
    invVB = ops.invVB
    _, Dr, Ds, Dt = _nodal_at(N, invVB, ops.r, ops.s, ops.t)
    rq, sq, tq, wq = basis.tet_cubature(min(21, 2 * N + 1))
    Vq, Vrq, Vsq, Vtq = _nodal_at(N, invVB, rq, sq, tq)
    invM = np.linalg.inv(Vq.T @ (wq[:, None] * Vq))
    Pq, Prq, Psq, Ptq = (invM @ (A.T * wq[None, :]) for A in (Vq, Vrq, Vsq, Vtq))
    rf, sf, tf, wfq, Nfq = face_cubature(N)
    Vfq, Vrf, Vsf, Vtf = _nodal_at(N, invVB, rf, sf, tf)
    Pfq = invM @ (Vfq.T * wfq[None, :])
    # a nodal basis function of a node off face f vanishes on f (degree N, zero at the face's
    # Nfp unisolvent nodes), so the face interpolation is Vfq's face block at Fmask's columns
    faces = [Vfq[f * Nfq:(f + 1) * Nfq][:, ops.Fmask[f]] for f in range(4)]
    VfqFace = faces[0]
    LIFT = np.zeros((Np, 4 * Nfp))
    for f in range(4):
        LIFT[:, f * Nfp:(f + 1) * Nfp] = Pfq[:, f * Nfq:(f + 1) * Nfq] @ faces[f]
    return NodalOps(N=N, Np=Np, Nfp=Nfp, Nq=len(wq), Nfq=Nfq, Fmask=ops.Fmask, Dr=Dr, Ds=Ds,
                    Dt=Dt, LIFT=LIFT, Vq=Vq, Vrq=Vrq, Vsq=Vsq, Vtq=Vtq, wq=wq, rq=rq, sq=sq,
                    tq=tq, Pq=Pq, Prq=Prq, Psq=Psq, Ptq=Ptq, Vfq=Vfq, Vrf=Vrf, Vsf=Vsf, Vtf=Vtf,
                    wfq=wfq, VfqFace=VfqFace, Pfq=Pfq, invM=invM)


# ----------------------------------------------------------------- Gordon-Hall sphere map
_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]   # vertex pairs (lam_1..lam_4)


def _blend_spaces(N, r, s, t):
    """Values at (r, s, t) of a basis of the edge space {lam_a lam_b q(lam_a - lam_b), deg q <=
    N-2} of each edge and of the face space {lam_a lam_b lam_c q, deg q <= N-3} of each face
    (``VandermondeHier``, cpp/src/Basis.cpp:358-422, spans the same spaces).  Returns
    (edge, face) lists of (npts, dim) arrays in the order of _EDGES and of the faces (vertex
    triples opposite lam_4, lam_3, lam_1, lam_2)."""
    lam = basis.tet_barycentric(r, s, t)
    edge = []
    for a, b in _EDGES:
        d = lam[a] - lam[b]
        cols = [lam[a] * lam[b] * d ** i for i in range(N - 1)]
        edge.append(np.stack(cols, axis=1) if cols else np.zeros((lam.shape[1], 0)))
    face = []
    for tri in ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)):
        a, b, c = tri
        cols = []
        for i in range(N - 2):
            for j in range(N - 2 - i):
                cols.append(lam[a] * lam[b] * lam[c] * lam[a] ** i * lam[b] ** j)
        face.append(np.stack(cols, axis=1) if cols else np.zeros((lam.shape[1], 0)))
    return edge, face


def _face_vertices(f):
    """Vertex ids (0..3, barycentric index) of face f (faces t = -1, s = -1, r+s+t = -1, r = -1)."""
    return [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)][f]


@dataclass
class CurvedGeometry:
    x: np.ndarray            # (K, Np) nodal coordinates after the map
    y: np.ndarray
    z: np.ndarray
    curved: np.ndarray       # (K,) bool: an edge or face was moved
    Klist_curved: np.ndarray
    Klist_planar: np.ndarray


def gordon_hall_sphere(mesh, ops, radius=1.0, tol=SPHERE_TOL):
    """Curve the elements of a mesh of the ball of ``radius`` centred at the origin: the nodes
    of every edge whose two vertices lie on the sphere are moved radially onto it and the
    displacement is blended into the element with the edge space; for N > 2 the nodes of every
    face whose three vertices lie on the sphere are then moved onto it and, on faces with
    EToF[e][f] == f (the reference's boundary-face test), blended with the face space
    (``GordonHallSphere``, cpp/src/WaveOKL3d.cpp:929-1297).  ``mesh`` needs EToE/EToF."""
    N, Np = ops.N, ops.Np
    lamn = basis.tet_barycentric(ops.r, ops.s, ops.t)           # (4, Np)
    on = np.abs(lamn) < tol
    verts = [int(np.nonzero(np.abs(lamn[v] - 1) < tol)[0][0]) for v in range(4)]
    edge_nodes, edge_int = [], []
    for a, b in _EDGES:
        others = [v for v in range(4) if v not in (a, b)]
        ids = np.nonzero(on[others[0]] & on[others[1]])[0]
        assert len(ids) == N + 1
        edge_nodes.append(ids)
        edge_int.append(np.array([i for i in ids if i not in (verts[a], verts[b])], dtype=int))
    face_int = []
    for f in range(4):
        opp = [v for v in range(4) if v not in _face_vertices(f)][0]
        fv = _face_vertices(f)
        ids = np.nonzero(on[opp] & ~on[fv[0]] & ~on[fv[1]] & ~on[fv[2]])[0]
        assert len(ids) == max(0, (N - 1) * (N - 2) // 2)
        face_int.append(ids)
    Ef, Ff = _blend_spaces(N, ops.r, ops.s, ops.t)
    # blending: all nodes <- interior edge (face) node displacements, through the span
    eint_all = np.concatenate(edge_int)
    EV = np.concatenate(Ef, axis=1)                                # (Np, 6(N-1))
    edge_blend = EV @ np.linalg.inv(EV[eint_all]) if N >= 2 else np.zeros((Np, 0))
    eint_all = eint_all.astype(int)
    if N > 2:
        fint_all = np.concatenate(face_int)
        FV = np.concatenate(Ff, axis=1)
        face_blend = FV @ np.linalg.inv(FV[fint_all])

    x, y, z = (np.ascontiguousarray(a.T) for a in M.affine_points(mesh, ops.r, ops.s, ops.t))
    K = mesh.K

    def on_sphere(ids, X, Y, Z):   # (K,) all nodes ids of each element within tol of the sphere
        r = np.sqrt(X[:, ids] ** 2 + Y[:, ids] ** 2 + Z[:, ids] ** 2)
        return np.abs(r - radius).max(axis=1) < tol

    def snap(sel, ids, X, Y, Z, Xc, Yc, Zc):
        e = np.nonzero(sel)[0][:, None]
        rb = np.sqrt(X[e, ids] ** 2 + Y[e, ids] ** 2 + Z[e, ids] ** 2) / radius
        Xc[e, ids], Yc[e, ids], Zc[e, ids] = X[e, ids] / rb, Y[e, ids] / rb, Z[e, ids] / rb

    xc, yc, zc = x.copy(), y.copy(), z.copy()
    edge_curved = np.zeros(K, bool)
    for k, (a, b) in enumerate(_EDGES):
        sel = on_sphere([verts[a], verts[b]], x, y, z)
        edge_curved |= sel
        if sel.any():
            snap(sel, edge_nodes[k], x, y, z, xc, yc, zc)
    ec = np.nonzero(edge_curved)[0]
    for X, Xc in ((x, xc), (y, yc), (z, zc)):
        X[ec] = X[ec] + (Xc[ec][:, eint_all] - X[ec][:, eint_all]) @ edge_blend.T
    face_curved = np.zeros(K, bool)
    if N > 2:
        xc, yc, zc = x.copy(), y.copy(), z.copy()
        for f in range(4):
            sel = on_sphere([verts[v] for v in _face_vertices(f)], x, y, z)
            face_curved |= sel
            if sel.any():
                snap(sel, ops.Fmask[f], x, y, z, xc, yc, zc)
        fc = np.nonzero(face_curved)[0]
        keep = np.concatenate([np.repeat((mesh.EToF[fc, f] == f)[:, None], len(face_int[f]), axis=1)
                               for f in range(4)], axis=1)
        for X, Xc in ((x, xc), (y, yc), (z, zc)):
            d = (Xc[fc][:, fint_all] - X[fc][:, fint_all]) * keep
            X[fc] = X[fc] + d @ face_blend.T
    curved = edge_curved | face_curved
    return CurvedGeometry(x=x, y=y, z=z, curved=curved,
                          Klist_curved=np.nonzero(curved)[0].astype(np.int32),
                          Klist_planar=np.nonzero(~curved)[0].astype(np.int32))


# ----------------------------------------------------------------- geometric factors
def vgeofacs(x, y, z, Vr, Vs, Vt):
    """rx, sx, tx, ry, sy, ty, rz, sz, tz, J at the points of the derivative matrices, for
    nodal coordinates (K, Np) (``vgeofacs3d``, cpp/src/Basis.cpp:625-661).  Returns (10, K, npts)."""
    xr, xs, xt = x @ Vr.T, x @ Vs.T, x @ Vt.T
    yr, ys, yt = y @ Vr.T, y @ Vs.T, y @ Vt.T
    zr, zs, zt = z @ Vr.T, z @ Vs.T, z @ Vt.T
    J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt)
    rx = (ys * zt - zs * yt) / J
    ry = -(xs * zt - zs * xt) / J
    rz = (xs * yt - ys * xt) / J
    sx = -(yr * zt - zr * yt) / J
    sy = (xr * zt - zr * xt) / J
    sz = -(xr * yt - yr * xt) / J
    tx = (yr * zs - zr * ys) / J
    ty = -(xr * zs - zr * xs) / J
    tz = (xr * ys - yr * xs) / J
    return np.stack([rx, sx, tx, ry, sy, ty, rz, sz, tz, J])


def sgeofacs(x, y, z, nops):
    """Unit normals and surface Jacobians at the face cubature points (``sgeofacs3d``,
    cpp/src/Basis.cpp:664-723): returns nx, ny, nz, sJ, each (K, 4 Nfq)."""
    g = vgeofacs(x, y, z, nops.Vrf, nops.Vsf, nops.Vtf)
    rx, sx, tx, ry, sy, ty, rz, sz, tz, J = g
    Nfq = nops.Nfq
    nx, ny, nz = np.empty_like(J), np.empty_like(J), np.empty_like(J)
    for f, (cx, cy, cz) in enumerate(((-tx, -ty, -tz), (-sx, -sy, -sz),
                                      (rx + sx + tx, ry + sy + ty, rz + sz + tz), (-rx, -ry, -rz))):
        sl = slice(f * Nfq, (f + 1) * Nfq)
        nx[:, sl], ny[:, sl], nz[:, sl] = cx[:, sl], cy[:, sl], cz[:, sl]
    sJ = np.sqrt(nx * nx + ny * ny + nz * nz)
    return nx / sJ, ny / sJ, nz / sJ, sJ * J


@dataclass
class CurvedArrays:
    """Per-element arrays of the curvilinear path in the reference's device layouts
    (``InitWADG_curved``, cpp/src/WaveOKL3d.cpp:196-232, :354-367)."""
    vgeoq: np.ndarray    # (K, 9, Nq): rx sx tx ry sy ty rz sz tz at the volume cubature
    Jq: np.ndarray       # (K, Nq)
    fgeoq: np.ndarray    # (K, 4, 4, Nfq): per face nx, ny, nz, sJ at the face cubature
    Jplanar: np.ndarray  # (K,) J at the first cubature point
    xq: np.ndarray       # (K, Nq) cubature point coordinates
    yq: np.ndarray
    zq: np.ndarray


def curved_arrays(geo, nops):
    g = vgeofacs(geo.x, geo.y, geo.z, nops.Vrq, nops.Vsq, nops.Vtq)
    if np.any(g[9] <= 0):
        raise M.MeshError("non-positive Jacobian at a cubature point after the curved map")
    vgeoq = np.ascontiguousarray(np.transpose(g[:9], (1, 0, 2)))
    nx, ny, nz, sJ = sgeofacs(geo.x, geo.y, geo.z, nops)
    K, Nfq = geo.x.shape[0], nops.Nfq
    fgeoq = np.stack([a.reshape(K, 4, Nfq) for a in (nx, ny, nz, sJ)], axis=2)
    return CurvedArrays(vgeoq=vgeoq, Jq=np.ascontiguousarray(g[9]), fgeoq=np.ascontiguousarray(fgeoq),
                        Jplanar=np.ascontiguousarray(g[9][:, 0]), xq=geo.x @ nops.Vq.T,
                        yq=geo.y @ nops.Vq.T, zq=geo.z @ nops.Vq.T)


def check_curved_geo(mesh, geo, nops, arr):
    """``checkCurvedGeo`` (cpp/src/WaveOKL3d.cpp:1299-1376): the volume sum J w, the error of
    the derivatives of 2x + 3y + 4z at the cubature points, and the largest distance from a
    face node to the nearest node of the neighbour's matching face."""
    vol = float((arr.Jq * nops.wq[None, :]).sum())
    u = 2 * geo.x + 3 * geo.y + 4 * geo.z
    ur, us, ut = u @ nops.Vrq.T, u @ nops.Vsq.T, u @ nops.Vtq.T
    rx, sx, tx, ry, sy, ty, rz, sz, tz = (arr.vgeoq[:, i] for i in range(9))
    err = (np.linalg.norm(rx * ur + sx * us + tx * ut - 2) + np.linalg.norm(ry * ur + sy * us + ty * ut - 3)
           + np.linalg.norm(rz * ur + sz * us + tz * ut - 4))
    worst = 0.0
    F = nops.Fmask
    for f in range(4):
        en, fn = mesh.EToE[:, f], mesh.EToF[:, f]
        a = np.stack([geo.x[:, F[f]], geo.y[:, F[f]], geo.z[:, F[f]]], axis=2)          # (K, Nfp, 3)
        b = np.stack([geo.x[en[:, None], F[fn]], geo.y[en[:, None], F[fn]], geo.z[en[:, None], F[fn]]], axis=2)
        d = np.sqrt(((a[:, :, None, :] - b[:, None, :, :]) ** 2).sum(-1)).min(axis=2)
        worst = max(worst, float(d.max()))
    return vol, float(err), worst


def bessel(x, y, z, t):
    """The spherical Bessel standing mode of the unit ball, p = 0 on its boundary
    (``BesselSolution``, cpp/main.cpp:9-14)."""
    rad = np.sqrt(x * x + y * y + z * z)
    with np.errstate(invalid="ignore", divide="ignore"):
        u = np.where(rad < 1e-10, 1.0, np.sin(np.pi * rad) / (np.pi * rad))
    return u * np.cos(np.pi * t)


@dataclass
class CurvedMesh:
    """Everything the curvilinear path needs besides the operators: the curved nodal
    coordinates, the element lists and the per-element arrays of both element kinds."""
    K: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    Klist_curved: np.ndarray
    Klist_planar: np.ndarray
    vgeoq: np.ndarray     # (K, 9, Nq)
    Jq: np.ndarray        # (K, Nq)
    fgeoq: np.ndarray     # (K, 4, 4, Nfq)
    vgeo: np.ndarray      # (K, 9) affine factors (rx ry rz sx sy sz tx ty tz)
    fgeo: np.ndarray      # (K, 4, 4) affine Fscale nx ny nz
    Jplanar: np.ndarray   # (K,)
    vmapP: np.ndarray     # (K, 4 Nfp) field-strided trace index
    xq: np.ndarray
    yq: np.ndarray
    zq: np.ndarray
    FscaleMax: float


def setup_curved(mesh, ops, nops, sphere=True):
    """Curve a ball mesh with ``gordon_hall_sphere`` (``sphere``; else keep it affine, every
    element planar) and build the arrays of the curvilinear path."""
    if mesh.EToE is None:
        M.face_pair(mesh)
    dm = M.device_arrays(mesh, ops, 4)
    if sphere:
        geo = gordon_hall_sphere(mesh, ops)
    else:
        x, y, z = (np.ascontiguousarray(a.T) for a in M.affine_points(mesh, ops.r, ops.s, ops.t))
        flat = np.zeros(mesh.K, bool)
        geo = CurvedGeometry(x=x, y=y, z=z, curved=flat, Klist_curved=np.zeros(0, np.int32),
                             Klist_planar=np.arange(mesh.K, dtype=np.int32))
    arr = curved_arrays(geo, nops)
    # affine factors in FP64 (the device arrays hold them in FP32)
    g = M.geometric_factors(mesh)
    nx, ny, nz, sJ = M.normals(g)
    vgeo = np.stack([g[k] for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")], axis=1)
    fgeo = np.stack([sJ / g["J"][:, None], nx, ny, nz], axis=2)
    return CurvedMesh(K=mesh.K, x=geo.x, y=geo.y, z=geo.z, Klist_curved=geo.Klist_curved,
                      Klist_planar=geo.Klist_planar, vgeoq=arr.vgeoq, Jq=arr.Jq, fgeoq=arr.fgeoq,
                      vgeo=vgeo, fgeo=fgeo,
                      Jplanar=arr.Jplanar, vmapP=dm.vmapP, xq=arr.xq, yq=arr.yq, zq=arr.zq,
                      FscaleMax=dm.FscaleMax)


def l2_error(nops, cm, Q, uex, time):
    """``compute_error`` on the curved mesh (cpp/src/WaveOKL3d.cpp:1683-1726, with J at the
    cubature points): (L2 error, relative L2 error) of field 0 (the reference's relL2err takes
    the square root twice; this is the plain ratio)."""
    pq = Q[:, 0, :] @ nops.Vq.T
    ue = uex(cm.xq, cm.yq, cm.zq, time)
    w = nops.wq[None, :] * cm.Jq
    err2 = float((w * (pq - ue) ** 2).sum())
    nrm2 = float((w * ue ** 2).sum())
    return np.sqrt(err2), np.sqrt(err2 / nrm2)
