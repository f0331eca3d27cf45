"""Hybrid meshes (tetrahedra, prisms, hexahedra, pyramids) and the wedge start-up of the
reference's acoustic-elastic directory (acous_elas_cpp/src/MeshHybrid.cpp:69-901), host side, FP64.

That directory reads a GMSH 2.2 hybrid mesh (``ReadGmshHybrid``), pairs faces of 3 or 4 vertices
(``FacePairHybrid``) and builds the prism ("wedge") reference element: nodes as triangle nodes
times 1D Gauss-Lobatto nodes, face points (the triangle nodes on the two triangular faces, a
Gauss tensor grid on the three quadrilateral faces), the interpolation to them, the triangle and
1D derivative matrices, the physical nodes of each prism and the face-point connectivity by
coordinates (``StartUpWedge``).  Its ``main`` (acous_elas_cpp/main.cpp) then runs the tetrahedral
acoustic solver; the directory has no coupled kernels, so this module is its set-up only.

Differences from the reference, each fixing a defect there:

* vertex coordinates are read as floating point (``ReadGmshHybrid`` reads them into ints, so a
  mesh like pri2.msh with fractional coordinates is misread);
* faces are matched on all of their vertices (``FacePairHybrid`` sorts a 4-slot vertex list whose
  unused slot is -1 and then copies only the first Nv entries, so a triangular face keeps -1 and
  its two smallest vertices and loses the largest; it also reads one face past the end of the
  sorted list).
"""
from dataclasses import dataclass

import numpy as np
from scipy.special import roots_jacobi, eval_legendre

from . import basis

HEX, PRI, PYR, TET = 0, 1, 2, 3      # element kinds, in the reference's storage order
_NV = {HEX: 8, PRI: 6, PYR: 5, TET: 4}
_GMSH_TYPE = {5: HEX, 6: PRI, 7: PYR, 4: TET}

# face -> vertex (0-based) tables of ``getVertexOnFace`` (MeshHybrid.cpp:39-66)
_FACE_VERTS = {
    TET: [(0, 2, 1), (0, 1, 3), (2, 0, 3), (1, 2, 3)],
    PYR: [(0, 1, 4), (3, 0, 4), (1, 2, 4), (2, 3, 4), (0, 3, 2, 1)],
    PRI: [(0, 2, 1), (3, 4, 5), (0, 1, 4, 3), (2, 0, 3, 5), (1, 2, 5, 4)],
    HEX: [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)],
}


@dataclass
class HybridMesh:
    VX: np.ndarray
    VY: np.ndarray
    VZ: np.ndarray
    EToV: np.ndarray        # (K, 8) int, -1 past the element's vertices
    kind: np.ndarray        # (K,) HEX / PRI / PYR / TET
    gmsh_ids: np.ndarray    # (K,) element numbers in the file
    tags: np.ndarray        # (K,) physical group
    EToE: np.ndarray = None  # (K, 6), -1 past the element's faces
    EToF: np.ndarray = None

    @property
    def K(self):
        return self.EToV.shape[0]

    def nfaces(self, e):
        return len(_FACE_VERTS[int(self.kind[e])])

    def face_vertices(self, e, f):
        return [int(self.EToV[e, v]) for v in _FACE_VERTS[int(self.kind[e])][f]]


def read_gmsh_hybrid(path):
    """GMSH 2.2 ASCII mesh with any of tets (4), hexes (5), prisms (6), pyramids (7); elements
    stored hexes first, then prisms, pyramids, tets, each in file order (``ReadGmshHybrid``,
    ``firstElement``, MeshHybrid.cpp:8-17, :69-300)."""
    with open(path) as f:
        lines = f.read().split("\n")
    i = lines.index("$Nodes")
    nv = int(lines[i + 1])
    ids, xyz = [], []
    for ln in lines[i + 2:i + 2 + nv]:
        t = ln.split()
        ids.append(int(t[0]))
        xyz.append([float(v) for v in t[1:4]])
    node_index = {g: j for j, g in enumerate(ids)}
    xyz = np.array(xyz)
    i = lines.index("$Elements")
    ne = int(lines[i + 1])
    groups = {k: [] for k in (HEX, PRI, PYR, TET)}
    for ln in lines[i + 2:i + 2 + ne]:
        t = [int(v) for v in ln.split()]
        gtype, ntags = t[1], t[2]
        if gtype not in _GMSH_TYPE:
            continue
        kind = _GMSH_TYPE[gtype]
        verts = [node_index[v] for v in t[3 + ntags:3 + ntags + _NV[kind]]]
        groups[kind].append((t[0], t[3] if ntags else 0, verts))
    rows = [(k, g) for k in (HEX, PRI, PYR, TET) for g in groups[k]]
    K = len(rows)
    EToV = -np.ones((K, 8), dtype=np.int64)
    kind = np.zeros(K, dtype=np.int32)
    gids = np.zeros(K, dtype=np.int64)
    tags = np.zeros(K, dtype=np.int64)
    for e, (k, (g, tag, verts)) in enumerate(rows):
        EToV[e, :len(verts)] = verts
        kind[e], gids[e], tags[e] = k, g, tag
    return HybridMesh(VX=xyz[:, 0].copy(), VY=xyz[:, 1].copy(), VZ=xyz[:, 2].copy(), EToV=EToV,
                      kind=kind, gmsh_ids=gids, tags=tags)


def face_pair_hybrid(m):
    """EToE / EToF (K, 6) from faces that share all of their vertices; boundary faces connect to
    themselves, slots past an element's faces hold -1 (``FacePairHybrid``, MeshHybrid.cpp:325-437)."""
    K = m.K
    EToE = -np.ones((K, 6), dtype=np.int64)
    EToF = -np.ones((K, 6), dtype=np.int64)
    seen = {}
    for e in range(K):
        for f in range(m.nfaces(e)):
            key = tuple(sorted(m.face_vertices(e, f)))
            if key in seen:
                e2, f2 = seen.pop(key)
                EToE[e, f], EToF[e, f] = e2, f2
                EToE[e2, f2], EToF[e2, f2] = e, f
            else:
                seen[key] = (e, f)
    for (e, f) in seen.values():
        EToE[e, f], EToF[e, f] = e, f
    m.EToE, m.EToF = EToE, EToF
    return EToE, EToF


# ----------------------------------------------------------------- wedge reference element
def _gauss_lobatto(N):
    """N+1 Gauss-Lobatto-Legendre points on [-1, 1] (``JacobiGL(N, 0, 0)``)."""
    if N == 1:
        return np.array([-1.0, 1.0])
    x, _ = roots_jacobi(N - 1, 1.0, 1.0)
    return np.concatenate([[-1.0], x, [1.0]])


def _legendre_1d(N, t):
    """Values and t-derivatives of the Legendre polynomials P_0..P_N at t: (len t, N+1) each."""
    t = np.asarray(t, dtype=np.float64)
    V = np.stack([eval_legendre(i, t) for i in range(N + 1)], axis=1)
    dV = np.zeros_like(V)
    for i in range(1, N + 1):   # P_i' = sum over j < i, i - j odd, of (2j + 1) P_j
        for j in range(i - 1, -1, -2):
            dV[:, i] += (2 * j + 1) * V[:, j]
    return V, dV


def wedge_basis(N, r, s, t):
    """A basis of P_N(triangle) x P_N(t) and its r, s, t derivatives at the points: triangle
    Bernstein polynomials times Legendre polynomials (``WedgeBasis``, Basis.cpp:1067-1095, takes
    the orthonormal triangle basis; every interpolation and derivative matrix built from it,
    V_X V^-1, is the same for any basis of the space)."""
    r, s, t = (np.asarray(a, dtype=np.float64) for a in (r, s, t))
    Bt = basis.bern_tri(N, r, s)
    # d/dr, d/ds of the triangle Bernstein basis by differences of one degree lower
    mi = basis.tri_multi_indices(N)
    lo = {tuple(a): j for j, a in enumerate(basis.tri_multi_indices(N - 1))} if N >= 1 else {}
    Bl = basis.bern_tri(N - 1, r, s) if N >= 1 else None
    dr = np.zeros_like(Bt)
    ds = np.zeros_like(Bt)
    # barycentrics of bern_tri: l0 = -(r + s)/2, l1 = (1 + r)/2, l2 = (1 + s)/2
    for c, a in enumerate(mi):
        for m, (cr, cs) in enumerate(((-0.5, -0.5), (0.5, 0.0), (0.0, 0.5))):
            if a[m] == 0:
                continue
            b = list(a)
            b[m] -= 1
            d = N * Bl[:, lo[tuple(b)]]
            dr[:, c] += cr * d
            ds[:, c] += cs * d
    L, dL = _legendre_1d(N, t)
    NpT = len(mi)
    V = np.einsum("pj,pi->pij", Bt, L).reshape(len(r), NpT * (N + 1))
    Vr = np.einsum("pj,pi->pij", dr, L).reshape(V.shape)
    Vs = np.einsum("pj,pi->pij", ds, L).reshape(V.shape)
    Vt = np.einsum("pj,pi->pij", Bt, dL).reshape(V.shape)
    return V, Vr, Vs, Vt


@dataclass
class WedgeOps:
    N: int
    Np: int
    NpTri: int
    NpQuad: int
    r: np.ndarray
    s: np.ndarray
    t: np.ndarray
    rf: np.ndarray          # face points: t = -1 tri, t = 1 tri, s = -1, r = -1, r + s = 0 quads
    sf: np.ndarray
    tf: np.ndarray
    face_offsets: np.ndarray  # (6,) start of each face's points
    V: np.ndarray
    Vfq: np.ndarray         # (NfpNfaces, Np) nodes -> face points
    Dr: np.ndarray          # (NpTri, NpTri) triangle nodal derivatives
    Ds: np.ndarray
    Dt: np.ndarray          # (N+1, N+1) 1D nodal derivative on the Gauss-Lobatto nodes


def wedge_ops(N):
    """``StartUpWedge``'s reference element (MeshHybrid.cpp:438-560, :745-791)."""
    NpTri, NpQuad = (N + 1) * (N + 2) // 2, (N + 1) ** 2
    Np = (N + 1) * NpTri
    r3, s3, t3 = basis.nodes3d(N)
    rtri, stri = r3[:NpTri], s3[:NpTri]        # the face t = -1 of the tetrahedral nodes
    assert np.allclose(t3[:NpTri], -1.0)
    r1 = _gauss_lobatto(N)
    g1, _ = roots_jacobi(N + 1, 0.0, 0.0)      # Gauss points of the quadrilateral faces
    r = np.tile(rtri, N + 1)
    s = np.tile(stri, N + 1)
    t = np.repeat(r1, NpTri)
    I, J = np.meshgrid(np.arange(N + 1), np.arange(N + 1), indexing="ij")
    I, J = I.ravel(), J.ravel()
    one = np.ones(NpTri)
    rf = np.concatenate([rtri, rtri, g1[I], -np.ones(NpQuad), -g1[J]])
    sf = np.concatenate([stri, stri, -np.ones(NpQuad), g1[I], g1[J]])
    tf = np.concatenate([-one, one, g1[J], g1[J], g1[I]])
    offs = np.cumsum([0, NpTri, NpTri, NpQuad, NpQuad, NpQuad])
    V, _, _, _ = wedge_basis(N, r, s, t)
    Vf, _, _, _ = wedge_basis(N, rf, sf, tf)
    Vfq = np.linalg.solve(V.T, Vf.T).T
    Bt = basis.bern_tri(N, rtri, stri)
    _, Vr, Vs, _ = wedge_basis(N, rtri, stri, -one)
    # triangle derivatives: the t = -1 layer of the wedge basis restricted to the triangle part
    # (its Legendre factor is constant along the layer), equivalently d(B) VB_tri^-1
    L0 = _legendre_1d(N, -one)[0][:, 0]
    Dr = np.linalg.solve(Bt.T, (Vr[:, :NpTri] / L0[:, None]).T).T
    Ds = np.linalg.solve(Bt.T, (Vs[:, :NpTri] / L0[:, None]).T).T
    L1, dL1 = _legendre_1d(N, r1)
    Dt = np.linalg.solve(L1.T, dL1.T).T
    return WedgeOps(N=N, Np=Np, NpTri=NpTri, NpQuad=NpQuad, r=r, s=s, t=t, rf=rf, sf=sf, tf=tf,
                    face_offsets=offs, V=V, Vfq=Vfq, Dr=Dr, Ds=Ds, Dt=Dt)


@dataclass
class WedgeMesh:
    elems: np.ndarray       # (Kp,) prism element ids of the hybrid mesh
    x: np.ndarray           # (Kp, Np) physical nodes
    y: np.ndarray
    z: np.ndarray
    mapM: np.ndarray        # (Kp, NfpNfaces) flat face-point index (own)
    mapP: np.ndarray        # (Kp, NfpNfaces) flat face-point index of the neighbour's match


def start_up_wedge(m, wo, tol=1e-7):
    """Physical nodes of every prism (the linear wedge map of its six vertices) and the
    face-point maps mapM / mapP by coordinates (MeshHybrid.cpp:562-703).  Face points of a prism
    face shared with another prism are matched to the neighbour's points at the same location;
    boundary points map to themselves."""
    pr = np.nonzero(m.kind == PRI)[0]
    r1 = np.array([-1., 1., -1., -1., 1., -1.])
    s1 = np.array([-1., -1., 1., -1., -1., 1.])
    t1 = np.array([-1., -1., -1., 1., 1., 1.])
    V1 = wedge_basis(1, r1, s1, t1)[0]
    V1N = np.linalg.solve(V1.T, wedge_basis(1, wo.r, wo.s, wo.t)[0].T).T   # (Np, 6)
    vid = m.EToV[pr, :6]
    x = m.VX[vid] @ V1N.T
    y = m.VY[vid] @ V1N.T
    z = m.VZ[vid] @ V1N.T
    xf, yf, zf = x @ wo.Vfq.T, y @ wo.Vfq.T, z @ wo.Vfq.T
    Kp, NF = len(pr), wo.Vfq.shape[0]
    pos = {int(e): i for i, e in enumerate(pr)}
    mapM = (np.arange(Kp)[:, None] * NF + np.arange(NF)[None, :]).astype(np.int64)
    mapP = mapM.copy()
    offs = wo.face_offsets
    for i, e in enumerate(pr):
        for f in range(5):
            en, fn = int(m.EToE[e, f]), int(m.EToF[e, f])
            if en == e or en not in pos:
                continue
            j = pos[en]
            a = np.stack([xf[i, offs[f]:offs[f + 1]], yf[i, offs[f]:offs[f + 1]], zf[i, offs[f]:offs[f + 1]]], 1)
            b = np.stack([xf[j, offs[fn]:offs[fn + 1]], yf[j, offs[fn]:offs[fn + 1]], zf[j, offs[fn]:offs[fn + 1]]], 1)
            d = ((a[:, None, :] - b[None, :, :]) ** 2).sum(-1)
            k = d.argmin(axis=1)
            if np.any(d[np.arange(len(k)), k] > tol):
                raise RuntimeError(f"face points of element {e} face {f} have no match")
            mapP[i, offs[f]:offs[f + 1]] = j * NF + offs[fn] + k
    return WedgeMesh(elems=pr, x=x, y=y, z=z, mapM=mapM, mapP=mapP)
