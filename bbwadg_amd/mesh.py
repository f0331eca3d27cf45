"""Tetrahedral meshes, connectivity, affine geometric factors and trace maps (FP64 host).

Restates the mesh layer of the reference:

* ``read_gmsh``       -- GMSH 2.2 ASCII reader, ``ReadGmsh3d`` (``cpp/src/Mesh3d.cpp:9-200``);
* ``write_gmsh_node_data`` -- per-element nodal values as a GMSH ``$ElementNodeData`` view,
  ``writeVisToGMSH`` (``cpp/src/WaveOKL3d.cpp:2560-2642``);
* ``cube_mesh``       -- structured n^3 cubes x 6 tetrahedra on [lo, hi]^3 (new: the
  reference reads cube meshes from ``cpp/meshes/cube*.msh``), cubes numbered x-fastest or
  along a Z-order curve;
* ``sfc_order``/``renumber`` -- Z-order renumbering of any mesh's elements (new: a
  cache-locality step for the fused stages' neighbour-trace reads);
* ``face_pair``       -- EToE/EToF by sorted vertex triples, ``FacePair3d``
  (``cpp/src/Mesh3d.cpp:335-410``) with the all-three-vertices comparison of
  ``elas_cpp/src/Mesh3d.cpp:392`` (the acoustic copy compares only two);
* ``geometric_factors``/``normals`` -- ``GeometricFactors3d``/``Normals3d``
  (``cpp/src/Mesh3d.cpp:218-287``);
* ``build_maps``      -- vmapM/vmapP by face-node coordinate match, ``BuildMaps3d``
  (``cpp/src/StartUp3d.cpp:643-741``), with the match tolerance scaled by the
  element size (the reference's absolute 1e-6 on squared distance fails once
  node spacing drops below 1e-3);
* ``device_arrays``   -- vgeo [K][9], fgeo [K][4][4] and field-strided vmapP
  (``cpp/src/WaveOKL3d.cpp:1528-1603``).
"""
from dataclasses import dataclass

import numpy as np

from .operators import NFACES

# local vertex numbers of each face (cpp/src/Mesh3d.cpp:345)
FACE_VERTS = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]], dtype=np.int64)


class MeshError(RuntimeError):
    pass


@dataclass
class Mesh:
    VX: np.ndarray
    VY: np.ndarray
    VZ: np.ndarray
    EToV: np.ndarray  # (K, 4) int64
    EToE: np.ndarray = None
    EToF: np.ndarray = None
    gmsh_ids: np.ndarray = None  # (K,) element numbers of the tets in the GMSH file (EToGmshE)

    @property
    def K(self):
        return self.EToV.shape[0]

    @property
    def GX(self):
        return self.VX[self.EToV]

    @property
    def GY(self):
        return self.VY[self.EToV]

    @property
    def GZ(self):
        return self.VZ[self.EToV]


def _orient(VX, VY, VZ, EToV):
    """Swap local vertices 2 and 3 of negatively oriented tets (``Mesh3d.cpp:140-161``).

    The reference swaps EToV but keeps the vertex coordinates it had already
    gathered; here the coordinates always follow EToV, so every J is positive.
    """
    J = jacobian(VX[EToV], VY[EToV], VZ[EToV])
    neg = J < 0
    EToV = EToV.copy()
    EToV[neg, 2], EToV[neg, 3] = EToV[neg, 3].copy(), EToV[neg, 2].copy()
    return EToV


def morton_codes(ix, iy, iz, bits=21):
    """Z-order (Morton) codes of non-negative integer coordinates below 2^bits: the bits of
    x, y, z interleaved, x lowest."""
    ix, iy, iz = (np.asarray(a, dtype=np.int64) for a in (ix, iy, iz))
    code = np.zeros(ix.shape, np.int64)
    for b in range(bits):
        code |= (((ix >> b) & 1) << (3 * b)) | (((iy >> b) & 1) << (3 * b + 1)) | \
            (((iz >> b) & 1) << (3 * b + 2))
    return code


def cube_mesh(n, lo=-0.5, hi=0.5, order="x"):
    """n^3 cubes, each split into 6 tets along its main diagonal (conforming Kuhn split).

    Elements are numbered cube by cube, 6 per cube.  ``order`` = "x": cubes x fastest, then
    y, then z; "morton": cubes along the Z-order curve of their indices, so that most of an
    element's face neighbours are a few elements away in memory in every direction (their
    traces are then read while still cached).  The two numberings are the same mesh.
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    if order not in ("x", "morton"):
        raise ValueError("order must be 'x' or 'morton'")
    g = np.linspace(lo, hi, n + 1)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    VX, VY, VZ = X.ravel(), Y.ravel(), Z.ravel()

    def vid(i, j, k):
        return i + (n + 1) * (j + (n + 1) * k)

    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # x fastest ordering of cubes
    ii, jj, kk = (a.transpose(2, 1, 0).ravel() for a in (ii, jj, kk))
    if order == "morton":
        zo = np.argsort(morton_codes(ii, jj, kk), kind="stable")
        ii, jj, kk = ii[zo], jj[zo], kk[zo]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for p in perms:
        c = [ii.copy(), jj.copy(), kk.copy()]
        v = [vid(*c)]
        for axis in p:
            c[axis] = c[axis] + 1
            v.append(vid(*c))
        tets.append(np.stack(v, axis=1))
    EToV = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int64)
    EToV = _orient(VX, VY, VZ, EToV)
    return Mesh(VX=VX, VY=VY, VZ=VZ, EToV=EToV)


def sfc_order(mesh, bits=21):
    """Element permutation along the Z-order curve of the element centroids (quantised to
    2^bits cells per axis of the bounding box): ``renumber(mesh, sfc_order(mesh))`` numbers
    face neighbours close together in memory, so that the fused stages read most neighbour
    traces from cache.  A mesh-preparation step; the solvers accept any numbering."""
    c = np.stack([mesh.GX.mean(1), mesh.GY.mean(1), mesh.GZ.mean(1)])
    lo, hi = c.min(1, keepdims=True), c.max(1, keepdims=True)
    q = np.floor((c - lo) / np.maximum(hi - lo, 1e-300) * ((1 << bits) - 1)).astype(np.int64)
    return np.argsort(morton_codes(q[0], q[1], q[2], bits), kind="stable")


def renumber(mesh, perm):
    """The same mesh with its elements renumbered: new element e is old element perm[e].
    Per-element arrays map as a_new = a_old[perm]; connectivity is rebuilt by ``face_pair``."""
    perm = np.asarray(perm, dtype=np.int64)
    if perm.shape != (mesh.K,) or not np.array_equal(np.sort(perm), np.arange(mesh.K)):
        raise MeshError("perm must be a permutation of the elements")
    return Mesh(VX=mesh.VX, VY=mesh.VY, VZ=mesh.VZ, EToV=mesh.EToV[perm].copy(),
                gmsh_ids=None if mesh.gmsh_ids is None else mesh.gmsh_ids[perm].copy())


def read_gmsh(path):
    """Read tetrahedra (element type 4) from a GMSH 2.2 ASCII file (``ReadGmsh3d``)."""
    try:
        f = open(path)
    except OSError as e:
        raise MeshError(f"mesh '{path}' not opened") from e
    with f:
        lines = f.read().splitlines()
    VX = VY = VZ = None
    tets, gids = [], []
    i = 0
    while i < len(lines):
        line = lines[i]
        if "$Nodes" in line:
            nv = int(lines[i + 1].split()[0])
            xyz = np.array([[float(v) for v in lines[i + 2 + a].split()[1:4]] for a in range(nv)])
            VX, VY, VZ = xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()
            i += 2 + nv
            continue
        if "$Elements" in line:
            ne = int(lines[i + 1].split()[0])
            for a in range(ne):
                parts = [int(v) for v in lines[i + 2 + a].split()]
                if parts[1] == 4:
                    ntags = parts[2]
                    tets.append(parts[3 + ntags:3 + ntags + 4])
                    gids.append(parts[0])
            i += 2 + ne
            continue
        i += 1
    if VX is None or not tets:
        raise MeshError(f"no tetrahedra in '{path}'")
    EToV = np.array(tets, dtype=np.int64) - 1
    EToV = _orient(VX, VY, VZ, EToV)
    return Mesh(VX=VX, VY=VY, VZ=VZ, EToV=EToV, gmsh_ids=np.array(gids, dtype=np.int64))


def write_gmsh_node_data(path, mesh, r, s, t, values, name="Field 0", time=0.0, step=0):
    """Write per-element nodal values as a GMSH 2.2 ``$ElementNodeData`` view with its own
    interpolation scheme (``writeVisToGMSH``, cpp/src/WaveOKL3d.cpp:2560-2642).

    ``values`` (K, Np) are the field at the reference nodes (r, s, t) of every element (the
    reference writes Q after WaveGetData3d, i.e. nodal values).  The scheme's monomials are
    u^a v^b w^c in u = (r+1)/2, v = (s+1)/2, w = (t+1)/2 with a+b+c <= N (a outer, c inner), and
    its coefficient matrix is the inverse of their Vandermonde matrix at the nodes, so GMSH
    reproduces the nodal values.  Elements are numbered as in the GMSH file the mesh came from
    (``EToGmshE``), else 1..K.
    """
    values = np.asarray(values, dtype=np.float64)
    K, Np = values.shape
    N = 0
    while (N + 1) * (N + 2) * (N + 3) // 6 < Np:
        N += 1
    if (N + 1) * (N + 2) * (N + 3) // 6 != Np or K != mesh.K or len(r) != Np:
        raise MeshError("values must be (K, Np) at the Np reference nodes of an order-N tetrahedron")
    monom = np.array([(a, b, c) for a in range(N + 1) for b in range(N + 1) for c in range(N + 1)
                      if a + b + c <= N], dtype=np.int64)
    u, v, w = (np.asarray(r) + 1) / 2, (np.asarray(s) + 1) / 2, (np.asarray(t) + 1) / 2
    vdm = (u[None, :] ** monom[:, 0:1]) * (v[None, :] ** monom[:, 1:2]) * (w[None, :] ** monom[:, 2:3])
    coeff = np.linalg.inv(vdm)
    ids = mesh.gmsh_ids if mesh.gmsh_ids is not None else np.arange(1, K + 1)
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        f.write('$InterpolationScheme\n"MyInterpScheme"\n1\n5 2\n')
        f.write(f"{Np} {Np}\n")
        for row in coeff:
            f.write(" ".join(repr(float(x)) for x in row) + " \n")
        f.write(f"{Np} 3\n")
        for row in monom:
            f.write(" ".join(str(int(x)) for x in row) + " \n")
        f.write("$EndInterpolationScheme\n")
        f.write(f'$ElementNodeData\n2\n"{name}"\n"MyInterpScheme"\n1\n{time!r}\n3\n{int(step)}\n1\n{K}\n')
        for k in range(K):
            f.write(f"{int(ids[k])} {Np} " + " ".join(repr(float(x)) for x in values[k]) + "\n")
        f.write("$EndElementNodeData\n")


def read_gmsh_node_data(path):
    """Read back a view written by ``write_gmsh_node_data``: (element numbers, values (K, Np),
    interpolation coefficient matrix, monomial exponents)."""
    with open(path) as f:
        lines = f.read().splitlines()
    i = lines.index("$InterpolationScheme")
    nr, nc = (int(x) for x in lines[i + 4].split())
    coeff = np.array([[float(x) for x in lines[i + 5 + a].split()] for a in range(nr)])
    j = i + 5 + nr
    nm = int(lines[j].split()[0])
    monom = np.array([[int(x) for x in lines[j + 1 + a].split()] for a in range(nm)])
    d = lines.index("$ElementNodeData")
    ntags = int(lines[d + 1])
    nreal = int(lines[d + 2 + ntags])
    nint_at = d + 3 + ntags + nreal
    nint = int(lines[nint_at])
    K = int(lines[nint_at + nint])
    rows = [lines[nint_at + nint + 1 + k].split() for k in range(K)]
    ids = np.array([int(rw[0]) for rw in rows])
    vals = np.array([[float(x) for x in rw[2:]] for rw in rows])
    return ids, vals, coeff, monom


def face_pair(mesh):
    """Fill EToE/EToF; boundary faces point to themselves (``FacePair3d``)."""
    K = mesh.K
    fv = np.sort(mesh.EToV[:, FACE_VERTS], axis=2).reshape(K * NFACES, 3)
    order = np.lexsort((fv[:, 0], fv[:, 1], fv[:, 2]))  # vc major, then vb, then va
    s = fv[order]
    same = np.all(s[1:] == s[:-1], axis=1)
    if np.any(same[1:] & same[:-1]):
        raise MeshError("a face is shared by more than two elements")
    EToE = np.repeat(np.arange(K)[:, None], NFACES, axis=1)
    EToF = np.repeat(np.arange(NFACES)[None, :], K, axis=0)
    a = order[:-1][same]
    b = order[1:][same]
    ka, fa = a // NFACES, a % NFACES
    kb, fb = b // NFACES, b % NFACES
    EToE[ka, fa], EToF[ka, fa] = kb, fb
    EToE[kb, fb], EToF[kb, fb] = ka, fa
    mesh.EToE, mesh.EToF = EToE, EToF
    return mesh


def jacobian(GX, GY, GZ):
    xr, yr, zr = (GX[:, 1] - GX[:, 0]) / 2, (GY[:, 1] - GY[:, 0]) / 2, (GZ[:, 1] - GZ[:, 0]) / 2
    xs, ys, zs = (GX[:, 2] - GX[:, 0]) / 2, (GY[:, 2] - GY[:, 0]) / 2, (GZ[:, 2] - GZ[:, 0]) / 2
    xt, yt, zt = (GX[:, 3] - GX[:, 0]) / 2, (GY[:, 3] - GY[:, 0]) / 2, (GZ[:, 3] - GZ[:, 0]) / 2
    return xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt)


def geometric_factors(mesh):
    """Affine-map factors per element: dict of rx..tz (K,), J (K,) (``GeometricFactors3d``)."""
    GX, GY, GZ = mesh.GX, mesh.GY, mesh.GZ
    xr, yr, zr = (GX[:, 1] - GX[:, 0]) / 2, (GY[:, 1] - GY[:, 0]) / 2, (GZ[:, 1] - GZ[:, 0]) / 2
    xs, ys, zs = (GX[:, 2] - GX[:, 0]) / 2, (GY[:, 2] - GY[:, 0]) / 2, (GZ[:, 2] - GZ[:, 0]) / 2
    xt, yt, zt = (GX[:, 3] - GX[:, 0]) / 2, (GY[:, 3] - GY[:, 0]) / 2, (GZ[:, 3] - GZ[:, 0]) / 2
    J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt)
    if np.any(J <= 0):
        raise MeshError("non-positive Jacobian")
    g = {
        "rx": (ys * zt - zs * yt) / J, "ry": -(xs * zt - zs * xt) / J, "rz": (xs * yt - ys * xt) / J,
        "sx": -(yr * zt - zr * yt) / J, "sy": (xr * zt - zr * xt) / J, "sz": -(xr * yt - yr * xt) / J,
        "tx": (yr * zs - zr * ys) / J, "ty": -(xr * zs - zr * xs) / J, "tz": (xr * ys - yr * xs) / J,
        "J": J,
    }
    return g


def normals(g):
    """Outward unit normals and surface Jacobians, arrays (K, 4) (``Normals3d``)."""
    nx = np.stack([-g["tx"], -g["sx"], g["rx"] + g["sx"] + g["tx"], -g["rx"]], axis=1)
    ny = np.stack([-g["ty"], -g["sy"], g["ry"] + g["sy"] + g["ty"], -g["ry"]], axis=1)
    nz = np.stack([-g["tz"], -g["sz"], g["rz"] + g["sz"] + g["tz"], -g["rz"]], axis=1)
    sJ = np.sqrt(nx * nx + ny * ny + nz * nz)
    return nx / sJ, ny / sJ, nz / sJ, sJ * g["J"][:, None]


def affine_points(mesh, r, s, t):
    """Physical coordinates (npts, K) of reference points under each element's affine map."""
    r, s, t = (np.asarray(a)[:, None] for a in (r, s, t))
    GX, GY, GZ = mesh.GX, mesh.GY, mesh.GZ
    out = []
    for G in (GX, GY, GZ):
        out.append(0.5 * (-G[None, :, 0] * (r + s + t + 1) + G[None, :, 1] * (1 + r)
                          + G[None, :, 2] * (1 + s) + G[None, :, 3] * (1 + t)))
    return out


def _face_weights(ops):
    """(4, Nfp, 3) barycentric weights of each face node on its face's three vertices."""
    r, s, t = ops.r, ops.s, ops.t
    W = np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2], axis=1)
    return np.stack([W[ops.Fmask[f]][:, FACE_VERTS[f]] for f in range(NFACES)])


def face_node_match_table(ops):
    """match[f1, f2, p, n1] = n2: face node n1 of face f1 meets node n2 of face f2
    when the neighbour's face vertices are the own face vertices permuted by
    ``PERMS[p]`` (neighbour vertex PERMS[p][i] is own vertex i)."""
    Wf = _face_weights(ops)
    Nfp = ops.Nfp
    tab = np.zeros((NFACES, NFACES, len(PERMS), Nfp), dtype=np.int64)
    for f1 in range(NFACES):
        for f2 in range(NFACES):
            for pi, p in enumerate(PERMS):
                v = np.zeros((Nfp, 3))
                v[:, list(p)] = Wf[f1]
                d = np.abs(v[:, None, :] - Wf[f2][None, :, :]).max(axis=2)
                j = np.argmin(d, axis=1)
                if d[np.arange(Nfp), j].max() > 1e-10:
                    raise MeshError("face nodes are not symmetric under vertex permutations")
                tab[f1, f2, pi] = j
    return tab


PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def build_mapP(mesh, ops):
    """mapP: (K, 4*Nfp) global face-node index ``k2*4Nfp + f2*Nfp + n2`` of the exterior
    trace of each face node (itself on boundary faces); the face-node counterpart of vmapP
    (the reference keeps it as ``mapP``, ``cpp/src/StartUp3d.cpp:659-712``)."""
    if mesh.EToE is None:
        face_pair(mesh)
    K, Nfp = mesh.K, ops.Nfp
    NfpNf = NFACES * Nfp
    mapP = np.arange(K * NfpNf, dtype=np.int64).reshape(K, NfpNf)
    tab = face_node_match_table(ops)
    k1, f1 = np.nonzero(mesh.EToE != np.arange(K)[:, None])
    k2, f2 = mesh.EToE[k1, f1], mesh.EToF[k1, f1]
    gv1 = mesh.EToV[k1[:, None], FACE_VERTS[f1]]
    gv2 = mesh.EToV[k2[:, None], FACE_VERTS[f2]]
    pidx = np.full(len(k1), -1, dtype=np.int64)
    for pi, p in enumerate(PERMS):
        pidx[np.all(gv2[:, list(p)] == gv1, axis=1)] = pi
    n2 = tab[f1, f2, pidx]
    mapP[k1[:, None], f1[:, None] * Nfp + np.arange(Nfp)[None, :]] = (k2 * NfpNf + f2 * Nfp)[:, None] + n2
    return mapP


def build_maps(mesh, ops):
    """vmapM, vmapP: (K, 4*Nfp) node ids ``k*Np + n`` of the interior / exterior trace.

    ``BuildMaps3d`` matches face-node coordinates pairwise; here the match is
    read from a table indexed by the two local face numbers and the permutation
    between the shared face's vertex lists (the face node set is symmetric, so
    the coordinate match depends on nothing else).  ``build_maps_by_coordinates``
    keeps the coordinate match for cross-checking.  Boundary faces map to
    themselves.
    """
    if mesh.EToE is None:
        face_pair(mesh)
    K, Np, Nfp = mesh.K, ops.Np, ops.Nfp
    Fm = ops.Fmask
    vmapM = (np.arange(K)[:, None] * Np + Fm.reshape(1, -1)).astype(np.int64)
    vmapP = vmapM.copy()
    tab = face_node_match_table(ops)
    k1, f1 = np.nonzero(mesh.EToE != np.arange(K)[:, None])
    k2, f2 = mesh.EToE[k1, f1], mesh.EToF[k1, f1]
    gv1 = mesh.EToV[k1[:, None], FACE_VERTS[f1]]  # (n, 3)
    gv2 = mesh.EToV[k2[:, None], FACE_VERTS[f2]]
    pidx = np.full(len(k1), -1, dtype=np.int64)
    for pi, p in enumerate(PERMS):
        ok = np.all(gv2[:, list(p)] == gv1, axis=1)
        pidx[ok] = pi
    if np.any(pidx < 0):
        raise MeshError("paired faces do not share their vertices")
    n2 = tab[f1, f2, pidx]  # (n, Nfp)
    vmapP[k1[:, None], f1[:, None] * Nfp + np.arange(Nfp)[None, :]] = k2[:, None] * Np + Fm[f2[:, None], n2]
    return vmapM, vmapP


def build_maps_by_coordinates(mesh, ops, tol=1e-6, chunk=1 << 16):
    """Coordinate-matching vmapP, as ``BuildMaps3d`` computes it (tolerance relative
    to the squared shortest element edge)."""
    if mesh.EToE is None:
        face_pair(mesh)
    K, Np, Nfp = mesh.K, ops.Np, ops.Nfp
    Fm = ops.Fmask
    vmapM = (np.arange(K)[:, None] * Np + Fm.reshape(1, -1)).astype(np.int64)
    vmapP = vmapM.copy()
    r, s, t = ops.r, ops.s, ops.t
    W = np.stack([-(1 + r + s + t) / 2, (1 + r) / 2, (1 + s) / 2, (1 + t) / 2], axis=1)
    GX, GY, GZ = mesh.GX, mesh.GY, mesh.GZ
    edge = np.min([(GX[:, a] - GX[:, b]) ** 2 + (GY[:, a] - GY[:, b]) ** 2 + (GZ[:, a] - GZ[:, b]) ** 2
                   for a in range(4) for b in range(a + 1, 4)], axis=0)
    k1, f1 = np.nonzero(mesh.EToE != np.arange(K)[:, None])
    k2, f2 = mesh.EToE[k1, f1], mesh.EToF[k1, f1]
    for c0 in range(0, len(k1), chunk):
        a1, b1 = k1[c0:c0 + chunk], f1[c0:c0 + chunk]
        a2, b2 = k2[c0:c0 + chunk], f2[c0:c0 + chunk]
        P1 = np.stack([np.einsum("nij,nj->ni", W[Fm[b1]], G[a1]) for G in (GX, GY, GZ)], axis=-1)
        P2 = np.stack([np.einsum("nij,nj->ni", W[Fm[b2]], G[a2]) for G in (GX, GY, GZ)], axis=-1)
        d2 = np.sum((P1[:, :, None, :] - P2[:, None, :, :]) ** 2, axis=-1)
        j = np.argmin(d2, axis=2)
        dmin = np.take_along_axis(d2, j[:, :, None], axis=2)[:, :, 0]
        bad = dmin >= tol * edge[a1][:, None]
        if np.any(bad):
            n, _ = np.argwhere(bad)[0]
            raise MeshError(f"lost node on element {a1[n]}, face {b1[n]}")
        vmapP[a1[:, None], b1[:, None] * Nfp + np.arange(Nfp)[None, :]] = a2[:, None] * Np + Fm[b2[:, None], j]
    return vmapM, vmapP


@dataclass
class DeviceMesh:
    """Element-local geometry in the reference's device layouts (``WaveInitOCCA3d``)."""
    K: int
    vgeo: np.ndarray   # (K, 9) float32: rx ry rz sx sy sz tx ty tz
    fgeo: np.ndarray   # (K, 4, 4) float32: Fscale nx ny nz per face
    vmapP: np.ndarray  # (K, 4*Nfp) int32 flat indices into Q[K][Nfields][Np]
    J: np.ndarray      # (K,) float64
    sJ: np.ndarray     # (K, 4) float64
    FscaleMax: float
    hmin: float


def device_arrays(mesh, ops, nfields, vmapP=None):
    """vgeo, fgeo and the field-strided vmapP (``cpp/src/WaveOKL3d.cpp:1528-1603``)."""
    g = geometric_factors(mesh)
    nx, ny, nz, sJ = normals(g)
    Fscale = sJ / g["J"][:, None]
    K, Np, Nfp = mesh.K, ops.Np, ops.Nfp
    vgeo = np.stack([g[k] for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")], axis=1)
    fgeo = np.stack([Fscale, nx, ny, nz], axis=2)
    if vmapP is None:
        _, vmapP = build_maps(mesh, ops)
    nbr = vmapP // Np
    vP = vmapP - Np * nbr + Np * nfields * nbr
    if vP.max() >= 2 ** 31:
        raise MeshError("mesh too large for int32 trace indices")
    GX, GY, GZ = mesh.GX, mesh.GY, mesh.GZ
    edge = np.min([np.sqrt((GX[:, a] - GX[:, b]) ** 2 + (GY[:, a] - GY[:, b]) ** 2 + (GZ[:, a] - GZ[:, b]) ** 2)
                   for a in range(4) for b in range(a + 1, 4)], axis=0)
    return DeviceMesh(K=K, vgeo=np.ascontiguousarray(vgeo, dtype=np.float32),
                      fgeo=np.ascontiguousarray(fgeo, dtype=np.float32),
                      vmapP=np.ascontiguousarray(vP, dtype=np.int32), J=g["J"], sJ=sJ,
                      FscaleMax=float(np.float32(Fscale.max())), hmin=float(edge.min()))
