"""Pins the hybrid-mesh set-up (bbwadg_amd/hybrid.py, restating acous_elas_cpp/src/MeshHybrid.cpp:
``ReadGmshHybrid``, ``FacePairHybrid``, ``StartUpWedge``) on the reference's prism meshes
(tests/golden/pri1.msh, pri2.msh):

* reading: element kinds and counts, fractional vertex coordinates;
* face pairing: the connectivity is an involution, paired faces share their vertex set, and the
  boundary faces close the cube's surface (their areas sum to 24); keying triangular faces as the
  reference does (the slot -1 and the two smallest vertices) would pair distinct triangles;
* the wedge element: interpolation to the face points exact for P_N(triangle) x P_N(t), the
  triangle and 1D derivative matrices exact on polynomials, the face points on their faces;
* the prism nodes and face maps: matched face points coincide and the map is an involution."""
import os

import numpy as np
import pytest

from bbwadg_amd import hybrid as H

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    m = H.read_gmsh_hybrid(os.path.join(GOLDEN, name))
    H.face_pair_hybrid(m)
    return m


def test_read():
    m1, m2 = load("pri1.msh"), load("pri2.msh")
    assert m1.K == 2 and m2.K == 16
    assert np.all(m1.kind == H.PRI) and np.all(m2.kind == H.PRI)
    assert list(m1.EToV[0, :6]) == [3, 0, 2, 7, 4, 6]      # 4 1 3 8 5 7 in the file, 0-based
    assert np.all(m1.EToV[:, 6:] == -1)
    assert abs(m2.VX[19] - 0.033587) < 1e-12 and abs(m2.VZ[19] - 0.033701) < 1e-12


def _face_area(m, e, f):
    v = m.face_vertices(e, f)
    P = np.stack([m.VX[v], m.VY[v], m.VZ[v]], 1)
    if len(v) == 3:
        return 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    # planar quadrilateral: two triangles
    return 0.5 * (np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
                  + np.linalg.norm(np.cross(P[2] - P[0], P[3] - P[0])))


@pytest.mark.parametrize("name", ["pri1.msh", "pri2.msh"])
def test_face_pairing(name):
    m = load(name)
    area = 0.0
    for e in range(m.K):
        for f in range(m.nfaces(e)):
            e2, f2 = m.EToE[e, f], m.EToF[e, f]
            assert m.EToE[e2, f2] == e and m.EToF[e2, f2] == f
            assert sorted(m.face_vertices(e, f)) == sorted(m.face_vertices(e2, f2))
            if e2 == e:
                area += _face_area(m, e, f)
        assert np.all(m.EToE[e, m.nfaces(e):] == -1)
    assert abs(area - 24.0) < 1e-9


def test_reference_triangle_face_key_would_pair_distinct_triangles():
    """The reference keys a triangular face by [-1, v0, v1] of its sorted vertices (MeshHybrid.cpp:
    348-351): two triangles with the same two smallest vertices collide.  Two prisms stacked on
    a common quad face, whose bottom triangles share an edge: full-vertex keys keep them apart."""
    m = H.HybridMesh(VX=np.array([0., 1, 0, 0, 0, 1, 0, 0]), VY=np.array([0., 0, 1, -1, 0, 0, 1, -1]),
                     VZ=np.array([0., 0, 0, 0, 1, 1, 1, 1]),
                     EToV=np.array([[0, 1, 2, 4, 5, 6, -1, -1], [1, 0, 3, 5, 4, 7, -1, -1]]),
                     kind=np.array([H.PRI, H.PRI]), gmsh_ids=np.array([1, 2]), tags=np.zeros(2))
    t0, t1 = sorted(m.face_vertices(0, 0)), sorted(m.face_vertices(1, 0))
    assert t0 != t1 and t0[:2] == t1[:2]      # the reference's keys would be equal
    H.face_pair_hybrid(m)
    assert m.EToE[0, 0] == 0 and m.EToE[1, 0] == 1   # both bottom triangles stay boundary
    assert m.EToE[0, 2] == 1 and m.EToE[1, 2] == 0   # the shared quadrilateral is paired


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_wedge_operators(N):
    wo = H.wedge_ops(N)
    assert wo.Np == (N + 1) ** 2 * (N + 2) // 2
    assert np.linalg.cond(wo.V) < 1e6
    for a in range(N + 1):
        for b in range(N + 1 - a):
            for c in range(N + 1):
                f = lambda r, s, t: r ** a * s ** b * t ** c
                assert np.abs(wo.Vfq @ f(wo.r, wo.s, wo.t) - f(wo.rf, wo.sf, wo.tf)).max() < 1e-12
    rt, st = wo.r[:wo.NpTri], wo.s[:wo.NpTri]
    for a in range(N + 1):
        for b in range(N + 1 - a):
            u = rt ** a * st ** b
            assert np.abs(wo.Dr @ u - a * rt ** max(a - 1, 0) * st ** b).max() < 1e-11
            assert np.abs(wo.Ds @ u - b * rt ** a * st ** max(b - 1, 0)).max() < 1e-11
    t1 = wo.t[::wo.NpTri]
    for c in range(N + 1):
        assert np.abs(wo.Dt @ t1 ** c - c * t1 ** max(c - 1, 0)).max() < 1e-11
    o = wo.face_offsets
    assert np.allclose(wo.tf[o[0]:o[1]], -1) and np.allclose(wo.tf[o[1]:o[2]], 1)
    assert np.allclose(wo.sf[o[2]:o[3]], -1) and np.allclose(wo.rf[o[3]:o[4]], -1)
    assert np.allclose(wo.rf[o[4]:o[5]] + wo.sf[o[4]:o[5]], 0)


@pytest.mark.parametrize("name", ["pri1.msh", "pri2.msh"])
@pytest.mark.parametrize("N", [1, 3])
def test_wedge_face_maps(name, N):
    m = load(name)
    wo = H.wedge_ops(N)
    wm = H.start_up_wedge(m, wo)
    NF = wo.Vfq.shape[0]
    xf = (wm.x @ wo.Vfq.T).ravel()
    yf = (wm.y @ wo.Vfq.T).ravel()
    zf = (wm.z @ wo.Vfq.T).ravel()
    P, Mm = wm.mapP.ravel(), wm.mapM.ravel()
    assert np.abs(xf[P] - xf[Mm]).max() < 1e-12
    assert np.abs(yf[P] - yf[Mm]).max() < 1e-12 and np.abs(zf[P] - zf[Mm]).max() < 1e-12
    assert np.array_equal(P[P], Mm)
    # interior faces map off the element, boundary faces to themselves
    inner = sum((m.EToE[e, f] != e) * (wo.face_offsets[f + 1] - wo.face_offsets[f])
                for e in range(m.K) for f in range(5))
    assert (P != Mm).sum() == inner
    # the wedge map of each prism reproduces its vertices
    assert np.allclose(wm.x[:, [0, N, wo.NpTri - 1]], m.VX[m.EToV[wm.elems][:, :3]])
