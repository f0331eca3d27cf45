"""GMSH I/O of the host layer: the reader keeps the file's element numbers (EToGmshE,
cpp/src/Mesh3d.cpp:55-87) and skips non-tetrahedral elements; the $ElementNodeData writer
(writeVisToGMSH, cpp/src/WaveOKL3d.cpp:2560-2642) round-trips the values, and its interpolation
scheme reproduces any polynomial of degree <= N from its nodal values at arbitrary points."""
import numpy as np
import pytest

from bbwadg_amd import mesh as M, operators as O

MSH = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
5
1 0 0 0
2 1 0 0
3 0 1 0
4 0 0 1
5 1 1 1
$EndNodes
$Elements
4
7 2 2 0 1 1 2 3
11 4 2 0 1 1 2 3 4
12 15 2 0 1 5
13 4 2 0 1 2 3 5 4
$EndElements
"""


def test_reader_keeps_gmsh_element_numbers(tmp_path):
    p = tmp_path / "two.msh"
    p.write_text(MSH)
    m = M.read_gmsh(str(p))
    assert m.K == 2
    assert list(m.gmsh_ids) == [11, 13]


@pytest.mark.parametrize("N", [1, 3, 5])
def test_node_data_round_trip_and_interpolation(tmp_path, N):
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(1)
    rng = np.random.default_rng(N)
    vals = rng.standard_normal((m.K, ops.Np))
    path = tmp_path / "view.msh"
    M.write_gmsh_node_data(str(path), m, ops.r, ops.s, ops.t, vals, name="Field 2", time=0.25, step=7)
    ids, back, coeff, monom = M.read_gmsh_node_data(str(path))
    assert list(ids) == list(range(1, m.K + 1))
    assert np.array_equal(back, vals)
    assert monom.shape == (ops.Np, 3) and monom.sum(axis=1).max() == N
    # a random polynomial of degree N in (u, v, w) = ((r+1)/2, (s+1)/2, (t+1)/2), sampled at the
    # nodes, is reproduced at other points by the written scheme: c = vals @ coeff
    cpoly = rng.standard_normal(ops.Np)
    u, v, w = (ops.r + 1) / 2, (ops.s + 1) / 2, (ops.t + 1) / 2
    at_nodes = sum(cpoly[i] * u ** a * v ** b * w ** c for i, (a, b, c) in enumerate(monom))
    pu, pv, pw = rng.uniform(0, 0.3, (3, 9))
    want = sum(cpoly[i] * pu ** a * pv ** b * pw ** c for i, (a, b, c) in enumerate(monom))
    coef = at_nodes @ coeff
    got = sum(coef[i] * pu ** a * pv ** b * pw ** c for i, (a, b, c) in enumerate(monom))
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9)


def test_writer_uses_file_element_numbers(tmp_path):
    p = tmp_path / "two.msh"
    p.write_text(MSH)
    m = M.read_gmsh(str(p))
    ops = O.build_ref_ops(2)
    out = tmp_path / "v.msh"
    M.write_gmsh_node_data(str(out), m, ops.r, ops.s, ops.t, np.ones((m.K, ops.Np)))
    ids, back, _, _ = M.read_gmsh_node_data(str(out))
    assert list(ids) == [11, 13] and np.all(back == 1.0)


@pytest.mark.parametrize("name,K,ntri", [("cube1", 24, 24), ("cube2", 100, 84), ("cube3", 196, 156)])
def test_reference_meshes(name, K, ntri):
    """The reference's own cube meshes: tets counted and numbered as in the file, positive
    Jacobians, unit volume, one boundary face per boundary triangle of the file, symmetric face
    pairing, linears differentiated exactly (the checks of cpp/src/WaveOKL3d.cpp:1314-1345), and
    the face-node trace map is the same by permutation table and by coordinates."""
    import os
    m = M.read_gmsh(os.path.join(os.path.dirname(__file__), "golden", f"{name}.msh"))
    assert m.K == K and m.gmsh_ids[0] == ntri + 1 and list(m.gmsh_ids) == list(range(ntri + 1, ntri + K + 1))
    M.face_pair(m)
    g = M.geometric_factors(m)
    assert g["J"].min() > 0
    assert abs(float(np.sum(g["J"])) * 4.0 / 3.0 - 1.0) < 1e-12
    k = np.arange(m.K)[:, None]
    assert int(np.sum(m.EToE == k)) == ntri
    assert np.array_equal(m.EToE[m.EToE, m.EToF], np.broadcast_to(k, m.EToE.shape))
    ops = O.build_ref_ops(3)
    x, y, z = M.affine_points(m, ops.r, ops.s, ops.t)
    u = 2 * x + 3 * y + 4 * z                                  # (Np, K) nodal
    # d/dr = (d/dl2 - d/dl1) / 2 etc. in the barycentric Bernstein derivatives DB[0..3]
    Dr, Ds, Dt = (ops.VB @ (0.5 * (ops.DB[i] - ops.DB[0])) @ ops.invVB for i in (1, 2, 3))
    ur, us, ut = Dr @ u, Ds @ u, Dt @ u
    for comp, want in zip("xyz", (2.0, 3.0, 4.0)):
        d = g["r" + comp] * ur + g["s" + comp] * us + g["t" + comp] * ut
        assert np.abs(d - want).max() < 1e-9
    vM, vP = M.build_maps(m, ops)
    vMc, vPc = M.build_maps_by_coordinates(m, ops)
    assert np.array_equal(vP, vPc)
