"""Element numbering: the Z-order cube numbering and ``sfc_order``/``renumber`` give the same
mesh, every per-element array of the renumbered mesh is the permuted array, and the FP64
oracle's LSRK run on the renumbered mesh is the permuted run (the element numbering carries no
arithmetic)."""
import numpy as np
import pytest

from bbwadg_amd import media, mesh as M, operators as O
import oracle as orc


def _tets(m):
    return {tuple(sorted(zip(m.VX[v], m.VY[v], m.VZ[v]))) for v in m.EToV}


def test_morton_cube_mesh_is_the_same_mesh():
    a, b = M.cube_mesh(4), M.cube_mesh(4, order="morton")
    assert a.K == b.K == 384 and _tets(a) == _tets(b)
    assert not np.array_equal(a.EToV, b.EToV)
    with pytest.raises(ValueError):
        M.cube_mesh(2, order="hilbert")


def test_morton_codes_interleave_bits():
    assert M.morton_codes([1, 0, 0, 3], [0, 1, 0, 0], [0, 0, 1, 1]).tolist() == [1, 2, 4, 13]
    c = M.morton_codes(*np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"))
    assert sorted(c.ravel().tolist()) == list(range(64))


def test_sfc_order_keeps_neighbours_close():
    m = M.cube_mesh(8)
    M.face_pair(m)
    r = M.renumber(m, M.sfc_order(m))
    M.face_pair(r)

    def near(mm):   # interior face pairs at most 16 tiles (256 elements) apart
        e = np.arange(mm.K)[:, None]
        return float((np.abs(mm.EToE - e) <= 256)[mm.EToE != e].mean())
    # x-fastest cubes: the z neighbours, 384 elements away, are the far sixth of the pairs
    assert near(m) == pytest.approx(0.844, abs=1e-3) and near(r) > 0.93


def test_renumber_permutes_every_per_element_array():
    ops = O.build_ref_ops(3)
    m = M.cube_mesh(3)
    M.face_pair(m)
    perm = M.sfc_order(m)
    r = M.renumber(m, perm)
    M.face_pair(r)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.K)
    # connectivity: r's element e is m's perm[e]; neighbours map through inv
    assert np.array_equal(r.EToE, inv[m.EToE[perm]])
    assert np.array_equal(r.EToF, m.EToF[perm])
    da, db = M.device_arrays(m, ops, 4), M.device_arrays(r, ops, 4)
    assert np.array_equal(db.vgeo, da.vgeo[perm]) and np.array_equal(db.fgeo, da.fgeo[perm])
    row = 4 * ops.Np
    va = da.vmapP.reshape(m.K, -1)[perm]
    assert np.array_equal(db.vmapP.reshape(m.K, -1), inv[va // row] * row + va % row)
    qa, qb = media.cubature_points(m, ops), media.cubature_points(r, ops)
    for x, y in zip(qa, qb):
        assert np.array_equal(y, x[perm])
    with pytest.raises(M.MeshError):
        M.renumber(m, perm[:-1])
    with pytest.raises(M.MeshError):
        M.renumber(m, np.zeros(m.K, dtype=np.int64))


def test_oracle_run_on_renumbered_mesh_is_the_permuted_run():
    ops = O.build_ref_ops(3)
    m = M.cube_mesh(3)
    M.face_pair(m)
    perm = M.sfc_order(m)
    r = M.renumber(m, perm)
    M.face_pair(r)
    oo = orc.OracleOps(ops, np.float64)
    out = []
    for mm, p in ((m, np.arange(m.K)), (r, perm)):
        dm = M.device_arrays(mm, ops, 4)
        xq, yq, zq = media.cubature_points(mm, ops)
        c2q = media.c2_fourier(2)(xq, yq, zq)
        Q = np.random.default_rng(1).uniform(-1, 1, (m.K, 4, ops.Np))[p].copy()
        orc.rk_run(oo, mm.K, dm.vgeo, dm.fgeo, dm.vmapP, c2q, Q, 0.01, 2)
        out.append(Q)
    assert np.abs(out[0]).max() > 0
    assert np.array_equal(out[1], out[0][perm])
