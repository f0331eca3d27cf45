"""Pins the curvilinear host setup (bbwadg_amd/curved.py) and the curved-path oracle
(oracle/bbw_oracle_curved.c, oracle/curved.py) without sharing their steps:

* the Gordon-Hall sphere map (``GordonHallSphere``, cpp/src/WaveOKL3d.cpp:929-1297) on the
  reference's sphere meshes: the cubature volume approaches 4 pi / 3 as N grows, derivatives of
  2x + 3y + 4z are exact at the cubature points, face nodes of neighbours coincide
  (``checkCurvedGeo``, :1299-1376); at N = 1 nothing moves;
* the face interpolation: one matrix serves every face (the reference's ``VfqFace``);
* the skew-form kernels equal the nodal strong-form planar kernels on affine elements, and the
  planar kernels equal J times the change of basis of the Bernstein oracle's right-hand side
  (itself pinned in tests/test_operators.py and tests/test_oracle.py);
* the whole curved driver (``RK_step_WADG``, :2111-2185) against the spherical Bessel standing
  mode of the reference's ``main`` (cpp/main.cpp:9-14): the error falls under mesh refinement at
  about order N + 1 and with N; a planted sign error in the skew volume term breaks it.
"""
import math

import numpy as np
import pytest

from bbwadg_amd import mesh as M, operators as O, curved as C
import oracle as orc
import curved as OC

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def sphere(name):
    m = M.read_gmsh(f"{GOLDEN}/{name}")
    M.face_pair(m)
    return m


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_gordon_hall_geometry(N):
    m = sphere("sphere385.msh")
    ops = O.build_ref_ops(N)
    nops = C.build_nodal_ops(ops)
    geo = C.gordon_hall_sphere(m, ops)
    arr = C.curved_arrays(geo, nops)
    vol, dlin, gap = C.check_curved_geo(m, geo, nops, arr)
    exact = 4.0 / 3.0 * math.pi
    assert dlin < 1e-10
    assert gap < 1e-12
    assert geo.curved.sum() == 321 and len(geo.Klist_planar) == 385 - 321
    if N == 1:
        # no interior edge nodes: the map moves nothing, elements on the sphere are still flagged
        x, y, z = (a.T for a in M.affine_points(m, ops.r, ops.s, ops.t))
        assert np.abs(geo.x - x).max() == 0
        assert abs(vol - exact) > 0.1
    else:
        assert abs(vol - exact) < {2: 2e-3, 3: 5e-4, 4: 1e-5, 5: 2e-6}[N]
    # boundary nodes lie on the unit sphere for N >= 3 (faces snapped)
    if N >= 3:
        bnd = [(e, f) for e in range(m.K) for f in range(4) if m.EToE[e, f] == e]
        r = [np.sqrt(geo.x[e, ops.Fmask[f]] ** 2 + geo.y[e, ops.Fmask[f]] ** 2
                     + geo.z[e, ops.Fmask[f]] ** 2) for e, f in bnd]
        assert np.abs(np.concatenate(r) - 1).max() < 1e-12


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_face_interpolation_is_one_matrix(N):
    nops = C.build_nodal_ops(O.build_ref_ops(N))
    Nfq = nops.Nfq
    for f in range(4):
        Vf = nops.Vfq[f * Nfq:(f + 1) * Nfq][:, nops.Fmask[f]]
        assert np.abs(Vf - nops.VfqFace).max() < 1e-13
    # Pq, Prq reproduce the projection and the weak derivative
    assert np.abs(nops.Pq @ nops.Vq - np.eye(nops.Np)).max() < 1e-11


def _affine_curved_mesh(N, n=2):
    m = M.cube_mesh(n)
    M.face_pair(m)
    ops = O.build_ref_ops(N)
    nops = C.build_nodal_ops(ops)
    cm = C.setup_curved(m, ops, nops, sphere=False)
    return m, ops, nops, cm


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_skew_form_equals_planar_strong_form(N):
    m, ops, nops, cm = _affine_curved_mesh(N)
    Q = np.random.default_rng(N).uniform(-1, 1, (m.K, 4, ops.Np))
    co = OC.CurvedOracleOps(nops, np.float64)
    allk = np.arange(m.K)
    a = np.zeros_like(Q)
    OC.volume_skew(co, allk, cm.vgeoq, cm.Jq, Q, a)
    OC.surface_skew(co, allk, cm.fgeoq, cm.vmapP, Q, a)
    b = np.zeros_like(Q)
    OC.volume_planar(co, allk, cm.vgeo, cm.Jplanar, Q, b)
    OC.surface_planar(co, allk, cm.fgeo, cm.Jplanar, cm.vmapP, Q, b)
    assert np.abs(a - b).max() < 1e-11 * np.abs(b).max()


@pytest.mark.parametrize("N", [2, 3])
def test_planar_kernels_equal_bernstein_rhs(N):
    m, ops, nops, cm = _affine_curved_mesh(N)
    Qn = np.random.default_rng(N).uniform(-1, 1, (m.K, 4, ops.Np))
    co = OC.CurvedOracleOps(nops, np.float64)
    allk = np.arange(m.K)
    rn = np.zeros_like(Qn)
    OC.volume_planar(co, allk, cm.vgeo, cm.Jplanar, Qn, rn)
    OC.surface_planar(co, allk, cm.fgeo, cm.Jplanar, cm.vmapP, Qn, rn)
    Qb = np.einsum("ij,kfj->kfi", ops.invVB, Qn)
    oo = orc.OracleOps(ops, np.float64)
    rb = np.zeros_like(Qb)
    orc.volume(oo, m.K, cm.vgeo, Qb, rb)
    orc.surface(oo, m.K, cm.fgeo, cm.vmapP, Qb, rb)
    ref = np.einsum("ij,kfj->kfi", ops.VB, rb) * cm.Jplanar[:, None, None]
    assert np.abs(rn - ref).max() < 1e-10 * np.abs(ref).max()


def bessel_error(name, N, T=0.1, planted=False):
    m = sphere(name)
    ops = O.build_ref_ops(N)
    nops = C.build_nodal_ops(ops)
    cm = C.setup_curved(m, ops, nops)
    Q = np.zeros((m.K, 4, ops.Np))
    Q[:, 0] = C.bessel(cm.x, cm.y, cm.z, 0.0)
    co = OC.CurvedOracleOps(nops, np.float64)
    if planted:   # the weak divergence with the wrong sign
        co.Prq, co.Psq, co.Ptq = -co.Prq, -co.Psq, -co.Ptq
    dt = 0.25 / ((N + 1) ** 2 * cm.FscaleMax)
    ns = int(math.ceil(T / dt))
    OC.rk_run_curved(co, cm, np.ones((m.K, nops.Nq)), Q, T / ns, ns)
    return C.l2_error(nops, cm, Q, C.bessel, T)[0]


def test_curved_driver_converges_to_the_bessel_mode():
    e = {(name, N): bessel_error(name, N) for name in ("sphere48.msh", "sphere385.msh") for N in (2, 3)}
    # h halves from 48 to 385 elements: about order N + 1
    assert e[("sphere48.msh", 2)] / e[("sphere385.msh", 2)] > 2 ** 2.5
    assert e[("sphere48.msh", 3)] / e[("sphere385.msh", 3)] > 2 ** 3.5
    assert e[("sphere385.msh", 3)] < e[("sphere385.msh", 2)] / 10
    assert e[("sphere385.msh", 3)] < 5e-4


def test_planted_sign_error_breaks_the_bessel_check():
    assert not bessel_error("sphere48.msh", 3, planted=True) < 1e-2
