"""The fused and adjoint stages' tile order (BBW_OPT_XCD_CHUNK: runs of consecutive tiles per
XCD) and the element numbering (Z-order renumbering of the mesh) change where a tile's work
runs, not its arithmetic: each element's result is the same bits under every tile order, and a
renumbered mesh gives the renumbered result.  Meshes of 8^3 x 6 tets (192 tiles), so that runs
of up to 16 tiles (8 x 16 = 128 workgroups per span) are taken and a partial span is left."""
import numpy as np
import pytest

from bbwadg_amd import _lib, media, mesh as M
from bbwadg_amd.adjoint import AcousticInversion, ElasticInversion, Shot
from bbwadg_amd.solver import WaveSolver
from cases import Problem

pytestmark = pytest.mark.gpu

CHUNKS = (0, 1, 4, 16)


def _acoustic(P, Q0, chunk, nsteps=3, mode=1, src=None):
    s = WaveSolver(P.mesh, P.N, ops=P.ops, dmesh=P.dm)
    s.InitWADG_subelem(P.c2q.astype(np.float32))
    s.set_option(_lib.OPT_XCD_CHUNK, chunk)
    if src is not None:
        s.set_source(*src, f0=10.0, amp=5.0)
    s.WaveSetData3d(Q0)
    s.step(nsteps, P.dt_cfl(), mode=mode)
    return s.WaveGetData3d()


def _elastic(P, Q0, chunk, form, nsteps=3):
    s = WaveSolver(P.mesh, P.N, nfields=9, ops=P.ops, dmesh=P.dm)
    lam, mu, rho = media.lame_fourier(2)
    s.InitElastic(1.0 / rho(P.xq, P.yq, P.zq), lam(P.xq, P.yq, P.zq), mu(P.xq, P.yq, P.zq))
    s.set_option(_lib.OPT_ELAS_FUSED, form)
    s.set_option(_lib.OPT_XCD_CHUNK, chunk)
    s.WaveSetData3d(Q0)
    s.step(nsteps, s.dt_ref, mode=1)
    return s.WaveGetData3d()


def test_run_length_must_be_a_power_of_two_up_to_1024():
    print("case run-length option", flush=True)
    P = Problem(2, 2)
    s = WaveSolver(P.mesh, 2, ops=P.ops, dmesh=P.dm)
    for bad in (3, 12, 2048, -16):
        with pytest.raises(_lib.BBWError):
            s.set_option(_lib.OPT_XCD_CHUNK, bad)
    for good in (0, 1, 2, 1024):
        s.set_option(_lib.OPT_XCD_CHUNK, good)


@pytest.mark.parametrize("N", [3, 5])
def test_acoustic_fused_stage_same_bits_under_every_tile_order(N):
    print(f"case acoustic tile order N={N}", flush=True)
    P = Problem(N, 8, seed=3, order="morton")
    Q0 = P.random_state()
    ref = _acoustic(P, Q0, 0)
    assert np.isfinite(ref).all() and np.abs(ref - Q0).max() > 1e-3
    for c in CHUNKS[1:]:
        assert np.array_equal(_acoustic(P, Q0, c), ref), f"chunk {c}"


@pytest.mark.parametrize("form", [_lib.ELAS_LIFT_CUBATURE, _lib.ELAS_LIFT_NODES])
def test_elastic_fused_stage_same_bits_under_every_tile_order(form):
    print(f"case elastic tile order form={form}", flush=True)
    P = Problem(4, 8, nfields=9, seed=3, order="morton")
    Q0 = P.random_state()
    ref = _elastic(P, Q0, 0, form)
    assert np.isfinite(ref).all()
    for c in CHUNKS[1:]:
        assert np.array_equal(_elastic(P, Q0, c, form), ref), f"chunk {c}"


def test_renumbered_mesh_gives_the_renumbered_result():
    print("case renumbered mesh", flush=True)
    P = Problem(5, 8, seed=4)                       # x-fastest cubes
    perm = M.sfc_order(P.mesh)
    assert not np.array_equal(perm, np.arange(P.K))
    R = Problem(5, 8, seed=4)
    R.mesh = M.renumber(P.mesh, perm)
    M.face_pair(R.mesh)
    R.dm = M.device_arrays(R.mesh, R.ops, 4)
    R.xq, R.yq, R.zq = media.cubature_points(R.mesh, R.ops)
    R.c2q = media.c2_fourier(4)(R.xq, R.yq, R.zq)
    assert np.array_equal(R.c2q, P.c2q[perm])
    Q0 = P.random_state()
    a = _acoustic(P, Q0, 0)
    b = _acoustic(R, Q0[perm], 16)
    assert np.array_equal(b, a[perm])
    # with a source: its elements scatter over many tiles of the renumbered mesh (the per-tile
    # source table), and the plain stage's update reads the same table
    el, co = media.gaussian_source(P.mesh, P.ops, (0.1, -0.05, 0.2), a=12.0)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(P.K)
    o = np.argsort(inv[el])
    srcR = (inv[el][o].astype(np.int32), np.ascontiguousarray(co[o]))
    assert len(el) > 64 and len(np.unique(el // 16)) > 8 and len(np.unique(srcR[0] // 16)) > 8
    for mode in (1, 0):
        a = _acoustic(P, Q0, 0, mode=mode, src=(el, co))
        b = _acoustic(R, Q0[perm], 16, mode=mode, src=srcR)
        assert np.array_equal(b, a[perm]), f"mode {mode}"


def test_acoustic_adjoint_same_bits_under_every_tile_order():
    print("case acoustic adjoint tile order", flush=True)
    P = Problem(3, 8, seed=5, order="morton")
    s = WaveSolver(P.mesh, 3, ops=P.ops, dmesh=P.dm)
    el, co = media.gaussian_source(P.mesh, P.ops, (0.05, -0.1, 0.0), a=8.0)
    inv = AcousticInversion(s, [Shot(el, co, f0=8.0, amp=30.0)], media.surface_receivers(2, 9), 12,
                            P.dt_cfl(), ckpt_every=5)
    d = inv.observe(P.c2q)
    c2 = np.ones_like(P.c2q)
    out = []
    for c in (0, 16):
        s.set_option(_lib.OPT_XCD_CHUNK, c)
        out.append(inv.misfit_and_gradient(c2, d))
    assert out[0][0] == out[1][0] and out[0][0] > 0
    assert np.array_equal(out[0][1], out[1][1]) and np.abs(out[0][1]).max() > 0


def test_elastic_adjoint_same_bits_under_every_tile_order():
    print("case elastic adjoint tile order", flush=True)
    P = Problem(3, 8, nfields=9, seed=5, order="morton")
    s = WaveSolver(P.mesh, 3, nfields=9, ops=P.ops, dmesh=P.dm)
    lam, mu, rho = media.lame_fourier(4)
    m_true = (1.0 / rho(P.xq, P.yq, P.zq), lam(P.xq, P.yq, P.zq), mu(P.xq, P.yq, P.zq))
    s.InitElastic(*m_true)
    el, co = media.gaussian_source(P.mesh, P.ops, (0.05, -0.1, 0.2), a=8.0)
    w = np.array([0.2, 0, 0, 1, 1, 1, 0, 0.3, 0])
    inv = ElasticInversion(s, [Shot(el, co, f0=20.0, amp=30.0, weights=w)], media.surface_receivers(2, 9),
                           12, s.dt_ref, ckpt_every=5)
    d = inv.observe(m_true)
    m0 = tuple(np.full_like(a, float(np.mean(a))).astype(np.float32) for a in m_true)
    out = []
    for c in (0, 16):
        s.set_option(_lib.OPT_XCD_CHUNK, c)
        out.append(inv.misfit_and_gradient(m0, d))
    assert out[0][0] == out[1][0] and out[0][0] > 0
    assert np.array_equal(out[0][1], out[1][1]) and np.abs(out[0][1]).max() > 0
