"""GPU parity of the elastic kernels (BB volume, BB-lift penalty surface, Lame WADG update)
against the elastic oracle on the same seeded inputs; tolerances as the acoustic tests."""
import numpy as np
import pytest

from bbwadg_amd import _lib, media, mesh as M, operators as O
from bbwadg_amd.solver import WaveSolver, MODE_PLAIN
import oracle as orc
from cases import rel, rel_l2

pytestmark = pytest.mark.gpu


def make(N, n, seed=0, tau=(1.0, 1.0)):
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    lam, mu, rho = media.lame_fourier(seed)
    xq, yq, zq = media.cubature_points(m, ops)
    mats = (1.0 / rho(xq, yq, zq), lam(xq, yq, zq), mu(xq, yq, zq))
    s = WaveSolver(m, N, nfields=9, ops=ops, dmesh=dm)
    s.InitElastic(*mats, tau_v=tau[0], tau_s=tau[1])
    return ops, m, dm, mats, s


@pytest.mark.parametrize("impl", [_lib.PLAIN_FLAT, _lib.PLAIN_TILED])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_elastic_kernels(N, impl):
    print(f"case elastic kernels N={N} impl={impl}", flush=True)
    ops, m, dm, mats, s = make(N, 3, tau=(0.8, 1.2))
    s.set_option(_lib.OPT_PLAIN_IMPL, impl)
    rng = np.random.default_rng(N)
    Q = rng.uniform(-1, 1, (m.K, 9, ops.Np)).astype(np.float32)
    R0 = rng.uniform(-1, 1, Q.shape).astype(np.float32)
    res = rng.uniform(-1, 1, Q.shape).astype(np.float32)
    oo = orc.OracleOps(ops, np.float64)
    s.WaveSetData3d(Q)
    s.volume()
    rv = np.zeros(Q.shape)
    orc.volume(oo, m.K, dm.vgeo, Q.astype(np.float64), rv, nfields=9)
    assert rel(s.get_rhs(), rv) < 1e-5
    s.set_rhs(R0)
    s.surface()
    rs = R0.astype(np.float64).copy()
    orc.surface(oo, m.K, dm.fgeo, dm.vmapP, Q.astype(np.float64), rs, nfields=9, tau_v=0.8, tau_s=1.2)
    assert rel(s.get_rhs(), rs) < 1e-5
    s.set_rhs(R0); s.set_res(res)
    fa, fb, fdt = np.float32(ops.rk4a[3]), np.float32(ops.rk4b[3]), np.float32(1e-2)
    s.update_wadg(fa, fb, fdt)
    rQ, rres = Q.astype(np.float64).copy(), res.astype(np.float64).copy()
    orc.update_elas(oo, m.K, *[a.astype(np.float32) for a in mats], fa, fb, fdt,
                    R0.astype(np.float64), rres, rQ)
    assert rel(s.get_res(), rres) < 2e-5
    assert rel(s.WaveGetData3d(), rQ) < 2e-5


def test_elastic_run_with_moment_source():
    """Explosive moment-tensor Ricker source (sxx = syy = szz), 30 steps vs FP64 oracle."""
    print("case elastic run", flush=True)
    N = 4
    ops, m, dm, mats, s = make(N, 3)
    elems, coef = media.gaussian_source(m, ops, (0.02, -0.03, 0.05), a=15.0)
    w = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0], dtype=np.float64)
    f0, amp = 10.0, 5.0
    s.set_source(elems, coef, f0=f0, amp=amp, weights=w)
    Q0 = np.zeros((m.K, 9, ops.Np), dtype=np.float32)
    s.WaveSetData3d(Q0)
    dt = 0.25 / ((N + 1) * (N + 3) / 3.0 * dm.FscaleMax)
    s.step(30, dt, mode=MODE_PLAIN, graph=True)
    src = np.zeros((m.K, ops.Np)); src[elems] = coef
    ref = Q0.astype(np.float64).copy()
    orc.rk_run_elas(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP,
                    *[a.astype(np.float32) for a in mats], ref, dt, 30, src=src, srcw=w, f0=f0, amp=amp)
    assert np.abs(ref).max() > 0
    assert rel_l2(s.WaveGetData3d(), ref) < 1e-5


LIFTS = [_lib.ELAS_LIFT_CUBATURE, _lib.ELAS_LIFT_NODES]


@pytest.mark.parametrize("lift", LIFTS)
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_fused_elastic_stage(N, lift):
    """One fused elastic stage (volume + penalty surface + Lame WADG update + LSRK in one launch),
    in both forms (lift folded into the cubature product, or at the nodes), against the FP64
    oracle's three kernels on the same FP32 inputs."""
    print(f"case fused elastic stage N={N} lift={lift}", flush=True)
    ops, m, dm, mats, s = make(N, 3, tau=(0.8, 1.2))
    s.set_option(_lib.OPT_ELAS_FUSED, lift)
    rng = np.random.default_rng(10 + N)
    Q = rng.uniform(-1, 1, (m.K, 9, ops.Np)).astype(np.float32)
    res = rng.uniform(-1, 1, Q.shape).astype(np.float32)
    s.WaveSetData3d(Q)
    s.set_res(res)
    fa, fb, fdt = np.float32(ops.rk4a[2]), np.float32(ops.rk4b[2]), np.float32(1e-2)
    s.stage_fused(fa, fb, fdt)
    oo = orc.OracleOps(ops, np.float64)
    rhs = np.zeros(Q.shape)
    Q64 = Q.astype(np.float64)
    orc.volume(oo, m.K, dm.vgeo, Q64, rhs, nfields=9)
    orc.surface(oo, m.K, dm.fgeo, dm.vmapP, Q64, rhs, nfields=9, tau_v=0.8, tau_s=1.2)
    rQ, rres = Q64.copy(), res.astype(np.float64).copy()
    orc.update_elas(oo, m.K, *[a.astype(np.float32) for a in mats], fa, fb, fdt, rhs, rres, rQ)
    assert rel(s.get_res(), rres) < 2e-5
    assert rel(s.WaveGetData3d(), rQ) < 2e-5
    s.close()


@pytest.mark.parametrize("lift", LIFTS)
@pytest.mark.parametrize("N", [3, 4])
def test_fused_elastic_run_with_moment_source(N, lift):
    """Fused elastic steps with the explosive moment-tensor source: graph replay bitwise equal to
    eager stepping, within 1e-5 of the FP64 oracle and of the three-launch path."""
    print(f"case fused elastic run N={N} lift={lift}", flush=True)
    ops, m, dm, mats, s = make(N, 3)
    s.set_option(_lib.OPT_ELAS_FUSED, lift)
    elems, coef = media.gaussian_source(m, ops, (0.02, -0.03, 0.05), a=15.0)
    w = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0], dtype=np.float64)
    f0, amp, nsteps = 10.0, 5.0, 23
    s.set_source(elems, coef, f0=f0, amp=amp, weights=w)
    Q0 = (0.1 * np.random.default_rng(N).uniform(-1, 1, (m.K, 9, ops.Np))).astype(np.float32)
    dt = s.dt_ref
    out = {}
    for name, mode, graph in (("fused_graph", _lib.MODE_FUSED, True), ("fused_eager", _lib.MODE_FUSED, False),
                              ("plain", MODE_PLAIN, True)):
        s.WaveSetData3d(Q0)
        s.set_res(np.zeros_like(Q0))
        s.step(nsteps, dt, mode=mode, graph=graph, t0=0.0)
        out[name] = s.WaveGetData3d()
    assert np.array_equal(out["fused_graph"], out["fused_eager"])
    src = np.zeros((m.K, ops.Np)); src[elems] = coef
    ref = Q0.astype(np.float64).copy()
    orc.rk_run_elas(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP,
                    *[a.astype(np.float32) for a in mats], ref, dt, nsteps, src=src, srcw=w, f0=f0, amp=amp)
    assert rel_l2(out["fused_graph"], ref) < 1e-5
    assert rel_l2(out["fused_graph"], out["plain"]) < 1e-5
    s.close()


@pytest.mark.parametrize("mode", [_lib.MODE_FUSED, MODE_PLAIN])
def test_reference_elastic_driver_switches(mode):
    """The reference's elastic run (elas_cpp/src/WaveOKL3d.cpp:453-506, 1675-1693): its media
    (lambda, mu jump across z = 0, mu averaged per element), the smoothed source at (0, 0, 0.1)
    on v_x with amplitude 1e4, the Ricker pulse evaluated at each step's start for all five
    stages; 20 steps against the FP64 oracle with the same step-start source time."""
    print(f"case reference elastic driver mode={mode}", flush=True)
    N = 3
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(4)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    weight = lambda x, y, z: 0.1 * np.sin(3 * x) * np.cos(2 * y)
    mats = tuple(np.asarray(a, dtype=np.float32) for a in media.elastic_reference_media(m, ops, weight))
    s = WaveSolver(m, N, nfields=9, ops=ops, dmesh=dm)
    s.InitElastic(*mats)
    elems, coef = media.gaussian_source(m, ops, (0.0, 0.0, 0.1), a=100.0)
    w = np.array([1, 0, 0, 0, 0, 0, 0, 0, 0], dtype=np.float64)
    s.set_source(elems, coef, f0=10.0, amp=1e4, weights=w)
    s.set_source_step_time(True)
    Q0 = np.zeros((m.K, 9, ops.Np), dtype=np.float32)
    s.WaveSetData3d(Q0)
    dt, nsteps = s.dt_ref, 20
    s.step(nsteps, dt, mode=mode, graph=True, t0=0.05)
    got = s.WaveGetData3d()
    src = np.zeros((m.K, ops.Np)); src[elems] = coef
    ref = Q0.astype(np.float64)
    orc.rk_run_elas(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, *mats, ref, dt, nsteps,
                    t0=0.05, src=src, srcw=w, f0=10.0, amp=1e4, src_time="step")
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) < 1e-5
    # the stage-time source differs from it at first order in dt
    ref2 = Q0.astype(np.float64)
    orc.rk_run_elas(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, *mats, ref2, dt, nsteps,
                    t0=0.05, src=src, srcw=w, f0=10.0, amp=1e4)
    assert rel_l2(got, ref2) > 1e-4
    s.close()
