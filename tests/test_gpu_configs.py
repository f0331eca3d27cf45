"""Each benchmark configuration's workload on the GPU at its own order and size, against the FP64
oracle (oracle/bbw_oracle.c, pinned in tests/test_oracle.py) on the same seeded inputs:

* cavity, 8^3 x 6 = 3072 tets, N = 3, c^2 = 1, 200 steps at dt = 0.5 h / N^2: fused and plain
  within 1e-5 relative L2 of the FP64 oracle, and their error against the exact standing mode
  equal to the oracle's (cpp/main_subelem.cpp:3-5);
* heterogeneous single shot at N = 5 with the Ricker source (small mesh, 40 steps) and at its
  full size (48^3 x 6 = 663 552 tets, 2 steps from a random state);
* elastic, 32^3 x 6 = 196 608 tets, N = 4: 2 steps at full size against the oracle; 500 steps
  at the reference's elastic time step with the explosive moment-tensor source stay finite, and
  without a source the FP64 energy does not grow over any 100-step window;
* batched shots at N = 5: uneven meshes and durations, one or two shots per tile, persistent
  and per-item grids, bitwise equal to single-shot runs and within 1e-5 of the oracle;
* adjoint gradient at N = 5: within 1e-4 of the FP64 host adjoint, and against central
  differences of the FP64 oracle's misfit (eps = 1e-4) along 20 random directions.

Tolerances are written in each test; the states are FP32 on the GPU."""
import math

import numpy as np
import pytest

from bbwadg_amd import _lib, media, mesh as M, operators as O
from bbwadg_amd.solver import WaveSolver, MODE_FUSED, MODE_PLAIN
import oracle as orc
from cases import Problem, cavity, rel_l2
from test_elastic_oracle import elastic_energy

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("mode", [MODE_FUSED, MODE_PLAIN])
def test_c1_cavity_full_size(mode):
    say(f"case C1 cavity 8^3x6 N=3 200 steps mode={mode}")
    P = Problem(3, 8, c2="one")
    s = WaveSolver(P.mesh, 3, ops=P.ops, dmesh=P.dm)
    s.InitWADG_subelem(P.c2q)
    Q0 = P.cavity_state()
    dt, nsteps = P.dt_cfl(), 200
    s.WaveSetData3d(Q0)
    s.step(nsteps, dt, mode=mode, graph=True)
    got = s.WaveGetData3d()
    ref = Q0.astype(np.float64).copy()
    orc.rk_run(orc.OracleOps(P.ops, np.float64), P.K, P.dm.vgeo, P.dm.fgeo, P.dm.vmapP, P.c2q, ref,
               dt, nsteps, nproj=1 if mode == MODE_FUSED else 4)
    assert rel_l2(got, ref) < 1e-5
    T = nsteps * dt
    e_gpu = s.compute_error(T, cavity, Q=got)[1]
    e_ref = s.compute_error(T, cavity, Q=ref)[1]
    assert e_ref < 2e-4                       # the scheme's own error at this resolution
    assert abs(e_gpu - e_ref) < 1e-2 * e_ref   # FP32 rounding is far below it
    E = s.energy()
    assert abs(E - orc.energy_quadrature(P.ops, P.dm.J, P.c2q, got)) < 1e-10 * E
    s.close()


@pytest.mark.parametrize("mode", [MODE_FUSED, MODE_PLAIN])
def test_c2_source_at_N5(mode):
    say(f"case C2 source N=5 mode={mode}")
    P = Problem(5, 3, seed=4)
    s = WaveSolver(P.mesh, 5, ops=P.ops, dmesh=P.dm)
    s.InitWADG_subelem(P.c2q)
    el, co = media.gaussian_source(P.mesh, P.ops, (0.07, -0.11, 0.04), a=20.0)
    s.set_source(el, co, f0=10.0, amp=1.0)
    relems, B = media.receivers(P.mesh, P.ops, media.surface_receivers(3, 16))
    nsteps = 40
    s.set_receivers(relems, B, nsteps)
    Q0 = np.zeros((P.K, 4, P.ops.Np), dtype=np.float32)
    s.WaveSetData3d(Q0)
    dt = P.dt_cfl()
    s.step(nsteps, dt, mode=mode, graph=True, t0=0.0)
    got, tr = s.WaveGetData3d(), s.get_traces(nsteps)
    src = np.zeros((P.K, P.ops.Np))
    src[el] = co
    ref, rec = Q0.astype(np.float64).copy(), []
    orc.rk_run(orc.OracleOps(P.ops, np.float64), P.K, P.dm.vgeo, P.dm.fgeo, P.dm.vmapP,
               P.c2q.astype(np.float32), ref, dt, nsteps, nproj=1 if mode == MODE_FUSED else 4,
               src=src, f0=10.0, amp=1.0,
               record=lambda st, Q: rec.append(np.einsum("rj,rj->r", B, Q[relems, 0])))
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) < 1e-5
    assert rel_l2(tr, np.array(rec)) < 1e-5
    s.close()


def _full_size_mesh(n, N, nfields):
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n)
    M.face_pair(m)
    return ops, m, M.device_arrays(m, ops, nfields)


def _gpu_eval(fn, cub):
    """fn(x, y, z) at the cubature points, evaluated in FP64 on the GPU through torch."""
    import torch
    t = [torch.as_tensor(c, dtype=torch.float64, device="cuda") for c in cub]
    return fn(*t).cpu().numpy()


def test_c2_full_size_two_steps():
    say("case C2 full size 48^3x6 N=5, 2 steps")
    N, n, nsteps = 5, 48, 2
    ops, m, dm = _full_size_mesh(n, N, 4)
    cub = media.cubature_points(m, ops)
    c2q = _gpu_eval(media.c2_fourier(1), cub).astype(np.float32)
    el, co = media.gaussian_source(m, ops, (0.12, -0.2, 0.05), a=100.0, cub=cub)
    Q0 = np.random.default_rng(3).uniform(-1, 1, (m.K, 4, ops.Np)).astype(np.float32)
    s = WaveSolver(m, N, ops=ops, dmesh=dm)
    s.InitWADG_subelem(c2q)
    s.set_source(el, co, f0=10.0, amp=1.0)
    s.WaveSetData3d(Q0)
    dt = 0.5 / n / N ** 2
    t0 = 0.08   # the Ricker pulse near its peak
    s.step(nsteps, dt, mode=MODE_FUSED, graph=False, t0=t0)
    got = s.WaveGetData3d()
    s.close()
    say("C2 full size: GPU done, FP64 oracle running")
    src = np.zeros((m.K, ops.Np))
    src[el] = co
    ref = Q0.astype(np.float64)
    orc.rk_run(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, c2q, ref, dt, nsteps,
               t0=t0, nproj=1, src=src, f0=10.0, amp=1.0)
    assert rel_l2(got, ref) < 1e-5


def _elastic_problem(n, N, seed=1):
    ops, m, dm = _full_size_mesh(n, N, 9)
    cub = media.cubature_points(m, ops)
    lam, mu, rho = media.lame_fourier(seed)
    mats = tuple(np.asarray(a, dtype=np.float32) for a in (1.0 / rho(*cub), lam(*cub), mu(*cub)))
    s = WaveSolver(m, N, nfields=9, ops=ops, dmesh=dm)
    s.InitElastic(*mats)
    return ops, m, dm, mats, s, cub


def test_c3_full_size_two_steps():
    say("case C3 full size 32^3x6 N=4 elastic, 2 steps")
    N, n, nsteps = 4, 32, 2
    ops, m, dm, mats, s, cub = _elastic_problem(n, N)
    el, co = media.gaussian_source(m, ops, (0.02, -0.03, 0.05), a=100.0, cub=cub)
    w = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0], dtype=np.float64)
    s.set_source(el, co, f0=10.0, amp=1.0, weights=w)
    Q0 = np.random.default_rng(4).uniform(-1, 1, (m.K, 9, ops.Np)).astype(np.float32)
    s.WaveSetData3d(Q0)
    dt, t0 = s.dt_ref, 0.09
    s.step(nsteps, dt, mode=MODE_PLAIN, graph=False, t0=t0)
    got = s.WaveGetData3d()
    s.close()
    say("C3 full size: GPU done, FP64 oracle running")
    src = np.zeros((m.K, ops.Np))
    src[el] = co
    ref = Q0.astype(np.float64)
    orc.rk_run_elas(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, *mats, ref, dt,
                    nsteps, t0=t0, src=src, srcw=w, f0=10.0, amp=1.0)
    assert rel_l2(got, ref) < 1e-5


def test_c3_500_steps_stable():
    say("case C3 500 steps at the reference elastic dt")
    N, n = 4, 32
    ops, m, dm, mats, s, cub = _elastic_problem(n, N)
    dt = s.dt_ref
    # the configuration: explosive moment-tensor Ricker source from rest, 500 steps
    el, co = media.gaussian_source(m, ops, (0.02, -0.03, 0.05), a=100.0, cub=cub)
    w = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0], dtype=np.float64)
    s.set_source(el, co, f0=10.0, amp=1.0, weights=w)
    s.WaveSetData3d(np.zeros((m.K, 9, ops.Np), dtype=np.float32))
    s.step(500, dt, mode=MODE_PLAIN, graph=True, t0=0.0)
    Q = s.WaveGetData3d()
    assert np.all(np.isfinite(Q)) and np.abs(Q).max() > 0
    # without a source, from a smooth start: the FP64 energy does not grow over 100-step windows
    s.set_source([], np.zeros((0, ops.Np)), f0=10.0)
    x, y, z = M.affine_points(m, ops.r, ops.s, ops.t)
    bump = (ops.invVB @ np.exp(-40 * ((x - 0.1) ** 2 + y ** 2 + (z + 0.05) ** 2))).T
    Q0 = np.stack([c * bump for c in (1.0, -0.5, 0.3, 0.8, 0.2, -0.4, 0.1, 0.3, -0.2)], axis=1)
    s.WaveSetData3d(Q0.astype(np.float32))
    s.set_res(np.zeros_like(Q0, dtype=np.float32))
    E = [elastic_energy(ops, dm.J, *mats, Q0)]
    for _ in range(5):
        s.step(100, dt, mode=MODE_PLAIN, graph=True)
        E.append(elastic_energy(ops, dm.J, *mats, s.WaveGetData3d()))
    say(f"C3 energies {E}")
    assert all(np.isfinite(E)) and all(E[i + 1] <= E[i] for i in range(len(E) - 1)), E
    s.close()


def _batch_case(N=5):
    ops = O.build_ref_ops(N)
    rng = np.random.default_rng(21)
    meshes = []
    for n in (3, 4):
        m = M.cube_mesh(n)
        M.face_pair(m)
        meshes.append((m, n, M.device_arrays(m, ops, 4)))
    shots = []
    for i, (mi, nsteps) in enumerate([(0, 13), (1, 30), (1, 22), (0, 47), (1, 11)]):
        m, n, dm = meshes[mi]
        c2q = media.c2_fourier(300 + i)(*media.cubature_points(m, ops))
        el, co = media.gaussian_source(m, ops, rng.uniform(-0.3, 0.3, 3), a=15.0)
        shots.append(dict(mi=mi, c2q=c2q, src=(el, co), f0=float(rng.uniform(5, 15)), amp=1.5,
                          dt=0.5 / n / N ** 2, nsteps=nsteps,
                          Q0=rng.uniform(-1, 1, (m.K, 4, ops.Np)).astype(np.float32)))
    return ops, meshes, shots


@pytest.mark.parametrize("spt,bpc", [(1, 0), (2, 0), (1, 2), (2, 3)])
def test_c4_batch_at_N5(spt, bpc):
    from bbwadg_amd.batch import ShotBatch
    say(f"case C4 batch N=5 shots_per_tile={spt} blocks_per_cu={bpc}")
    ops, meshes, shots = _batch_case()
    b = ShotBatch(5, ops=ops)
    mids = [b.add_mesh(m, dm) for m, n, dm in meshes]
    sids = [b.add_shot(mids[s["mi"]], s["c2q"], s["Q0"], src=s["src"], f0=s["f0"], amp=s["amp"],
                       dt=s["dt"], nsteps=s["nsteps"]) for s in shots]
    b.run(shots_per_tile=spt, blocks_per_cu=bpc)
    for sid, sh in zip(sids, shots):
        m, n, dm = meshes[sh["mi"]]
        sv = WaveSolver(m, 5, ops=ops, dmesh=dm)
        sv.InitWADG_subelem(sh["c2q"])
        sv.set_source(*sh["src"], f0=sh["f0"], amp=sh["amp"])
        sv.WaveSetData3d(sh["Q0"])
        sv.step(sh["nsteps"], sh["dt"], mode=MODE_FUSED, graph=True, t0=0.0)
        assert np.array_equal(b.state(sid), sv.WaveGetData3d()), sid
        sv.close()
    sh = shots[1]
    m, n, dm = meshes[sh["mi"]]
    src = np.zeros((m.K, ops.Np))
    src[sh["src"][0]] = sh["src"][1]
    ref = sh["Q0"].astype(np.float64).copy()
    orc.rk_run(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP,
               sh["c2q"].astype(np.float32), ref, sh["dt"], sh["nsteps"], nproj=1, src=src,
               f0=sh["f0"], amp=sh["amp"])
    assert rel_l2(b.state(sids[1]), ref) < 1e-5
    b.close()


def _adjoint_case(N=5, n=3, nsteps=40, nshots=2):
    from bbwadg_amd.adjoint import AcousticInversion, Shot
    P = Problem(N, n, seed=8)
    s = WaveSolver(P.mesh, N, ops=P.ops, dmesh=P.dm)
    shots = []
    for i in range(nshots):
        el, co = media.gaussian_source(P.mesh, P.ops, (0.12 * (i - 0.5), -0.08, 0.06 * i), a=8.0)
        shots.append(Shot(el, co, f0=9.0 + 3 * i, amp=30.0))
    inv = AcousticInversion(s, shots, media.surface_receivers(6, 32), nsteps, P.dt_cfl(), ckpt_every=9)
    d = inv.observe(P.c2q)
    return P, s, inv, d, np.ones_like(P.c2q), shots


def test_c5_adjoint_at_N5_against_fp64_adjoint_and_fd():
    import adjoint as adjm
    say("case C5 adjoint N=5")
    P, s, inv, d, c2, shots = _adjoint_case()
    J, g = inv.misfit_and_gradient(c2, d)
    oo = orc.OracleOps(P.ops, np.float64)
    adj = adjm.AdjointOps(P.ops, P.dm)
    Q0 = np.zeros((P.K, 4, P.ops.Np))
    srcs = []
    Jr, gr = 0.0, np.zeros_like(g)
    for sh, dd in zip(shots, d):
        src = np.zeros((P.K, P.ops.Np))
        src[sh.elems] = sh.coef
        srcs.append(src)
        gi, Ji = adjm.gradient(oo, P.ops, P.dm, adj, c2, Q0, inv.dt, inv.nsteps, inv.rec_elems, inv.rec_B,
                               dd.astype(np.float64), src, sh.f0, sh.amp)
        Jr += Ji
        gr += gi
    assert abs(J - Jr) < 1e-5 * Jr
    assert rel_l2(g, gr) < 1e-4

    def J64(c):
        return sum(adjm.misfit(adjm.forward(oo, P.ops, P.dm, c, Q0, inv.dt, inv.nsteps, inv.rec_elems,
                                            inv.rec_B, src, sh.f0, sh.amp)[1], dd.astype(np.float64))
                   for sh, dd, src in zip(shots, d, srcs))

    rng = np.random.default_rng(17)
    eps = 1e-4
    scale = np.linalg.norm(g) / math.sqrt(g.size)
    errs = []
    for _ in range(20):
        v = rng.standard_normal(c2.shape)
        fd = (J64(c2 + eps * v) - J64(c2 - eps * v)) / (2 * eps)
        ad = float(np.sum(g * v))
        # FP32 gradient against FP64 central differences: the error is measured against the
        # size of a directional derivative along a random direction, |g| |v| / sqrt(n)
        errs.append(abs(fd - ad) / (scale * np.linalg.norm(v)))
    say(f"C5 FD errors max {max(errs):.2e}")
    assert max(errs) < 1e-4, errs   # measured 1.0e-5 on MI355X
    s.close()
