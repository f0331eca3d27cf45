"""GPU parity: each HIP kernel and the step loop against the oracle (oracle/bbw_oracle.c)
on the same seeded inputs.  Tolerances (FP32 arithmetic, summation order differs from the
oracle's): single kernels 2e-5 relative to the output's max-norm against the FP32 oracle
and 1e-5 against the FP64 oracle evaluated on the same FP32 inputs; whole runs 1e-5
relative L2 against the FP64 oracle."""
import numpy as np
import pytest

from bbwadg_amd import _lib
from bbwadg_amd.solver import WaveSolver, MODE_FUSED, MODE_PLAIN
import oracle as orc
from cases import Problem, rel, rel_l2

pytestmark = pytest.mark.gpu


def make(N, n, seed=0, c2="fourier", **kw):
    P = Problem(N, n, seed=seed, c2=c2)
    s = WaveSolver(P.mesh, N, ops=P.ops, dmesh=P.dm, **kw)
    s.InitWADG_subelem(P.c2q)
    return P, s


def oracle_ops(P, dtype):
    return orc.OracleOps(P.ops, dtype)


@pytest.mark.parametrize("impl", [_lib.PLAIN_FLAT, _lib.PLAIN_TILED])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_volume(N, impl):
    print(f"case volume N={N} impl={impl}", flush=True)
    P, s = make(N, 3)
    s.set_option(_lib.OPT_PLAIN_IMPL, impl)
    Q = P.random_state()
    s.WaveSetData3d(Q)
    s.volume()
    got = s.get_rhs()
    for dt, tol in ((np.float32, 2e-5), (np.float64, 1e-5)):
        ref = np.zeros(Q.shape, dtype=dt)
        orc.volume(oracle_ops(P, dt), P.K, P.dm.vgeo.astype(dt), Q.astype(dt), ref)
        assert rel(got, ref) < tol


@pytest.mark.parametrize("impl", [_lib.PLAIN_FLAT, _lib.PLAIN_TILED])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_surface(N, impl):
    print(f"case surface N={N} impl={impl}", flush=True)
    P, s = make(N, 3)
    s.set_option(_lib.OPT_PLAIN_IMPL, impl)
    Q = P.random_state()
    R0 = P.random_state()
    s.WaveSetData3d(Q)
    s.set_rhs(R0)
    s.surface()
    got = s.get_rhs()
    for dt, tol in ((np.float32, 2e-5), (np.float64, 1e-5)):
        ref = R0.astype(dt).copy()
        orc.surface(oracle_ops(P, dt), P.K, P.dm.fgeo.astype(dt), P.dm.vmapP, Q.astype(dt), ref)
        assert rel(got, ref) < tol


@pytest.mark.parametrize("impl", [_lib.PLAIN_FLAT, _lib.PLAIN_TILED])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_update_wadg(N, impl):
    print(f"case update_wadg N={N} impl={impl}", flush=True)
    P, s = make(N, 3)
    s.set_option(_lib.OPT_PLAIN_IMPL, impl)
    Q, R, res = P.random_state(), P.random_state(), P.random_state()
    s.WaveSetData3d(Q)
    s.set_rhs(R)
    s.set_res(res)
    fa, fb, fdt = np.float32(P.ops.rk4a[2]), np.float32(P.ops.rk4b[2]), np.float32(1e-2)
    s.update_wadg(fa, fb, fdt)
    gQ, gres = s.WaveGetData3d(), s.get_res()
    for dt, tol in ((np.float32, 2e-5), (np.float64, 2e-5)):
        rQ, rres = Q.astype(dt).copy(), res.astype(dt).copy()
        orc.update_wadg(oracle_ops(P, dt), P.K, P.c2q, fa, fb, fdt, R.astype(dt), rres, rQ)
        assert rel(gres, rres) < tol
        assert rel(gQ, rQ) < tol


def test_update_plain():
    P, s = make(3, 2)
    Q, R, res = P.random_state(), P.random_state(), P.random_state()
    s.WaveSetData3d(Q); s.set_rhs(R); s.set_res(res)
    s.update(np.float32(-0.5), np.float32(0.25), np.float32(0.1))
    rres = np.float32(-0.5) * res + np.float32(0.1) * R
    rQ = Q + np.float32(0.25) * rres
    assert rel(s.get_res(), rres) < 1e-6 and rel(s.WaveGetData3d(), rQ) < 1e-6


@pytest.mark.parametrize("impl", [_lib.FUSED_VECTOR, _lib.FUSED_MFMA])
@pytest.mark.parametrize("N,proj", [(1, False), (3, False), (4, True), (5, False), (5, True), (6, False), (2, True)])
def test_fused_stage(N, proj, impl):
    print(f"case fused N={N} proj={proj} impl={impl}", flush=True)
    P, s = make(N, 3)
    s.set_option(_lib.OPT_FUSED_IMPL, impl)
    Q, res = P.random_state(), P.random_state()
    s.WaveSetData3d(Q); s.set_res(res)
    fa, fb, fdt = np.float32(P.ops.rk4a[1]), np.float32(P.ops.rk4b[1]), np.float32(1e-2)
    s.stage_fused(fa, fb, fdt, project_velocity=proj)
    gQ, gres = s.WaveGetData3d(), s.get_res()
    dt = np.float64
    oo = oracle_ops(P, dt)
    rhs = np.zeros(Q.shape)
    rQ, rres = Q.astype(dt).copy(), res.astype(dt).copy()
    orc.volume(oo, P.K, P.dm.vgeo, Q.astype(dt), rhs)
    orc.surface(oo, P.K, P.dm.fgeo, P.dm.vmapP, Q.astype(dt), rhs)
    orc.update_wadg(oo, P.K, P.c2q, fa, fb, fdt, rhs, rres, rQ, nproj=4 if proj else 1)
    assert rel(gres, rres) < 2e-5
    assert rel(gQ, rQ) < 2e-5


@pytest.mark.parametrize("mode", [MODE_PLAIN, MODE_FUSED])
def test_cavity_run_matches_fp64_oracle(mode):
    """CPU-runnable configuration shape (N=3, c2=1, cavity mode, dt = 0.5h/N^2), 40 steps."""
    print(f"case cavity mode={mode}", flush=True)
    P, s = make(3, 4, c2="one")
    Q0 = P.cavity_state()
    s.WaveSetData3d(Q0)
    dt, nsteps = P.dt_cfl(), 40
    s.step(nsteps, dt, mode=mode, graph=True)
    got = s.WaveGetData3d()
    oo = oracle_ops(P, np.float64)
    ref = Q0.astype(np.float64).copy()
    orc.rk_run(oo, P.K, P.dm.vgeo, P.dm.fgeo, P.dm.vmapP, P.c2q, ref, dt, nsteps)
    assert rel_l2(got, ref) < 1e-5
    # and the plain path against the FP32 oracle, stage by stage the same arithmetic order class
    if mode == MODE_PLAIN:
        oo32 = oracle_ops(P, np.float32)
        r32 = Q0.copy()
        orc.rk_run(oo32, P.K, P.dm.vgeo, P.dm.fgeo, P.dm.vmapP, P.c2q, r32, dt, nsteps)
        assert rel_l2(got, r32) < 1e-5


def test_heterogeneous_run_fused_vs_plain_vs_oracle():
    print("case heterogeneous N=5", flush=True)
    P, s = make(5, 2)
    Q0 = P.random_state()
    dt, nsteps = P.dt_cfl() * 0.5, 20
    s.WaveSetData3d(Q0)
    s.step(nsteps, dt, mode=MODE_PLAIN, graph=False)
    plain = s.WaveGetData3d()
    s.WaveSetData3d(Q0); s.set_res(np.zeros_like(Q0))
    s.step(nsteps, dt, mode=MODE_FUSED, project_velocity=True, graph=True)
    fused_p = s.WaveGetData3d()
    s.WaveSetData3d(Q0); s.set_res(np.zeros_like(Q0))
    s.step(nsteps, dt, mode=MODE_FUSED, graph=False)
    fused = s.WaveGetData3d()
    s.set_option(_lib.OPT_FUSED_IMPL, _lib.FUSED_VECTOR)
    s.WaveSetData3d(Q0); s.set_res(np.zeros_like(Q0))
    s.step(nsteps, dt, mode=MODE_FUSED, graph=True)
    fused_v = s.WaveGetData3d()
    oo = oracle_ops(P, np.float64)
    ref = Q0.astype(np.float64).copy()
    orc.rk_run(oo, P.K, P.dm.vgeo, P.dm.fgeo, P.dm.vmapP, P.c2q, ref, dt, nsteps)
    for got in (plain, fused_p, fused, fused_v):
        assert rel_l2(got, ref) < 1e-5


def test_graph_replay_is_bitwise_eager():
    P, s = make(4, 2)
    Q0 = P.random_state()
    dt = P.dt_cfl()
    outs = []
    for graph in (False, True, True):
        s.WaveSetData3d(Q0); s.set_res(np.zeros_like(Q0))
        s.step(30, dt, mode=MODE_FUSED, graph=graph)
        outs.append(s.WaveGetData3d())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


def test_energy_matches_oracle():
    P, s = make(4, 3)
    Q = P.random_state()
    s.WaveSetData3d(Q)
    E = s.energy()
    Eref = orc.energy_quadrature(P.ops, P.dm.J, P.c2q.astype(np.float32), Q)
    assert abs(E - Eref) < 1e-10 * abs(Eref)


def test_source_and_receivers_match_oracle():
    from bbwadg_amd import media
    print("case source+receivers", flush=True)
    P, s = make(3, 3)
    elems, coef = media.gaussian_source(P.mesh, P.ops, (0.05, -0.1, 0.02), a=20.0)
    f0, amp = 10.0, 3.0
    s.set_source(elems, coef, f0=f0, amp=amp)
    pts = media.surface_receivers(1, 7)
    relems, B = media.receivers(P.mesh, P.ops, pts)
    nsteps = 30
    s.set_receivers(relems, B, nsteps)
    Q0 = np.zeros((P.K, 4, P.ops.Np), dtype=np.float32)
    s.WaveSetData3d(Q0)
    dt = P.dt_cfl()
    s.step(nsteps, dt, mode=MODE_FUSED, project_velocity=True, graph=True)
    got = s.WaveGetData3d()
    tr = s.get_traces(nsteps)
    src = np.zeros((P.K, P.ops.Np))
    src[elems] = coef
    ref = Q0.astype(np.float64).copy()
    rec = []
    orc.rk_run(orc.OracleOps(P.ops, np.float64), P.K, P.dm.vgeo, P.dm.fgeo, P.dm.vmapP, P.c2q,
               ref, dt, nsteps, src=src, f0=f0, amp=amp,
               record=lambda st, Qs: rec.append(np.einsum("rj,rj->r", B, Qs[relems, 0])))
    assert rel_l2(got, ref) < 1e-5
    assert rel_l2(tr, np.array(rec)) < 1e-5


def test_tune_single_restores_state_and_applies_choice():
    """The single-shot autotuner times every fused and plain variant, applies the fastest and
    leaves the state as it was."""
    from bbwadg_amd import autotune
    print("case tune_single", flush=True)
    P, s = make(3, 3)
    Q, res = P.random_state(), P.random_state()
    s.WaveSetData3d(Q)
    s.set_res(res)
    r = autotune.tune_single(s, P.dt_cfl(), nsteps=2, kblk_vector=(4,), kblk_plain=(2,))
    assert {t[0] for t in r["plain_trials"]} == {"tiled", 2}
    assert len(r["fused_trials"]) == 2
    assert np.array_equal(s.WaveGetData3d(), Q) and np.array_equal(s.get_res(), res)


@pytest.mark.parametrize("kind", ["fused", "plain_tiled", "plain_flat"])
def test_first_stage_does_not_depend_on_res(kind):
    """a_0 = 0: the first LSRK45 stage computes res = dt rhs whatever res held (the kernels skip
    reading it), against the FP64 oracle that multiplies the old res by 0."""
    print(f"case first stage {kind}", flush=True)
    P, s = make(5, 3)
    assert P.ops.rk4a[0] == 0.0
    Q, res = P.random_state(), 1e3 * P.random_state()
    fa, fb, fdt = np.float32(P.ops.rk4a[0]), np.float32(P.ops.rk4b[0]), np.float32(1e-2)
    s.WaveSetData3d(Q); s.set_res(res)
    oo = oracle_ops(P, np.float64)
    rhs = np.zeros(Q.shape)
    orc.volume(oo, P.K, P.dm.vgeo, Q.astype(np.float64), rhs)
    orc.surface(oo, P.K, P.dm.fgeo, P.dm.vmapP, Q.astype(np.float64), rhs)
    if kind == "fused":
        s.stage_fused(fa, fb, fdt)
        nproj = 1
    else:
        s.set_option(_lib.OPT_PLAIN_IMPL, _lib.PLAIN_TILED if kind == "plain_tiled" else _lib.PLAIN_FLAT)
        s.set_rhs(rhs.astype(np.float32))
        s.update_wadg(fa, fb, fdt)
        nproj = 4
    rQ, rres = Q.astype(np.float64).copy(), res.astype(np.float64).copy()
    orc.update_wadg(oo, P.K, P.c2q, fa, fb, fdt, rhs, rres, rQ, nproj=nproj)
    assert rel(s.get_res(), rres) < 2e-5
    assert rel(s.WaveGetData3d(), rQ) < 2e-5


@pytest.mark.parametrize("graph,nsteps", [(False, 7), (True, 10), (True, 23)])
def test_run_leaves_reference_residual(graph, nsteps):
    """Inside a multi-step call the fused stage drops the dead store of res after the last stage
    of a step; the residual left at the end of the call is still the reference's."""
    print(f"case residual after run graph={graph} nsteps={nsteps}", flush=True)
    P, s = make(4, 2)
    Q0 = P.random_state()
    dt = P.dt_cfl()
    s.WaveSetData3d(Q0); s.set_res(P.random_state())
    s.step(nsteps, dt, mode=MODE_FUSED, graph=graph)
    got_Q, got_res = s.WaveGetData3d(), s.get_res()
    # one step per call stores res after every step: same arithmetic, so bitwise equal
    s.WaveSetData3d(Q0); s.set_res(P.random_state())
    for _ in range(nsteps):
        s.step(1, dt, mode=MODE_FUSED, graph=False)
    assert np.array_equal(s.WaveGetData3d(), got_Q) and np.array_equal(s.get_res(), got_res)
    ref = Q0.astype(np.float64).copy()
    _, rres = orc.rk_run(oracle_ops(P, np.float64), P.K, P.dm.vgeo, P.dm.fgeo, P.dm.vmapP, P.c2q,
                         ref, dt, nsteps, nproj=1)
    assert rel_l2(got_Q, ref) < 1e-5
    # res is an increment (dt-scaled right-hand side): FP32 rounding relative to it is larger
    # than relative to the state; a stale residual would be off by O(1)
    assert rel_l2(got_res, rres) < 1e-4


@pytest.mark.parametrize("mode", [MODE_PLAIN, MODE_FUSED])
def test_unstructured_reference_mesh(mode):
    """The reference's cube3.msh (196 tets, unstructured, every face orientation and partial last
    tile): cavity start, Fourier c2, N = 4, 25 steps against the FP64 oracle; the GMSH view of the
    result holds the nodal values."""
    import os
    from bbwadg_amd import media, mesh as M, operators as O
    from cases import cavity
    print(f"case unstructured cube3.msh mode={mode}", flush=True)
    N = 4
    m = M.read_gmsh(os.path.join(os.path.dirname(__file__), "golden", "cube3.msh"))
    M.face_pair(m)
    ops = O.build_ref_ops(N)
    dm = M.device_arrays(m, ops, 4)
    xq, yq, zq = media.cubature_points(m, ops)
    c2q = media.c2_fourier(2)(xq, yq, zq).astype(np.float32)
    x, y, z = M.affine_points(m, ops.r, ops.s, ops.t)
    Q0 = np.zeros((m.K, 4, ops.Np))
    Q0[:, 0] = (ops.invVB @ cavity(x, y, z, 0.0)).T
    Q0 = Q0.astype(np.float32)
    dt, nsteps = 0.25 / (N ** 2 * dm.FscaleMax), 25
    s = WaveSolver(m, N, ops=ops, dmesh=dm)
    s.InitWADG_subelem(c2q)
    s.WaveSetData3d(Q0)
    s.step(nsteps, dt, mode=mode, graph=True)
    got = s.WaveGetData3d()
    ref = Q0.astype(np.float64).copy()
    orc.rk_run(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, c2q, ref, dt, nsteps)
    assert rel_l2(got, ref) < 1e-5
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"bbw_view_{os.getpid()}_{mode}.msh")
    s.writeVisToGMSH(path, Q=got, iField=0)
    ids, vals, _, _ = M.read_gmsh_node_data(path)
    os.remove(path)
    assert list(ids) == list(m.gmsh_ids)
    assert np.allclose(vals, s.to_nodal(got)[:, 0], rtol=0, atol=1e-6)
    s.close()


@pytest.mark.parametrize("kind", ["fused", "plain_tiled", "plain_flat"])
@pytest.mark.parametrize("N", [3, 5])
def test_wadg_update_against_exact_bb_product(N, kind):
    """Second oracle for the WADG projection: with c2 linear the quadrature update must equal the
    exact BB product projection P_N(c2 p) (oracle/bb_exact.py, the rk_update_BBWADG algorithm)."""
    import bb_exact as X
    print(f"case exact BB-WADG N={N} {kind}", flush=True)
    P = Problem(N, 3, seed=2)
    c2q = 1.0 + 0.3 * P.xq - 0.2 * P.yq + 0.25 * P.zq
    s = WaveSolver(P.mesh, N, ops=P.ops, dmesh=P.dm)
    s.InitWADG_subelem(c2q)
    Q, res = P.random_state(), P.random_state()
    fa, fb, fdt = np.float32(P.ops.rk4a[3]), np.float32(P.ops.rk4b[3]), np.float32(1e-2)
    oo = oracle_ops(P, np.float64)
    rhs = np.zeros(Q.shape)
    orc.volume(oo, P.K, P.dm.vgeo, Q.astype(np.float64), rhs)
    orc.surface(oo, P.K, P.dm.fgeo, P.dm.vmapP, Q.astype(np.float64), rhs)
    s.WaveSetData3d(Q); s.set_res(res)
    if kind == "fused":
        s.stage_fused(fa, fb, fdt)
    else:
        s.set_option(_lib.OPT_PLAIN_IMPL, _lib.PLAIN_TILED if kind == "plain_tiled" else _lib.PLAIN_FLAT)
        s.set_rhs(rhs.astype(np.float32))
        s.update_wadg(fa, fb, fdt)
    rQ, rres = X.update_bbwadg(P.ops, X.c2_bb(P.ops, s.c2q), fa, fb, fdt, rhs,
                               res.astype(np.float64), Q.astype(np.float64))
    assert rel(s.get_res(), rres) < 2e-5
    assert rel(s.WaveGetData3d(), rQ) < 2e-5
    s.close()


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_exact_bbwadg_update(N):
    """bbw_update_bbwadg (rk_update_BBWADG) against the exact BB-WADG oracle: rhs_p <- P_N(c2 rhs_p)
    with the Fourier c2 in BB coefficients; the velocity rows are not touched."""
    import bb_exact as X
    print(f"case exact BB-WADG kernel N={N}", flush=True)
    P, s = make(N, 3)
    s.InitBBWADG(P.c2q)
    R = P.random_state()
    s.set_rhs(R)
    s.update_bbwadg()
    got = s.get_rhs()
    ref = X.wadg_exact(P.ops, s.c2bb.astype(np.float64), R[:, 0].astype(np.float64))
    assert rel(got[:, 0], ref) < 2e-5
    assert np.array_equal(got[:, 1:], R[:, 1:])
    s.close()


def test_receivers_set_after_graph_steps():
    """Graph steps without receivers, then receivers set on the same solver: the graph's window
    grows with the receiver count (its rows are written by captured nodes) and the traces equal
    an eager run's."""
    from bbwadg_amd import media
    print("case receivers after graph steps", flush=True)
    P, s = make(3, 3)
    Q0 = P.random_state()
    dt = P.dt_cfl()
    s.WaveSetData3d(Q0); s.set_res(np.zeros_like(Q0))
    s.step(10, dt, mode=MODE_FUSED, graph=True)          # captures a graph with no receiver rows
    relems, B = media.receivers(P.mesh, P.ops, media.surface_receivers(5, 40))
    traces = []
    for graph in (True, False):
        s.set_receivers(relems, B, 20)
        s.WaveSetData3d(Q0); s.set_res(np.zeros_like(Q0))
        s.step(20, dt, mode=MODE_FUSED, graph=graph)
        traces.append(s.get_traces(20))
    assert np.array_equal(traces[0], traces[1])
    assert np.abs(traces[0]).max() > 0
    s.close()


def test_prepared_graph_steps_equal_eager_steps():
    """bbw_step's prepare flag captures and uploads the graph (and sizes the stage-time buffers)
    without stepping; the steps that follow equal eager steps bitwise, with a source."""
    from bbwadg_amd import media
    print("case prepared graph", flush=True)
    P, s = make(3, 3)
    el, co = media.gaussian_source(P.mesh, P.ops, (0.1, 0.0, -0.1), a=10.0)
    s.set_source(el, co, f0=8.0, amp=2.0)
    Q0 = P.random_state()
    dt = P.dt_cfl()
    mid, end = [], []
    for graph in (False, True):
        s.WaveSetData3d(Q0); s.set_res(np.zeros_like(Q0))
        s.time = 0.0
        s.step(3, dt, mode=MODE_FUSED, graph=graph)
        mid.append(s.WaveGetData3d())
        if graph:
            s.prepare_graph(20, dt, mode=MODE_FUSED)
            assert np.array_equal(s.WaveGetData3d(), mid[-1])   # prepare does not step
        s.step(20, dt, mode=MODE_FUSED, graph=graph)
        end.append(s.WaveGetData3d())
    assert np.array_equal(mid[0], mid[1]) and np.array_equal(end[0], end[1])
    s.close()
