"""Benchmark: DOF-stage updates/s of the BB-WADG acoustic solver on one MI355X.

Workload (default): heterogeneous acoustic single shot, cube [-0.5,0.5]^3 split into
48^3 x 6 = 663,552 tets, N = 5 (Np = 56, Nq = 95), seeded Fourier c^2 in [0.6, 1.4] at the
cubature points, Ricker point-like source f0 = 10 at a seeded location, dt = 0.5 h / N^2,
FP32.  A step is one LSRK45 step (5 stages) over the whole mesh, inputs resident in HBM.

Reported: the fused path (one kernel per stage, HIP-graph replay) as ``value``; the plain
three-launch path of the reference (volume, surface, WADG update) beside it; the
roofline of the fused stage kernel; the adjoint-gradient / forward ratio; the batched
multi-shot run; the elastic configuration's stage rate and its adjoint-gradient / forward
ratio; the curvilinear run of the reference's sphere mesh (49 748 tets, N = 5); the CPU oracle
timed on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--mesh 48] [--N 5]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GRAPH_STEPS = 10   # steps per captured HIP graph in bbw_step


def stage_bytes(N, Nf=4, Nm=1):
    """Compulsory HBM bytes per element-stage of the fused stage: Q read + write, res read +
    write (4 Nf Np), neighbour traces (Nf 4 Nfp), geometry (9 + 16), connectivity (4 Nfp),
    media at cubature (Nm Nq); 4 bytes per word."""
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    Nfp = (N + 1) * (N + 2) // 2
    Nq = {1: 6, 2: 14, 3: 31, 4: 57, 5: 95, 6: 146}[N]
    return 4 * (4 * Nf * Np + Nf * 4 * Nfp + 25 + 4 * Nfp + Nm * Nq)


def run_bytes(N, nsteps, graph_steps=10, Nf=4, Nm=1):
    """Average compulsory bytes per element-stage over one bbw_step call of nsteps steps: the
    first LSRK45 stage (a_0 = 0) reads no res, and the last stage of a step that another step of
    the same captured graph (of graph_steps steps) or call follows stores no res."""
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    row = 4 * Nf * Np
    full = stage_bytes(N, Nf, Nm)
    chunks = [graph_steps] * (nsteps // graph_steps) if nsteps >= graph_steps else []
    rest = nsteps - sum(chunks)
    if rest:
        chunks.append(rest)
    skipped_writes = sum(c - 1 for c in chunks)
    return full - row * (nsteps + skipped_writes) / (5.0 * nsteps)


def build_problem(n, N, seed=1, order="morton"):
    from bbwadg_amd import media, mesh as M, operators as O
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n, order=order)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 4)
    xq, yq, zq = media.cubature_points(m, ops)
    c2q = media.c2_fourier(seed)(xq, yq, zq).astype(np.float32)
    rng = np.random.default_rng(seed + 1)
    x0 = rng.uniform(-0.3, 0.3, 3)
    src = media.gaussian_source(m, ops, x0, a=100.0)
    return ops, m, dm, c2q, src, x0


def cpu_baseline(ops, dm, c2q, K, sample_elems, threads):
    """Oracle FP32 (OpenMP) over the first ``sample_elems`` elements for one LSRK step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import ctypes
    import oracle as orc
    orc.select_build("liborc_fast.so")   # -O3, AVX2/FMA build of the same C sources
    orc.lib()
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(threads))
    oo = orc.OracleOps(ops, np.float32)
    Ks = min(sample_elems, K)
    rng = np.random.default_rng(0)
    Q = rng.uniform(-1, 1, (K, 4, ops.Np)).astype(np.float32)
    res = np.zeros_like(Q)
    rhs = np.zeros_like(Q)
    vgeo = np.ascontiguousarray(dm.vgeo)
    fgeo = np.ascontiguousarray(dm.fgeo)
    t0 = time.perf_counter()
    for i in range(5):
        orc.volume(oo, Ks, vgeo, Q, rhs)
        orc.surface(oo, Ks, fgeo, dm.vmapP, Q, rhs)
        orc.update_wadg(oo, Ks, c2q, np.float32(ops.rk4a[i]), np.float32(ops.rk4b[i]),
                        np.float32(1e-4), rhs, res, Q)
    el = time.perf_counter() - t0
    return Ks * ops.Np * 4 * 5 / el, el, Ks


def _host_cpu():
    """CPU model and logical CPU count of this host (for the CPU baseline line)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"


def adjoint_ratio(nsteps, mesh_n=32, N=5, ckpt=50):
    """Time of one adjoint gradient (forward with checkpoints + segment recompute + reverse
    stages) over the time of one forward misfit evaluation, on the adjoint-gradient
    configuration's shape (32^3 x 6 tets, N = 5, 256 surface receivers, c2 = 1 start model,
    seeded Fourier true model, one Ricker shot)."""
    from bbwadg_amd import media, mesh as M, operators as O
    from bbwadg_amd.adjoint import AcousticInversion, Shot
    from bbwadg_amd.solver import WaveSolver
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(mesh_n, order="morton")
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 4)
    s = WaveSolver(m, N, ops=ops, dmesh=dm)
    xq, yq, zq = media.cubature_points(m, ops)
    c2_true = media.c2_fourier(3)(xq, yq, zq)
    el, co = media.gaussian_source(m, ops, (0.05, -0.08, 0.1), a=100.0)
    dt = 0.5 / mesh_n / N ** 2
    inv = AcousticInversion(s, [Shot(el, co, f0=10.0, amp=1.0)], media.surface_receivers(4, 256),
                            nsteps, dt, ckpt_every=ckpt)
    d = inv.observe(c2_true)
    c2 = np.ones_like(c2_true)
    inv.misfit(c2, d)  # warm-up (graph capture, allocations)
    inv.misfit_and_gradient(c2, d)
    t0 = time.perf_counter()
    Jf = inv.misfit(c2, d)
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    Jg, g = inv.misfit_and_gradient(c2, d)
    tg = time.perf_counter() - t0
    s.close()
    return {"ratio": tg / tf, "forward_s": tf, "gradient_s": tg, "steps": nsteps, "K": m.K,
            "ckpt_every": ckpt, "misfit": Jf, "grad_norm": float(np.linalg.norm(g)),
            "misfit_agree": abs(Jf - Jg) / max(Jf, 1e-300)}


def elastic_adjoint_ratio(nsteps, mesh_n=32, N=4, ckpt=50):
    """The elastic counterpart of adjoint_ratio on the elastic configuration's shape (32^3 x 6
    tets, N = 4, 9 fields): gradient in (1/rho, lambda, mu) of the misfit of v_z at 256 surface
    receivers, start model the means of the seeded Lame/density fields, one explosive
    moment-tensor Ricker shot (f0 = 40), the reference's elastic time step."""
    from bbwadg_amd import media, mesh as M, operators as O
    from bbwadg_amd.adjoint import ElasticInversion, Shot
    from bbwadg_amd.solver import WaveSolver
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(mesh_n, order="morton")
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    s = WaveSolver(m, N, nfields=9, ops=ops, dmesh=dm)
    xq, yq, zq = media.cubature_points(m, ops)
    lam, mu, rho = media.lame_fourier(1)
    m_true = (1.0 / rho(xq, yq, zq), lam(xq, yq, zq), mu(xq, yq, zq))
    s.InitElastic(*m_true)
    # the shot 0.05 below the receiver plane, so that the 200-step window records its arrivals
    el, co = media.gaussian_source(m, ops, (0.05, -0.08, 0.45), a=100.0)
    w = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0], dtype=np.float64)
    inv = ElasticInversion(s, [Shot(el, co, f0=40.0, amp=1.0, weights=w)],
                           media.surface_receivers(4, 256), nsteps, s.dt_ref, ckpt_every=ckpt,
                           rec_field=2)
    d = inv.observe(m_true)
    m0 = tuple(np.full_like(a, float(np.mean(a))) for a in m_true)
    inv.misfit(m0, d)  # warm-up (graph capture, allocations)
    inv.misfit_and_gradient(m0, d)
    t0 = time.perf_counter()
    Jf = inv.misfit(m0, d)
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    Jg, g = inv.misfit_and_gradient(m0, d)
    tg = time.perf_counter() - t0
    s.close()
    return {"ratio": tg / tf, "forward_s": tf, "gradient_s": tg, "steps": nsteps, "K": m.K, "N": N,
            "ckpt_every": ckpt, "misfit": Jf, "grad_norm": [float(np.linalg.norm(x)) for x in g],
            "misfit_agree": abs(Jf - Jg) / max(Jf, 1e-300)}


def elastic_throughput(nsteps, mesh_n=32, N=4):
    """Elastic velocity-stress configuration (32^3 x 6 tets, N = 4, 9 fields, seeded Lame and
    density fields, random start, the reference's elastic time step): the fused stage (one
    launch per stage) and the three-launch stage, nsteps steps each from HIP graphs, timed with
    HIP events on the solver's stream; the roofline fraction prices the fused rate at the
    compulsory bytes of an elastic element-stage."""
    import torch
    from bbwadg_amd import _lib
    from bbwadg_amd.drive import setup_elastic
    from bbwadg_amd.solver import MODE_FUSED, MODE_PLAIN
    s = setup_elastic(mesh_n, N)
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    dt = s.dt_ref   # the reference's elastic time step estimate
    Q0 = s.WaveGetData3d()
    out = {}
    form0 = _lib.ELAS_LIFT_CUBATURE if N <= 4 else _lib.ELAS_LIFT_NODES   # the library's default
    for name, mode, form in (("fused", MODE_FUSED, form0), ("nodal", MODE_FUSED, _lib.ELAS_LIFT_NODES),
                             ("plain", MODE_PLAIN, form0)):
        s.set_option(_lib.OPT_ELAS_FUSED, form)
        s.WaveSetData3d(Q0)
        s.set_res(np.zeros_like(Q0))
        s.step(2, dt, mode=mode)
        s.prepare_graph(nsteps, dt, mode=mode)   # graph capture and instantiation
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        s.step(nsteps, dt, mode=mode)
        e1.record(stream)
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1)
        if not np.all(np.isfinite(s.WaveGetData3d())):
            raise RuntimeError(f"non-finite elastic state after the timed {name} run")
    dofs = s.K * s.ops.Np * 9 * 5 * nsteps
    K = s.K
    s.close()
    ms_stage = out["fused"] / (5 * nsteps)
    Bst = run_bytes(N, nsteps, Nf=9, Nm=3)
    return {"workload": f"elastic velocity-stress, {mesh_n}^3x6={K} tets, N={N}, 9 fields, seeded "
                        "Lame/density fields, fused stage + HIP graph",
            "value": dofs / (out["fused"] / 1e3), "unit": "DOF-stage/s", "steps": nsteps, "dt": dt,
            "ms_per_stage": ms_stage,
            "fused_form": "lift folded into the cubature product" if form0 == _lib.ELAS_LIFT_CUBATURE
                          else "lift at the nodes",
            "nodal_form_ms_per_stage": out["nodal"] / (5 * nsteps),
            "plain_value": dofs / (out["plain"] / 1e3), "plain_ms_per_stage": out["plain"] / (5 * nsteps),
            "fused_over_plain": out["plain"] / out["fused"],
            "bytes_per_element_stage": Bst, "bytes_per_element_stage_full": stage_bytes(N, 9, 3),
            "achieved_GBs": Bst * K / (ms_stage / 1e3) / 1e9,
            "frac_of_hbm_peak": Bst * K / (ms_stage / 1e3) / 1e9 / HBM_PEAK_GBS,
            "frac_of_stage_roofline": stage_bytes(N, 9, 3) * K / (ms_stage / 1e3) / 1e9 / HBM_PEAK_GBS}


def curved_bytes(N, curved):
    """Compulsory HBM bytes per element-stage of the curvilinear stage (right-hand side kernel
    plus WADG update): state read (4 Np) and rhs written and read again (2 x 4 Np), res and Q
    read and written (4 x 4 Np), neighbour traces (4 x 4 Nfp), vmapP (4 Nfp), c2 and 1/J at the
    Nq points (2 Nq), and the element's geometry: at the cubature points for a curved element
    (vgeoq 9 Nq, Jq Nq, fgeoq 16 Nfq), constants for a planar one (9 + 1 + 16); 4 bytes a word."""
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    Nfp = (N + 1) * (N + 2) // 2
    Nq = {1: 6, 2: 14, 3: 31, 4: 57, 5: 95, 6: 146}[N]
    Nfq = {1: 6, 2: 7, 3: 15, 4: 19, 5: 28, 6: 40}[N]
    geo = 10 * Nq + 16 * Nfq if curved else 26
    return 4 * (4 * Np + 2 * 4 * Np + 4 * 4 * Np + 4 * 4 * Nfp + 4 * Nfp + 2 * Nq + geo)


def curved_throughput(nsteps, N=5):
    """The reference's curvilinear run (cpp/main.cpp with runCurvedBlkTimings' mesh): the
    49 748-tet ball mesh sphere49748.msh, Gordon-Hall sphere map, Bessel standing mode, c^2 = 1,
    the reference's time step; nsteps RK_step_WADG steps from a HIP graph, HIP events on the
    solver's stream."""
    import torch
    from bbwadg_amd import curved, mesh as M
    from bbwadg_amd.solver import CurvedWaveSolver
    m = M.read_gmsh(os.path.join(ROOT, "bbwadg_amd", "data", "meshes", "sphere49748.msh"))
    M.face_pair(m)
    s = CurvedWaveSolver(m, N)
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    dt = s.dt_ref
    s.WaveSetData3d(s.WaveSetU0(curved.bessel))
    s.step(2, dt)
    s.prepare_graph(nsteps, dt)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    s.step(nsteps, dt)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    T = s.time
    err = s.compute_error(T, curved.bessel)
    if not np.isfinite(err[0]):
        raise RuntimeError("non-finite curved state after the timed run")
    K, Kc = s.K, len(s.cm.Klist_curved)
    dofs = K * s.ops.Np * 4 * 5 * nsteps
    s.close()
    ms_stage = ms / (5 * nsteps)
    B = (Kc * curved_bytes(N, True) + (K - Kc) * curved_bytes(N, False))
    return {"workload": f"curvilinear acoustic, sphere49748.msh ({K} tets, {Kc} curved), N={N}, "
                        "Gordon-Hall sphere map, Bessel mode, c^2 = 1",
            "value": dofs / (ms / 1e3), "unit": "DOF-stage/s", "steps": nsteps, "dt": dt,
            "ms_per_stage": ms_stage, "bytes_per_stage": B,
            "achieved_GBs": B / (ms_stage / 1e3) / 1e9,
            "frac_of_hbm_peak": B / (ms_stage / 1e3) / 1e9 / HBM_PEAK_GBS,
            "l2_error_vs_exact": err[0], "rel_l2_error_vs_exact": err[1], "time": T}


def batched_throughput(step_scale=1.0, tune_steps=20):
    """Batched configuration (64 shots on 16^3..40^3 x 6 meshes, N = 5, durations 300-1200
    steps): one persistent launch per stage with the autotuned shots-per-tile and
    blocks-per-CU, against the single-shot plain (three launches per stage) and fused paths
    run shot after shot, whose times are the per-mesh single-shot rates applied to every
    shot's DOF-stages."""
    from bbwadg_amd import autotune, operators as O
    from bbwadg_amd.batch import c4_workload
    ops = O.build_ref_ops(5)
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    b, desc = c4_workload(seed=7, step_scale=step_scale, ops=ops, log=log, order="morton")
    best, trials = autotune.tune_batch(b, step_cap=tune_steps)
    log(f"batch autotune: {best} from {trials}")
    est = min(t for _, t in trials) / 1e3 * sum(b.meshes[m][0].K * n for m, n in b.shots) / \
        (tune_steps * sum(b.meshes[m][0].K for m, _ in b.shots))
    log(f"batch run: estimated {est:.0f} s")
    ms = b.run(shots_per_tile=best[0], blocks_per_cu=best[1])
    log(f"batch run {ms:.1f} ms")
    total = b.dof_stages()
    rate = total / (ms / 1e3)
    # single-shot rates per mesh (plain and fused), one representative shot per mesh
    from bbwadg_amd.solver import WaveSolver, MODE_FUSED, MODE_PLAIN
    import torch
    per_mesh = {}
    stream = torch.cuda.Stream()
    for mid, (m, dm) in enumerate(b.meshes):
        s = WaveSolver(m, 5, ops=ops, dmesh=dm)
        s.InitWADG_subelem(1.0)
        s.set_stream(stream.cuda_stream)
        s.WaveSetData3d(np.random.default_rng(mid).uniform(-1, 1, (m.K, 4, ops.Np)).astype(np.float32))
        n = round((m.K / 6) ** (1 / 3))
        dt = 0.5 / n / 25
        r = {}
        for name, mode, nst in (("plain", MODE_PLAIN, 40), ("fused", MODE_FUSED, 80)):
            s.step(GRAPH_STEPS, dt, mode=mode)   # graph capture and instantiation
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            s.step(nst, dt, mode=mode)
            e1.record(stream)
            torch.cuda.synchronize()
            r[name] = m.K * ops.Np * 4 * 5 * nst / (e0.elapsed_time(e1) / 1e3)
        per_mesh[mid] = r
        s.close()
    t_plain = sum(b.meshes[mid][0].K * ops.Np * 20 * nst / per_mesh[mid]["plain"] for mid, nst in b.shots)
    t_fused = sum(b.meshes[mid][0].K * ops.Np * 20 * nst / per_mesh[mid]["fused"] for mid, nst in b.shots)
    b.close()
    return {"workload": desc, "value": rate, "ms": ms, "dof_stages": total,
            "shots_per_tile": best[0], "blocks_per_cu": best[1],
            "tune_trials": [[list(k), v] for k, v in trials],
            "plain_sequential_value": total / t_plain, "fused_sequential_value": total / t_fused,
            "batched_over_plain": rate / (total / t_plain),
            "batched_over_fused_sequential": rate / (total / t_fused),
            "frac_of_stage_roofline": rate * stage_bytes(5) / ops.Np / 4 / (HBM_PEAK_GBS * 1e9),
            "per_mesh_single_shot": {str(k): v for k, v in per_mesh.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mesh", type=int, default=48)
    ap.add_argument("--N", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=200)
    ap.add_argument("--cpu-sample", type=int, default=1 << 30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=800)
    ap.add_argument("--curved-steps", type=int, default=50,
                    help="steps of the curvilinear leg (0 skips it)")
    ap.add_argument("--elastic-steps", type=int, default=50,
                    help="steps of the elastic configuration (0 skips it)")
    ap.add_argument("--elastic-adjoint-steps", type=int, default=200,
                    help="steps of the elastic adjoint leg (0 skips it)")
    ap.add_argument("--batched-scale", type=float, default=1.0,
                    help="duration scale of the batched configuration (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(shots and element partitions shard across GPUs by design, not measured here)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from bbwadg_amd import _lib
    from bbwadg_amd.solver import WaveSolver, MODE_FUSED, MODE_PLAIN

    N, n = args.N, args.mesh
    t_setup = time.perf_counter()
    ops, m, dm, c2q, (selems, scoef), x0 = build_problem(n, N)
    K = m.K
    dt = 0.5 * (1.0 / n) / N ** 2
    s = WaveSolver(m, N, ops=ops, dmesh=dm)
    s.InitWADG_subelem(c2q)
    s.set_source(selems, scoef, f0=10.0, amp=1.0)
    # A stream of our own (default priority) for the solver, so that the HIP events below are
    # recorded on the stream the kernels are launched on.  torch's default stream has handle 0,
    # which bbw_set_stream takes as "the context's own stream".
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    # timed from a non-zero state (seeded, uniform in [-1, 1]) with the source active
    Q0 = np.random.default_rng(11).uniform(-1, 1, (K, 4, ops.Np)).astype(np.float32)
    s.WaveSetData3d(Q0)
    setup_s = time.perf_counter() - t_setup
    print(f"setup {setup_s:.1f} s, K={K}, N={N}, dt={dt:.3e}", file=sys.stderr, flush=True)

    def timed(nsteps, mode, sol=None):
        sol = sol or s
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        sol.step(nsteps, dt, mode=mode, graph=True, t0=0.0)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    # warm-up steps, then the step graph is captured, instantiated and uploaded and the
    # stage-time buffers sized for the timed call, so the timed region holds steps only
    warm = args.warmup
    s.step(warm, dt, mode=MODE_FUSED, graph=True, t0=0.0)
    s.prepare_graph(args.steps, dt, mode=MODE_FUSED)
    ms = timed(args.steps, MODE_FUSED)
    dof_stage = K * ops.Np * 4 * 5
    value = dof_stage * args.steps / (ms / 1e3)
    stage_ms = ms / (5 * args.steps)
    E = s.energy()
    if not np.isfinite(E):
        raise RuntimeError("non-finite energy after the timed run")

    # plain path (the reference's three launches per stage)
    s.WaveSetData3d(Q0)
    s.set_res(np.zeros_like(Q0))
    s.step(warm, dt, mode=MODE_PLAIN, graph=True, t0=0.0)
    s.prepare_graph(args.plain_steps, dt, mode=MODE_PLAIN)
    ms_plain = timed(args.plain_steps, MODE_PLAIN)
    plain_value = dof_stage * args.plain_steps / (ms_plain / 1e3)

    Bst = run_bytes(N, args.steps)
    achieved = Bst * K / (stage_ms / 1e3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_fused_N{N}_n{n}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    s.close()
    # the same workload on the x-fastest cube numbering in the dispatch tile order (the element
    # numbering and tile order of the earlier measurements), for comparison
    ops_x, m_x, dm_x, c2q_x, (se_x, sc_x), _ = build_problem(n, N, order="x")
    sx = WaveSolver(m_x, N, ops=ops_x, dmesh=dm_x)
    sx.InitWADG_subelem(c2q_x)
    sx.set_source(se_x, sc_x, f0=10.0, amp=1.0)
    sx.set_stream(stream.cuda_stream)
    sx.set_option(_lib.OPT_XCD_CHUNK, 0)
    sx.WaveSetData3d(np.random.default_rng(11).uniform(-1, 1, (K, 4, ops.Np)).astype(np.float32))
    xsteps = min(args.steps, 200)
    sx.step(warm, dt, mode=MODE_FUSED, graph=True, t0=0.0)
    sx.prepare_graph(xsteps, dt, mode=MODE_FUSED)
    ms_x = timed(xsteps, MODE_FUSED, sx)
    sx.close()
    del m_x, dm_x, c2q_x
    fused_x = {"steps": xsteps, "ms_per_step": ms_x / xsteps, "value": dof_stage * xsteps / (ms_x / 1e3),
               "element_order": "x-fastest cubes", "tile_order": "dispatch"}
    adj = adjoint_ratio(args.adjoint_steps) if args.adjoint_steps > 0 else None
    elas = elastic_throughput(args.elastic_steps) if args.elastic_steps > 0 else None
    eadj = elastic_adjoint_ratio(args.elastic_adjoint_steps) if args.elastic_adjoint_steps > 0 else None
    curv = curved_throughput(args.curved_steps) if args.curved_steps > 0 else None
    bat = batched_throughput(args.batched_scale) if args.batched_scale > 0 else None

    cpu = None
    if not args.no_cpu:
        threads = max(1, min(os.cpu_count() or 1, 16))
        rate, el, Ks = cpu_baseline(ops, dm, c2q, K, args.cpu_sample, threads)
        rate1, el1, Ks1 = cpu_baseline(ops, dm, c2q, K, max(1, min(args.cpu_sample, K) // 16), 1)
        cpu = {"value": rate, "unit": "DOF-stage/s", "cores": threads, "kind": "port",
               "sample": f"oracle FP32 C restatement (gcc -O3 -march=x86-64-v3, OpenMP), 1 LSRK45 "
                         f"step (5 stages) over the "
                         f"first {Ks} of {K} elements of the same mesh/order, {el:.1f} s",
               "value_1_thread": rate1,
               "sample_1_thread": f"the same on 1 thread over the first {Ks1} elements, {el1:.1f} s",
               "host": _host_cpu()}

    out = {
        "metric": "DOF-stage updates/s N=5 FP32, batched vs plain; adjoint-grad/forward time ratio",
        "value": value,
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Fourier c^2, seeded Ricker source position)",
        "config": {"workload": f"acoustic heterogeneous single shot, {n}^3x6={K} tets, N={N}, "
                               f"{args.steps} steps, fused stage kernel + HIP graph",
                   "K": K, "N": N, "Np": ops.Np, "Nfields": 4, "dt": dt,
                   "element_order": "cubes along the Z-order curve", "tile_order": "runs of 16 tiles per XCD"},
        "fused_x_order": fused_x,
        "plain": {"value": plain_value, "steps": args.plain_steps,
                  "ms_per_step": ms_plain / args.plain_steps},
        "fused_over_plain": value / plain_value,
        "frac_of_stage_roofline": value * stage_bytes(N) / ops.Np / 4 / (HBM_PEAK_GBS * 1e9),
        "adjoint_over_forward": adj["ratio"] if adj else None,
        "adjoint": adj,
        "batched": bat,
        "elastic": elas,
        "elastic_adjoint_over_forward": eadj["ratio"] if eadj else None,
        "elastic_adjoint": eadj,
        "curved": curv,
        "energy_after": E,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_fused_mfma<5,false>", "bytes_per_element_stage": Bst,
                     "bytes_per_element_stage_full": stage_bytes(N),
                     "avg_launch_ms": stage_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
