"""Run the acoustic hot path on one GPU for profiling or timing (the analogue of the
reference's ``runBlkTimings`` / ``profileKernels`` runs, ``cpp/runBlkTimings:4-24``).

    python -m bbwadg_amd.drive --mesh 32 --N 5 --steps 10 --mode fused --impl mfma

Builds a seeded cube mesh (cubes along the Z-order curve unless ``--order x``) with Fourier c^2 (or, with ``--elastic``, seeded Lame fields and
9 fields), runs ``--steps`` LSRK45 steps in the chosen mode and prints one
line with the time per stage.
"""
import argparse
import time

import numpy as np

from . import _lib, media, mesh as M, operators as O
from .solver import WaveSolver, MODE_FUSED, MODE_PLAIN


def setup(n, N, seed=1, order="morton", **kw):
    # N = 8, 9: the collapsed form of the tensor-product WADG set (the reference's own set
    # diverges with a variable c^2; the kernels and their cost are the same)
    ops = O.build_ref_ops(N, tp_rule="collapsed")
    m = M.cube_mesh(n, order=order)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 4)
    xq, yq, zq = media.cubature_points(m, ops)
    c2q = media.c2_fourier(seed)(xq, yq, zq).astype(np.float32)
    s = WaveSolver(m, N, ops=ops, dmesh=dm, **kw)
    s.InitWADG_subelem(c2q)
    rng = np.random.default_rng(seed)
    s.WaveSetData3d(rng.uniform(-1, 1, (m.K, 4, ops.Np)).astype(np.float32))
    return s


def setup_elastic(n, N, seed=1, order="morton"):
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n, order=order)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    xq, yq, zq = media.cubature_points(m, ops)
    lam, mu, rho = media.lame_fourier(seed)
    s = WaveSolver(m, N, nfields=9, ops=ops, dmesh=dm)
    s.InitElastic(1.0 / rho(xq, yq, zq), lam(xq, yq, zq), mu(xq, yq, zq))
    rng = np.random.default_rng(seed)
    s.WaveSetData3d(rng.uniform(-1, 1, (m.K, 9, ops.Np)).astype(np.float32))
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mesh", type=int, default=32)
    ap.add_argument("--N", type=int, default=5)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--mode", choices=["fused", "plain"], default="fused")
    ap.add_argument("--impl", choices=["mfma", "vector"], default="mfma")
    ap.add_argument("--plain-impl", choices=["tiled", "flat", "slice"], default=None,
                    help="plain kernels (default: tiled for N <= 6, flat with the slice lift above)")
    ap.add_argument("--project-velocity", action="store_true")
    ap.add_argument("--order", choices=["morton", "x"], default="morton",
                    help="cube numbering: Z-order curve (default) or x fastest")
    ap.add_argument("--xcd-chunk", type=int, default=None,
                    help="tiles per XCD run of the fused stages (0: dispatch order; default: the library's)")
    ap.add_argument("--elastic", action="store_true", help="elastic (9 fields), fused or plain stage")
    ap.add_argument("--elas-lift", choices=["cubature", "nodes"], default="cubature",
                    help="fused elastic stage: lift folded into the cubature product (default) or at the nodes")
    ap.add_argument("--curved", default=None, metavar="MSH",
                    help="curvilinear run of a ball mesh (GMSH file; 'sphere49748' is the packaged "
                         "mesh): Gordon-Hall map, Bessel start, RK_step_WADG stages")
    ap.add_argument("--test-rk", action="store_true",
                    help="instead: time the WADG update variants per launch (test_RK): the plain "
                         "MFMA update (all four fields + LSRK) and the exact BB-WADG pressure update")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="instead: one device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        import torch
        n = args.copy_bytes // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied {n * 4} bytes", flush=True)
        return
    if args.curved:
        import os
        from . import curved
        from .solver import CurvedWaveSolver
        path = args.curved
        if not os.path.exists(path):
            path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "meshes",
                                args.curved + ".msh")
        m = M.read_gmsh(path)
        M.face_pair(m)
        s = CurvedWaveSolver(m, args.N)
        s.WaveSetData3d(s.WaveSetU0(curved.bessel))
        dt = s.dt_ref
        s.step(2, dt)
        s.sync()
        t0 = time.perf_counter()
        s.step(args.steps, dt)
        s.sync()
        el = time.perf_counter() - t0
        print(f"K={s.K} (curved {len(s.cm.Klist_curved)}) N={args.N} curved steps={args.steps} "
              f"ms/stage={el / (5 * args.steps) * 1e3:.4f} "
              f"DOF-stage/s={s.K * s.ops.Np * 4 * 5 * args.steps / el:.4e}", flush=True)
        return
    if args.test_rk:
        s = setup(args.mesh, args.N)
        xq, yq, zq = s.cubature_points()
        s.InitBBWADG(media.c2_fourier(1)(xq, yq, zq))
        s.set_rhs(np.random.default_rng(2).uniform(-1, 1, (s.K, 4, s.ops.Np)).astype(np.float32))
        for name, fn in (("update_wadg (MFMA, 4 fields + LSRK)", lambda: s.update_wadg(0.5, 0.1, 1e-4)),
                         ("update_bbwadg (exact BB product, pressure)", s.update_bbwadg)):
            fn()
            s.sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn()
            s.sync()
            el = (time.perf_counter() - t0) / args.steps
            print(f"K={s.K} N={args.N} {name}: {el * 1e6:.1f} us per launch", flush=True)
        return
    if args.elastic:
        s = setup_elastic(args.mesh, args.N, order=args.order)
        s.set_option(_lib.OPT_ELAS_FUSED, _lib.ELAS_LIFT_CUBATURE if args.elas_lift == "cubature"
                     else _lib.ELAS_LIFT_NODES)
        mode = MODE_FUSED if args.mode == "fused" else MODE_PLAIN
    else:
        s = setup(args.mesh, args.N, order=args.order)
        if args.N <= 6:
            s.set_option(_lib.OPT_FUSED_IMPL, _lib.FUSED_MFMA if args.impl == "mfma" else _lib.FUSED_VECTOR)
        mode = MODE_FUSED if args.mode == "fused" and args.N <= 6 else MODE_PLAIN
    if args.xcd_chunk is not None:
        s.set_option(_lib.OPT_XCD_CHUNK, args.xcd_chunk)
    if args.plain_impl:
        s.set_option(_lib.OPT_PLAIN_IMPL, {"tiled": _lib.PLAIN_TILED, "flat": _lib.PLAIN_FLAT,
                                           "slice": _lib.PLAIN_SLICE}[args.plain_impl])
    # acoustic: CFL 0.5 h / N^2 at c ~ 1; elastic: the reference's estimate (faster P waves)
    dt = s.dt_ref if args.elastic else 0.5 / args.mesh / args.N ** 2
    s.step(2, dt, mode=mode, project_velocity=args.project_velocity)
    s.sync()
    t0 = time.perf_counter()
    s.step(args.steps, dt, mode=mode, project_velocity=args.project_velocity)
    s.sync()
    el = time.perf_counter() - t0
    dofs = s.K * s.ops.Np * s.nfields * 5 * args.steps
    print(f"K={s.K} N={args.N} fields={s.nfields} mode={'plain' if mode == MODE_PLAIN else 'fused'} "
          f"impl={args.impl} plain_impl={args.plain_impl} steps={args.steps} "
          f"ms/stage={el / (5 * args.steps) * 1e3:.4f} DOF-stage/s={dofs / el:.4e}", flush=True)


if __name__ == "__main__":
    main()
