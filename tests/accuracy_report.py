"""Accuracy of the FP32 GPU solver on each benchmark workload, against the FP64 oracle
(oracle/bbw_oracle.c through oracle/oracle.py) on the same mesh, order, inputs and step count.

For every row: relative L2 error (coefficient norm) of the pressure field and of the velocity
fields (elastic: velocity and stress), the relative difference of the discrete FP64 energy
(1/2 sum_K J sum_q w_q (p^2/c^2 + |u|^2), elastic: rho |v|^2 + sigma : C^-1 sigma) between the
GPU and oracle states, and the energy change over the run.  The cavity row also gives the L2
error of p and u against the exact standing mode (cpp/main_subelem.cpp:3-5 for p; u from
u_t = -grad p).  Full-size workloads whose FP64 oracle run would take minutes are checked over
a few steps at their full size and over a long run on a smaller mesh of the same generator.

    python tests/accuracy_report.py [--json accuracy.json] [--skip-large]

Test infrastructure (it runs the oracle); prints one markdown row per check as it finishes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, os.path.join(ROOT, "oracle"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as orc  # noqa: E402
from bbwadg_amd import media, mesh as M, operators as O  # noqa: E402
from bbwadg_amd.solver import WaveSolver, MODE_FUSED, MODE_PLAIN  # noqa: E402
from cases import cavity, rel_l2  # noqa: E402
from test_elastic_oracle import elastic_energy  # noqa: E402

ROWS = []


def start(name):
    print(f"<!-- running: {name} -->", flush=True)


def emit(row):
    ROWS.append(row)
    keys = ["case", "mode", "steps", "err_p", "err_u", "err_E", "dE_gpu", "dE_ref", "exact_p", "exact_u"]
    print("| " + " | ".join(_fmt(row.get(k)) for k in keys) + " |", flush=True)


def _fmt(v):
    if v is None:
        return "-"
    if isinstance(v, float):
        return f"{v:.2e}"
    return str(v)


def acoustic_case(name, n, N, nsteps, c2="fourier", seed=1, cavity_init=False, source=None,
                  random_init=False, modes=(MODE_FUSED, MODE_PLAIN)):
    start(name)
    t0 = time.perf_counter()
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 4)
    xq, yq, zq = media.cubature_points(m, ops)
    c2q = (media.c2_fourier(seed)(xq, yq, zq) if c2 == "fourier" else np.ones_like(xq)).astype(np.float32)
    if cavity_init:
        x, y, z = M.affine_points(m, ops.r, ops.s, ops.t)
        Q0 = np.zeros((m.K, 4, ops.Np))
        Q0[:, 0] = (ops.invVB @ cavity(x, y, z, 0.0)).T
        Q0 = Q0.astype(np.float32)
    elif random_init:
        Q0 = np.random.default_rng(seed).uniform(-1, 1, (m.K, 4, ops.Np)).astype(np.float32)
    else:
        Q0 = np.zeros((m.K, 4, ops.Np), dtype=np.float32)
    dt = 0.5 / n / N ** 2
    src = None
    if source is not None:
        x0, f0, amp = source
        el, co = media.gaussian_source(m, ops, x0, a=100.0 if n >= 24 else 20.0)
        src = (el, co, f0, amp)
    ref = Q0.astype(np.float64).copy()
    sarr = None
    if src is not None:
        sarr = np.zeros((m.K, ops.Np))
        sarr[src[0]] = src[1]
    orc.rk_run(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, c2q, ref, dt, nsteps,
               src=sarr, f0=None if src is None else src[2], amp=1.0 if src is None else src[3])
    E0 = orc.energy_quadrature(ops, dm.J, c2q, Q0)
    Eref = orc.energy_quadrature(ops, dm.J, c2q, ref)
    s = WaveSolver(m, N, ops=ops, dmesh=dm)
    s.InitWADG_subelem(c2q)
    if src is not None:
        s.set_source(src[0], src[1], f0=src[2], amp=src[3])
    for mode in modes:
        s.WaveSetData3d(Q0)
        s.set_res(np.zeros_like(Q0))
        s.step(nsteps, dt, mode=mode, graph=True, t0=0.0)
        got = s.WaveGetData3d()
        Eg = s.energy()
        row = dict(case=name, mode="fused" if mode == MODE_FUSED else "plain", steps=nsteps, K=m.K, N=N,
                   err_p=rel_l2(got[:, 0], ref[:, 0]), err_u=rel_l2(got[:, 1:], ref[:, 1:]),
                   err_E=abs(Eg - Eref) / Eref, dE_gpu=(Eg / E0 - 1.0) if E0 > 0 else None,
                   dE_ref=(Eref / E0 - 1.0) if E0 > 0 else None)
        if cavity_init:
            T = nsteps * dt
            row["exact_p"] = s.compute_error(T, cavity, Q=got)[1]
            row["exact_u"] = _cavity_u_error(s, got, T)
            row["exact_p_ref"] = s.compute_error(T, cavity, Q=ref)[1]
            row["exact_u_ref"] = _cavity_u_error(s, ref, T)
        row["seconds"] = time.perf_counter() - t0
        emit(row)
    s.close()


def _cavity_u_error(s, Q, T):
    """Relative L2 error of the velocity against u = grad-dual of the cavity mode:
    u_i = sin(pi x_i) prod_{j != i} cos(pi x_j) sin(sqrt(3) pi t) / sqrt(3)."""
    xq, yq, zq = s.cubature_points()
    w = s.ops.wq[None, :] * s.dmesh.J[:, None]
    c = np.cos(np.pi * np.array([xq, yq, zq]))
    sn = np.sin(np.pi * np.array([xq, yq, zq]))
    amp = np.sin(np.sqrt(3) * np.pi * T) / np.sqrt(3)
    e2 = n2 = 0.0
    for i in range(3):
        ue = amp * sn[i] * c[(i + 1) % 3] * c[(i + 2) % 3]
        uq = np.asarray(Q, dtype=np.float64)[:, 1 + i] @ s.ops.Vq.T
        e2 += float(np.sum((uq - ue) ** 2 * w))
        n2 += float(np.sum(ue ** 2 * w))
    return float(np.sqrt(e2 / n2))


def elastic_case(name, n, N, nsteps, seed=1, random_init=False):
    start(name)
    t0 = time.perf_counter()
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    lam, mu, rho = media.lame_fourier(seed)
    xq, yq, zq = media.cubature_points(m, ops)
    mats = [np.asarray(a, dtype=np.float32) for a in (1.0 / rho(xq, yq, zq), lam(xq, yq, zq), mu(xq, yq, zq))]
    el, co = media.gaussian_source(m, ops, (0.02, -0.03, 0.05), a=100.0 if n >= 24 else 15.0)
    w = np.array([0, 0, 0, 1, 1, 1, 0, 0, 0], dtype=np.float64)   # explosive moment tensor
    f0, amp = 10.0, 5.0
    if random_init:
        Q0 = np.random.default_rng(seed).uniform(-1, 1, (m.K, 9, ops.Np)).astype(np.float32)
    else:
        Q0 = np.zeros((m.K, 9, ops.Np), dtype=np.float32)
    dt = 0.25 / ((N + 1) * (N + 3) / 3.0 * dm.FscaleMax)
    src = np.zeros((m.K, ops.Np))
    src[el] = co
    ref = Q0.astype(np.float64).copy()
    orc.rk_run_elas(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, *mats, ref, dt,
                    nsteps, src=src, srcw=w, f0=f0, amp=amp)
    s = WaveSolver(m, N, nfields=9, ops=ops, dmesh=dm)
    s.InitElastic(*mats)
    s.set_source(el, co, f0=f0, amp=amp, weights=w)
    s.WaveSetData3d(Q0)
    s.step(nsteps, dt, mode=MODE_PLAIN, graph=True, t0=0.0)
    got = s.WaveGetData3d()
    E0 = elastic_energy(ops, dm.J, *[a.astype(np.float64) for a in mats], Q0)
    Eg = elastic_energy(ops, dm.J, *[a.astype(np.float64) for a in mats], got)
    Er = elastic_energy(ops, dm.J, *[a.astype(np.float64) for a in mats], ref)
    emit(dict(case=name, mode="plain", steps=nsteps, K=m.K, N=N,
              err_p=rel_l2(got[:, 3:], ref[:, 3:]), err_u=rel_l2(got[:, :3], ref[:, :3]),
              err_E=abs(Eg - Er) / Er, dE_gpu=(Eg / E0 - 1.0) if E0 > 0 else None,
              dE_ref=(Er / E0 - 1.0) if E0 > 0 else None, seconds=time.perf_counter() - t0))
    s.close()


def batch_case(name):
    """Batched shots (two meshes, uneven durations) against single-shot fused runs (bitwise) and
    one shot against the FP64 oracle."""
    from bbwadg_amd.batch import ShotBatch
    start(name)
    t0 = time.perf_counter()
    N = 5
    ops = O.build_ref_ops(N)
    rng = np.random.default_rng(3)
    meshes = []
    for n in (6, 8):
        m = M.cube_mesh(n)
        M.face_pair(m)
        meshes.append((m, n, M.device_arrays(m, ops, 4)))
    shots = []
    for i, (mi, nsteps) in enumerate([(0, 60), (1, 150), (1, 90), (0, 120)]):
        m, n, dm = meshes[mi]
        xq, yq, zq = media.cubature_points(m, ops)
        c2q = media.c2_fourier(100 + i)(xq, yq, zq).astype(np.float32)
        el, co = media.gaussian_source(m, ops, rng.uniform(-0.3, 0.3, 3), a=15.0)
        shots.append(dict(mi=mi, c2q=c2q, src=(el, co), f0=float(rng.uniform(5, 15)), amp=2.0,
                          dt=0.5 / n / N ** 2, nsteps=nsteps, Q0=np.zeros((m.K, 4, ops.Np), np.float32)))
    b = ShotBatch(N, ops=ops)
    mids = [b.add_mesh(m, dm) for m, n, dm in meshes]
    sids = [b.add_shot(mids[s["mi"]], s["c2q"], s["Q0"], src=s["src"], f0=s["f0"], amp=s["amp"],
                       dt=s["dt"], nsteps=s["nsteps"]) for s in shots]
    b.run(shots_per_tile=2, blocks_per_cu=0)
    maxdiff = 0.0
    for sid, s in zip(sids, shots):
        m, n, dm = meshes[s["mi"]]
        sv = WaveSolver(m, N, ops=ops, dmesh=dm)
        sv.InitWADG_subelem(s["c2q"])
        sv.set_source(*s["src"], f0=s["f0"], amp=s["amp"])
        sv.WaveSetData3d(s["Q0"])
        sv.step(s["nsteps"], s["dt"], mode=MODE_FUSED, graph=True, t0=0.0)
        maxdiff = max(maxdiff, float(np.abs(b.state(sid) - sv.WaveGetData3d()).max()))
        sv.close()
    s = shots[1]
    m, n, dm = meshes[s["mi"]]
    src = np.zeros((m.K, ops.Np))
    src[s["src"][0]] = s["src"][1]
    ref = s["Q0"].astype(np.float64).copy()
    orc.rk_run(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP, s["c2q"], ref, s["dt"],
               s["nsteps"], src=src, f0=s["f0"], amp=s["amp"])
    got = b.state(sids[1])
    emit(dict(case=name, mode=f"batch (max abs diff to single-shot runs {maxdiff:.1e})", steps=s["nsteps"],
              K=m.K, N=N, err_p=rel_l2(got[:, 0], ref[:, 0]), err_u=rel_l2(got[:, 1:], ref[:, 1:]),
              seconds=time.perf_counter() - t0))
    b.close()
    return maxdiff


def adjoint_case(name):
    """c2 gradient of the receiver misfit: GPU adjoint vs the FP64 host adjoint and vs central
    differences of the GPU forward misfit along random directions."""
    from bbwadg_amd.adjoint import AcousticInversion, Shot
    import adjoint as adjm
    start(name)
    t0 = time.perf_counter()
    N, n, nsteps = 5, 3, 40
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 4)
    s = WaveSolver(m, N, ops=ops, dmesh=dm)
    shots = []
    for i in range(2):
        el, co = media.gaussian_source(m, ops, (0.15 * (i - 0.5), -0.1, 0.05 * i), a=8.0)
        shots.append(Shot(el, co, f0=8.0 + 2 * i, amp=30.0))
    dt = 0.5 / n / N ** 2
    inv = AcousticInversion(s, shots, media.surface_receivers(2, 16), nsteps, dt, ckpt_every=10)
    xq, yq, zq = media.cubature_points(m, ops)
    d = inv.observe(media.c2_fourier(5)(xq, yq, zq))
    c2 = np.ones((m.K, ops.Nq))
    J, g = inv.misfit_and_gradient(c2, d)
    oo = orc.OracleOps(ops, np.float64)
    adj = adjm.AdjointOps(ops, dm)
    gr = np.zeros_like(g)
    for sh, dd in zip(shots, d):
        src = np.zeros((m.K, ops.Np))
        src[sh.elems] = sh.coef
        gi, _ = adjm.gradient(oo, ops, dm, adj, c2, np.zeros((m.K, 4, ops.Np)), dt, nsteps, inv.rec_elems,
                              inv.rec_B, dd.astype(np.float64), src, sh.f0, sh.amp)
        gr += gi
    rng = np.random.default_rng(4)
    eps, errs = 1e-3, []
    for _ in range(20):
        v = rng.standard_normal(c2.shape)
        fd = (inv.misfit(c2 + eps * v, d) - inv.misfit(c2 - eps * v, d)) / (2 * eps)
        ad = float(np.sum(g * v))
        errs.append(abs(fd - ad) / abs(ad))
    row = dict(case=name, mode=f"adjoint: c2 gradient vs FP64 host adjoint (FD max rel {max(errs):.1e}, 20 directions, eps 1e-3)",
               steps=nsteps, K=m.K, N=N, err_p=rel_l2(g, gr), seconds=time.perf_counter() - t0)
    emit(row)
    s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--json", default=None)
    ap.add_argument("--skip-large", action="store_true", help="skip the full-size few-step checks")
    args = ap.parse_args()
    print("| case | mode | steps | p (elastic: sigma) rel L2 vs FP64 oracle | u (elastic: v) rel L2 "
          "| energy rel diff | energy change GPU | energy change oracle | p vs exact | u vs exact |")
    print("|---|---|---|---|---|---|---|---|---|---|", flush=True)
    acoustic_case("cavity 8^3x6 N=3 c2=1", 8, 3, 200, c2="one", cavity_init=True)
    acoustic_case("heterogeneous + Ricker 8^3x6 N=5", 8, 5, 250, source=((0.1, -0.05, 0.12), 10.0, 50.0))
    if not args.skip_large:
        acoustic_case("heterogeneous + Ricker 48^3x6 N=5 (full size)", 48, 5, 2, random_init=True,
                      source=((0.1, -0.05, 0.12), 10.0, 50.0), modes=(MODE_FUSED,))
    elastic_case("elastic moment source 8^3x6 N=4", 8, 4, 150)
    if not args.skip_large:
        elastic_case("elastic 32^3x6 N=4 (full size)", 32, 4, 2, random_init=True)
    batch_case("batched 6^3/8^3x6 N=5, 4 shots")
    adjoint_case("adjoint 3^3x6 N=5, 2 shots, 32 receivers")
    if args.json:
        os.makedirs(os.path.dirname(os.path.abspath(args.json)), exist_ok=True)
        with open(args.json, "w") as f:
            json.dump(ROWS, f, indent=1)


if __name__ == "__main__":
    main()
