"""Batched multi-shot runs (persistent launch, uneven meshes and durations, shots-per-tile
grouping, HIP graphs) against single-shot runs of the same shots: the batch runs the same
per-tile stage code, so every shot's final state must be bitwise identical; one shot is also
checked against the FP64 oracle."""
import numpy as np
import pytest

from bbwadg_amd import autotune, media, mesh as M, operators as O
from bbwadg_amd.batch import ShotBatch
from bbwadg_amd.solver import WaveSolver, MODE_FUSED
import oracle as orc
from cases import rel_l2

pytestmark = pytest.mark.gpu


def build(N=3):
    ops = O.build_ref_ops(N)
    rng = np.random.default_rng(9)
    meshes = []
    for n in (2, 3):
        m = M.cube_mesh(n)
        M.face_pair(m)
        meshes.append((m, n, M.device_arrays(m, ops, 4)))
    shots = []
    for i, (mi, nsteps) in enumerate([(0, 7), (1, 12), (1, 23), (0, 12), (1, 3)]):
        m, n, dm = meshes[mi]
        xq, yq, zq = media.cubature_points(m, ops)
        c2q = media.c2_fourier(100 + i)(xq, yq, zq)
        el, co = media.gaussian_source(m, ops, rng.uniform(-0.3, 0.3, 3), a=10.0)
        Q0 = rng.uniform(-1, 1, (m.K, 4, ops.Np)).astype(np.float32)
        shots.append(dict(mi=mi, c2q=c2q, src=(el, co), f0=float(rng.uniform(5, 15)), amp=2.0,
                          dt=0.5 / n / N ** 2, nsteps=nsteps, Q0=Q0))
    return ops, meshes, shots


@pytest.mark.parametrize("spt", [1, 2])
def test_batch_equals_single_shot_runs(spt):
    print(f"case batch spt={spt}", flush=True)
    ops, meshes, shots = build()
    b = ShotBatch(3, ops=ops)
    mids = [b.add_mesh(m, dm) for m, n, dm in meshes]
    sids = [b.add_shot(mids[s["mi"]], s["c2q"], s["Q0"], src=s["src"], f0=s["f0"], amp=s["amp"],
                       dt=s["dt"], nsteps=s["nsteps"]) for s in shots]
    ms = b.run(shots_per_tile=spt, blocks_per_cu=2)
    assert ms > 0
    for sid, s in zip(sids, shots):
        m, n, dm = meshes[s["mi"]]
        sv = WaveSolver(m, 3, ops=ops, dmesh=dm)
        sv.InitWADG_subelem(s["c2q"])
        sv.set_source(*s["src"], f0=s["f0"], amp=s["amp"])
        sv.WaveSetData3d(s["Q0"])
        sv.step(s["nsteps"], s["dt"], mode=MODE_FUSED, graph=True, t0=0.0)
        assert np.array_equal(b.state(sid), sv.WaveGetData3d()), sid
        sv.close()
    # shot 2 against the FP64 oracle
    s = shots[2]
    m, n, dm = meshes[s["mi"]]
    src = np.zeros((m.K, ops.Np)); src[s["src"][0]] = s["src"][1]
    ref = s["Q0"].astype(np.float64).copy()
    orc.rk_run(orc.OracleOps(ops, np.float64), m.K, dm.vgeo, dm.fgeo, dm.vmapP,
               s["c2q"].astype(np.float32), ref, s["dt"], s["nsteps"], nproj=1, src=src,
               f0=s["f0"], amp=s["amp"])
    assert rel_l2(b.state(sids[2]), ref) < 1e-5
    # reset restores the initial states; a second run reproduces the first
    first = [b.state(i) for i in sids]
    b.reset()
    b.run(shots_per_tile=spt, blocks_per_cu=3, graph=False)
    for i, f in zip(sids, first):
        assert np.array_equal(b.state(i), f)
    best, trials = autotune.tune_batch(b, shots_per_tile=(1, 2), blocks_per_cu=(1, 3))
    assert len(trials) == 4 and best in [t[0] for t in trials]
    b.close()
