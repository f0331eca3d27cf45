"""Autotuner for launch shapes (the analogue of the reference's Kblk sweeps,
``cpp/runBlkTimings:4-24`` and the Kblk argv of ``cpp/main.cpp:43-52``).

* ``tune_single(solver, dt)``  times the fused implementations (vector kernel with several
  elements-per-block, MFMA kernel with its 16-element tile) and the plain kernels (tiled, or
  flat with several Kblk), keeps the fastest and sets it on the solver;
* ``tune_batch(batch)``       times shots-per-tile x persistent blocks-per-CU on a batch and
  returns the fastest (the batch is reset after every trial).

Timings are device time (HIP events around graph-replayed steps).
"""
import time

from . import _lib
from .solver import MODE_FUSED, MODE_PLAIN


def _time_steps(solver, dt, nsteps, mode):
    solver.step(2, dt, mode=mode)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(nsteps, dt, mode=mode)
    solver.sync()
    return (time.perf_counter() - t0) / nsteps


def tune_single(solver, dt, nsteps=10, kblk_vector=(4, 8, 16), kblk_plain=(1, 2, 4, 8)):
    """Pick the fused implementation and the plain kernels (16-element tiles, or the flat
    kernels with their elements-per-block); returns a dict of the trials and the choice (also
    applied to the solver).  The state, residual and solver time are restored."""
    Q0, res0, time0 = solver.WaveGetData3d(), solver.get_res(), solver.time
    trials = []
    for kb in kblk_vector:
        solver.set_option(_lib.OPT_FUSED_IMPL, _lib.FUSED_VECTOR)
        solver.set_option(_lib.OPT_KBLK_F, kb)
        trials.append((("vector", kb), _time_steps(solver, dt, nsteps, MODE_FUSED)))
    solver.set_option(_lib.OPT_FUSED_IMPL, _lib.FUSED_MFMA)
    trials.append((("mfma", 16), _time_steps(solver, dt, nsteps, MODE_FUSED)))
    best = min(trials, key=lambda t: t[1])
    impl, kb = best[0]
    solver.set_option(_lib.OPT_FUSED_IMPL, _lib.FUSED_MFMA if impl == "mfma" else _lib.FUSED_VECTOR)
    if impl == "vector":
        solver.set_option(_lib.OPT_KBLK_F, kb)
    plain = []
    solver.set_option(_lib.OPT_PLAIN_IMPL, _lib.PLAIN_TILED)
    plain.append(("tiled", _time_steps(solver, dt, nsteps, MODE_PLAIN)))
    solver.set_option(_lib.OPT_PLAIN_IMPL, _lib.PLAIN_FLAT)
    for kb in kblk_plain:
        for opt in (_lib.OPT_KBLK_V, _lib.OPT_KBLK_S, _lib.OPT_KBLK_U):
            solver.set_option(opt, kb)
        plain.append((kb, _time_steps(solver, dt, nsteps, MODE_PLAIN)))
    best_plain = min(plain, key=lambda t: t[1])
    if best_plain[0] == "tiled":
        solver.set_option(_lib.OPT_PLAIN_IMPL, _lib.PLAIN_TILED)
    else:
        for opt in (_lib.OPT_KBLK_V, _lib.OPT_KBLK_S, _lib.OPT_KBLK_U):
            solver.set_option(opt, best_plain[0])
    solver.WaveSetData3d(Q0)
    solver.set_res(res0)
    solver.time = time0
    return {"fused_trials": trials, "fused_choice": best[0], "plain_trials": plain,
            "plain_choice": best_plain[0]}


def tune_batch(batch, shots_per_tile=(1, 2, 4), blocks_per_cu=(0, 2, 3, 4), step_cap=0):
    """Time every (shots-per-tile, blocks-per-CU) pair on the batch (each shot capped at
    ``step_cap`` steps when > 0; blocks-per-CU 0 = one workgroup per work item); returns (best,
    trials).  The batch is reset afterwards."""
    trials = []
    batch.set_step_cap(step_cap)
    for s in shots_per_tile:
        for bpc in blocks_per_cu:
            batch.reset()
            ms = batch.run(shots_per_tile=s, blocks_per_cu=bpc)
            trials.append(((s, bpc), ms))
    batch.set_step_cap(0)
    batch.reset()
    best = min(trials, key=lambda t: t[1])
    return best[0], trials
