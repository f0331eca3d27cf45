"""GPU parity of the curvilinear WADG path (BBW_MODE_CURVED: the split-form right-hand side on
MFMA tiles for curved and planar elements, then the WADG update with 1/J) against the oracle of
oracle/curved.py (pinned in tests/test_curved_oracle.py), on the reference's sphere meshes with
the Gordon-Hall map (cpp/main.cpp).

Tolerances: the right-hand side 5e-5 of its max-norm against the FP64 oracle (FP32 products of
nodal operators, whose entries reach ~10, summed in another order; with BBW_OPT_PLANAR_SPLIT the
planar elements take the split form where the oracle takes the reference's strong form, equal
in exact arithmetic),
the update 2e-5, whole runs 1e-5 relative L2 against the FP64 oracle.
"""
import math
import os

import numpy as np
import pytest

from bbwadg_amd import _lib, mesh as M, curved as C
from bbwadg_amd.solver import CurvedWaveSolver, MODE_FUSED
import oracle as orc
import curved as OC
from cases import rel, rel_l2

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def make(N, name="sphere385.msh", sphere=True):
    if name.startswith("cube"):
        m = M.cube_mesh(int(name[4:]))
    else:
        m = M.read_gmsh(os.path.join(GOLDEN, name))
    M.face_pair(m)
    return CurvedWaveSolver(m, N, sphere=sphere)


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_curved_rhs_matches_oracle(N, split):
    """Planar elements through the strong nodal form (default) and through the split form with
    constant factors (BBW_OPT_PLANAR_SPLIT)."""
    print(f"case curved rhs N={N} planar_split={split}", flush=True)
    s = make(N)
    s.set_option(_lib.OPT_PLANAR_SPLIT, split)
    assert len(s.cm.Klist_curved) == 321 and len(s.cm.Klist_planar) == 64
    Q = np.random.default_rng(N).uniform(-1, 1, (s.K, 4, s.ops.Np)).astype(np.float32)
    s.WaveSetData3d(Q)
    s.rhs()
    got = s.get_rhs()
    co = OC.CurvedOracleOps(s.nops, np.float64)
    ref = np.zeros(Q.shape)
    OC.rhs_curved(co, s.cm, Q.astype(np.float64), ref)
    assert rel(got, ref) < 5e-5
    # curved and planar elements separately
    for kl in (s.cm.Klist_curved, s.cm.Klist_planar):
        assert rel(got[kl], ref[kl]) < 5e-5
    s.close()


@pytest.mark.parametrize("N", [2, 4])
def test_curved_update_matches_oracle(N):
    print(f"case curved update N={N}", flush=True)
    s = make(N)
    rng = np.random.default_rng(10 + N)
    Q, R, res = (rng.uniform(-1, 1, (s.K, 4, s.ops.Np)).astype(np.float32) for _ in range(3))
    c2 = (1 + 0.3 * np.sin(3 * s.cm.xq) * np.cos(2 * s.cm.yq)).astype(np.float32)
    s.set_c2(c2)
    s.WaveSetData3d(Q)
    s.set_rhs(R)
    s.set_res(res)
    fa, fb, fdt = np.float32(s.ops.rk4a[3]), np.float32(s.ops.rk4b[3]), np.float32(1e-2)
    s.update_wadg(fa, fb, fdt)
    co = OC.CurvedOracleOps(s.nops, np.float64)
    rQ, rres = Q.astype(np.float64), res.astype(np.float64)
    orc.update_wadg(co, s.K, c2, fa, fb, fdt, R.astype(np.float64), rres, rQ, Jq=s.cm.Jq)
    assert rel(s.get_res(), rres) < 2e-5
    assert rel(s.WaveGetData3d(), rQ) < 2e-5
    s.close()


@pytest.mark.parametrize("N,name,sphere", [(3, "sphere385.msh", True), (5, "sphere48.msh", True),
                                            (3, "cube2", False)])
def test_curved_run_matches_oracle(N, name, sphere):
    """The Bessel standing mode (cube: the cavity-free initial pulse of the same function),
    20 steps through the step graph."""
    print(f"case curved run N={N} mesh={name}", flush=True)
    s = make(N, name, sphere)
    Q0 = s.WaveSetU0(C.bessel)
    s.WaveSetData3d(Q0)
    dt, ns = s.dt_ref, 20
    s.step(ns, dt, graph=True)
    got = s.WaveGetData3d()
    co = OC.CurvedOracleOps(s.nops, np.float64)
    ref = Q0.astype(np.float64)
    OC.rk_run_curved(co, s.cm, np.ones((s.K, s.nops.Nq)), ref, dt, ns)
    assert np.abs(ref).max() > 0.1
    assert rel_l2(got, ref) < 1e-5
    s.close()


def test_curved_bessel_error_equals_the_oracle_error():
    """The reference's run (main.cpp: Bessel mode, FinalTime 0.1) on sphere385 at N = 3: the
    GPU's L2 error against the exact mode equals the FP64 oracle's."""
    print("case curved bessel N=3", flush=True)
    s = make(3)
    T = 0.1
    ns = int(math.ceil(T / s.dt_ref))
    s.WaveSetData3d(s.WaveSetU0(C.bessel))
    s.Wave_RK(T, T / ns)
    e_gpu = s.compute_error(T, C.bessel)[0]
    co = OC.CurvedOracleOps(s.nops, np.float64)
    ref = s.WaveSetU0(C.bessel).astype(np.float64)
    OC.rk_run_curved(co, s.cm, np.ones((s.K, s.nops.Nq)), ref, T / ns, ns)
    e_ref = C.l2_error(s.nops, s.cm, ref, C.bessel, T)[0]
    assert e_ref < 5e-4
    assert abs(e_gpu - e_ref) < 1e-2 * e_ref
    s.close()


def test_curved_context_limits():
    print("case curved limits", flush=True)
    s = make(2, "sphere48.msh")
    s.WaveSetData3d(s.WaveSetU0(C.bessel))
    for call in (lambda: s.volume(), lambda: s.surface(), lambda: s.energy(),
                 lambda: s.lib.bbw_step(s.ctx, 1, 1e-3, 0.0, MODE_FUSED, 0)):
        try:
            rc = call()
        except _lib.BBWError as e:
            rc = e.code
        assert rc in (_lib.BBW_ERR_UNSUPPORTED, _lib.BBW_ERR_ARG), rc
    s.step(1, s.dt_ref, graph=False)
    s.close()
