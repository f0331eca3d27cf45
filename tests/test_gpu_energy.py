"""GPU FP64 energies (bbw_energy, bbw_energy_wadg) against the oracle's energies
(oracle/oracle.py: energy_quadrature, energy_wadg and their elastic forms, pinned by closed
forms and decay in tests/test_oracle.py) on the same FP32 state and media: within 1e-10
relative (FP64 sums in a different order; the WADG form's conjugate gradients stop at a
residual of 1e-13 relative, which moves the energy at its square)."""
import numpy as np
import pytest

from bbwadg_amd import media, mesh as M, operators as O
from bbwadg_amd.solver import WaveSolver
import oracle as orc
from cases import Problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,n", [(1, 3), (2, 3), (3, 12), (4, 3), (5, 3), (6, 2)])
def test_acoustic_energies(N, n):
    print(f"case acoustic energies N={N} n={n}", flush=True)
    P = Problem(N, n, seed=N)
    s = WaveSolver(P.mesh, N, ops=P.ops, dmesh=P.dm)
    c2 = P.c2q.astype(np.float32)
    s.InitWADG_subelem(c2)
    Q = P.random_state()
    s.WaveSetData3d(Q)
    Eq, Ew = s.energy(), s.energy("wadg")
    rq = orc.energy_quadrature(P.ops, P.dm.J, c2, Q)
    rw = orc.energy_wadg(P.ops, P.dm.J, c2, Q)
    assert abs(Eq - rq) < 1e-10 * rq
    assert abs(Ew - rw) < 1e-10 * rw
    assert Ew != Eq
    s.close()


@pytest.mark.parametrize("N,n", [(1, 3), (2, 2), (3, 3), (4, 4), (5, 2), (6, 2)])
def test_elastic_energies(N, n):
    print(f"case elastic energies N={N} n={n}", flush=True)
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(n)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    lam, mu, rho = media.lame_fourier(N)
    cub = media.cubature_points(m, ops)
    mats = tuple(np.asarray(a, dtype=np.float32) for a in (1.0 / rho(*cub), lam(*cub), mu(*cub)))
    s = WaveSolver(m, N, nfields=9, ops=ops, dmesh=dm)
    s.InitElastic(*mats)
    Q = np.random.default_rng(N).uniform(-1, 1, (m.K, 9, ops.Np)).astype(np.float32)
    s.WaveSetData3d(Q)
    Eq, Ew = s.energy(), s.energy("wadg")
    rq = orc.energy_quadrature_elastic(ops, dm.J, *mats, Q)
    rw = orc.energy_wadg_elastic(ops, dm.J, *mats, Q)
    assert abs(Eq - rq) < 1e-10 * rq
    assert abs(Ew - rw) < 1e-10 * rw
    s.close()
