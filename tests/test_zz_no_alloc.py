"""bbw_energy_noalloc -- the FP64 energy reduction through the library's static device scratch
array, with no allocation -- against bbw_energy (the same reduction into a buffer the library
allocates) on two calls in a row, the second on a smaller input: the same kernels on the same
grid, so the results are bitwise equal; both also match the FP64 oracle's energy."""
import ctypes

import numpy as np
import pytest

from bbwadg_amd import _lib, media, mesh as M, operators as O
from bbwadg_amd.solver import WaveSolver
import oracle as orc
from cases import Problem

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    return lib.bbw_energy_noalloc


def _both(s, fn):
    E = ctypes.c_double()
    _lib.check(s.ctx, fn(s.ctx, ctypes.byref(E)))
    return E.value, s.energy()


def test_energy_noalloc_matches_energy_on_two_calls():
    lib = _lib.load()
    fn = _noalloc(lib)
    # 1: acoustic, N = 5, 14^3 x 6 = 16464 elements: the reduction's grid is capped at its
    # 4096 partial sums (the grid-stride loop covers the rest)
    print("case no-alloc energy: acoustic N=5 14^3x6", flush=True)
    P = Problem(5, 14, seed=2)
    s = WaveSolver(P.mesh, 5, ops=P.ops, dmesh=P.dm)
    c2 = P.c2q.astype(np.float32)
    s.InitWADG_subelem(c2)
    Q = P.random_state()
    s.WaveSetData3d(Q)
    E0, Eref = _both(s, fn)
    assert E0 == Eref
    assert abs(E0 - orc.energy_quadrature(P.ops, P.dm.J, c2, Q)) < 1e-10 * Eref
    s.close()
    # 2: a smaller input on the same device (elastic, N = 3, 2^3 x 6 = 48 elements, 6 partial
    # sums): the partial sums left by the first call are not read
    print("case no-alloc energy: elastic N=3 2^3x6", flush=True)
    ops = O.build_ref_ops(3)
    m = M.cube_mesh(2)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    lam, mu, rho = media.lame_fourier(1)
    cub = media.cubature_points(m, ops)
    mats = tuple(np.asarray(a, dtype=np.float32) for a in (1.0 / rho(*cub), lam(*cub), mu(*cub)))
    s = WaveSolver(m, 3, nfields=9, ops=ops, dmesh=dm)
    s.InitElastic(*mats)
    Q = np.random.default_rng(5).uniform(-1, 1, (m.K, 9, ops.Np)).astype(np.float32)
    s.WaveSetData3d(Q)
    E1, Eref1 = _both(s, fn)
    assert E1 == Eref1
    assert abs(E1 - orc.energy_quadrature_elastic(ops, dm.J, *mats, Q)) < 1e-10 * Eref1
    assert E1 != E0
    s.close()
