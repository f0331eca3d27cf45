"""Pins the elastic oracle (BB restatement of rk_volume_elas / rk_surface_elas /
rk_update_elas, elas_cpp/okl/ElasKernelsWADG.okl): the BB right-hand side equals the
reference's nodal formulation through the change of basis, and the semi-discrete scheme
dissipates the elastic energy (penalty flux, Q+ = 0 boundary) in a short run."""
import numpy as np
import pytest

from bbwadg_amd import media, mesh as M, operators as O
import oracle as orc
import ref_ops


@pytest.mark.parametrize("N", [2, 3, 4])
def test_bb_elastic_rhs_equals_nodal(N):
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(2)
    M.face_pair(m)
    vmapM, vmapP = M.build_maps(m, ops)
    dm = M.device_arrays(m, ops, 9, vmapP=vmapP)
    g = M.geometric_factors(m)
    nx, ny, nz, sJ = M.normals(g)
    vgeo = np.stack([g[k] for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz")], axis=1)
    fgeo = np.stack([sJ / g["J"][:, None], nx, ny, nz], axis=2)
    rng = np.random.default_rng(N)
    Q = rng.standard_normal((m.K, 9, ops.Np))
    oo = orc.OracleOps(ops, np.float64)
    rhs = np.zeros_like(Q)
    orc.volume(oo, m.K, vgeo, Q, rhs, nfields=9)
    orc.surface(oo, m.K, fgeo, dm.vmapP, Q, rhs, nfields=9, tau_v=0.7, tau_s=1.3)
    Dr, Ds, Dt, LIFT = ref_ops.nodal_operators(N, ops.r, ops.s, ops.t, ops.Fmask)
    qn = np.einsum("ij,kfj->kfi", ops.VB, Q)
    rn = ref_ops.nodal_rhs_elastic(Dr, Ds, Dt, LIFT, vgeo, fgeo, vmapM, vmapP, qn, 0.7, 1.3)
    rbb = np.einsum("ij,kfj->kfi", ops.invVB, rn)
    assert np.abs(rbb - rhs).max() < 1e-8 * np.abs(rhs).max()


def elastic_energy(ops, J, rhoinvq, lamq, muq, Q):
    """1/2 sum_K J sum_q w_q (rho |v|^2 + sigma : C^-1 sigma) at the WADG cubature (the
    oracle's quadrature form; pinned in tests/test_oracle.py)."""
    return orc.energy_quadrature_elastic(ops, J, rhoinvq, lamq, muq, Q)


def test_elastic_energy_decays_constant_media():
    N = 3
    ops = O.build_ref_ops(N)
    m = M.cube_mesh(3)
    M.face_pair(m)
    dm = M.device_arrays(m, ops, 9)
    K = m.K
    rhoinv = np.ones((K, ops.Nq)); lam = np.ones((K, ops.Nq)); mu = np.ones((K, ops.Nq))
    rng = np.random.default_rng(0)
    x, y, z = M.affine_points(m, ops.r, ops.s, ops.t)
    Q = np.zeros((K, 9, ops.Np))
    bump = np.exp(-40 * (x ** 2 + y ** 2 + z ** 2))
    for f in range(9):
        Q[:, f] = (ops.invVB @ (rng.uniform(-1, 1) * bump)).T
    oo = orc.OracleOps(ops, np.float64)
    dt = 0.25 / ((N + 1) * (N + 3) / 3.0 * dm.FscaleMax)
    E = [elastic_energy(ops, dm.J, rhoinv, lam, mu, Q)]
    for _ in range(4):
        orc.rk_run_elas(oo, K, dm.vgeo, dm.fgeo, dm.vmapP, rhoinv, lam, mu, Q, dt, 5)
        E.append(elastic_energy(ops, dm.J, rhoinv, lam, mu, Q))
    assert all(E[i + 1] <= E[i] * (1 + 1e-12) for i in range(len(E) - 1))
    assert E[-1] > 0.5 * E[0]
