"""Oracle for the BB-WADG hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, and only as the checker (or as the timed CPU
baseline); the product never imports it.

It wraps ``liborc.so`` (``bbw_oracle.c``: a loop-for-loop C restatement of the
reference's OKL kernels in FP64 and FP32) and adds, in numpy FP64:

* ``rk_run``     -- the LSRK45 driver of ``Wave_RK``/``RK_step_WADG_subelem``
  (``cpp/src/WaveOKL3d.cpp:2014-2109``): per stage volume -> surface ->
  WADG update, stage coefficients cast to the working precision as the
  reference does (``:2052-2053``);
* ``energy_quadrature`` -- 1/2 sum_K J sum_q w_q (p_q^2 / c^2_q + |u_q|^2);
* ``energy_wadg``       -- the exact WADG energy 1/2 sum_K [p^T M (M_{c^2})^-1 M p + u^T M u]
  (weighted mass of ``matlab/WADG/check_conservation_loss.m:51-54``);
* ``energy_quadrature_elastic`` / ``energy_wadg_elastic`` -- the same two forms for the
  elastic fields (rho |v|^2 + sigma : C^-1 sigma; weighted mass of the elastic WADG update);
* ``l2_error``          -- ``compute_error`` (``cpp/src/WaveOKL3d.cpp:1683-1726``) with the
  relative error computed as sqrt(err2)/sqrt(norm2) (the reference takes an
  extra square root, ``:1722``).

Parity pins for each function are in ``tests/test_oracle.py`` (drivers, energies,
errors, elastic update), ``tests/test_operators.py`` (operators) and
``tests/test_elastic_oracle.py`` (elastic right-hand side); see the module docstrings there.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIBNAME = "liborc.so"


def select_build(name):
    """Load ``name`` (``liborc.so``, the checker, or ``liborc_fast.so``, the same sources
    optimised for the CPU baseline timing) on first use; call before anything else."""
    global _LIBNAME
    if _LIB is not None and name != _LIBNAME:
        raise RuntimeError("the oracle library is already loaded")
    _LIBNAME = name


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, _LIBNAME)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle/{_LIBNAME} is not built (run make -C oracle)")
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _sfx(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


class OracleOps:
    """Operator arrays cast to one working precision, in the reference's layouts."""

    def __init__(self, ops, dtype):
        dt = np.dtype(dtype)
        self.dtype = dt
        self.N, self.Np, self.Nfp, self.Nq = ops.N, ops.Np, ops.Nfp, ops.Nq
        # per-element arrays of bbw_oracle.c: Np <= 256 (512 volume), 4 Nfp <= 256, Nq <= 1024
        if ops.Np > 256 or 4 * ops.Nfp > 256 or ops.Nq > 1024:
            raise ValueError(f"order {ops.N} exceeds the oracle's per-element arrays")
        self.Dids = np.ascontiguousarray(np.transpose(ops.D_ids, (2, 1, 0)), dtype=np.int32)
        self.Dvals = np.ascontiguousarray(ops.D_vals, dtype=dt)
        self.Fmask = np.ascontiguousarray(ops.Fmask.reshape(-1), dtype=np.int32)
        self.EEL_ids = np.ascontiguousarray(ops.EEL_ids.T, dtype=np.int32)
        self.EEL_vals = np.ascontiguousarray(ops.EEL_vals.T, dtype=dt)
        self.EEL_nnz = ops.EEL_nnz
        self.L0_ids = np.ascontiguousarray(ops.L0_ids.T, dtype=np.int32)
        self.L0_vals = np.ascontiguousarray(ops.L0_vals.T, dtype=dt)
        self.L0_nnz = ops.L0_nnz
        self.Vq = np.ascontiguousarray(ops.Vq.T, dtype=dt)
        self.Pq = np.ascontiguousarray(ops.Pq.T, dtype=dt)
        self.rk4a, self.rk4b, self.rk4c = ops.rk4a, ops.rk4b, ops.rk4c


def volume(oo, K, vgeo, Q, rhs, nfields=4):
    dt = oo.dtype
    vgeo = np.ascontiguousarray(vgeo, dtype=dt)
    name = "orc_volume_" if nfields == 4 else "orc_volume_elas_"
    getattr(lib(), name + _sfx(dt))(ctypes.c_int(K), ctypes.c_int(oo.Np), _p(vgeo), _p(oo.Dids),
                                    _p(oo.Dvals), _p(Q), _p(rhs))


def surface(oo, K, fgeo, vmapP, Q, rhs, nfields=4, tau_v=1.0, tau_s=1.0):
    dt = oo.dtype
    fgeo = np.ascontiguousarray(fgeo, dtype=dt)
    vmapP = np.ascontiguousarray(vmapP, dtype=np.int32)
    R = ctypes.c_double if dt == np.float64 else ctypes.c_float
    if nfields == 4:
        getattr(lib(), "orc_surface_" + _sfx(dt))(
            ctypes.c_int(K), ctypes.c_int(oo.Np), ctypes.c_int(oo.Nfp), _p(fgeo), _p(oo.Fmask),
            _p(vmapP), _p(oo.EEL_ids), _p(oo.EEL_vals), ctypes.c_int(oo.EEL_nnz), _p(oo.L0_ids),
            _p(oo.L0_vals), ctypes.c_int(oo.L0_nnz), _p(Q), _p(rhs))
    else:
        getattr(lib(), "orc_surface_elas_" + _sfx(dt))(
            ctypes.c_int(K), ctypes.c_int(oo.Np), ctypes.c_int(oo.Nfp), _p(fgeo), _p(oo.Fmask),
            _p(vmapP), _p(oo.EEL_ids), _p(oo.EEL_vals), ctypes.c_int(oo.EEL_nnz), _p(oo.L0_ids),
            _p(oo.L0_vals), ctypes.c_int(oo.L0_nnz), R(tau_v), R(tau_s), _p(Q), _p(rhs))


def update_wadg(oo, K, c2q, fa, fb, fdt, rhs, res, Q, Jq=None, nproj=4, ftime=0.0, src=None):
    dt = oo.dtype
    R = ctypes.c_double if dt == np.float64 else ctypes.c_float
    c2q = np.ascontiguousarray(c2q, dtype=dt)
    Jq = None if Jq is None else np.ascontiguousarray(Jq, dtype=dt)
    src = None if src is None else np.ascontiguousarray(src, dtype=dt)
    getattr(lib(), "orc_update_wadg_" + _sfx(dt))(
        ctypes.c_int(K), ctypes.c_int(oo.Np), ctypes.c_int(oo.Nq), ctypes.c_int(nproj), _p(oo.Vq),
        _p(oo.Pq), _p(Jq), _p(c2q), R(fa), R(fb), R(fdt), R(ftime), _p(src), _p(rhs), _p(res), _p(Q))


def update_elas(oo, K, rhoinvq, lamq, muq, fa, fb, fdt, rhs, res, Q, ftime=0.0, src=None, srcw=None):
    dt = oo.dtype
    R = ctypes.c_double if dt == np.float64 else ctypes.c_float
    rhoinvq, lamq, muq = (np.ascontiguousarray(a, dtype=dt) for a in (rhoinvq, lamq, muq))
    src = None if src is None else np.ascontiguousarray(src, dtype=dt)
    srcw = np.ascontiguousarray(np.zeros(9) if srcw is None else srcw, dtype=dt)
    getattr(lib(), "orc_update_elas_" + _sfx(dt))(
        ctypes.c_int(K), ctypes.c_int(oo.Np), ctypes.c_int(oo.Nq), _p(oo.Vq), _p(oo.Pq), _p(rhoinvq),
        _p(lamq), _p(muq), R(ftime), _p(src), _p(srcw), R(fa), R(fb), R(fdt), _p(rhs), _p(res), _p(Q))


def update(oo, fa, fb, fdt, rhs, res, Q):
    dt = oo.dtype
    R = ctypes.c_double if dt == np.float64 else ctypes.c_float
    getattr(lib(), "orc_update_" + _sfx(dt))(ctypes.c_long(Q.size), R(fa), R(fb), R(fdt), _p(rhs),
                                             _p(res), _p(Q))


def ricker(t, f0):
    """Ricker wavelet (1 - 2a^2) exp(-a^2), a = pi f0 (t - 1/f0) (``elas_cpp/src/WaveOKL3d.cpp:1677-1680``)."""
    a = np.pi * f0 * (t - 1.0 / f0)
    return (1.0 - 2.0 * a * a) * np.exp(-a * a)


def stage_time(t, c_i, dt, src_time):
    """Time at which a stage evaluates the source: t + c_i dt ("stage"), or the step start t
    for every stage ("step", as the reference's elastic driver does,
    ``elas_cpp/src/WaveOKL3d.cpp:1675-1680``)."""
    return t + c_i * dt if src_time == "stage" else t


def rk_run(oo, K, vgeo, fgeo, vmapP, c2q, Q, dt, nsteps, t0=0.0, nproj=4, Jq=None,
           src=None, f0=None, amp=1.0, res=None, record=None, src_time="stage"):
    """Advance Q (modified in place) by ``nsteps`` LSRK45 steps of the acoustic WADG scheme.

    Each stage: rhs = volume(Q); rhs += surface(Q); WADG update (``RK_step_WADG_subelem``).
    ``src`` (K, Np) BB coefficients are added to the p right-hand side with
    amplitude amp*ricker(stage_time(t, c_i, dt, src_time), f0).  ``record(step, Q)`` is called
    after each step.
    """
    wd = oo.dtype
    rhs = np.zeros_like(Q)
    if res is None:
        res = np.zeros_like(Q)
    for step in range(nsteps):
        t = t0 + step * dt
        for i in range(5):
            fa, fb, fdt = wd.type(oo.rk4a[i]), wd.type(oo.rk4b[i]), wd.type(dt)
            volume(oo, K, vgeo, Q, rhs)
            surface(oo, K, fgeo, vmapP, Q, rhs)
            ft = 0.0 if src is None else amp * ricker(stage_time(t, oo.rk4c[i], dt, src_time), f0)
            update_wadg(oo, K, c2q, fa, fb, fdt, rhs, res, Q, Jq=Jq, nproj=nproj, ftime=ft, src=src)
        if record is not None:
            record(step, Q)
    return Q, res


def rk_run_elas(oo, K, vgeo, fgeo, vmapP, rhoinvq, lamq, muq, Q, dt, nsteps, t0=0.0,
                src=None, srcw=None, f0=None, amp=1.0, tau_v=1.0, tau_s=1.0, res=None,
                src_time="stage"):
    wd = oo.dtype
    rhs = np.zeros_like(Q)
    if res is None:
        res = np.zeros_like(Q)
    for step in range(nsteps):
        t = t0 + step * dt
        for i in range(5):
            fa, fb, fdt = wd.type(oo.rk4a[i]), wd.type(oo.rk4b[i]), wd.type(dt)
            volume(oo, K, vgeo, Q, rhs, nfields=9)
            surface(oo, K, fgeo, vmapP, Q, rhs, nfields=9, tau_v=tau_v, tau_s=tau_s)
            ft = 0.0 if src is None else amp * ricker(stage_time(t, oo.rk4c[i], dt, src_time), f0)
            update_elas(oo, K, rhoinvq, lamq, muq, fa, fb, fdt, rhs, res, Q, ftime=ft, src=src, srcw=srcw)
    return Q, res


def energy_quadrature(ops, J, c2q, Q):
    """1/2 sum_K J_K sum_q w_q (p_q^2/c^2_q + u_q^2 + v_q^2 + w_q^2), FP64."""
    K, Np = len(J), ops.Np
    Qr = np.asarray(Q, dtype=np.float64).reshape(K, -1, Np)
    Uq = np.einsum("qj,kfj->kfq", ops.Vq, Qr)
    dens = Uq[:, 0] ** 2 / np.asarray(c2q, dtype=np.float64) + (Uq[:, 1:] ** 2).sum(axis=1)
    return 0.5 * float(np.sum(J[:, None] * ops.wq[None, :] * dens))


def energy_wadg(ops, J, c2q, Q):
    """Exact WADG energy 1/2 sum_K [p^T M (M_{c2})^-1 M p + sum_i u_i^T M u_i], M = J * MB."""
    K, Np = len(J), ops.Np
    Qr = np.asarray(Q, dtype=np.float64).reshape(K, -1, Np)
    MB = ops.Vq.T @ (ops.wq[:, None] * ops.Vq)
    Mc = np.einsum("qi,kq,qj->kij", ops.Vq, ops.wq[None, :] * np.asarray(c2q, dtype=np.float64), ops.Vq)
    Mp = Qr[:, 0] @ MB  # (K, Np) = (MB p) per element (MB symmetric)
    y = np.linalg.solve(Mc, Mp[:, :, None])[:, :, 0]
    Ep = np.sum(Mp * y, axis=1) * J
    Eu = np.einsum("kfi,ij,kfj->k", Qr[:, 1:], MB, Qr[:, 1:]) * J
    return 0.5 * float(np.sum(Ep + Eu))


def elastic_density(ops, rhoinvq, lamq, muq, Uq):
    """rho |v|^2 + sigma : C^-1 sigma at the cubature points (C^-1 the isotropic compliance;
    shear stresses with engineering strains, sigma_yz = mu gamma_yz)."""
    rho = 1.0 / np.asarray(rhoinvq, dtype=np.float64)
    lam = np.asarray(lamq, dtype=np.float64)
    mu = np.asarray(muq, dtype=np.float64)
    kin = rho * (Uq[:, 0] ** 2 + Uq[:, 1] ** 2 + Uq[:, 2] ** 2)
    tr = Uq[:, 3] + Uq[:, 4] + Uq[:, 5]
    normal = (Uq[:, 3] ** 2 + Uq[:, 4] ** 2 + Uq[:, 5] ** 2 - lam / (2 * mu + 3 * lam) * tr ** 2) / (2 * mu)
    shear = (Uq[:, 6] ** 2 + Uq[:, 7] ** 2 + Uq[:, 8] ** 2) / mu
    return kin + normal + shear


def energy_quadrature_elastic(ops, J, rhoinvq, lamq, muq, Q):
    """1/2 sum_K J sum_q w_q (rho |v|^2 + sigma : C^-1 sigma), FP64."""
    K = len(J)
    Uq = np.einsum("qj,kfj->kfq", ops.Vq, np.asarray(Q, dtype=np.float64).reshape(K, 9, -1))
    dens = elastic_density(ops, rhoinvq, lamq, muq, Uq)
    return 0.5 * float(np.sum(np.asarray(J)[:, None] * ops.wq[None, :] * dens))


def energy_wadg_elastic(ops, J, rhoinvq, lamq, muq, Q):
    """Weighted-mass energy of the elastic WADG scheme: the update applies
    T = Pq diag(m) Vq (m = 1/rho on v, the isotropic stiffness on the normal stresses, mu on the
    shear stresses), so the scheme's mass matrix is M A_m^-1 M with A_m = Vq^T W m Vq and
    E = 1/2 sum_K J (MB x)^T A_m^-1 (MB x) over the velocity, normal-stress and shear-stress
    blocks (dense FP64 solves per element)."""
    K, Np = len(J), ops.Np
    Qr = np.asarray(Q, dtype=np.float64).reshape(K, 9, Np)
    V, w = ops.Vq, ops.wq
    MB = V.T @ (w[:, None] * V)
    rhoinv, lam, mu = (np.asarray(a, dtype=np.float64) for a in (rhoinvq, lamq, muq))
    E = 0.0
    for k in range(K):
        def scalar_block(c, fields):
            A = V.T @ ((w * c)[:, None] * V)
            B = MB @ Qr[k, fields].T
            return float(np.sum(B * np.linalg.solve(A, B)))
        e = scalar_block(rhoinv[k], [0, 1, 2]) + scalar_block(mu[k], [6, 7, 8])
        # normal block: C = lambda 1 1^T + 2 mu I couples the three normal stresses per point
        A = np.zeros((3 * Np, 3 * Np))
        for a in range(3):
            for b in range(3):
                cab = lam[k] + (2 * mu[k] if a == b else 0.0)
                A[a * Np:(a + 1) * Np, b * Np:(b + 1) * Np] = V.T @ ((w * cab)[:, None] * V)
        B = np.concatenate([MB @ Qr[k, 3 + a] for a in range(3)])
        e += float(B @ np.linalg.solve(A, B))
        E += J[k] * e
    return 0.5 * E


def l2_error(ops, xq, yq, zq, J, Q, exact, time, field=0, nfields=4):
    """(L2err, relL2err) of one field against exact(x, y, z, t) at the WADG cubature points
    xq, yq, zq, each (K, Nq)."""
    K, Np = len(J), ops.Np
    Qr = np.asarray(Q, dtype=np.float64).reshape(K, nfields, Np)
    uq = Qr[:, field] @ ops.Vq.T  # (K, Nq)
    ue = exact(xq, yq, zq, time)
    w = ops.wq[None, :] * J[:, None]
    err2 = float(np.sum((uq - ue) ** 2 * w))
    nrm2 = float(np.sum(ue ** 2 * w))
    return np.sqrt(err2), np.sqrt(err2) / np.sqrt(nrm2)
