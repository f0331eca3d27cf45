"""Oracle for the curvilinear WADG path -- TEST INFRASTRUCTURE ONLY.

Wraps the curved-path kernels of ``liborc.so`` (``bbw_oracle_curved.c``, a loop-for-loop C
restatement of ``rk_volume_WADG_skew_combine``, ``rk_surface_WADG_skew_combine``,
``rk_volume_planar`` and ``rk_surface_planar``, cpp/okl/WaveKernelsWADG.okl) and adds the LSRK
driver of ``RK_step_WADG`` (cpp/src/WaveOKL3d.cpp:2111-2185): per stage the skew-form kernels
on the curved elements, the nodal planar kernels on the others, then ``rk_update_WADG`` with the
nodal Vq/Pq of the degree 2N+1 rule and 1/J at its points.  Only ``tests/`` use it.
"""
import ctypes

import numpy as np

import oracle as orc

_p = orc._p


class CurvedOracleOps:
    """Nodal operators (bbwadg_amd.curved.NodalOps) in one working precision, in the reference's
    column-major layouts."""

    def __init__(self, nops, dtype):
        dt = np.dtype(dtype)
        self.dtype = dt
        self.N, self.Np, self.Nfp, self.Nq, self.Nfq = nops.N, nops.Np, nops.Nfp, nops.Nq, nops.Nfq
        if self.Nq > 512 or self.Nfp > 256 or self.Nfq > 256 or self.Np > 256:
            raise ValueError("order exceeds the curved oracle's per-element arrays")

        def cm(a):  # column-major copy
            return np.ascontiguousarray(np.asarray(a).T, dtype=dt)

        self.Vq, self.Vrq, self.Vsq, self.Vtq = (cm(a) for a in (nops.Vq, nops.Vrq, nops.Vsq, nops.Vtq))
        self.Pq, self.Prq, self.Psq, self.Ptq = (cm(a) for a in (nops.Pq, nops.Prq, nops.Psq, nops.Ptq))
        self.VfqFace, self.Pfq = cm(nops.VfqFace), cm(nops.Pfq)
        self.Dr, self.Ds, self.Dt, self.LIFT = (cm(a) for a in (nops.Dr, nops.Ds, nops.Dt, nops.LIFT))
        self.Fmask = np.ascontiguousarray(nops.Fmask.reshape(-1), dtype=np.int32)
        from bbwadg_amd import operators as O
        self.rk4a, self.rk4b, self.rk4c = O.lsrk45_coefficients()


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def volume_skew(co, Klist, vgeoq, Jq, Q, rhs):
    dt = co.dtype
    Klist = _i32(Klist)
    vgeoq, Jq = (np.ascontiguousarray(a, dtype=dt) for a in (vgeoq, Jq))
    getattr(orc.lib(), "orc_volume_skew_" + orc._sfx(dt))(
        ctypes.c_int(len(Klist)), _p(Klist), ctypes.c_int(co.Np), ctypes.c_int(co.Nq), _p(vgeoq),
        _p(Jq), _p(co.Vq), _p(co.Vrq), _p(co.Vsq), _p(co.Vtq), _p(co.Pq), _p(co.Prq), _p(co.Psq),
        _p(co.Ptq), _p(Q), _p(rhs))


def surface_skew(co, Klist, fgeoq, vmapP, Q, rhs):
    dt = co.dtype
    Klist, vmapP = _i32(Klist), _i32(vmapP)
    fgeoq = np.ascontiguousarray(fgeoq, dtype=dt)
    getattr(orc.lib(), "orc_surface_skew_" + orc._sfx(dt))(
        ctypes.c_int(len(Klist)), _p(Klist), ctypes.c_int(co.Np), ctypes.c_int(co.Nfp),
        ctypes.c_int(co.Nfq), _p(fgeoq), _p(co.VfqFace), _p(co.Pfq), _p(co.Fmask), _p(vmapP),
        _p(Q), _p(rhs))


def volume_planar(co, Klist, vgeo, J, Q, rhs):
    dt = co.dtype
    Klist = _i32(Klist)
    vgeo, J = (np.ascontiguousarray(a, dtype=dt) for a in (vgeo, J))
    getattr(orc.lib(), "orc_volume_planar_" + orc._sfx(dt))(
        ctypes.c_int(len(Klist)), _p(Klist), ctypes.c_int(co.Np), _p(vgeo), _p(J), _p(co.Dr),
        _p(co.Ds), _p(co.Dt), _p(Q), _p(rhs))


def surface_planar(co, Klist, fgeo, J, vmapP, Q, rhs):
    dt = co.dtype
    Klist, vmapP = _i32(Klist), _i32(vmapP)
    fgeo, J = (np.ascontiguousarray(a, dtype=dt) for a in (fgeo, J))
    getattr(orc.lib(), "orc_surface_planar_" + orc._sfx(dt))(
        ctypes.c_int(len(Klist)), _p(Klist), ctypes.c_int(co.Np), ctypes.c_int(co.Nfp), _p(fgeo),
        _p(J), _p(co.Fmask), _p(vmapP), _p(co.LIFT), _p(Q), _p(rhs))


def rhs_curved(co, cm, Q, rhs):
    """The right-hand side of one stage on a curved mesh ``cm`` (attributes Klist_curved,
    Klist_planar, vgeoq, Jq, fgeoq, vgeo, fgeo, Jplanar, vmapP)."""
    if len(cm.Klist_curved):
        volume_skew(co, cm.Klist_curved, cm.vgeoq, cm.Jq, Q, rhs)
        surface_skew(co, cm.Klist_curved, cm.fgeoq, cm.vmapP, Q, rhs)
    if len(cm.Klist_planar):
        volume_planar(co, cm.Klist_planar, cm.vgeo, cm.Jplanar, Q, rhs)
        surface_planar(co, cm.Klist_planar, cm.fgeo, cm.Jplanar, cm.vmapP, Q, rhs)


def rk_run_curved(co, cm, c2q, Q, dt, nsteps, res=None, nproj=4):
    """Advance the nodal state Q (K, 4, Np; modified in place) by ``nsteps`` LSRK45 steps of
    ``RK_step_WADG`` (cpp/src/WaveOKL3d.cpp:2111-2185)."""
    wd = co.dtype
    K = Q.shape[0]
    rhs = np.zeros_like(Q)
    if res is None:
        res = np.zeros_like(Q)
    for _ in range(nsteps):
        for i in range(5):
            rhs_curved(co, cm, Q, rhs)
            orc.update_wadg(co, K, c2q, wd.type(co.rk4a[i]), wd.type(co.rk4b[i]), wd.type(dt), rhs,
                            res, Q, Jq=cm.Jq, nproj=nproj)
    return Q, res
