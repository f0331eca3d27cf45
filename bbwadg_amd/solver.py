"""Host-side driver mirroring the reference's solver interface, over the HIP C ABI.

Reference call sequence (``cpp/main_subelem.cpp:19-104``)::

    mesh = ReadGmsh3d(file); FacePair3d(mesh); StartUp3d(mesh)
    dt = WaveInitOCCA3d(mesh, KblkV, KblkS, KblkU, ...)
    InitWADG_subelem(mesh, c2); WaveSetU0(mesh, Q, 0, uex); WaveSetData3d(Q)
    Wave_RK(mesh, FinalTime, dt, 1); WaveGetData3d(mesh, Q); compute_error(...)

Here::

    s = WaveSolver(mesh, N)                 # FacePair3d + StartUp3d + WaveInitOCCA3d
    s.InitWADG_subelem(c2)                  # c^2 at the WADG cubature -> device
    Q = s.WaveSetU0(uex); s.WaveSetData3d(Q)
    s.Wave_RK(FinalTime, dt)                # or s.step(nsteps, dt, mode=...)
    Q = s.WaveGetData3d()                   # BB coefficients, [K][Nf][Np]
    L2err, relL2err = s.compute_error(FinalTime, uex)

Every numerical call goes to ``libbbwadg_hip.so``; there is no CPU path.
``WaveGetData3d`` returns BB coefficients (the reference converts to nodal
values with VB on read-back; ``to_nodal`` does that).
"""
import ctypes

import numpy as np

from . import _lib, media, mesh as meshmod, operators
from ._lib import MODE_FUSED, MODE_PLAIN, check, f32, f64, i32, ptr


def l2_error(ops, cub, J, Q, uex, time, field=0):
    """(L2err, relL2err) of field ``field`` of the BB coefficients Q [K][Nf][Np] against
    uex(x, y, z, t) at the WADG cubature ``cub`` = (xq, yq, zq), each (K, Nq): the sums
    sum_K J_K sum_q w_q (u_q - uex_q)^2 and sum_K J_K sum_q w_q uex_q^2 of ``compute_error``
    (cpp/src/WaveOKL3d.cpp:1683-1726), FP64.  The relative error is sqrt(err2) / sqrt(norm2);
    the reference takes one more square root of it (:1722)."""
    xq, yq, zq = cub
    uq = np.asarray(Q, dtype=np.float64)[:, field] @ ops.Vq.T
    ue = uex(xq, yq, zq, time)
    w = ops.wq[None, :] * np.asarray(J, dtype=np.float64)[:, None]
    e2 = float(np.sum((uq - ue) ** 2 * w))
    n2 = float(np.sum(ue ** 2 * w))
    return np.sqrt(e2), (np.sqrt(e2) / np.sqrt(n2) if n2 > 0 else np.inf)


class WaveSolver:
    def __init__(self, mesh, N, nfields=4, KblkV=0, KblkS=0, KblkU=0, KblkF=0, device=0,
                 ops=None, dmesh=None):
        if nfields not in (4, 9):
            raise ValueError("nfields must be 4 (acoustic) or 9 (elastic)")
        self.lib = _lib.load()
        self.mesh = mesh
        if mesh.EToE is None:
            meshmod.face_pair(mesh)
        self.ops = ops if ops is not None else operators.build_ref_ops(N)
        self.N, self.nfields = N, nfields
        self.dmesh = dmesh if dmesh is not None else meshmod.device_arrays(mesh, self.ops, nfields)
        self.K = mesh.K
        cfg = _lib.Config(N=N, Nfields=nfields, K=self.K, Nq=self.ops.Nq, KblkV=KblkV,
                          KblkS=KblkS, KblkU=KblkU, KblkF=KblkF, device=device)
        h = ctypes.c_void_p()
        rc = self.lib.bbw_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise _lib.BBWError(rc, "bbw_create failed")
        self.ctx = h
        self._upload_ops()
        dm = self.dmesh
        self._keep = [f32(dm.vgeo), f32(dm.fgeo), i32(dm.vmapP), f64(dm.J)]
        check(self.ctx, self.lib.bbw_upload_mesh(self.ctx, *[ptr(a) for a in self._keep]))
        # reference time step estimate (cpp/src/WaveOKL3d.cpp:1677; elastic :1296-1297)
        if nfields == 4:
            self.dt_ref = 0.25 / ((N + 1) ** 2 * dm.FscaleMax)
        else:
            self.dt_ref = 0.25 / ((N + 1) * (N + 3) / 3.0 * dm.FscaleMax)
        self.time = 0.0

    # -- setup ---------------------------------------------------------------------
    def _upload_ops(self):
        o = operators.packed_device_ops(self.ops)
        self._ops_keep = dict(o, wq=f64(self.ops.wq), Vq64=f64(self.ops.Vq.T), rk4a=f64(self.ops.rk4a),
                              rk4b=f64(self.ops.rk4b), rk4c=f64(self.ops.rk4c))
        k = self._ops_keep
        s = _lib.Ops(Fmask=ptr(k["Fmask"]).value, D_ids=ptr(k["D_ids"]).value,
                     Dvals=ptr(k["Dvals"]).value, EEL_ids=ptr(k["EEL_ids"]).value,
                     EEL_vals=ptr(k["EEL_vals"]).value, EEL_nnz=self.ops.EEL_nnz,
                     L0_ids=ptr(k["L0_ids"]).value, L0_vals=ptr(k["L0_vals"]).value,
                     L0_nnz=self.ops.L0_nnz, Vq=ptr(k["Vq"]).value, Pq=ptr(k["Pq"]).value,
                     wq=ptr(k["wq"]).value, Vq64=ptr(k["Vq64"]).value, rk4a=ptr(k["rk4a"]).value, rk4b=ptr(k["rk4b"]).value,
                     rk4c=ptr(k["rk4c"]).value)
        check(self.ctx, self.lib.bbw_upload_operators(self.ctx, ctypes.byref(s)))
        if self.nfields == 4:
            # slice-by-slice lift tables (the surface kernel for N >= 7, BBW_PLAIN_SLICE at any N)
            sl = operators.slice_lift_ops(self.ops)
            self._slice_keep = [i32(sl["slice_ids"]), i32(sl["red_ids"]), f32(sl["red_vals"]),
                                f32(sl["cEL"])]
            check(self.ctx, self.lib.bbw_upload_slice_lift(self.ctx, *[ptr(a) for a in self._slice_keep]))

    def cubature_points(self):
        return media.cubature_points(self.mesh, self.ops)

    def InitWADG_subelem(self, c2, Jq=None):
        """c2: callable c2(x, y, z) or array (K, Nq) at the WADG cubature."""
        if callable(c2):
            xq, yq, zq = self.cubature_points()
            c2q = c2(xq, yq, zq)
        else:
            c2q = np.asarray(c2)
        self.c2q = f32(np.broadcast_to(c2q, (self.K, self.ops.Nq)))
        Jq = None if Jq is None else f32(Jq)
        check(self.ctx, self.lib.bbw_upload_media(self.ctx, ptr(self.c2q), ptr(Jq)))

    def InitBBWADG(self, c2q):
        """Operators and c2 for the exact Bernstein-Bezier WADG update (the BB branch of
        ``InitWADG_subelem``, cpp/src/WaveOKL3d.cpp:486-690): c2 in degree-N BB coefficients per
        element, c2_bb = Pq c2q (FP64), with the product and projection tables."""
        d = operators.bbwadg_exact_ops(self.N)
        self.c2bb = f32(np.asarray(c2q, dtype=np.float64) @ self.ops.Pq.T)
        self._bb_keep = {k: (i32(v) if v.dtype.kind == "i" else f64(v)) for k, v in d.items()
                         if isinstance(v, np.ndarray)}
        k = self._bb_keep
        check(self.ctx, self.lib.bbw_upload_bbwadg(self.ctx, ptr(k["subtet_ids"]), ptr(k["CNscale"]),
                                                   ptr(k["invC2Nscale"]), ptr(k["Ei_ids4"]),
                                                   ptr(k["Ei_vals4"]), ptr(k["cj"]), ptr(self.c2bb)))

    def update_bbwadg(self):
        """rhs_p <- P_N(c2 rhs_p) in place (``rk_update_BBWADG``)."""
        check(self.ctx, self.lib.bbw_update_bbwadg(self.ctx))

    def InitElastic(self, rhoinvq, lamq, muq, tau_v=1.0, tau_s=1.0, average_mu=False):
        """Elastic media at the WADG cubature (``InitWADG_subelem`` of elas_cpp): 1/rho, lambda,
        mu, each (K, Nq); ``average_mu`` replaces mu by its element averages as the reference
        does (elas_cpp/src/WaveOKL3d.cpp:501-506)."""
        if average_mu:
            muq = media.element_average(self.ops, muq)
        self.rhoinvq, self.lamq, self.muq = f32(rhoinvq), f32(lamq), f32(muq)
        check(self.ctx, self.lib.bbw_upload_elastic_media(self.ctx, ptr(self.rhoinvq), ptr(self.lamq),
                                                          ptr(self.muq), tau_v, tau_s))

    def set_source_step_time(self, on=True):
        """Evaluate the source at each step's start time for all five stages (the reference's
        elastic driver, elas_cpp/src/WaveOKL3d.cpp:1675-1680) instead of at the stage times."""
        self.set_option(_lib.OPT_SRC_STEP_TIME, 1 if on else 0)

    def set_source(self, elems, coef, f0, amp=1.0, weights=None):
        w = np.zeros(9)
        if weights is None:
            w[0] = 1.0
        else:
            w[:len(weights)] = weights
        elems, coef = i32(elems), f32(coef)
        check(self.ctx, self.lib.bbw_set_source(self.ctx, len(elems), ptr(elems), ptr(coef),
                                                ptr(f32(w)), float(f0), float(amp)))

    def set_receivers(self, elems, basis_vals, max_steps):
        elems, basis_vals = i32(elems), f32(basis_vals)
        self.nrec = len(elems)
        self._rec = (elems, basis_vals, int(max_steps))
        check(self.ctx, self.lib.bbw_set_receivers(self.ctx, len(elems), ptr(elems), ptr(basis_vals),
                                                   int(max_steps)))

    def rec_reset(self):
        """Restart receiver recording at trace row 0."""
        self.set_receivers(*self._rec)

    def get_traces(self, nsteps):
        out = np.zeros((nsteps, self.nrec), dtype=np.float32)
        check(self.ctx, self.lib.bbw_get_traces(self.ctx, ptr(out), nsteps))
        return out

    def WaveSetU0(self, uex, field=0, time=0.0):
        """Interpolate uex at the nodes and convert to BB coefficients (``WaveSetU0``)."""
        x, y, z = meshmod.affine_points(self.mesh, self.ops.r, self.ops.s, self.ops.t)
        Q = np.zeros((self.K, self.nfields, self.ops.Np))
        Q[:, field] = (self.ops.invVB @ uex(x, y, z, time)).T
        return f32(Q)

    # -- state ---------------------------------------------------------------------
    def _state(self, fn, arr=None):
        if arr is None:
            out = np.empty((self.K, self.nfields, self.ops.Np), dtype=np.float32)
            check(self.ctx, fn(self.ctx, ptr(out)))
            return out
        a = f32(arr).reshape(self.K, self.nfields, self.ops.Np)
        check(self.ctx, fn(self.ctx, ptr(a)))

    def WaveSetData3d(self, Q):
        self._state(self.lib.bbw_set_state, Q)

    def WaveGetData3d(self):
        return self._state(self.lib.bbw_get_state)

    def set_res(self, r):
        self._state(self.lib.bbw_set_res, r)

    def get_res(self):
        return self._state(self.lib.bbw_get_res)

    def set_rhs(self, r):
        self._state(self.lib.bbw_set_rhs, r)

    def get_rhs(self):
        return self._state(self.lib.bbw_get_rhs)

    def to_nodal(self, Q):
        return np.einsum("ij,kfj->kfi", self.ops.VB, Q)

    def writeVisToGMSH(self, fileName, Q=None, iField=0, time=0.0, step=0):
        """Field iField of Q (BB coefficients; default the current state) at the nodes, as a GMSH
        ``$ElementNodeData`` view (``writeVisToGMSH``, cpp/src/WaveOKL3d.cpp:2560-2642)."""
        Q = self.WaveGetData3d() if Q is None else np.asarray(Q)
        nodal = self.to_nodal(Q)[:, iField]
        meshmod.write_gmsh_node_data(fileName, self.mesh, self.ops.r, self.ops.s, self.ops.t, nodal,
                                     name=f"Field {iField}", time=time, step=step)

    # -- kernels -------------------------------------------------------------------
    def volume(self):
        check(self.ctx, self.lib.bbw_volume(self.ctx))

    def surface(self):
        check(self.ctx, self.lib.bbw_surface(self.ctx))

    def update_wadg(self, fa, fb, fdt, ftime=0.0):
        check(self.ctx, self.lib.bbw_update_wadg(self.ctx, fa, fb, fdt, ftime))

    def update(self, fa, fb, fdt):
        check(self.ctx, self.lib.bbw_update(self.ctx, fa, fb, fdt))

    def stage_fused(self, fa, fb, fdt, ftime=0.0, project_velocity=False):
        check(self.ctx, self.lib.bbw_stage_fused(self.ctx, fa, fb, fdt, ftime, int(project_velocity)))

    def RK_step_WADG_subelem(self, fa, fb, fdt, ftime=0.0):
        """One stage of the plain path: volume, surface, WADG update (three launches)."""
        self.volume()
        self.surface()
        self.update_wadg(fa, fb, fdt, ftime)

    def _mode(self, mode):
        # the fused stage is compiled for N <= 6; higher orders run the plain stage
        if mode is None:
            return MODE_FUSED if self.N <= 6 else MODE_PLAIN
        return mode

    def step(self, nsteps, dt, mode=None, project_velocity=False, graph=True, t0=None):
        """``nsteps`` LSRK45 steps; ``mode`` None picks the fused stage for N <= 6, else plain."""
        mode = self._mode(mode)
        t0 = self.time if t0 is None else t0
        flags = (1 if project_velocity else 0) | (2 if graph else 0)
        check(self.ctx, self.lib.bbw_step(self.ctx, int(nsteps), float(dt), float(t0), int(mode), flags))
        self.time = t0 + nsteps * dt

    def prepare_graph(self, nsteps, dt, mode=None, project_velocity=False):
        """Capture, instantiate and upload the step graph for these arguments and size the
        stage-time buffers for a call of ``nsteps`` steps, without stepping."""
        mode = self._mode(mode)
        flags = (1 if project_velocity else 0) | 2 | 4
        check(self.ctx, self.lib.bbw_step(self.ctx, int(nsteps), float(dt), float(self.time), int(mode), flags))

    def Wave_RK(self, FinalTime, dt, mode=MODE_PLAIN, **kw):
        """Steps of size dt to FinalTime, the last one shortened (``Wave_RK``)."""
        nfull = int(np.floor(FinalTime / dt + 1e-12))
        self.step(nfull, dt, mode=mode, **kw)
        rem = FinalTime - nfull * dt
        if rem > 1e-14 * max(1.0, FinalTime):
            self.step(1, rem, mode=mode, graph=False)

    def sync(self):
        check(self.ctx, self.lib.bbw_sync(self.ctx))

    def energy(self, form="quadrature"):
        """FP64 energy of the current state: ``"quadrature"`` (p^2/c^2 + |u|^2, elastic
        rho |v|^2 + sigma : C^-1 sigma at the cubature points) or ``"wadg"`` (the weighted-mass
        energy the WADG scheme dissipates)."""
        E = ctypes.c_double()
        fn = {"quadrature": self.lib.bbw_energy, "wadg": self.lib.bbw_energy_wadg}[form]
        check(self.ctx, fn(self.ctx, ctypes.byref(E)))
        return E.value

    def compute_error(self, time, uex, Q=None, field=0):
        """(L2err, relL2err) of one field at the WADG cubature (``compute_error``)."""
        Q = self.WaveGetData3d() if Q is None else Q
        return l2_error(self.ops, self.cubature_points(), self.dmesh.J, Q, uex, time, field)

    def set_option(self, option, value):
        check(self.ctx, self.lib.bbw_set_option(self.ctx, int(option), int(value)))

    def device_ptr(self, which):
        p = ctypes.c_void_p()
        check(self.ctx, self.lib.bbw_device_ptr(self.ctx, which, ctypes.byref(p)))
        return p.value

    def set_stream(self, stream_handle):
        check(self.ctx, self.lib.bbw_set_stream(self.ctx, ctypes.c_void_p(stream_handle)))

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.bbw_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CurvedWaveSolver(WaveSolver):
    """The curvilinear WADG run of the reference's ``main`` (cpp/main.cpp:40-130): ball mesh,
    Gordon-Hall sphere map (``sphere``; else every element stays planar), nodal state,
    ``RK_step_WADG`` stages (BBW_MODE_CURVED).

        s = CurvedWaveSolver(mesh, N)            # StartUp3d + WaveInitOCCA3d + GordonHallSphere
                                                 # + InitQuadratureArrays + InitWADG_curved
        s.WaveSetData3d(s.WaveSetU0(curved.bessel))
        s.step(nsteps, dt)                       # Wave_RK(mesh, FinalTime, dt, 2)
        s.compute_error(T, curved.bessel)

    ``WaveSetU0`` interpolates at the curved nodes and returns nodal values (the reference builds
    this path with the nodal basis); c^2 defaults to 1, as in ``InitWADG_curved``."""

    def __init__(self, mesh, N, sphere=True, device=0, ops=None):
        from . import curved
        self.lib = _lib.load()
        self.mesh = mesh
        if mesh.EToE is None:
            meshmod.face_pair(mesh)
        self.ops = ops if ops is not None else operators.build_ref_ops(N)
        self.nops = curved.build_nodal_ops(self.ops)
        self.N, self.nfields, self.K = N, 4, mesh.K
        self.cm = curved.setup_curved(mesh, self.ops, self.nops, sphere=sphere)
        self.dmesh = meshmod.device_arrays(mesh, self.ops, 4)
        cfg = _lib.Config(N=N, Nfields=4, K=self.K, Nq=self.nops.Nq, device=device)
        h = ctypes.c_void_p()
        rc = self.lib.bbw_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise _lib.BBWError(rc, "bbw_create failed")
        self.ctx = h
        dm, cm, no = self.dmesh, self.cm, self.nops
        self._keep = [f32(dm.vgeo), f32(dm.fgeo), i32(dm.vmapP), f64(dm.J)]
        check(self.ctx, self.lib.bbw_upload_mesh(self.ctx, *[ptr(a) for a in self._keep]))

        def cm32(a):   # column-major FP32
            return np.ascontiguousarray(np.asarray(a).T, dtype=np.float32)

        k = dict(Vq=cm32(no.Vq), Vrq=cm32(no.Vrq), Vsq=cm32(no.Vsq), Vtq=cm32(no.Vtq),
                 Pq=cm32(no.Pq), Prq=cm32(no.Prq), Psq=cm32(no.Psq), Ptq=cm32(no.Ptq),
                 VfqFace=cm32(no.VfqFace), Pfq=cm32(no.Pfq), Dr=cm32(no.Dr), Ds=cm32(no.Ds),
                 Dt=cm32(no.Dt), LIFT=cm32(no.LIFT), Fmask=i32(no.Fmask.reshape(-1)),
                 wq=f64(no.wq), rk4a=f64(self.ops.rk4a), rk4b=f64(self.ops.rk4b), rk4c=f64(self.ops.rk4c))
        geo = dict(vgeoq=f32(cm.vgeoq), Jq=f32(cm.Jq), fgeoq=f32(cm.fgeoq), Jplanar=f32(cm.Jplanar),
                   Kc=i32(cm.Klist_curved), Kp=i32(cm.Klist_planar))
        self._curved_keep = (k, geo)
        s = _lib.CurvedOps(Nfq=no.Nfq, **{n: ptr(v).value for n, v in k.items()})
        check(self.ctx, self.lib.bbw_upload_curved(
            self.ctx, ctypes.byref(s), ptr(geo["vgeoq"]), ptr(geo["Jq"]), ptr(geo["fgeoq"]),
            ptr(geo["Jplanar"]), len(geo["Kc"]), ptr(geo["Kc"]) if len(geo["Kc"]) else None,
            len(geo["Kp"]), ptr(geo["Kp"]) if len(geo["Kp"]) else None))
        # reference time step (cpp/src/WaveOKL3d.cpp:1677) from the affine face scales
        self.dt_ref = 0.25 / ((N + 1) ** 2 * dm.FscaleMax)
        self.time = 0.0
        self.set_c2(np.ones((self.K, no.Nq)))

    def set_c2(self, c2q):
        """c^2 at the volume cubature points (K, Nq) and 1/J there for the WADG update
        (``InitWADG_curved``'s c2q and Jq_reduced, cpp/src/WaveOKL3d.cpp:385-408)."""
        self.c2q, self._Jq = f32(c2q), f32(self.cm.Jq)
        check(self.ctx, self.lib.bbw_upload_media(self.ctx, ptr(self.c2q), ptr(self._Jq)))

    def cubature_points(self):
        return self.cm.xq, self.cm.yq, self.cm.zq

    def WaveSetU0(self, uex, field=0, time=0.0):
        """uex at the curved nodes, nodal values (``WaveSetU0`` with the nodal basis)."""
        Q = np.zeros((self.K, 4, self.ops.Np))
        Q[:, field] = uex(self.cm.x, self.cm.y, self.cm.z, time)
        return f32(Q)

    def rhs(self):
        """rhsQ = the split-form right-hand side times J (``RK_step_WADG``'s kernels)."""
        check(self.ctx, self.lib.bbw_curved_rhs(self.ctx))

    def _mode(self, mode):
        if mode not in (None, _lib.MODE_CURVED):
            raise ValueError("a curvilinear solver steps with MODE_CURVED")
        return _lib.MODE_CURVED

    def Wave_RK(self, FinalTime, dt, mode=None, **kw):
        super().Wave_RK(FinalTime, dt, mode=_lib.MODE_CURVED, **kw)

    def compute_error(self, time, uex, Q=None, field=0):
        """(L2err, relL2err) at the volume cubature with J there (``compute_error``)."""
        from . import curved
        Q = self.WaveGetData3d() if Q is None else Q
        return curved.l2_error(self.nops, self.cm, np.asarray(Q, dtype=np.float64)[:, field:field + 1],
                               uex, time)

    def to_nodal(self, Q):
        return np.asarray(Q)
