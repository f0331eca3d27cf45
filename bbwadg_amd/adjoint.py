"""Receiver-misfit gradients w.r.t. the media through the hand-written discrete adjoint.

No counterpart in the reference (its solvers have no adjoint); the forward schemes are the
reference's WADG LSRK45 paths (acoustic ``cpp/src/WaveOKL3d.cpp:2014-2109``, elastic
``elas_cpp/src/WaveOKL3d.cpp:1605-1693``).

    inv = AcousticInversion(solver, shots, rec_points, nsteps, dt, ckpt_every=50)
    J, g = inv.misfit_and_gradient(c2q, d_obs)          # numpy, summed over shots
    J = c2_misfit(c2q_tensor, inv, d_obs)               # torch.autograd.Function, on the GPU
    J.backward(); c2q_tensor.grad                       # dJ/dc2q at the cubature points

    inv = ElasticInversion(solver, shots, rec_points, nsteps, dt, rec_field=0)
    J, g = inv.misfit_and_gradient((rhoinvq, lamq, muq), d_obs)   # g: (3, K, Nq)

Each shot is a source (element ids, BB coefficients, Ricker f0, amplitude, field weights).
All shots share the receivers.  Every call goes to libbbwadg_hip.so (bbw_misfit,
bbw_adjoint_gradient); the gradient is that of the FP32 discrete scheme.
"""
import ctypes

import numpy as np

from . import _lib, media, mesh as meshmod
from ._lib import check, f32, i32, ptr


class Shot:
    def __init__(self, elems, coef, f0, amp=1.0, weights=None):
        self.elems, self.coef, self.f0, self.amp, self.weights = elems, coef, f0, amp, weights


class _Inversion:
    NFIELDS = 4
    NPARAM = 1

    def __init__(self, solver, shots, rec_points, nsteps, dt, ckpt_every=50, t0=0.0, rec_field=0):
        if solver.nfields != self.NFIELDS:
            raise ValueError(f"{type(self).__name__} needs a solver with {self.NFIELDS} fields")
        self.s = solver
        self.shots = list(shots)
        self.nsteps, self.dt, self.ckpt_every, self.t0 = int(nsteps), float(dt), int(ckpt_every), t0
        relems, B = media.receivers(solver.mesh, solver.ops, rec_points)
        self.rec_elems, self.rec_B = relems, B
        self.rec_field = int(rec_field)
        solver.set_option(_lib.OPT_REC_FIELD, self.rec_field)
        solver.set_receivers(relems, B, self.nsteps)
        mapP = i32(meshmod.build_mapP(solver.mesh, solver.ops))
        check(solver.ctx, solver.lib.bbw_upload_adjoint_maps(solver.ctx, ptr(mapP)))
        self.Q0 = np.zeros((solver.K, solver.nfields, solver.ops.Np), dtype=np.float32)
        self._stream = None

    def _grad_shape(self):
        K, Nq = self.s.K, self.s.ops.Nq
        return (K, Nq) if self.NPARAM == 1 else (self.NPARAM, K, Nq)

    def _set_shot(self, shot):
        self.s.set_source(shot.elems, shot.coef, shot.f0, shot.amp, shot.weights)
        self.s.WaveSetData3d(self.Q0)

    def observe(self, model):
        """Synthetic observations: receiver traces (nshots, nsteps, nrec) for the model."""
        self.set_model(model)
        out = []
        for sh in self.shots:
            self._set_shot(sh)
            self.s.rec_reset()
            self.s.step(self.nsteps, self.dt, t0=self.t0)
            out.append(self.s.get_traces(self.nsteps))
        return np.stack(out)

    def misfit(self, model, d_obs):
        self.set_model(model)
        J = 0.0
        for sh, d in zip(self.shots, d_obs):
            self._set_shot(sh)
            v = ctypes.c_double()
            d = f32(d)
            check(self.s.ctx, self.s.lib.bbw_misfit(self.s.ctx, self.nsteps, self.dt, self.t0, ptr(d),
                                                    ctypes.byref(v)))
            J += v.value
        return J

    def misfit_and_gradient(self, model, d_obs):
        self.set_model(model)
        J = 0.0
        g = np.zeros(self._grad_shape(), dtype=np.float64)
        gs = np.zeros(self._grad_shape(), dtype=np.float32)
        for sh, d in zip(self.shots, d_obs):
            self._set_shot(sh)
            v = ctypes.c_double()
            d = f32(d)
            check(self.s.ctx, self.s.lib.bbw_adjoint_gradient(self.s.ctx, self.nsteps, self.dt, self.t0,
                                                              ptr(d), self.ckpt_every, ptr(gs),
                                                              ctypes.byref(v)))
            J += v.value
            g += gs
        return J, g


    # -- device-resident path (torch tensors on the GPU) -------------------------------------
    def torch_stream(self):
        """The torch stream (default priority) the solver's work goes to on the tensor path."""
        import torch
        if self._stream is None:
            self._stream = torch.cuda.Stream()
            self.s.set_stream(self._stream.cuda_stream)
        return self._stream

    def misfit_and_gradient_device(self, model, d_obs, need_grad=True):
        """model: float32 CUDA tensor of the gradient's shape ((K, Nq) c2, or (3, K, Nq)
        1/rho, lambda, mu); d_obs: (nshots, nsteps, nrec) float32 CUDA tensor.  Returns (J, grad)
        with grad a CUDA tensor shaped like the model (None when need_grad is false: then only
        the forward misfit runs).  Everything is enqueued on ``torch_stream()``, which the caller
        orders against its own stream; J is reduced in FP64 on the device."""
        import torch
        lib, ctx = self.s.lib, self.s.ctx
        st = self.torch_stream()
        with torch.cuda.stream(st):
            model = model.detach().to(torch.float32).contiguous()
            if tuple(model.shape) != self._grad_shape():
                raise ValueError(f"model shape {tuple(model.shape)}, expected {self._grad_shape()}")
            d_obs = d_obs.detach().to(torch.float32).contiguous()
            self._upload_model_device(model)
            g = torch.zeros_like(model) if need_grad else None
            gs = torch.empty_like(model) if need_grad else None
            J = 0.0
            for sh, d in zip(self.shots, d_obs):
                self.s.set_source(sh.elems, sh.coef, sh.f0, sh.amp, sh.weights)
                check(ctx, lib.bbw_set_state(ctx, None))
                v = ctypes.c_double()
                dp = ctypes.c_void_p(d.data_ptr())
                if need_grad:
                    check(ctx, lib.bbw_adjoint_gradient_device(ctx, self.nsteps, self.dt, self.t0, dp,
                                                               self.ckpt_every, ctypes.c_void_p(gs.data_ptr()),
                                                               ctypes.byref(v)))
                    g += gs
                else:
                    check(ctx, lib.bbw_misfit_device(ctx, self.nsteps, self.dt, self.t0, dp, ctypes.byref(v)))
                J += v.value
        return J, g


class ElasticInversion(_Inversion):
    """Misfit of one receiver field (``rec_field``: 0..2 velocities, 3..8 stresses) and its
    gradient in (1/rho, lambda, mu) at the cubature points, shape (3, K, Nq); the model is the
    triple (rhoinvq, lamq, muq) (or one (3, K, Nq) array), the penalties those of the
    constructor."""
    NFIELDS = 9
    NPARAM = 3

    def __init__(self, solver, shots, rec_points, nsteps, dt, ckpt_every=50, t0=0.0, rec_field=0,
                 tau_v=1.0, tau_s=1.0):
        super().__init__(solver, shots, rec_points, nsteps, dt, ckpt_every, t0, rec_field)
        self.tau_v, self.tau_s = tau_v, tau_s

    def set_model(self, model):
        rhoinvq, lamq, muq = model
        self.s.InitElastic(rhoinvq, lamq, muq, tau_v=self.tau_v, tau_s=self.tau_s)

    def _upload_model_device(self, m):
        lib, ctx = self.s.lib, self.s.ctx
        check(ctx, lib.bbw_upload_elastic_media_device(ctx, *(ctypes.c_void_p(m[i].data_ptr()) for i in range(3)),
                                                       float(self.tau_v), float(self.tau_s)))


class AcousticInversion(_Inversion):
    """Misfit of the pressure (or the field ``rec_field``) at the receivers and its gradient in c2
    at the cubature points, shape (K, Nq)."""

    def set_model(self, c2q):
        self.s.InitWADG_subelem(np.asarray(c2q, dtype=np.float32))

    def set_c2(self, c2q):
        self.set_model(c2q)

    def _upload_model_device(self, c2q):
        check(self.s.ctx, self.s.lib.bbw_upload_media_device(self.s.ctx, ctypes.c_void_p(c2q.data_ptr())))


def _torch_function():
    import torch

    class ModelMisfit(torch.autograd.Function):
        """J(m) = sum over shots of 1/2 ||receiver traces - d_obs||^2 for a CUDA tensor m (c2 at
        the cubature points, or 1/rho, lambda, mu); backward returns the adjoint gradient,
        computed with the forward (checkpoint and recompute inside) only when m needs a
        gradient.  m, the observations and the gradient stay on the GPU; the work is ordered
        after, and its results before, the caller's current stream."""

        @staticmethod
        def forward(ctx, m, inv, d_obs):
            if not m.is_cuda:
                raise ValueError("the misfit takes the model on the GPU")
            need = bool(ctx.needs_input_grad[0])
            cur = torch.cuda.current_stream(m.device)
            d = torch.as_tensor(d_obs, device=m.device)   # on the caller's stream
            st = inv.torch_stream()
            st.wait_stream(cur)
            J, g = inv.misfit_and_gradient_device(m, d, need)
            cur.wait_stream(st)
            if need:
                g.record_stream(cur)
                ctx.save_for_backward(g.to(m.dtype))
            return torch.tensor(J, dtype=torch.float64, device=m.device)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return (gJ.to(g.dtype) * g, None, None)

    return ModelMisfit


_MISFIT = None


def model_misfit(m, inv, d_obs):
    """torch entry point: ``model_misfit(m_tensor, inv, d_obs)`` -> scalar tensor with grad, for
    an AcousticInversion (m: (K, Nq) c2) or an ElasticInversion (m: (3, K, Nq))."""
    global _MISFIT
    if _MISFIT is None:
        _MISFIT = _torch_function()
    return _MISFIT.apply(m, inv, d_obs)


def c2_misfit(c2q, inv, d_obs):
    """torch entry point: ``c2_misfit(c2q_tensor, inv, d_obs)`` -> scalar tensor with grad."""
    return model_misfit(c2q, inv, d_obs)


def elastic_misfit(m, inv, d_obs):
    """torch entry point for an ElasticInversion: m (3, K, Nq) = (1/rho, lambda, mu) on the GPU."""
    return model_misfit(m, inv, d_obs)
