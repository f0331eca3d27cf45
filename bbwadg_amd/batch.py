"""Batched multi-shot acoustic runs: uneven meshes and durations, one persistent launch per stage.

No reference counterpart (the reference runs one shot per process); each shot is the
reference's WADG LSRK45 solve (``cpp/src/WaveOKL3d.cpp:2014-2109``).

    b = ShotBatch(N=5)
    m = b.add_mesh(mesh)                      # geometry shared by every shot on that mesh
    s = b.add_shot(m, c2q, Q0, src=(elems, coef), f0=10, amp=1, dt=dt, nsteps=300)
    ms = b.run(shots_per_tile=2, blocks_per_cu=3)
    Q = b.state(s)

``c4_workload`` builds the batched configuration: 64 shots on 16^3, 24^3, 32^3 and 40^3 cube
meshes (x 6 tets), N = 5, seeded Fourier c^2 per shot, seeded source positions, Ricker
f0 in [5, 15], durations 300..1200 steps.
"""
import ctypes

import numpy as np

from . import _lib, media, mesh as meshmod, operators
from ._lib import f32, f64, i32, ptr


class BatchError(RuntimeError):
    pass


class ShotBatch:
    def __init__(self, N, device=0, ops=None):
        self.lib = _lib.load()
        self.N = N
        self.ops = ops if ops is not None else operators.build_ref_ops(N)
        h = ctypes.c_void_p()
        rc = self.lib.bbw_batch_create(N, device, ctypes.byref(h))
        if rc:
            raise BatchError(f"bbw_batch_create failed: {rc}")
        self.b = h
        o = operators.packed_device_ops(self.ops)
        self._keep = dict(o, wq=f64(self.ops.wq), Vq64=f64(self.ops.Vq.T), rk4a=f64(self.ops.rk4a),
                          rk4b=f64(self.ops.rk4b), rk4c=f64(self.ops.rk4c))
        k = self._keep
        s = _lib.Ops(Fmask=ptr(k["Fmask"]).value, D_ids=ptr(k["D_ids"]).value,
                     Dvals=ptr(k["Dvals"]).value, EEL_ids=ptr(k["EEL_ids"]).value,
                     EEL_vals=ptr(k["EEL_vals"]).value, EEL_nnz=self.ops.EEL_nnz,
                     L0_ids=ptr(k["L0_ids"]).value, L0_vals=ptr(k["L0_vals"]).value,
                     L0_nnz=self.ops.L0_nnz, Vq=ptr(k["Vq"]).value, Pq=ptr(k["Pq"]).value,
                     wq=ptr(k["wq"]).value, Vq64=ptr(k["Vq64"]).value, rk4a=ptr(k["rk4a"]).value,
                     rk4b=ptr(k["rk4b"]).value, rk4c=ptr(k["rk4c"]).value)
        self._check(self.lib.bbw_batch_upload_operators(self.b, ctypes.byref(s)))
        self.meshes, self.shots = [], []

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.bbw_batch_last_error(self.b)
            raise BatchError(f"bbw batch error {rc}: {msg.decode() if msg else ''}")

    def add_mesh(self, mesh, dmesh=None):
        if mesh.EToE is None:
            meshmod.face_pair(mesh)
        dm = dmesh if dmesh is not None else meshmod.device_arrays(mesh, self.ops, 4)
        mid = ctypes.c_int()
        self._check(self.lib.bbw_batch_add_mesh(self.b, mesh.K, ptr(f32(dm.vgeo)), ptr(f32(dm.fgeo)),
                                                ptr(i32(dm.vmapP)), ctypes.byref(mid)))
        self.meshes.append((mesh, dm))
        return mid.value

    def add_shot(self, mesh_id, c2q, Q0, src=None, f0=10.0, amp=1.0, dt=1e-3, nsteps=100, weights=None):
        K = self.meshes[mesh_id][0].K
        c2q = f32(np.broadcast_to(c2q, (K, self.ops.Nq)))
        Q0 = f32(Q0).reshape(K, 4, self.ops.Np)
        w = np.zeros(9, dtype=np.float32)
        if weights is None:
            w[0] = 1.0
        else:
            w[:len(weights)] = weights
        if src is None:
            el, co, nsrc = None, None, 0
        else:
            el, co = i32(src[0]), f32(src[1])
            nsrc = len(el)
        sid = ctypes.c_int()
        self._check(self.lib.bbw_batch_add_shot(self.b, mesh_id, ptr(c2q), ptr(Q0), nsrc, ptr(el), ptr(co),
                                                ptr(w), float(f0), float(amp), float(dt), int(nsteps),
                                                ctypes.byref(sid)))
        self.shots.append((mesh_id, int(nsteps)))
        return sid.value

    def run(self, shots_per_tile=1, blocks_per_cu=3, graph=True):
        """Advance every shot to its step count; returns the device time in ms."""
        self._check(self.lib.bbw_batch_run(self.b, int(shots_per_tile), int(blocks_per_cu), int(bool(graph))))
        return self.lib.bbw_batch_last_ms(self.b)

    def set_step_cap(self, cap):
        """Run every shot at most ``cap`` steps (0: full durations), e.g. for tuning trials."""
        self._check(self.lib.bbw_batch_set_step_cap(self.b, int(cap)))

    def reset(self):
        self._check(self.lib.bbw_batch_reset(self.b))

    def state(self, shot_id):
        mid = self.shots[shot_id][0]
        K = self.meshes[mid][0].K
        out = np.empty((K, 4, self.ops.Np), dtype=np.float32)
        self._check(self.lib.bbw_batch_get_state(self.b, shot_id, ptr(out)))
        return out

    def dof_stages(self):
        """Total DOF-stage updates of one run: sum over shots of K Np 4 fields 5 stages steps."""
        return sum(self.meshes[m][0].K * self.ops.Np * 4 * 5 * n for m, n in self.shots)

    def close(self):
        if getattr(self, "b", None):
            self.lib.bbw_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _c2_evaluator(m, ops):
    """Cubature points of a mesh and an evaluator of seeded Fourier c^2 there (FP64; on the
    GPU through torch when a device is present, which keeps 64-shot setups short)."""
    cub = media.cubature_points(m, ops)
    try:
        import torch
        if torch.cuda.is_available():
            tx, ty, tz = (torch.as_tensor(c, dtype=torch.float64, device="cuda") for c in cub)
            return cub, lambda seed: media.c2_fourier(seed)(tx, ty, tz).to(torch.float32).cpu().numpy()
    except ImportError:
        pass
    return cub, lambda seed: media.c2_fourier(seed)(*cub)


def c4_workload(seed=0, nshots=64, mesh_sizes=(16, 24, 32, 40), durations=(300, 1200), N=5,
                step_scale=1.0, ops=None, log=None, order="x"):
    """The batched configuration; returns (batch, description).  ``step_scale`` < 1 shortens
    every duration by that factor (the description states it).  ``log(msg)`` reports progress.
    ``order``: the cube meshes' element numbering (``mesh.cube_mesh``)."""
    rng = np.random.default_rng(seed)
    ops = ops if ops is not None else operators.build_ref_ops(N)
    b = ShotBatch(N, ops=ops)
    meshes = []
    for n in mesh_sizes:
        m = meshmod.cube_mesh(n, order=order)
        meshmod.face_pair(m)
        mid = b.add_mesh(m)
        cub, c2eval = _c2_evaluator(m, ops)
        meshes.append((mid, m, n, cub, c2eval))
        if log:
            log(f"batch mesh {n}^3 x 6 = {m.K} tets ready")
    per_mesh = nshots // len(mesh_sizes)
    total_steps = 0
    for i in range(nshots):
        mid, m, n, cub, c2eval = meshes[min(i // per_mesh, len(meshes) - 1)]
        c2q = c2eval(int(rng.integers(1 << 30)))
        x0 = rng.uniform(-0.3, 0.3, 3)
        f0 = float(rng.uniform(5.0, 15.0))
        nsteps = max(1, int(round(rng.integers(durations[0], durations[1] + 1) * step_scale)))
        total_steps += nsteps
        if log and i % 8 == 7:
            log(f"batch shot {i + 1}/{nshots}")
        src = media.gaussian_source(m, ops, x0, a=100.0, cub=cub)
        dt = 0.5 / n / N ** 2
        Q0 = np.zeros((m.K, 4, ops.Np), dtype=np.float32)
        b.add_shot(mid, c2q, Q0, src=src, f0=f0, amp=1.0, dt=dt, nsteps=nsteps)
    desc = (f"{nshots} shots on {'/'.join(f'{n}^3' for n in mesh_sizes)} x 6 tet cube meshes, N={N}, "
            f"Fourier c2 per shot, Ricker f0 in [5,15], durations {durations[0]}-{durations[1]} steps"
            + (f" scaled by {step_scale}" if step_scale != 1.0 else ""))
    return b, desc
