"""FP64 host adjoint of the elastic BB-WADG LSRK45 scheme -- TEST INFRASTRUCTURE ONLY.

Discrete adjoint of the elastic forward scheme restated in ``oracle.py`` (``rk_run_elas``: the
stage of ``elas_cpp/src/WaveOKL3d.cpp:1605-1693`` with the kernels of
``elas_cpp/okl/ElasKernelsWADG.okl:25-433``), fields v1 v2 v3 sxx syy szz syz sxz sxy:

    stage i:  r = W(m) A Q + f_i s;  res' = a_i res + dt r;  Q' = Q + b_i res'

A = volume + penalty surface operator (linear in Q, ``orc_volume_elas`` + ``orc_surface_elas``),
W(m) = Pq M(m) Vq applied per field with M(m) the pointwise map at the cubature points: 1/rho on
the velocities, the isotropic stiffness (lambda, mu) on the normal stresses, mu on the shear
stresses.  Misfit J = 1/2 sum_n sum_r (u_r(Q_{n+1}) - d_{n,r})^2 with u_r the receiver sample of
one field.  Reverse stage i, given the adjoints (lQ', lres') of (Q', res'):

    g = lres' + b_i lQ';   lQ = lQ' + dt A^T W^T g;   lres = a_i g;
    t = Pq^T g, y = Vq (A Q_i) at the cubature points, per field:
    dJ/d(1/rho) += dt sum_{v} t_v y_v;   dJ/dlambda += dt (t_xx + t_yy + t_zz)(y_xx + y_yy + y_zz);
    dJ/dmu += dt (2 sum_{normal} t y + sum_{shear} t y)

A^T is built here from dense transposes: the volume from D^T with the element's constant
geometric factors, the surface from the lift transpose and, per face node, the transpose of the
9 x 9 flux matrix P(n) (formed by applying the forward flux formula to unit vectors), scattered
to the own trace (-) and the neighbour's (+, with the neighbour's own normal and scale).
``tests/test_adjoint_elas_oracle.py`` pins it with the dot-product identity against the C
forward kernels and with central finite differences of the misfit in all three media fields.
"""
import numpy as np

import oracle as orc

NF = 9


def flux_matrix(n, tau_v, tau_s):
    """P(n) (9 x 9): the penalty flux of ``rk_surface_elas`` (ElasKernelsWADG.okl:163-240) as a
    matrix acting on the jump dQ = Q+ - Q-, before the 1/2 Fscale factor."""
    nx, ny, nz = n
    P = np.zeros((NF, NF))
    for j in range(NF):
        dQ = np.zeros(NF)
        dQ[j] = 1.0
        fc = np.array([dQ[3] * nx + dQ[8] * ny + dQ[7] * nz,
                       dQ[8] * nx + dQ[4] * ny + dQ[6] * nz,
                       dQ[7] * nx + dQ[6] * ny + dQ[5] * nz,
                       dQ[0] * nx, dQ[1] * ny, dQ[2] * nz,
                       dQ[2] * ny + dQ[1] * nz, dQ[2] * nx + dQ[0] * nz, dQ[1] * nx + dQ[0] * ny])
        fl = fc.copy()
        fl[0] += tau_v * (fc[3] * nx + fc[8] * ny + fc[7] * nz)
        fl[1] += tau_v * (fc[8] * nx + fc[4] * ny + fc[6] * nz)
        fl[2] += tau_v * (fc[7] * nx + fc[6] * ny + fc[5] * nz)
        fl[3] += tau_s * fc[0] * nx
        fl[4] += tau_s * fc[1] * ny
        fl[5] += tau_s * fc[2] * nz
        fl[6] += tau_s * (fc[2] * ny + fc[1] * nz)
        fl[7] += tau_s * (fc[2] * nx + fc[0] * nz)
        fl[8] += tau_s * (fc[1] * nx + fc[0] * ny)
        P[:, j] = fl
    return P


class ElasAdjointOps:
    def __init__(self, ops, dm, tau_v=1.0, tau_s=1.0):
        self.ops = ops
        self.Np, self.Nfp = ops.Np, ops.Nfp
        self.Dr = 0.5 * (ops.DB[1] - ops.DB[0])
        self.Ds = 0.5 * (ops.DB[2] - ops.DB[0])
        self.Dt = 0.5 * (ops.DB[3] - ops.DB[0])
        self.LIFT = ops.LIFTB  # (Np, 4Nfp)
        self.vgeo = np.asarray(dm.vgeo, dtype=np.float64)
        self.fgeo = np.asarray(dm.fgeo, dtype=np.float64)
        self.tau_v, self.tau_s = tau_v, tau_s
        K = self.vgeo.shape[0]
        Np, Nfp = self.Np, self.Nfp
        vP = np.asarray(dm.vmapP, dtype=np.int64).reshape(K, 4 * Nfp)
        self.kP = vP // (NF * Np)
        self.nodeP = vP % (NF * Np)          # node id (field 0) of the exterior trace
        self.bnd = self.kP == np.arange(K)[:, None]
        self.Fm = ops.Fmask.reshape(-1)
        # per element and face: 1/2 Fscale P(n)^T
        self.PT = np.zeros((K, 4, NF, NF))
        for k in range(K):
            for f in range(4):
                Fs, nx, ny, nz = self.fgeo[k, f]
                self.PT[k, f] = 0.5 * Fs * flux_matrix((nx, ny, nz), tau_v, tau_s).T

    def AT(self, Z):
        """A^T Z for Z (K, 9, Np)."""
        K = Z.shape[0]
        Nfp = self.Nfp
        g = self.vgeo
        rx, ry, rz, sx, sy, sz, tx, ty, tz = (g[:, i:i + 1] for i in range(9))
        out = np.zeros_like(Z)
        # volume: rhs_o += d_x Q_a etc. (orc_volume_elas); (d_x)^T z = Dr^T (rx z) + ...
        dr = np.einsum("kjn,nm->kjm", Z, self.Dr)   # row form of Dr^T z_j
        ds = np.einsum("kjn,nm->kjm", Z, self.Ds)
        dtt = np.einsum("kjn,nm->kjm", Z, self.Dt)
        X = rx[:, None] * dr + sx[:, None] * ds + tx[:, None] * dtt
        Y = ry[:, None] * dr + sy[:, None] * ds + ty[:, None] * dtt
        Zd = rz[:, None] * dr + sz[:, None] * ds + tz[:, None] * dtt
        # (output field of the forward, input field, direction)
        terms = [(0, 3, X), (0, 8, Y), (0, 7, Zd), (1, 8, X), (1, 4, Y), (1, 6, Zd),
                 (2, 7, X), (2, 6, Y), (2, 5, Zd), (3, 0, X), (4, 1, Y), (5, 2, Zd),
                 (6, 2, Y), (6, 1, Zd), (7, 2, X), (7, 0, Zd), (8, 1, X), (8, 0, Y)]
        for o, a, D in terms:
            out[:, a] += D[:, o]
        # surface: rhs += LIFT flux, flux = 1/2 Fscale P(n) (Q+ - Q-) per face node
        w = np.einsum("kjn,nm->kjm", Z, self.LIFT)   # (K, 9, 4Nfp) = LIFT^T z
        face = np.repeat(np.arange(4), Nfp)
        h = np.einsum("kmij,kjm->kim", self.PT[:, face], w)   # (K, 9, 4Nfp)
        kk = np.arange(K)[:, None]
        for j in range(NF):
            np.add.at(out[:, j], (kk, self.Fm[None, :]), -h[:, j])
            hi = np.where(self.bnd, 0.0, h[:, j])
            np.add.at(out[:, j], (self.kP, self.nodeP), hi)
        return out


def forward_rhs(oo, K, dm, Q, tau_v, tau_s):
    """A Q with the C oracle (elastic volume + penalty surface), FP64."""
    rhs = np.zeros_like(Q)
    orc.volume(oo, K, dm.vgeo, Q, rhs, nfields=NF)
    orc.surface(oo, K, dm.fgeo, dm.vmapP, Q, rhs, nfields=NF, tau_v=tau_v, tau_s=tau_s)
    return rhs


def media_map(m, Rq):
    """M(m) at the cubature points on Rq (K, 9, Nq); m = (rhoinvq, lamq, muq)."""
    rho, lam, mu = m
    out = np.empty_like(Rq)
    out[:, 0:3] = rho[:, None] * Rq[:, 0:3]
    tr = Rq[:, 3] + Rq[:, 4] + Rq[:, 5]
    for j in (3, 4, 5):
        out[:, j] = 2.0 * mu * Rq[:, j] + lam * tr
    out[:, 6:9] = mu[:, None] * Rq[:, 6:9]
    return out


def W(ops, m, R):
    return np.einsum("kjq,nq->kjn", media_map(m, np.einsum("kjn,qn->kjq", R, ops.Vq)), ops.Pq)


def WT(ops, m, G):
    # M(m) is symmetric at each point
    return np.einsum("kjq,qn->kjn", media_map(m, np.einsum("kjn,nq->kjq", G, ops.Pq)), ops.Vq)


def forward(oo, ops, dm, m, Q0, dt, nsteps, rec_elems, rec_B, rec_field=0, src=None, srcw=None,
            f0=None, amp=1.0, tau_v=1.0, tau_s=1.0, keep_stages=False):
    """Forward LSRK45 (FP64) with W applied densely; returns the final state, the traces
    (nsteps, nrec) of field ``rec_field`` and optionally the stage inputs [step][stage]."""
    K = Q0.shape[0]
    Q = Q0.astype(np.float64).copy()
    res = np.zeros_like(Q)
    traces = np.zeros((nsteps, len(rec_elems)))
    stages = [] if keep_stages else None
    for n in range(nsteps):
        t = n * dt
        st = []
        for i in range(5):
            if keep_stages:
                st.append(Q.copy())
            r = W(ops, m, forward_rhs(oo, K, dm, Q, tau_v, tau_s))
            if src is not None:
                r += (amp * orc.ricker(t + ops.rk4c[i] * dt, f0)) * np.asarray(srcw)[None, :, None] * src[:, None, :]
            res = ops.rk4a[i] * res + dt * r
            Q = Q + ops.rk4b[i] * res
        if keep_stages:
            stages.append(st)
        traces[n] = np.einsum("rj,rj->r", rec_B, Q[rec_elems, rec_field])
    return Q, traces, stages


def misfit(traces, d):
    return 0.5 * float(np.sum((traces - d) ** 2))


def gradient(oo, ops, dm, adj, m, Q0, dt, nsteps, rec_elems, rec_B, d, rec_field=0, src=None,
             srcw=None, f0=None, amp=1.0):
    """(dJ/d(1/rho), dJ/dlambda, dJ/dmu) at the cubature points, (3, K, Nq), and J, by the
    discrete adjoint (FP64, all stage states kept in memory)."""
    tau_v, tau_s = adj.tau_v, adj.tau_s
    K = Q0.shape[0]
    _, traces, stages = forward(oo, ops, dm, m, Q0, dt, nsteps, rec_elems, rec_B, rec_field, src,
                                srcw, f0, amp, tau_v, tau_s, True)
    lQ = np.zeros((K, NF, ops.Np))
    lres = np.zeros_like(lQ)
    grad = np.zeros((3,) + m[0].shape)
    resid = traces - d
    for n in range(nsteps - 1, -1, -1):
        np.add.at(lQ[:, rec_field], rec_elems, resid[n][:, None] * rec_B)
        lres[:] = 0.0
        for i in range(4, -1, -1):
            g = lres + ops.rk4b[i] * lQ
            y = np.einsum("kjn,qn->kjq", forward_rhs(oo, K, dm, stages[n][i], tau_v, tau_s), ops.Vq)
            t = np.einsum("kjn,nq->kjq", g, ops.Pq)
            grad[0] += dt * np.sum(t[:, 0:3] * y[:, 0:3], axis=1)
            grad[1] += dt * (t[:, 3] + t[:, 4] + t[:, 5]) * (y[:, 3] + y[:, 4] + y[:, 5])
            grad[2] += dt * (2.0 * np.sum(t[:, 3:6] * y[:, 3:6], axis=1) + np.sum(t[:, 6:9] * y[:, 6:9], axis=1))
            lQ = lQ + dt * adj.AT(WT(ops, m, g))
            lres = ops.rk4a[i] * g
    return grad, misfit(traces, d)
