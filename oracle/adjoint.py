"""FP64 host adjoint of the acoustic BB-WADG LSRK45 scheme -- TEST INFRASTRUCTURE ONLY.

Discrete adjoint of the forward scheme restated in ``oracle.py`` (``Wave_RK`` /
``RK_step_WADG_subelem``, ``cpp/src/WaveOKL3d.cpp:2014-2109``):

    stage i:  r = W(c2) A Q + s(t_i);  res' = a_i res + dt r;  Q' = Q + b_i res'

with A = volume + surface operator (linear in Q), W(c2) = Pq diag(c2) Vq on p and the
identity (Pq Vq = I) on u, v, w.  Misfit J = 1/2 sum_n sum_r (p_r(Q_{n+1}) - d_{n,r})^2 with
p_r(Q) = B_r . Q[k_r, p] the receiver samples after each step.  Reverse stage i, given the
adjoints (lQ', lres') of (Q', res'):

    g = lres' + b_i lQ';   lQ = lQ' + dt A^T W^T g;   lres = a_i g;
    dJ/dc2[k, q] += dt (Pq^T g_p[k])[q] * (Vq (A Q_i)_p[k])[q]

A^T is built here from dense transposes of the operators (volume: D^T with the same
geometric factors; surface: lift transpose, flux transpose, neighbour scatter via the
trace map).  ``tests/test_adjoint_oracle.py`` pins it with the dot-product identity
<A x, y> = <x, A^T y> against the C forward kernels and with central finite differences of
the misfit.
"""
import numpy as np

import oracle as orc


class AdjointOps:
    def __init__(self, ops, dm):
        self.ops = ops
        self.Np, self.Nfp = ops.Np, ops.Nfp
        Dr = 0.5 * (ops.DB[1] - ops.DB[0])
        Ds = 0.5 * (ops.DB[2] - ops.DB[0])
        Dt = 0.5 * (ops.DB[3] - ops.DB[0])
        self.Dr, self.Ds, self.Dt = Dr, Ds, Dt
        self.LIFT = ops.LIFTB  # (Np, 4Nfp)
        self.vgeo = np.asarray(dm.vgeo, dtype=np.float64)
        self.fgeo = np.asarray(dm.fgeo, dtype=np.float64)
        K = dm.vgeo.shape[0]
        NfpNf = 4 * self.Nfp
        Np, Nf = self.Np, 4
        vP = np.asarray(dm.vmapP, dtype=np.int64).reshape(K, NfpNf)
        kP = vP // (Nf * Np)
        nodeP = vP % (Nf * Np)
        self.kP = kP
        self.nodeP = nodeP  # volume node id (field 0) of the exterior trace
        self.bnd = kP == np.arange(K)[:, None]
        self._Fm = ops.Fmask.reshape(-1)

    def AT(self, Z):
        """A^T Z for Z (K, 4, Np)."""
        K = Z.shape[0]
        Np, Nfp = self.Np, self.Nfp
        g = self.vgeo
        rx, ry, rz, sx, sy, sz, tx, ty, tz = (g[:, i:i + 1] for i in range(9))
        zp, zu, zv, zw = Z[:, 0], Z[:, 1], Z[:, 2], Z[:, 3]
        DrT, DsT, DtT = self.Dr.T, self.Ds.T, self.Dt.T
        out = np.zeros_like(Z)
        # volume: rhs_p = -(Dr Ur + Ds Us + Dt Ut), Ur = rx u + ry v + rz w ...
        Zr = rx * zu + ry * zv + rz * zw
        Zs = sx * zu + sy * zv + sz * zw
        Zt = tx * zu + ty * zv + tz * zw
        # transpose of the p-row: contributes to u, v, w through Ur, Us, Ut
        wr, ws, wt = zp @ DrT.T, zp @ DsT.T, zp @ DtT.T  # (K, Np): D^T z_p
        out[:, 1] -= rx * wr + sx * ws + tx * wt
        out[:, 2] -= ry * wr + sy * ws + ty * wt
        out[:, 3] -= rz * wr + sz * ws + tz * wt
        # rhs_u = -(rx Dr + sx Ds + tx Dt) p -> transpose into p
        out[:, 0] -= Zr @ DrT.T + Zs @ DsT.T + Zt @ DtT.T
        # surface: rhs_p += LIFT pflux, rhs_u += LIFT (Uflux nx) ...
        NfpNf = 4 * Nfp
        F = np.repeat(self.fgeo, Nfp, axis=1)  # (K, 4Nfp, 4)
        Fs, nx, ny, nz = F[..., 0], F[..., 1], F[..., 2], F[..., 3]
        gp = zp @ self.LIFT
        gU = nx * (zu @ self.LIFT) + ny * (zv @ self.LIFT) + nz * (zw @ self.LIFT)
        h = 0.5 * Fs * (gp - gU)  # (K, 4Nfp)
        Fm = self._Fm
        # own trace
        own_p = np.where(self.bnd, -2.0 * h, -h)
        own_u = np.where(self.bnd, 0.0, h)
        np.add.at(out[:, 0], (np.arange(K)[:, None], Fm[None, :]), own_p)
        for d, n in ((1, nx), (2, ny), (3, nz)):
            np.add.at(out[:, d], (np.arange(K)[:, None], Fm[None, :]), own_u * n)
        # neighbour trace: p+ gets +h, u+ gets -h n (own normal)
        hi = np.where(self.bnd, 0.0, h)
        kk, nn = self.kP, self.nodeP
        np.add.at(out[:, 0], (kk, nn), hi)
        for d, n in ((1, nx), (2, ny), (3, nz)):
            np.add.at(out[:, d], (kk, nn), -hi * n)
        return out


def forward_rhs(oo, K, dm, Q):
    """A Q with the C oracle (volume + surface), FP64."""
    rhs = np.zeros_like(Q)
    orc.volume(oo, K, dm.vgeo, Q, rhs)
    orc.surface(oo, K, dm.fgeo, dm.vmapP, Q, rhs)
    return rhs


def W(ops, c2q, R):
    out = R.copy()
    out[:, 0] = ((R[:, 0] @ ops.Vq.T) * c2q) @ ops.Pq.T
    return out


def WT(ops, c2q, G):
    out = G.copy()
    out[:, 0] = ((G[:, 0] @ ops.Pq) * c2q) @ ops.Vq
    return out


def forward(oo, ops, dm, c2q, Q0, dt, nsteps, rec_elems, rec_B, src=None, f0=None, amp=1.0,
            keep_stages=False):
    """Forward LSRK45 (FP64) with W applied densely; returns traces (nsteps, nrec) and
    optionally the stage inputs [step][stage] (for the adjoint)."""
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
            r = W(ops, c2q, forward_rhs(oo, K, dm, Q))
            if src is not None:
                r[:, 0] += amp * orc.ricker(t + ops.rk4c[i] * dt, f0) * src
            res = ops.rk4a[i] * res + dt * r
            Q = Q + ops.rk4b[i] * res
        if keep_stages:
            stages.append(st)
        traces[n] = np.einsum("rj,rj->r", rec_B, Q[rec_elems, 0])
    return Q, traces, stages


def misfit(traces, d):
    return 0.5 * float(np.sum((traces - d) ** 2))


def gradient(oo, ops, dm, adj, c2q, Q0, dt, nsteps, rec_elems, rec_B, d, src=None, f0=None, amp=1.0):
    """dJ/dc2q (K, Nq) by the discrete adjoint (FP64, all stage states kept in memory)."""
    K = Q0.shape[0]
    _, traces, stages = forward(oo, ops, dm, c2q, Q0, dt, nsteps, rec_elems, rec_B, src, f0, amp, True)
    lQ = np.zeros_like(Q0, dtype=np.float64)
    lres = np.zeros_like(lQ)
    grad = np.zeros_like(c2q, dtype=np.float64)
    resid = traces - d
    for n in range(nsteps - 1, -1, -1):
        np.add.at(lQ[:, 0], rec_elems, resid[n][:, None] * rec_B)
        lres[:] = 0.0
        for i in range(4, -1, -1):
            g = lres + ops.rk4b[i] * lQ
            y = forward_rhs(oo, K, dm, stages[n][i])[:, 0] @ ops.Vq.T  # Vq (A Q_i)_p
            grad += dt * (g[:, 0] @ ops.Pq) * y
            lQ = lQ + dt * adj.AT(WT(ops, c2q, g))
            lres = ops.rk4a[i] * g
    return grad, misfit(traces, d)
