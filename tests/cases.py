"""Seeded problem builders shared by the tests (host setup only)."""
import numpy as np

from bbwadg_amd import media, mesh as M, operators as O


def cavity(x, y, z, t):
    """Cavity mode on [-0.5, 0.5]^3, zero on the walls (cpp/main_subelem.cpp:3-5)."""
    return np.cos(np.pi * x) * np.cos(np.pi * y) * np.cos(np.pi * z) * np.cos(np.sqrt(3) * np.pi * t)


class Problem:
    def __init__(self, N, n, nfields=4, seed=0, c2="fourier", order="x"):
        self.N, self.n = N, n
        self.ops = O.build_ref_ops(N)
        self.mesh = M.cube_mesh(n, order=order)
        M.face_pair(self.mesh)
        self.dm = M.device_arrays(self.mesh, self.ops, nfields)
        self.K = self.mesh.K
        self.nfields = nfields
        xq, yq, zq = media.cubature_points(self.mesh, self.ops)
        self.xq, self.yq, self.zq = xq, yq, zq
        if c2 == "fourier":
            self.c2q = media.c2_fourier(seed)(xq, yq, zq)
        else:
            self.c2q = np.ones_like(xq)
        self.rng = np.random.default_rng(seed)

    def random_state(self, scale=1.0):
        return (scale * self.rng.uniform(-1, 1, (self.K, self.nfields, self.ops.Np))).astype(np.float32)

    def cavity_state(self):
        x, y, z = M.affine_points(self.mesh, self.ops.r, self.ops.s, self.ops.t)
        Q = np.zeros((self.K, self.nfields, self.ops.Np))
        Q[:, 0] = (self.ops.invVB @ cavity(x, y, z, 0.0)).T
        return Q.astype(np.float32)

    def dt_cfl(self):
        """dt = 0.5 h / N^2 with h the cube edge (the CPU-runnable configuration)."""
        return 0.5 * (1.0 / self.n) / self.N ** 2


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
