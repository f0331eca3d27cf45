"""Seeded media, sources and receivers (FP64 host setup).

* ``c2_fourier``   -- c^2(x) = 1 + 0.5 sum_{k=1..8} a_k sin(2 pi w_k . x + phi_k),
  a_k ~ U(-0.1, 0.1), w_k in {1..6}^3, phi_k ~ U(0, 2 pi), so c^2 in [0.6, 1.4];
* ``lame_fourier`` -- smooth lambda, mu, rho from random Fourier series with
  lambda ~= mu (Vp/Vs ~= sqrt 3);
* ``gaussian_source`` -- the reference's smoothed source: exp(-a^2 |x - x0|^2) at the
  WADG cubature, normalised to max 1, projected with Pq
  (``elas_cpp/src/WaveOKL3d.cpp:487-498, 535-538``);
* ``locate`` / ``receivers`` -- element containing a point and BB basis values there.

The reference evaluates media at cubature points through the nodal
interpolation of the element coordinates (``cpp/src/WaveOKL3d.cpp:463-477``);
for affine elements this equals the affine map used here.
"""
import numpy as np

from . import basis, mesh as meshmod


def c2_fourier(seed, nterms=8, amp=0.1):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-amp, amp, nterms)
    w = rng.integers(1, 7, size=(nterms, 3))
    phi = rng.uniform(0.0, 2.0 * np.pi, nterms)

    def c2(x, y, z):
        if not isinstance(x, np.ndarray) and hasattr(x, "sin"):  # torch tensors (same FP64 formula)
            acc = 0.0 * x
            for k in range(nterms):
                acc = acc + float(a[k]) * (2 * np.pi * (float(w[k, 0]) * x + float(w[k, 1]) * y
                                                        + float(w[k, 2]) * z) + float(phi[k])).sin()
            return 1.0 + 0.5 * acc
        x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
        acc = np.zeros(np.broadcast(x, y, z).shape)
        for k in range(nterms):
            acc += a[k] * np.sin(2 * np.pi * (w[k, 0] * x + w[k, 1] * y + w[k, 2] * z) + phi[k])
        return 1.0 + 0.5 * acc

    c2.params = {"a": a, "w": w, "phi": phi}
    return c2


def lame_fourier(seed, nterms=6, amp=0.1):
    """Return (lam, mu, rho) callables; lam = mu*(1+small), rho ~ 1, all positive."""
    rng = np.random.default_rng(seed)

    def series():
        a = rng.uniform(-amp, amp, nterms)
        w = rng.integers(1, 5, size=(nterms, 3))
        phi = rng.uniform(0.0, 2.0 * np.pi, nterms)

        def f(x, y, z):
            acc = np.zeros(np.broadcast(x, y, z).shape)
            for k in range(nterms):
                acc += a[k] * np.sin(2 * np.pi * (w[k, 0] * x + w[k, 1] * y + w[k, 2] * z) + phi[k])
            return acc
        return f

    s_mu, s_lam, s_rho = series(), series(), series()

    def mu(x, y, z):
        return 1.0 + 0.5 * s_mu(x, y, z)

    def lam(x, y, z):
        return mu(x, y, z) * (1.0 + 0.1 * s_lam(x, y, z))

    def rho(x, y, z):
        return 1.0 + 0.3 * s_rho(x, y, z)

    return lam, mu, rho


def element_average(ops, mq):
    """The cubature-weighted element average sum_q w_q m_q / sum_q w_q of a field at the WADG
    cubature (K, Nq), repeated at every point: the reference's smoothing of mu
    (elas_cpp/src/WaveOKL3d.cpp:501-506)."""
    mq = np.asarray(mq, dtype=np.float64)
    avg = (mq @ ops.wq) / ops.wq.sum()
    return np.repeat(avg[:, None], mq.shape[1], axis=1)


def elastic_reference_media(m, ops, weight, average_mu=True):
    """(1/rho, lambda, mu) at the WADG cubature as the reference's elastic driver builds them
    (elas_cpp/src/WaveOKL3d.cpp:453-506): rho = 1; lambda = 1, mu = 1 + weight(x) for z >= 0;
    lambda = 2, mu = 2 + weight(x) for z < 0 (isotropic, discontinuous across z = 0); mu then
    averaged per element (``average_mu``)."""
    xq, yq, zq = cubature_points(m, ops)
    w = weight(xq, yq, zq)
    below = zq < 0
    lam = np.where(below, 2.0, 1.0)
    mu = np.where(below, 2.0, 1.0) + w
    if average_mu:
        mu = element_average(ops, mu)
    return np.ones_like(lam), lam, mu


def cubature_points(m, ops):
    """Physical WADG cubature coordinates, each (K, Nq)."""
    xq, yq, zq = meshmod.affine_points(m, ops.rq, ops.sq, ops.tq)
    return xq.T.copy(), yq.T.copy(), zq.T.copy()


def gaussian_source(m, ops, x0, a=100.0, rel_tol=1e-10, cub=None):
    """(elems ascending int32, coef (nsrc, Np)) of the projected smoothed source.

    Only elements within exp(-a^2 r^2) > rel_tol of x0 (plus their size) are evaluated;
    ``cub`` may pass precomputed cubature points."""
    GX, GY, GZ = m.GX, m.GY, m.GZ
    cx, cy, cz = GX.mean(axis=1), GY.mean(axis=1), GZ.mean(axis=1)
    size = np.sqrt(((GX - cx[:, None]) ** 2 + (GY - cy[:, None]) ** 2 + (GZ - cz[:, None]) ** 2).max(axis=1))
    reach = np.sqrt(-np.log(rel_tol * 1e-3)) / a
    near = np.nonzero(np.sqrt((cx - x0[0]) ** 2 + (cy - x0[1]) ** 2 + (cz - x0[2]) ** 2) < reach + size)[0]
    if cub is None:
        xq, yq, zq = (v.T for v in meshmod.affine_points(_Sub(m, near), ops.rq, ops.sq, ops.tq))
    else:
        xq, yq, zq = (c[near] for c in cub)
    r2 = (xq - x0[0]) ** 2 + (yq - x0[1]) ** 2 + (zq - x0[2]) ** 2
    fq = np.exp(-a * a * r2)
    fq /= np.abs(fq).max()
    keep = np.nonzero(fq.max(axis=1) > rel_tol)[0]
    coef = fq[keep] @ ops.Pq.T  # (nsrc, Np)
    return near[keep].astype(np.int32), coef


class _Sub:
    """Element subset view with the vertex arrays a mesh exposes."""

    def __init__(self, m, idx):
        self.GX, self.GY, self.GZ = m.GX[idx], m.GY[idx], m.GZ[idx]


def locate(m, pts):
    """Element id and reference coordinates (r, s, t) of each point (npts, 3)."""
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    GX, GY, GZ = m.GX, m.GY, m.GZ
    v0 = np.stack([GX[:, 0], GY[:, 0], GZ[:, 0]], axis=1)
    A = np.stack([np.stack([G[:, i] - G[:, 0] for i in (1, 2, 3)], axis=1) for G in (GX, GY, GZ)], axis=1)
    Ainv = np.linalg.inv(A)  # (K, 3, 3): lam(2..4) = Ainv (x - v0)
    elems, rst = [], []
    for p in pts:
        lam = np.einsum("kij,kj->ki", Ainv, p[None, :] - v0)
        l1 = 1.0 - lam.sum(axis=1)
        mn = np.minimum(l1, lam.min(axis=1))
        k = int(np.argmax(mn))
        if mn[k] < -1e-10:
            raise ValueError(f"point {p} is outside the mesh")
        elems.append(k)
        rst.append(2.0 * lam[k] - 1.0)
    return np.array(elems, dtype=np.int32), np.array(rst)


def receivers(m, ops, pts):
    """(elems int32, basis (nrec, Np) BB values) for receiver points."""
    elems, rst = locate(m, pts)
    B = basis.bern_tet(ops.N, rst[:, 0], rst[:, 1], rst[:, 2])
    return elems, B


def surface_receivers(seed, n, lo=-0.5, hi=0.5, inset=1e-3):
    """n seeded points on the top face z = hi (inset slightly inside the domain)."""
    rng = np.random.default_rng(seed)
    xy = rng.uniform(lo + 0.05, hi - 0.05, size=(n, 2))
    return np.column_stack([xy, np.full(n, hi - inset)])
