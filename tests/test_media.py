"""Host setup pieces: the element-local source equals the source evaluated on the whole
mesh; receivers locate their element and reproduce linear fields; c^2 stays in [0.6, 1.4]."""
import numpy as np

from bbwadg_amd import basis, media, mesh as M, operators as O


def test_local_source_equals_global_evaluation():
    ops = O.build_ref_ops(3)
    m = M.cube_mesh(4)
    x0, a, tol = (0.05, -0.1, 0.02), 20.0, 1e-10
    xq, yq, zq = media.cubature_points(m, ops)
    fq = np.exp(-a * a * ((xq - x0[0]) ** 2 + (yq - x0[1]) ** 2 + (zq - x0[2]) ** 2))
    fq /= fq.max()
    keep = np.nonzero(fq.max(axis=1) > tol)[0]
    el, co = media.gaussian_source(m, ops, x0, a=a, rel_tol=tol)
    assert np.array_equal(el, keep)
    assert np.abs(co - fq[keep] @ ops.Pq.T).max() < 1e-13


def test_receivers_reproduce_linear_field():
    ops = O.build_ref_ops(4)
    m = M.cube_mesh(3)
    pts = media.surface_receivers(1, 6)
    el, B = media.receivers(m, ops, pts)
    # BB coefficients of f = 2x - y + 3z (affine fields are exact in the basis)
    x, y, z = M.affine_points(m, ops.r, ops.s, ops.t)
    coef = (ops.invVB @ (2 * x - y + 3 * z)).T
    vals = np.einsum("rj,rj->r", B, coef[el])
    assert np.abs(vals - (2 * pts[:, 0] - pts[:, 1] + 3 * pts[:, 2])).max() < 1e-12


def test_c2_range():
    g = media.c2_fourier(3)
    x = np.random.default_rng(0).uniform(-0.5, 0.5, (3, 10000))
    c = g(*x)
    assert c.min() >= 0.6 and c.max() <= 1.4


def test_element_average_and_reference_elastic_media():
    """media.element_average is the cubature-weighted mean (exact for a linear field: the mean
    of a linear function over a tetrahedron is its value at the centroid); the reference's
    elastic media (elas_cpp/src/WaveOKL3d.cpp:453-506) are rho = 1, (lambda, mu) = (1, 1 + w)
    above z = 0 and (2, 2 + w) below, mu averaged per element."""
    from bbwadg_amd import operators as O
    ops = O.build_ref_ops(3)
    m = M.cube_mesh(2)
    xq, yq, zq = media.cubature_points(m, ops)
    f = 1.0 + 2.0 * xq - yq + 0.5 * zq
    avg = media.element_average(ops, f)
    cx, cy, cz = (G.mean(axis=1) for G in (m.GX, m.GY, m.GZ))
    assert np.allclose(avg, (1.0 + 2.0 * cx - cy + 0.5 * cz)[:, None], atol=1e-13)
    w = lambda x, y, z: 0.1 * np.sin(3 * x) * np.cos(2 * y)
    rinv, lam, mu = media.elastic_reference_media(m, ops, w, average_mu=False)
    below = zq < 0
    assert np.all(rinv == 1.0)
    assert np.array_equal(lam, np.where(below, 2.0, 1.0))
    assert np.allclose(mu, np.where(below, 2.0, 1.0) + w(xq, yq, zq))
    _, _, mu_avg = media.elastic_reference_media(m, ops, w)
    assert np.allclose(mu_avg, media.element_average(ops, mu))
    assert np.allclose(mu_avg.std(axis=1), 0.0)
