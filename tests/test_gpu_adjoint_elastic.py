"""GPU discrete adjoint of the fused elastic stage against the FP64 host adjoint
(oracle/adjoint_elas.py, itself pinned by the dot-product identity and central differences) on
the same seeded problem, at N = 1..6 and for a velocity and a stress receiver field; checkpoint
placement changes the result only at roundoff; GPU central differences agree along random
directions of 1/rho, lambda and mu."""
import numpy as np
import pytest

from bbwadg_amd import _lib, media
from bbwadg_amd.adjoint import ElasticInversion, Shot
from bbwadg_amd.solver import WaveSolver
import oracle as orc
import adjoint_elas as ae
from cases import Problem, rel_l2

pytestmark = pytest.mark.gpu

TAU = (0.8, 1.3)
SRCW = np.array([0.2, 0, 0, 1, 1, 1, 0, 0.3, 0])


def setup(N=3, n=3, nsteps=40, nshots=2, rec_field=0, ckpt_every=7):
    P = Problem(N, n, nfields=9, seed=5)
    s = WaveSolver(P.mesh, N, nfields=9, ops=P.ops, dmesh=P.dm)
    lam, mu, rho = media.lame_fourier(4)
    m_true = (1.0 / rho(P.xq, P.yq, P.zq), lam(P.xq, P.yq, P.zq), mu(P.xq, P.yq, P.zq))
    s.InitElastic(*m_true, tau_v=TAU[0], tau_s=TAU[1])
    shots = []
    for i in range(nshots):
        x0 = (0.15 * (i - 0.5), -0.1, 0.3 - 0.05 * i)   # below the receivers on z = 0.5
        el, co = media.gaussian_source(P.mesh, P.ops, x0, a=8.0)
        shots.append(Shot(el, co, f0=20.0 + 4 * i, amp=30.0, weights=SRCW))
    rec = media.surface_receivers(2, 9)
    dt = s.dt_ref
    inv = ElasticInversion(s, shots, rec, nsteps, dt, ckpt_every=ckpt_every, rec_field=rec_field,
                           tau_v=TAU[0], tau_s=TAU[1])
    d = inv.observe(m_true)
    m0 = tuple(np.full_like(a, float(np.mean(a))).astype(np.float32) for a in m_true)
    return P, s, inv, d, m0, shots


def host_gradient(P, inv, d, m0, shots):
    oo = orc.OracleOps(P.ops, np.float64)
    adj = ae.ElasAdjointOps(P.ops, P.dm, *TAU)
    m64 = tuple(a.astype(np.float64) for a in m0)
    Jr, gr = 0.0, np.zeros((3,) + m0[0].shape)
    for sh, dd in zip(shots, d):
        src = np.zeros((P.K, P.ops.Np)); src[sh.elems] = sh.coef
        gi, Ji = ae.gradient(oo, P.ops, P.dm, adj, m64, np.zeros((P.K, 9, P.ops.Np)), inv.dt,
                             inv.nsteps, inv.rec_elems, inv.rec_B, dd.astype(np.float64),
                             inv.rec_field, src, SRCW, sh.f0, sh.amp)
        Jr += Ji
        gr += gi
    return Jr, gr


@pytest.mark.parametrize("N,rec_field", [(1, 0), (2, 4), (3, 0), (3, 8), (4, 1), (5, 0), (6, 3)])
def test_elastic_gradient_matches_fp64_host_adjoint(N, rec_field):
    print(f"case elastic adjoint N={N} rec_field={rec_field}", flush=True)
    P, s, inv, d, m0, shots = setup(N, n=2 if N >= 5 else 3, rec_field=rec_field)
    J, g = inv.misfit_and_gradient(m0, d)
    Jr, gr = host_gradient(P, inv, d, m0, shots)
    assert J > 0 and g.shape == gr.shape
    errs = [rel_l2(g[i], gr[i]) for i in range(3)]
    print(f"elastic adjoint N={N} field {rec_field}: J rel {abs(J - Jr) / Jr:.2e}, "
          f"gradient rel L2 {errs}", flush=True)
    # J is small against the traces (the two models differ little), so the FP32 forward's
    # roundoff in the residual shows at 1e-5 relative in J
    assert abs(J - Jr) < 1e-4 * Jr, (J, Jr)
    # measured 3e-6 .. 5e-5 on MI355X: the FP32 forward's roundoff against the small residuals,
    # the same size as J's relative error in each case
    assert max(errs) < 1e-4, errs
    s.close()


def test_elastic_checkpoints_exact_and_fd_agrees():
    print("case elastic adjoint checkpoints + FD", flush=True)
    P, s, inv, d, m0, shots = setup(N=3, nshots=1)
    J1, g1 = inv.misfit_and_gradient(m0, d)            # all steps cached (automatic budget)
    s.set_option(_lib.OPT_ADJ_CACHE_MB, -1)             # no cache: checkpoint every 7, recompute
    J2, g2 = inv.misfit_and_gradient(m0, d)
    inv.ckpt_every = 40
    J3, g3 = inv.misfit_and_gradient(m0, d)
    s.set_option(_lib.OPT_ADJ_CACHE_MB, 0)
    for J, g in ((J2, g2), (J3, g3)):
        assert abs(J1 - J) <= 1e-6 * J1 and rel_l2(g, g1) < 1e-5
    assert abs(inv.misfit(m0, d) - J1) < 1e-6 * J1
    rng = np.random.default_rng(4)
    eps = 1e-3   # FP32 forward: larger step than the FP64 check, same central formula
    errs = []
    for which in range(3):
        v = rng.standard_normal(m0[0].shape).astype(np.float32)
        mp, mm = list(m0), list(m0)
        mp[which] = m0[which] + eps * v
        mm[which] = m0[which] - eps * v
        fd = (inv.misfit(tuple(mp), d) - inv.misfit(tuple(mm), d)) / (2 * eps)
        ad = float(np.sum(g1[which] * v))
        errs.append(abs(fd - ad) / abs(ad))
    print(f"elastic adjoint FD rel errors {errs}", flush=True)
    assert max(errs) < 2e-2, errs
    s.close()


def test_elastic_autograd_device_resident():
    """elastic_misfit on a (3, K, Nq) CUDA tensor (1/rho, lambda, mu): the media, observations and
    gradient stay on the GPU; the gradient equals the host-array path's (to roundoff: the
    receiver injection adds atomically); without a gradient only the forward misfit runs."""
    import torch
    from bbwadg_amd.adjoint import elastic_misfit
    print("case elastic autograd device-resident", flush=True)
    P, s, inv, d, m0, shots = setup(N=3, nshots=2)
    J, g = inv.misfit_and_gradient(m0, d)
    dd = torch.as_tensor(d, device="cuda")
    t = torch.tensor(np.stack(m0), dtype=torch.float32, device="cuda", requires_grad=True)
    Jt = elastic_misfit(t, inv, dd)
    Jt.backward()
    torch.cuda.synchronize()
    assert t.grad.is_cuda and tuple(t.grad.shape) == g.shape
    assert abs(float(Jt.detach()) - J) <= 1e-12 * J
    assert rel_l2(t.grad.cpu().numpy(), g) < 1e-5
    with torch.no_grad():
        J0 = elastic_misfit(torch.tensor(np.stack(m0), dtype=torch.float32, device="cuda"), inv, dd)
    assert abs(float(J0) - inv.misfit(m0, d)) <= 1e-12 * J
    s.close()
