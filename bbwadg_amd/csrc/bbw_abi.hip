// C ABI of libbbwadg_hip.so: context, uploads, single launches and the LSRK45 step loop.
// The step loop restates Wave_RK / RK_step_WADG_subelem (cpp/src/WaveOKL3d.cpp:2014-2109,
// elas_cpp/src/WaveOKL3d.cpp:1605-1693) on one HIP stream, optionally replayed from a
// captured HIP graph.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <vector>

#include "bbw_internal.h"

using namespace bbw;

namespace {

int fail(bbw_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hipfail(bbw_ctx *c, hipError_t e, const char *what) {
  return fail(c, BBW_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, call)                          \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return hipfail(c, e_, #call); \
  } while (0)

template <typename T>
int dalloc(bbw_ctx *c, T **p, size_t n) {
  if (*p) return BBW_OK;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, BBW_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  return BBW_OK;
}

template <typename T>
int upload(bbw_ctx *c, T **dst, const T *src, size_t n) {
  int rc = dalloc(c, dst, n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BBW_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

size_t nstate(const bbw_ctx *c) { return (size_t)c->K * c->Nf * c->Np; }

// time at which stage i of the step starting at t evaluates the source: t + c_i dt, or t itself
// for every stage when the context reproduces the reference's elastic driver
// (elas_cpp/src/WaveOKL3d.cpp:1675-1680 evaluates the pulse at the step's start time)
double stage_time(const bbw_ctx *c, double t, int i, double dt) {
  return c->src_step_time ? t : t + c->rk4c[i] * dt;
}

double ricker(double t, double f0) {
  const double a = M_PI * f0 * (t - 1.0 / f0);
  return (1.0 - 2.0 * a * a) * std::exp(-a * a);
}

// MFMA A operand in lane order (see bbw_ctx::VqBlk): M x K matrix A, 16-row tiles, blocks of four
// K-steps of 4, 64 lanes x 4 floats per block, zero outside A.
void lane_blocks(int M, int K, const std::function<float(int, int)> &A, std::vector<float> &out) {
  const int MT = (M + 15) / 16, KT = (K + 3) / 4, KB = (KT + 3) / 4;
  out.assign((size_t)MT * KB * 256, 0.f);
  for (int t = 0; t < MT; ++t)
    for (int b = 0; b < KB; ++b)
      for (int lane = 0; lane < 64; ++lane)
        for (int sb = 0; sb < 4; ++sb) {
          const int i = 16 * t + lane % 16, k = 4 * (4 * b + sb) + lane / 16;
          if (i < M && k < K) out[(((size_t)t * KB + b) * 64 + lane) * 4 + sb] = A(i, k);
        }
}

int check_ready(bbw_ctx *c) {
  if (!c->have_ops) return fail(c, BBW_ERR_STATE, "operators not uploaded");
  if (!c->have_mesh) return fail(c, BBW_ERR_STATE, "mesh not uploaded");
  if (!c->have_media) return fail(c, BBW_ERR_STATE, "media not uploaded");
  if (c->plain_impl == BBW_PLAIN_SLICE && !c->have_slice)
    return fail(c, BBW_ERR_STATE, "slice-lift tables not uploaded");
  return BBW_OK;
}

// paths that need the Bernstein operators (not set up in a curvilinear context)
int check_not_curved(bbw_ctx *c, const char *what) {
  if (c->curved)
    return fail(c, BBW_ERR_UNSUPPORTED, std::string(what) + " is not available in a curvilinear context");
  return BBW_OK;
}

// paths whose kernels are instantiated for N <= BBW_MAX_ORDER_TILED only
int check_tiled_order(bbw_ctx *c, const char *what) {
  if (c->N > BBW_MAX_ORDER_TILED)
    return fail(c, BBW_ERR_UNSUPPORTED, std::string(what) + " is compiled for N <= 6");
  return BBW_OK;
}

void reset_graph(bbw_ctx *c) {
  if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
  c->graph = Graph();
}

// One LSRK stage; ftd (device) or ft (host value) is the source time function value.
// more_steps: another step of the same call follows, so when a_0 = 0 the residual left by the
// last stage is never read (the next stage scales it by a_0) and the fused stage skips its store.
int do_stage(bbw_ctx *c, int i, double dt, const float *ftd, float ft, int mode, int projU,
             bool more_steps) {
  const float fa = (float)c->rk4a[i], fb = (float)c->rk4b[i], fdt = (float)dt;
  hipError_t e;
  if (mode == BBW_MODE_CURVED) {
    if ((e = launch_curved_rhs(c)) != hipSuccess) return hipfail(c, e, "curved rhs launch");
    if ((e = launch_update_wadg(c, fa, fb, fdt, ftd, ft)) != hipSuccess)
      return hipfail(c, e, "update launch");
    return BBW_OK;
  }
  if (c->Nf == 4) {
    if (mode == BBW_MODE_FUSED) {
      const int store_res = !(i == 4 && more_steps && c->rk4a[0] == 0.0);
      e = c->fused_impl == BBW_FUSED_MFMA
              ? launch_fused_mfma(c, c->Q[c->cur], c->Q[1 - c->cur], fa, fb, fdt, ftd, ft, projU,
                                  nullptr, store_res)
              : launch_fused(c, c->Q[c->cur], c->Q[1 - c->cur], fa, fb, fdt, ftd, ft, projU);
      if (e != hipSuccess) return hipfail(c, e, "fused stage launch");
      c->cur = 1 - c->cur;
      return BBW_OK;
    }
    if ((e = launch_volume(c)) != hipSuccess) return hipfail(c, e, "volume launch");
    if ((e = launch_surface(c)) != hipSuccess) return hipfail(c, e, "surface launch");
    if ((e = launch_update_wadg(c, fa, fb, fdt, ftd, ft)) != hipSuccess)
      return hipfail(c, e, "update launch");
    return BBW_OK;
  }
  if (mode == BBW_MODE_FUSED) {
    const int store_res = !(i == 4 && more_steps && c->rk4a[0] == 0.0);
    e = launch_fused_elas(c, c->Q[c->cur], c->Q[1 - c->cur], fa, fb, fdt, ftd, ft, store_res);
    if (e != hipSuccess) return hipfail(c, e, "fused elastic stage launch");
    c->cur = 1 - c->cur;
    return BBW_OK;
  }
  if ((e = launch_volume_elas(c)) != hipSuccess) return hipfail(c, e, "elastic volume launch");
  if ((e = launch_surface_elas(c)) != hipSuccess) return hipfail(c, e, "elastic surface launch");
  if ((e = launch_update_elas(c, fa, fb, fdt, ftd, ft)) != hipSuccess)
    return hipfail(c, e, "elastic update launch");
  return BBW_OK;
}

}  // namespace

extern "C" {

int bbw_create(const bbw_config *cfg, bbw_ctx **out) {
  if (!cfg || !out) return BBW_ERR_ARG;
  *out = nullptr;
  if (!order_supported(cfg->N)) return BBW_ERR_UNSUPPORTED;
  if (cfg->Nfields != 4 && cfg->Nfields != 9) return BBW_ERR_UNSUPPORTED;
  if (cfg->Nfields == 9 && cfg->N > BBW_MAX_ORDER_TILED) return BBW_ERR_UNSUPPORTED;
  if (cfg->K <= 0 || cfg->Nq != nq_of(cfg->N)) return BBW_ERR_ARG;
  if ((double)cfg->K * cfg->Nfields * np_of(cfg->N) >= 2147483647.0) return BBW_ERR_ARG;
  bbw_ctx *c = new bbw_ctx();
  c->N = cfg->N;
  c->Np = np_of(cfg->N);
  c->Nfp = nfp_of(cfg->N);
  c->Nq = cfg->Nq;
  c->Nf = cfg->Nfields;
  // the folded elastic stage is the faster form up to N = 4; at N = 5 its operator blocks no
  // longer stay in registers and the nodal form is faster
  c->elas_fused = c->N <= 4 ? BBW_ELAS_LIFT_CUBATURE : BBW_ELAS_LIFT_NODES;
  c->K = cfg->K;
  c->dev = cfg->device;
  c->KblkV = cfg->KblkV;
  c->KblkS = cfg->KblkS;
  c->KblkU = cfg->KblkU;
  c->KblkF = cfg->KblkF;
  if (c->N > BBW_MAX_ORDER_TILED) c->plain_impl = BBW_PLAIN_SLICE;
  if (hipSetDevice(c->dev) != hipSuccess || hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return BBW_ERR_HIP;
  }
  c->stream = c->own_stream;
  if (energy_scratch_init(c->dev) != hipSuccess) {
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return BBW_ERR_HIP;
  }
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->dev) != hipSuccess || c->ncu <= 0) {
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return BBW_ERR_HIP;
  }
  const size_t n = nstate(c);
  int rc = dalloc(c, &c->Q[0], n);
  if (!rc) rc = dalloc(c, &c->Q[1], n);
  if (!rc) rc = dalloc(c, &c->res, n);
  if (!rc) rc = dalloc(c, &c->rhs, n);
  if (rc) {
    bbw_destroy(c);
    return rc;
  }
  for (float *p : {c->Q[0], c->Q[1], c->res, c->rhs})
    if (hipMemsetAsync(p, 0, n * sizeof(float), c->stream) != hipSuccess) {
      bbw_destroy(c);
      return BBW_ERR_HIP;
    }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    bbw_destroy(c);
    return BBW_ERR_HIP;
  }
  *out = c;
  return BBW_OK;
}

void bbw_destroy(bbw_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  reset_graph(c);
  dfree(c->Dids); dfree(c->Dvals); dfree(c->Fmask); dfree(c->EEL_ids); dfree(c->L0_ids);
  dfree(c->EEL_vals); dfree(c->L0_vals); dfree(c->Vq); dfree(c->Pq); dfree(c->Vq64); dfree(c->Pq64); dfree(c->VqPad); dfree(c->PqPad); dfree(c->EEL_ids4); dfree(c->LiftPad); dfree(c->VqBlk); dfree(c->PqBlk); dfree(c->LiftBlk); dfree(c->VqCBlk); dfree(c->PqTBlk); dfree(c->VqTBlk); dfree(c->LiftTBlk); dfree(c->PqTPad); dfree(c->VqTPad);
  dfree(c->bb_CN); dfree(c->bb_PBlk); dfree(c->c2bb);
  dfree(c->slice_inv); dfree(c->red_ids); dfree(c->red_vals); dfree(c->cEL);
  dfree(c->cv_Vq); dfree(c->cv_Vr); dfree(c->cv_Vs); dfree(c->cv_Vt); dfree(c->cv_Pq); dfree(c->cv_Pr);
  dfree(c->cv_Ps); dfree(c->cv_Pt); dfree(c->cv_Vf); dfree(c->cv_Pf); dfree(c->vgeoq); dfree(c->cv_Jq);
  dfree(c->fgeoq); dfree(c->Jplanar); dfree(c->Klist_curved); dfree(c->Klist_planar);
  dfree(c->cv_Dr); dfree(c->cv_Ds); dfree(c->cv_Dt); dfree(c->cv_Lift);
  dfree(c->DTids); dfree(c->DTvals); dfree(c->Sids); dfree(c->Eids); dfree(c->Evals); dfree(c->Swts); dfree(c->EELT_ids); dfree(c->EELT_vals); dfree(c->L0T_ids);
  dfree(c->L0T_vals); dfree(c->FmaskInv); dfree(c->mapP); dfree(c->lQ); dfree(c->lres);
  dfree(c->Zadj); dfree(c->Hadj); dfree(c->grad); dfree(c->Ybuf); dfree(c->ckpt);
  dfree(c->resid); dfree(c->Q0save); dfree(c->Ycache);
  dfree(c->wq64); dfree(c->vgeo); dfree(c->fgeo); dfree(c->vmapP); dfree(c->J);
  dfree(c->c2q); dfree(c->Jq); dfree(c->rhoinv); dfree(c->lambda); dfree(c->mu);
  dfree(c->Q[0]); dfree(c->Q[1]); dfree(c->res); dfree(c->rhs); dfree(c->src_elems);
  dfree(c->src_coef); dfree(c->src_tiles); dfree(c->rec_elems); dfree(c->rec_basis); dfree(c->rec_trace);
  dfree(c->partial); dfree(c->ftime_win); dfree(c->ft_all); dfree(c->dobs);
  if (c->ft_host) (void)hipHostFree(c->ft_host);
  if (c->ft_evt) (void)hipEventDestroy(c->ft_evt);
  if (c->side_stream) (void)hipStreamSynchronize(c->side_stream);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side_stream) (void)hipStreamDestroy(c->side_stream);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

const char *bbw_last_error(const bbw_ctx *c) { return c ? c->err.c_str() : "null context"; }

int bbw_set_stream(bbw_ctx *c, void *s) {
  if (!c) return BBW_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stream = s ? (hipStream_t)s : c->own_stream;
  reset_graph(c);
  return BBW_OK;
}

int bbw_sync(bbw_ctx *c) {
  if (!c) return BBW_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BBW_OK;
}

int bbw_upload_operators(bbw_ctx *c, const bbw_ops *o) {
  if (!c || !o) return BBW_ERR_ARG;
  if (!o->Fmask || !o->D_ids || !o->Dvals || !o->EEL_ids || !o->EEL_vals || !o->L0_ids ||
      !o->L0_vals || !o->Vq || !o->Pq || !o->wq || !o->rk4a || !o->rk4b || !o->rk4c)
    return fail(c, BBW_ERR_ARG, "null operator array");
  const int Np = c->Np, Nfp = c->Nfp, Nq = c->Nq;
  if (o->EEL_nnz != Nfp + 3 || o->L0_nnz != (Nfp < 7 ? Nfp : 7))
    return fail(c, BBW_ERR_ARG, "EEL_nnz must be Nfp+3 and L0_nnz min(Nfp,7)");
  for (int i = 0; i < 4 * Nfp; ++i)
    if (o->Fmask[i] < 0 || o->Fmask[i] >= Np) return fail(c, BBW_ERR_ARG, "Fmask out of range");
  for (int i = 0; i < 16 * Np; ++i)
    if (o->D_ids[i] < 0 || o->D_ids[i] >= Np) return fail(c, BBW_ERR_ARG, "D_ids out of range");
  for (int i = 0; i < o->EEL_nnz * Np; ++i)
    if (o->EEL_ids[i] < 0 || o->EEL_ids[i] >= 4 * Nfp) return fail(c, BBW_ERR_ARG, "EEL_ids out of range");
  for (int i = 0; i < o->L0_nnz * Nfp; ++i)
    if (o->L0_ids[i] < 0 || o->L0_ids[i] >= Nfp) return fail(c, BBW_ERR_ARG, "L0_ids out of range");
  int rc;
  if ((rc = upload(c, &c->Fmask, (const int *)o->Fmask, (size_t)4 * Nfp))) return rc;
  if ((rc = upload(c, &c->Dids, (const int4 *)o->D_ids, (size_t)4 * Np))) return rc;
  if ((rc = upload(c, &c->Dvals, (const float4 *)o->Dvals, (size_t)Np))) return rc;
  if ((rc = upload(c, &c->EEL_ids, (const int *)o->EEL_ids, (size_t)o->EEL_nnz * Np))) return rc;
  if ((rc = upload(c, &c->EEL_vals, o->EEL_vals, (size_t)o->EEL_nnz * Np))) return rc;
  if ((rc = upload(c, &c->L0_ids, (const int *)o->L0_ids, (size_t)o->L0_nnz * Nfp))) return rc;
  if ((rc = upload(c, &c->L0_vals, o->L0_vals, (size_t)o->L0_nnz * Nfp))) return rc;
  if ((rc = upload(c, &c->Vq, o->Vq, (size_t)Nq * Np))) return rc;
  if ((rc = upload(c, &c->Pq, o->Pq, (size_t)Nq * Np))) return rc;
  {
    const int KJ = 4 * ((Np + 3) / 4), NQP = 16 * ((Nq + 15) / 16), NPP = 16 * ((Np + 15) / 16);
    std::vector<float> vp((size_t)KJ * NQP, 0.f), pp((size_t)NQP * NPP, 0.f);
    std::vector<int> elev_ids;    // the degree elevation E of the factored derivative: node a
    std::vector<float> elev_w;    // takes sum_j elev_w[4a+j] F[elev_ids[4a+j]] (degree N-1 nodes)
    for (int j = 0; j < Np; ++j)
      for (int q = 0; q < Nq; ++q) vp[(size_t)j * NQP + q] = o->Vq[q + (size_t)j * Nq];
    for (int q = 0; q < Nq; ++q)
      for (int n = 0; n < Np; ++n) pp[(size_t)q * NPP + n] = o->Pq[n + (size_t)q * Np];
    dfree(c->VqPad);
    dfree(c->PqPad);
    if ((rc = upload(c, &c->VqPad, vp.data(), vp.size()))) return rc;
    if ((rc = upload(c, &c->PqPad, pp.data(), pp.size()))) return rc;
    std::vector<float> ptp((size_t)KJ * NQP, 0.f), vtp((size_t)NQP * NPP, 0.f);
    for (int nn = 0; nn < Np; ++nn)
      for (int q = 0; q < Nq; ++q) ptp[(size_t)nn * NQP + q] = o->Pq[nn + (size_t)q * Np];
    for (int q = 0; q < Nq; ++q)
      for (int nn = 0; nn < Np; ++nn) vtp[(size_t)q * NPP + nn] = o->Vq[q + (size_t)nn * Nq];
    dfree(c->PqTPad);
    dfree(c->VqTPad);
    if ((rc = upload(c, &c->PqTPad, ptp.data(), ptp.size()))) return rc;
    if ((rc = upload(c, &c->VqTPad, vtp.data(), vtp.size()))) return rc;
    // transposed operators for the adjoint
    std::vector<double> D(4 * (size_t)Np * Np, 0.0);
    for (int m = 0; m < 4; ++m)
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < 4; ++j)
          D[((size_t)m * Np + i) * Np + o->D_ids[(j * Np + i) * 4 + m]] += o->Dvals[i * 4 + j];
    std::vector<int> dti(16 * (size_t)Np, 0);
    std::vector<float> dtv(16 * (size_t)Np, 0.f);
    for (int m = 0; m < 4; ++m)
      for (int cc = 0; cc < Np; ++cc) {
        int w = 0;
        for (int i = 0; i < Np; ++i) {
          const double v = D[((size_t)m * Np + i) * Np + cc];
          if (v != 0.0) {
            if (w == 4) return fail(c, BBW_ERR_ARG, "transposed derivative row has more than 4 entries");
            dti[((size_t)m * Np + cc) * 4 + w] = i;
            dtv[((size_t)m * Np + cc) * 4 + w] = (float)v;
            ++w;
          }
        }
      }
    dfree(c->DTids);
    dfree(c->DTvals);
    if ((rc = upload(c, &c->DTids, (const int4 *)dti.data(), (size_t)4 * Np))) return rc;
    if ((rc = upload(c, &c->DTvals, (const float4 *)dtv.data(), (size_t)4 * Np))) return rc;
    // The Bernstein derivative in factored form for the fused stage: D_m u [a] =
    // sum_j a_j u[a - e_j + e_m] = (elevation of the shifted coefficients u[b + e_m], |b| = N - 1).
    // Multi-indices (i, j, k, l) in the reference's order (l outer, then k, then j; i = N - j - k - l;
    // cpp/src/Basis.cpp:539-566); e_m adds 1 to component m.  The uploaded D must be exactly this
    // operator (checked), which is what StartUp3d builds from VB \ GradBernTet.
    {
      const int N = c->N, Np1 = N * (N + 1) * (N + 2) / 6;
      auto order = [](int M, std::vector<std::array<int, 4>> &ix) {
        ix.clear();
        for (int l = 0; l <= M; ++l)
          for (int k = 0; k <= M - l; ++k)
            for (int j = 0; j <= M - l - k; ++j) ix.push_back({M - j - k - l, j, k, l});
      };
      std::vector<std::array<int, 4>> aN, aM;
      order(N, aN);
      order(N - 1, aM);
      auto find = [](const std::vector<std::array<int, 4>> &ix, const std::array<int, 4> &a) {
        for (size_t n = 0; n < ix.size(); ++n)
          if (ix[n] == a) return (int)n;
        return -1;
      };
      std::vector<int> sid(4 * (size_t)Np1), eid(4 * (size_t)Np);
      std::vector<float> ev(4 * (size_t)Np), sw(4 * (size_t)Np1);
      for (int b = 0; b < Np1; ++b)
        for (int m = 0; m < 4; ++m) {
          std::array<int, 4> a = aM[b];
          ++a[m];
          sid[4 * b + m] = find(aN, a);
          sw[4 * b + m] = (float)(aM[b][m] + 1);
        }
      for (int a = 0; a < Np; ++a) {
        int first = -1;
        for (int j = 0; j < 4; ++j)
          if (aN[a][j] > 0 && first < 0) {
            std::array<int, 4> b = aN[a];
            --b[j];
            first = find(aM, b);
          }
        for (int j = 0; j < 4; ++j) {
          std::array<int, 4> b = aN[a];
          --b[j];
          eid[4 * a + j] = aN[a][j] > 0 ? find(aM, b) : first;   // zero weight: any valid node
          ev[4 * a + j] = (float)aN[a][j];
        }
      }
      for (int m = 0; m < 4; ++m)
        for (int a = 0; a < Np; ++a)
          for (int g = 0; g < Np; ++g) {
            double want = 0.0;
            for (int j = 0; j < 4; ++j) {
              if (aN[a][j] == 0) continue;
              std::array<int, 4> t = aN[a];
              --t[j];
              ++t[m];
              if (find(aN, t) == g) want += aN[a][j];
            }
            if (std::fabs(D[((size_t)m * Np + a) * Np + g] - want) > 1e-4 * N)
              return fail(c, BBW_ERR_ARG, "D_ids/Dvals are not the Bernstein derivative operators of order N");
          }
      dfree(c->Sids);
      dfree(c->Eids);
      dfree(c->Evals);
      dfree(c->Swts);
      if ((rc = upload(c, &c->Swts, (const float4 *)sw.data(), (size_t)Np1))) return rc;
      if ((rc = upload(c, &c->Sids, (const int4 *)sid.data(), (size_t)Np1))) return rc;
      if ((rc = upload(c, &c->Eids, (const int4 *)eid.data(), (size_t)Np))) return rc;
      if ((rc = upload(c, &c->Evals, (const float4 *)ev.data(), (size_t)Np))) return rc;
      elev_ids = eid;
      elev_w = ev;
    }
    const int NfpNf = 4 * Nfp;
    auto transpose_sparse = [&](const int32_t *ids, const float *vals, int rows, int cols, int nnz,
                                std::vector<int> &tid, std::vector<float> &tval, int &width) {
      std::vector<std::vector<std::pair<int, float>>> colv(cols);
      for (int i = 0; i < rows; ++i)
        for (int j = 0; j < nnz; ++j) {
          const float v = vals[i + (size_t)j * rows];
          if (v != 0.f) colv[ids[i + (size_t)j * rows]].push_back({i, v});
        }
      width = 1;
      for (auto &cv : colv) width = (int)cv.size() > width ? (int)cv.size() : width;
      tid.assign((size_t)width * cols, 0);
      tval.assign((size_t)width * cols, 0.f);
      for (int r = 0; r < cols; ++r)
        for (size_t j = 0; j < colv[r].size(); ++j) {
          tid[r + j * cols] = colv[r][j].first;
          tval[r + j * cols] = colv[r][j].second;
        }
    };
    std::vector<int> ti;
    std::vector<float> tv;
    transpose_sparse(o->EEL_ids, o->EEL_vals, Np, NfpNf, o->EEL_nnz, ti, tv, c->WE);
    dfree(c->EELT_ids);
    dfree(c->EELT_vals);
    if ((rc = upload(c, &c->EELT_ids, ti.data(), ti.size()))) return rc;
    if ((rc = upload(c, &c->EELT_vals, tv.data(), tv.size()))) return rc;
    transpose_sparse(o->L0_ids, o->L0_vals, Nfp, Nfp, o->L0_nnz, ti, tv, c->WL);
    dfree(c->L0T_ids);
    dfree(c->L0T_vals);
    if ((rc = upload(c, &c->L0T_ids, ti.data(), ti.size()))) return rc;
    if ((rc = upload(c, &c->L0T_vals, tv.data(), tv.size()))) return rc;
    std::vector<int> finv(3 * (size_t)Np, -1);
    for (int r = 0; r < NfpNf; ++r) {
      const int n = o->Fmask[r];
      int a = 0;
      while (a < 3 && finv[3 * n + a] >= 0) ++a;
      if (a == 3) return fail(c, BBW_ERR_ARG, "a node lies on more than three faces");
      finv[3 * n + a] = r;
    }
    dfree(c->FmaskInv);
    if ((rc = upload(c, &c->FmaskInv, finv.data(), finv.size()))) return rc;
    std::vector<int> e4((size_t)o->EEL_nnz * Np);
    for (size_t i = 0; i < e4.size(); ++i) e4[i] = o->EEL_ids[i] * 4;  // float offset of a float4 slot
    dfree(c->EEL_ids4);
    if ((rc = upload(c, &c->EEL_ids4, e4.data(), e4.size()))) return rc;
    // dense lift per face: LIFT[n][f Nfp + t] = sum over EEL entries (n, f Nfp + s) of
    // EEL(n, f Nfp + s) L0(s, t), accumulated in double from the FP32 factors
    {
      const int KF = 4 * ((Nfp + 3) / 4);
      std::vector<double> lift((size_t)Np * NfpNf, 0.0);
      for (int n = 0; n < Np; ++n)
        for (int j = 0; j < o->EEL_nnz; ++j) {
          const int col = o->EEL_ids[n + (size_t)j * Np];
          const double ev = o->EEL_vals[n + (size_t)j * Np];
          if (ev == 0.0) continue;
          const int f = col / Nfp, s = col - f * Nfp;
          for (int jj = 0; jj < o->L0_nnz; ++jj) {
            const double lv = o->L0_vals[s + (size_t)jj * Nfp];
            if (lv != 0.0) lift[(size_t)n * NfpNf + f * Nfp + o->L0_ids[s + (size_t)jj * Nfp]] += ev * lv;
          }
        }
      std::vector<float> lp((size_t)4 * KF * NPP, 0.f);
      for (int f = 0; f < 4; ++f)
        for (int t = 0; t < Nfp; ++t)
          for (int n = 0; n < Np; ++n)
            lp[((size_t)f * KF + t) * NPP + n] = (float)lift[(size_t)n * NfpNf + f * Nfp + t];
      dfree(c->LiftPad);
      if ((rc = upload(c, &c->LiftPad, lp.data(), lp.size()))) return rc;
      // lane-ordered A operand blocks (see bbw_ctx::VqBlk)
      auto blocks = lane_blocks;
      std::vector<float> vb, pb, lb, lf;
      blocks(Nq, Np, [&](int q, int j) { return o->Vq[q + (size_t)j * Nq]; }, vb);
      blocks(Np, Nq, [&](int n, int q) { return o->Pq[n + (size_t)q * Np]; }, pb);
      for (int f = 0; f < 4; ++f) {
        blocks(Np, Nfp, [&](int n, int t) { return (float)lift[(size_t)n * NfpNf + f * Nfp + t]; }, lf);
        lb.insert(lb.end(), lf.begin(), lf.end());
      }
      // The folded elastic stage's A operand (M = cubature points): Vq (E F + sum_f LIFT_f flux_f)
      // as one product over K = [the degree-(N-1) nodes | face 0 | ... | face 3], each part padded
      // to whole K-steps of 4 (KT1 steps for the Np1 nodes; KF / 4 per face, zero past Nfp).
      // Formed in double from the FP64 Vq when given.
      {
        const int N = c->N, Np1 = N * (N + 1) * (N + 2) / 6, KT1 = (Np1 + 3) / 4, KTL = KF / 4;
        auto vq = [&](int q, int a) { return o->Vq64 ? o->Vq64[q + (size_t)a * Nq] : (double)o->Vq[q + (size_t)a * Nq]; };
        std::vector<double> vqe((size_t)Nq * Np1, 0.0), vql((size_t)Nq * NfpNf, 0.0);
        for (int q = 0; q < Nq; ++q)
          for (int a = 0; a < Np; ++a) {
            const double v = vq(q, a);
            for (int j = 0; j < 4; ++j) vqe[(size_t)q * Np1 + elev_ids[4 * a + j]] += v * elev_w[4 * a + j];
            for (int t = 0; t < NfpNf; ++t) vql[(size_t)q * NfpNf + t] += v * lift[(size_t)a * NfpNf + t];
          }
        std::vector<float> cb;
        blocks(Nq, 4 * (KT1 + 4 * KTL), [&](int q, int col) -> float {
          const int kt = col / 4;
          if (kt < KT1) return col < Np1 ? (float)vqe[(size_t)q * Np1 + col] : 0.f;
          const int f = (kt - KT1) / KTL, t = 4 * ((kt - KT1) % KTL) + col % 4;
          return t < Nfp ? (float)vql[(size_t)q * NfpNf + f * Nfp + t] : 0.f;
        }, cb);
        dfree(c->VqCBlk);
        if ((rc = upload(c, &c->VqCBlk, cb.data(), cb.size()))) return rc;
      }
      dfree(c->VqBlk);
      dfree(c->PqBlk);
      dfree(c->LiftBlk);
      if ((rc = upload(c, &c->VqBlk, vb.data(), vb.size()))) return rc;
      if ((rc = upload(c, &c->PqBlk, pb.data(), pb.size()))) return rc;
      if ((rc = upload(c, &c->LiftBlk, lb.data(), lb.size()))) return rc;
      std::vector<float> ptb, vtb, ltb;
      blocks(Nq, Np, [&](int q, int n) { return o->Pq[n + (size_t)q * Np]; }, ptb);
      blocks(Np, Nq, [&](int n, int q) { return o->Vq[q + (size_t)n * Nq]; }, vtb);
      for (int f = 0; f < 4; ++f) {
        blocks(Nfp, Np, [&](int t, int n) { return (float)lift[(size_t)n * NfpNf + f * Nfp + t]; }, lf);
        ltb.insert(ltb.end(), lf.begin(), lf.end());
      }
      dfree(c->PqTBlk);
      dfree(c->VqTBlk);
      dfree(c->LiftTBlk);
      if ((rc = upload(c, &c->PqTBlk, ptb.data(), ptb.size()))) return rc;
      if ((rc = upload(c, &c->VqTBlk, vtb.data(), vtb.size()))) return rc;
      if ((rc = upload(c, &c->LiftTBlk, ltb.data(), ltb.size()))) return rc;
    }
  }
  std::vector<double> v64((size_t)Nq * Np);
  for (size_t i = 0; i < v64.size(); ++i) v64[i] = o->Vq64 ? o->Vq64[i] : (double)o->Vq[i];
  // FP64 projection for the energy reductions: Pq64 = MB^-1 Vq^T W with MB = Vq^T W Vq (the
  // mass matrix, exact for the degree 2N + 1 rule), by a Cholesky factorisation of MB
  std::vector<double> p64((size_t)Np * Nq);
  if (c->N <= BBW_MAX_ORDER_TILED) {
    std::vector<double> L((size_t)Np * Np, 0.0);
    for (int i = 0; i < Np; ++i)
      for (int j = 0; j <= i; ++j) {
        double m = 0.0;
        for (int q = 0; q < Nq; ++q) m += v64[q + (size_t)i * Nq] * o->wq[q] * v64[q + (size_t)j * Nq];
        for (int t = 0; t < j; ++t) m -= L[(size_t)i * Np + t] * L[(size_t)j * Np + t];
        if (i == j) {
          if (!(m > 0.0)) return fail(c, BBW_ERR_ARG, "Vq^T W Vq is not positive definite");
          L[(size_t)i * Np + i] = std::sqrt(m);
        } else {
          L[(size_t)i * Np + j] = m / L[(size_t)j * Np + j];
        }
      }
    std::vector<double> x(Np);
    for (int q = 0; q < Nq; ++q) {
      for (int i = 0; i < Np; ++i) x[i] = v64[q + (size_t)i * Nq] * o->wq[q];
      for (int i = 0; i < Np; ++i) {
        for (int t = 0; t < i; ++t) x[i] -= L[(size_t)i * Np + t] * x[t];
        x[i] /= L[(size_t)i * Np + i];
      }
      for (int i = Np - 1; i >= 0; --i) {
        for (int t = i + 1; t < Np; ++t) x[i] -= L[(size_t)t * Np + i] * x[t];
        x[i] /= L[(size_t)i * Np + i];
      }
      for (int i = 0; i < Np; ++i) p64[i + (size_t)q * Np] = x[i];
    }
  }
  dfree(c->Vq64);
  dfree(c->Pq64);
  dfree(c->wq64);
  if ((rc = upload(c, &c->Vq64, v64.data(), v64.size()))) return rc;
  if ((rc = upload(c, &c->Pq64, p64.data(), p64.size()))) return rc;
  if ((rc = upload(c, &c->wq64, o->wq, (size_t)Nq))) return rc;
  std::memcpy(c->rk4a, o->rk4a, 5 * sizeof(double));
  std::memcpy(c->rk4b, o->rk4b, 5 * sizeof(double));
  std::memcpy(c->rk4c, o->rk4c, 6 * sizeof(double));
  c->have_ops = true;
  reset_graph(c);
  return BBW_OK;
}

int bbw_upload_slice_lift(bbw_ctx *c, const int32_t *slice_ids, const int32_t *red_ids,
                          const float *red_vals, const float *cEL) {
  if (!c || !slice_ids || !red_ids || !red_vals || !cEL) return BBW_ERR_ARG;
  if (c->Nf != 4) return fail(c, BBW_ERR_UNSUPPORTED, "the slice-by-slice lift is acoustic");
  if (!c->have_ops) return fail(c, BBW_ERR_STATE, "upload the operators first (Fmask)");
  const int N = c->N, Np = c->Np, Nfp = c->Nfp;
  std::vector<int> fm((size_t)4 * Nfp);
  HIPCHK(c, hipMemcpy(fm.data(), c->Fmask, fm.size() * sizeof(int), hipMemcpyDeviceToHost));
  // inverse table: node n, face f -> (row of n in f's slices) | (slice << 16); every face's
  // rows must be a permutation of the nodes, slice 0 the face in Fmask order
  std::vector<int> inv((size_t)4 * Np, -1);
  for (int f = 0; f < 4; ++f) {
    int sk = 0;
    for (int s = 0; s <= N; ++s) {
      const int nn = (N - s + 1) * (N - s + 2) / 2;
      for (int r = 0; r < nn; ++r, ++sk) {
        const int n = slice_ids[4 * sk + f];
        if (n < 0 || n >= Np || inv[4 * n + f] >= 0)
          return fail(c, BBW_ERR_ARG, "slice_ids: a face's rows are not a permutation of the nodes");
        if (s == 0 && n != fm[(size_t)f * Nfp + r])
          return fail(c, BBW_ERR_ARG, "slice_ids: slice 0 must be the face nodes in Fmask order");
        inv[4 * n + f] = sk | (s << 16);
      }
    }
  }
  // reduction rows of slice s index slice s - 1
  int row = 0;
  for (int s = 1; s <= N; ++s) {
    const int nn = (N - s + 1) * (N - s + 2) / 2, nprev = (N - s + 2) * (N - s + 3) / 2;
    for (int r = 0; r < nn; ++r, ++row)
      for (int j = 0; j < 3; ++j)
        if (red_ids[3 * row + j] < 0 || red_ids[3 * row + j] >= nprev)
          return fail(c, BBW_ERR_ARG, "red_ids out of range of the previous slice");
  }
  dfree(c->slice_inv);
  dfree(c->red_ids);
  dfree(c->red_vals);
  dfree(c->cEL);
  c->have_slice = false;
  int rc;
  if ((rc = upload(c, (int **)&c->slice_inv, inv.data(), inv.size()))) return rc;
  if ((rc = upload(c, &c->red_ids, red_ids, (size_t)3 * (Np - Nfp)))) return rc;
  if ((rc = upload(c, &c->red_vals, red_vals, (size_t)3 * (Np - Nfp)))) return rc;
  if ((rc = upload(c, &c->cEL, cEL, (size_t)N + 1))) return rc;
  c->have_slice = true;
  reset_graph(c);
  return BBW_OK;
}

int bbw_upload_curved(bbw_ctx *c, const bbw_curved_ops *o, const float *vgeoq, const float *Jq,
                      const float *fgeoq, const float *Jplanar, int Kcurved,
                      const int32_t *KlistCurved, int Kplanar, const int32_t *KlistPlanar) {
  if (!c || !o || !vgeoq || !Jq || !fgeoq || !Jplanar || Kcurved < 0 || Kplanar < 0 ||
      (Kcurved && !KlistCurved) || (Kplanar && !KlistPlanar))
    return BBW_ERR_ARG;
  if (!o->Vq || !o->Vrq || !o->Vsq || !o->Vtq || !o->Pq || !o->Prq || !o->Psq || !o->Ptq ||
      !o->VfqFace || !o->Pfq || !o->Dr || !o->Ds || !o->Dt || !o->LIFT || !o->Fmask || !o->wq ||
      !o->rk4a || !o->rk4b || !o->rk4c)
    return fail(c, BBW_ERR_ARG, "null curvilinear operator array");
  if (c->Nf != 4) return fail(c, BBW_ERR_UNSUPPORTED, "the curvilinear path is acoustic");
  if (int rc = check_tiled_order(c, "the curvilinear path")) return rc;
  if (c->have_ops && !c->curved)
    return fail(c, BBW_ERR_STATE, "Bernstein operators already uploaded: a curvilinear context is a new context");
  const int N = c->N, Np = c->Np, Nfp = c->Nfp, Nq = c->Nq, K = c->K;
  if (o->Nfq != curved_nfq(N)) return fail(c, BBW_ERR_ARG, "Nfq must be the degree 2N+1 triangle rule's size");
  const int Nfq = o->Nfq;
  if ((double)K * 4 * Np * 4 >= 2147483648.0 || (double)K * 9 * Nq * 4 >= 4294967296.0 ||
      (double)K * 16 * Nfq * 4 >= 4294967296.0)
    return fail(c, BBW_ERR_ARG, "mesh too large for 32-bit offsets of the curvilinear kernels");
  if (Kcurved + Kplanar != K) return fail(c, BBW_ERR_ARG, "the element lists must partition the mesh");
  std::vector<char> seen(K, 0);
  for (int i = 0; i < Kcurved + Kplanar; ++i) {
    const int e = i < Kcurved ? KlistCurved[i] : KlistPlanar[i - Kcurved];
    if (e < 0 || e >= K || seen[e]) return fail(c, BBW_ERR_ARG, "the element lists must partition the mesh");
    seen[e] = 1;
  }
  for (int i = 0; i < 4 * Nfp; ++i)
    if (o->Fmask[i] < 0 || o->Fmask[i] >= Np) return fail(c, BBW_ERR_ARG, "Fmask out of range");
  const int NQP = 16 * ((Nq + 15) / 16), NFQP = 16 * ((Nfq + 15) / 16);
  std::vector<float> b[14];
  const float *V4[4] = {o->Vq, o->Vrq, o->Vsq, o->Vtq};
  const float *P4[4] = {o->Pq, o->Prq, o->Psq, o->Ptq};
  for (int m = 0; m < 4; ++m) {
    const float *V = V4[m], *P = P4[m];
    lane_blocks(Nq, Np, [&](int q, int j) { return V[q + (size_t)j * Nq]; }, b[m]);
    lane_blocks(Np, NQP, [&](int n, int q) { return q < Nq ? P[n + (size_t)q * Np] : 0.f; }, b[4 + m]);
  }
  lane_blocks(Nfq, Nfp, [&](int q, int j) { return o->VfqFace[q + (size_t)j * Nfq]; }, b[8]);
  lane_blocks(Np, 4 * NFQP, [&](int n, int k) {
    const int f = k / NFQP, j = k - f * NFQP;
    return j < Nfq ? o->Pfq[n + (size_t)(f * Nfq + j) * Np] : 0.f;
  }, b[9]);
  const float *D3[3] = {o->Dr, o->Ds, o->Dt};
  for (int m = 0; m < 3; ++m) {
    const float *D = D3[m];
    lane_blocks(Np, Np, [&](int n, int j) { return D[n + (size_t)j * Np]; }, b[10 + m]);
  }
  lane_blocks(Np, 4 * Nfp, [&](int n, int j) { return o->LIFT[n + (size_t)j * Np]; }, b[13]);
  float **dst[14] = {&c->cv_Vq, &c->cv_Vr, &c->cv_Vs, &c->cv_Vt, &c->cv_Pq, &c->cv_Pr,
                     &c->cv_Ps, &c->cv_Pt, &c->cv_Vf, &c->cv_Pf, &c->cv_Dr, &c->cv_Ds,
                     &c->cv_Dt, &c->cv_Lift};
  int rc;
  for (int m = 0; m < 14; ++m) {
    dfree(*dst[m]);
    if ((rc = upload(c, dst[m], b[m].data(), b[m].size()))) return rc;
  }
  // the WADG update's operators: the same Vq and Pq (flat copies and lane-ordered blocks)
  std::vector<float> vb, pb;
  lane_blocks(Nq, Np, [&](int q, int j) { return o->Vq[q + (size_t)j * Nq]; }, vb);
  lane_blocks(Np, Nq, [&](int n, int q) { return o->Pq[n + (size_t)q * Np]; }, pb);
  dfree(c->Vq); dfree(c->Pq); dfree(c->VqBlk); dfree(c->PqBlk); dfree(c->Fmask);
  if ((rc = upload(c, &c->Vq, o->Vq, (size_t)Nq * Np))) return rc;
  if ((rc = upload(c, &c->Pq, o->Pq, (size_t)Np * Nq))) return rc;
  if ((rc = upload(c, &c->VqBlk, vb.data(), vb.size()))) return rc;
  if ((rc = upload(c, &c->PqBlk, pb.data(), pb.size()))) return rc;
  if ((rc = upload(c, &c->Fmask, o->Fmask, (size_t)4 * Nfp))) return rc;
  dfree(c->vgeoq); dfree(c->cv_Jq); dfree(c->fgeoq); dfree(c->Jplanar);
  dfree(c->Klist_curved); dfree(c->Klist_planar);
  if ((rc = upload(c, &c->vgeoq, vgeoq, (size_t)K * 9 * Nq))) return rc;
  if ((rc = upload(c, &c->cv_Jq, Jq, (size_t)K * Nq))) return rc;
  if ((rc = upload(c, &c->fgeoq, fgeoq, (size_t)K * 16 * Nfq))) return rc;
  if ((rc = upload(c, &c->Jplanar, Jplanar, (size_t)K))) return rc;
  if (Kcurved && (rc = upload(c, &c->Klist_curved, KlistCurved, (size_t)Kcurved))) return rc;
  if (Kplanar && (rc = upload(c, &c->Klist_planar, KlistPlanar, (size_t)Kplanar))) return rc;
  std::memcpy(c->rk4a, o->rk4a, 5 * sizeof(double));
  std::memcpy(c->rk4b, o->rk4b, 5 * sizeof(double));
  std::memcpy(c->rk4c, o->rk4c, 6 * sizeof(double));
  if (!c->side_stream) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  c->Nfq = Nfq;
  c->Kcurved = Kcurved;
  c->Kplanar = Kplanar;
  c->plain_impl = BBW_PLAIN_TILED;   // the WADG update on MFMA (with 1/J from bbw_upload_media)
  c->curved = true;
  c->have_ops = true;
  reset_graph(c);
  return BBW_OK;
}

int bbw_upload_mesh(bbw_ctx *c, const float *vgeo, const float *fgeo, const int32_t *vmapP,
                    const double *J) {
  if (!c || !vgeo || !fgeo || !vmapP || !J) return BBW_ERR_ARG;
  const size_t K = c->K, NfpNf = 4 * (size_t)c->Nfp;
  const long long lim = (long long)K * c->Nf * c->Np;
  for (size_t k = 0; k < K; ++k)
    for (size_t i = 0; i < NfpNf; ++i) {
      const long long v = vmapP[k * NfpNf + i];
      // every exterior trace index must address field 0 of a node inside Q
      if (v < 0 || v >= lim || (v % ((long long)c->Nf * c->Np)) >= c->Np)
        return fail(c, BBW_ERR_ARG, "vmapP entry out of range");
    }
  int rc;
  if ((rc = upload(c, &c->vgeo, vgeo, K * 9))) return rc;
  if ((rc = upload(c, &c->fgeo, fgeo, K * 16))) return rc;
  if ((rc = upload(c, &c->vmapP, (const int *)vmapP, K * NfpNf))) return rc;
  if ((rc = upload(c, &c->J, J, K))) return rc;
  c->have_mesh = true;
  reset_graph(c);
  return BBW_OK;
}

int bbw_upload_media(bbw_ctx *c, const float *c2q, const float *Jq) {
  if (!c || !c2q) return BBW_ERR_ARG;
  if (c->Nf != 4) return fail(c, BBW_ERR_ARG, "acoustic media on an elastic context");
  const size_t n = (size_t)c->K * c->Nq;
  int rc;
  if ((rc = upload(c, &c->c2q, c2q, n))) return rc;
  if (Jq) {
    if ((rc = upload(c, &c->Jq, Jq, n))) return rc;
  } else {
    dfree(c->Jq);
  }
  c->have_media = true;
  reset_graph(c);
  return BBW_OK;
}

int bbw_upload_media_device(bbw_ctx *c, const float *c2q_dev) {
  if (!c || !c2q_dev) return BBW_ERR_ARG;
  if (c->Nf != 4) return fail(c, BBW_ERR_ARG, "acoustic media on an elastic context");
  const size_t n = (size_t)c->K * c->Nq;
  int rc = dalloc(c, &c->c2q, n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->c2q, c2q_dev, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  dfree(c->Jq);
  c->have_media = true;
  return BBW_OK;
}

int bbw_upload_elastic_media(bbw_ctx *c, const float *rhoinvq, const float *lambdaq,
                             const float *muq, float tau_v, float tau_s) {
  if (!c || !rhoinvq || !lambdaq || !muq) return BBW_ERR_ARG;
  if (c->Nf != 9) return fail(c, BBW_ERR_ARG, "elastic media on an acoustic context");
  const size_t n = (size_t)c->K * c->Nq;
  int rc;
  if ((rc = upload(c, &c->rhoinv, rhoinvq, n))) return rc;
  if ((rc = upload(c, &c->lambda, lambdaq, n))) return rc;
  if ((rc = upload(c, &c->mu, muq, n))) return rc;
  c->tau_v = tau_v;
  c->tau_s = tau_s;
  c->have_media = true;
  reset_graph(c);
  return BBW_OK;
}

int bbw_upload_elastic_media_device(bbw_ctx *c, const float *rhoinvq_dev, const float *lambdaq_dev,
                                    const float *muq_dev, float tau_v, float tau_s) {
  if (!c || !rhoinvq_dev || !lambdaq_dev || !muq_dev) return BBW_ERR_ARG;
  if (c->Nf != 9) return fail(c, BBW_ERR_ARG, "elastic media on an acoustic context");
  const size_t n = (size_t)c->K * c->Nq;
  int rc;
  if ((rc = dalloc(c, &c->rhoinv, n)) || (rc = dalloc(c, &c->lambda, n)) || (rc = dalloc(c, &c->mu, n)))
    return rc;
  HIPCHK(c, hipMemcpyAsync(c->rhoinv, rhoinvq_dev, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->lambda, lambdaq_dev, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->mu, muq_dev, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  // the penalties are kernel arguments: captured graphs hold the old ones
  if (tau_v != c->tau_v || tau_s != c->tau_s) reset_graph(c);
  c->tau_v = tau_v;
  c->tau_s = tau_s;
  c->have_media = true;
  return BBW_OK;
}

int bbw_set_source(bbw_ctx *c, int nsrc, const int32_t *elems, const float *coef,
                   const float *weights, double f0, double amp) {
  if (!c || nsrc < 0) return BBW_ERR_ARG;
  dfree(c->src_elems);
  dfree(c->src_coef);
  dfree(c->src_tiles);
  c->nsrc = 0;
  c->src_first = 0;
  c->src_last = -1;
  c->src_tfirst = 0;
  c->src_tlast = -1;
  reset_graph(c);
  if (nsrc == 0) return BBW_OK;
  if (!elems || !coef || !weights || !(f0 > 0)) return fail(c, BBW_ERR_ARG, "bad source");
  for (int i = 0; i < nsrc; ++i) {
    if (elems[i] < 0 || elems[i] >= c->K) return fail(c, BBW_ERR_ARG, "source element out of range");
    if (i && elems[i] <= elems[i - 1]) return fail(c, BBW_ERR_ARG, "source elements must ascend");
  }
  int rc;
  if ((rc = upload(c, &c->src_elems, (const int *)elems, (size_t)nsrc))) return rc;
  if ((rc = upload(c, &c->src_coef, coef, (size_t)nsrc * c->Np))) return rc;
  const std::vector<int2> tiles = src_tile_table(elems, nsrc);
  if ((rc = upload(c, &c->src_tiles, tiles.data(), tiles.size()))) return rc;
  c->nsrc = nsrc;
  c->src_first = elems[0];
  c->src_last = elems[nsrc - 1];
  c->src_tfirst = elems[0] >> 4;
  c->src_tlast = elems[nsrc - 1] >> 4;
  for (int f = 0; f < 9; ++f) c->src_w[f] = f < c->Nf ? (float)(weights[f] * amp) : 0.f;
  c->src_f0 = f0;
  c->src_amp = amp;
  return BBW_OK;
}

int bbw_set_receivers(bbw_ctx *c, int nrec, const int32_t *elems, const float *basis,
                      int max_steps) {
  if (!c || nrec < 0 || max_steps < 0) return BBW_ERR_ARG;
  dfree(c->rec_elems);
  dfree(c->rec_basis);
  dfree(c->rec_trace);
  c->nrec = 0;
  c->rec_steps = 0;
  c->rec_max_steps = 0;
  reset_graph(c);
  if (nrec == 0) return BBW_OK;
  if (!elems || !basis || max_steps == 0) return fail(c, BBW_ERR_ARG, "bad receivers");
  for (int i = 0; i < nrec; ++i)
    if (elems[i] < 0 || elems[i] >= c->K) return fail(c, BBW_ERR_ARG, "receiver element out of range");
  int rc;
  if ((rc = upload(c, &c->rec_elems, (const int *)elems, (size_t)nrec))) return rc;
  if ((rc = upload(c, &c->rec_basis, basis, (size_t)nrec * c->Np))) return rc;
  if ((rc = dalloc(c, &c->rec_trace, (size_t)nrec * max_steps))) return rc;
  HIPCHK(c, hipMemsetAsync(c->rec_trace, 0, sizeof(float) * (size_t)nrec * max_steps, c->stream));
  c->nrec = nrec;
  c->rec_max_steps = max_steps;
  return BBW_OK;
}

int bbw_get_traces(bbw_ctx *c, float *out, int nsteps) {
  if (!c || !out || nsteps < 0 || nsteps > c->rec_max_steps) return BBW_ERR_ARG;
  if (c->nrec == 0 || nsteps == 0) return BBW_OK;
  HIPCHK(c, hipMemcpyAsync(out, c->rec_trace, sizeof(float) * (size_t)c->nrec * nsteps,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BBW_OK;
}

static int copy_state(bbw_ctx *c, float *dev, const float *hin, float *hout) {
  if (!c || !dev || (!hin && !hout)) return BBW_ERR_ARG;
  const size_t bytes = nstate(c) * sizeof(float);
  if (hin) HIPCHK(c, hipMemcpyAsync(dev, hin, bytes, hipMemcpyHostToDevice, c->stream));
  else HIPCHK(c, hipMemcpyAsync(hout, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BBW_OK;
}

int bbw_set_state(bbw_ctx *c, const float *Q) {
  if (!c) return BBW_ERR_ARG;
  if (!Q) {
    HIPCHK(c, hipMemsetAsync(c->Q[c->cur], 0, nstate(c) * sizeof(float), c->stream));
    return BBW_OK;
  }
  return copy_state(c, c->Q[c->cur], Q, nullptr);
}
int bbw_get_state(bbw_ctx *c, float *Q) { return c ? copy_state(c, c->Q[c->cur], nullptr, Q) : BBW_ERR_ARG; }
int bbw_set_res(bbw_ctx *c, const float *r) { return c ? copy_state(c, c->res, r, nullptr) : BBW_ERR_ARG; }
int bbw_get_res(bbw_ctx *c, float *r) { return c ? copy_state(c, c->res, nullptr, r) : BBW_ERR_ARG; }
int bbw_set_rhs(bbw_ctx *c, const float *r) { return c ? copy_state(c, c->rhs, r, nullptr) : BBW_ERR_ARG; }
int bbw_get_rhs(bbw_ctx *c, float *r) { return c ? copy_state(c, c->rhs, nullptr, r) : BBW_ERR_ARG; }

int bbw_curved_rhs(bbw_ctx *c) {
  if (!c) return BBW_ERR_ARG;
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->curved) return fail(c, BBW_ERR_STATE, "curvilinear operators not uploaded");
  hipError_t e = launch_curved_rhs(c);
  return e == hipSuccess ? BBW_OK : hipfail(c, e, "curved rhs launch");
}

int bbw_volume(bbw_ctx *c) {
  if (!c) return BBW_ERR_ARG;
  if (!c->have_ops || !c->have_mesh) return fail(c, BBW_ERR_STATE, "operators/mesh not uploaded");
  if (int rc = check_not_curved(c, "bbw_volume")) return rc;
  hipError_t e = c->Nf == 4 ? launch_volume(c) : launch_volume_elas(c);
  return e == hipSuccess ? BBW_OK : hipfail(c, e, "volume launch");
}

int bbw_surface(bbw_ctx *c) {
  if (!c) return BBW_ERR_ARG;
  if (!c->have_ops || !c->have_mesh) return fail(c, BBW_ERR_STATE, "operators/mesh not uploaded");
  if (int rc = check_not_curved(c, "bbw_surface")) return rc;
  hipError_t e = c->Nf == 4 ? launch_surface(c) : launch_surface_elas(c);
  return e == hipSuccess ? BBW_OK : hipfail(c, e, "surface launch");
}

int bbw_update_wadg(bbw_ctx *c, float fa, float fb, float fdt, float ftime) {
  if (!c) return BBW_ERR_ARG;
  int rc = check_ready(c);
  if (rc) return rc;
  hipError_t e = c->Nf == 4 ? launch_update_wadg(c, fa, fb, fdt, nullptr, ftime)
                            : launch_update_elas(c, fa, fb, fdt, nullptr, ftime);
  return e == hipSuccess ? BBW_OK : hipfail(c, e, "update launch");
}

int bbw_upload_bbwadg(bbw_ctx *c, const int32_t *subtet_ids, const double *CNscale,
                      const double *invC2Nscale, const int32_t *Ei_ids4, const double *Ei_vals4,
                      const double *cj, const float *c2_bb) {
  if (!c || !subtet_ids || !CNscale || !invC2Nscale || !Ei_ids4 || !Ei_vals4 || !cj || !c2_bb)
    return BBW_ERR_ARG;
  if (c->Nf != 4) return fail(c, BBW_ERR_UNSUPPORTED, "the exact BB-WADG update is acoustic");
  if (int rc = check_tiled_order(c, "the exact BB-WADG update")) return rc;
  if (int rc = check_not_curved(c, "the exact BB-WADG update")) return rc;
  const int N = c->N, Np = c->Np, N2 = 2 * N, N2p = np_of(N2);
  auto npd = [](int d) { return (d + 1) * (d + 2) * (d + 3) / 6; };
  // the kernel forms the degree-2N index of a + b arithmetically, Np(2N) - Np(2N - l) +
  // k (2N - l + 1) - k (k - 1) / 2 + j for (j, k, l) = (j, k, l)_a + (j, k, l)_b in the reference
  // order (l outer, k, j inner); subtet_ids must be that table
  {
    std::vector<std::array<int, 3>> jkl;
    for (int l = 0; l <= N; ++l)
      for (int k = 0; k <= N - l; ++k)
        for (int j = 0; j <= N - l - k; ++j) jkl.push_back({j, k, l});
    for (int a = 0; a < Np; ++a)
      for (int b = 0; b < Np; ++b) {
        const int j = jkl[a][0] + jkl[b][0], k = jkl[a][1] + jkl[b][1], l = jkl[a][2] + jkl[b][2];
        const int g = N2p - npd(N2 - l) + k * (N2 - l + 1) - k * (k - 1) / 2 + j;
        if (subtet_ids[(size_t)a * Np + b] != g)
          return fail(c, BBW_ERR_ARG, "subtet_ids is not the degree-2N index of a + b");
      }
  }
  for (int i = 0; i < N2; ++i)
    for (int r = 0; r < npd(i + 1); ++r)
      for (int w = 0; w < 4; ++w) {
        const int id = Ei_ids4[((size_t)i * N2p + r) * 4 + w];
        if (id < 0 || id >= npd(i)) return fail(c, BBW_ERR_ARG, "Ei_ids4 out of range");
      }
  // dense E_{a -> b} = E_{b-1} ... E_a from the one-step sparse elevations, FP64
  auto elev = [&](int a, int b) {
    std::vector<double> M((size_t)npd(a) * npd(a), 0.0);
    for (int i = 0; i < npd(a); ++i) M[(size_t)i * npd(a) + i] = 1.0;
    for (int d = a; d < b; ++d) {
      const int rows = npd(d + 1), cols = npd(a);
      std::vector<double> R((size_t)rows * cols, 0.0);
      for (int r = 0; r < rows; ++r)
        for (int w = 0; w < 4; ++w) {
          const size_t o = ((size_t)d * N2p + r) * 4 + w;
          const double v = Ei_vals4[o];
          if (v == 0.0) continue;
          const int m = Ei_ids4[o];
          for (int cc = 0; cc < cols; ++cc) R[(size_t)r * cols + cc] += v * M[(size_t)m * cols + cc];
        }
      M.swap(R);
    }
    return M;  // npd(b) x npd(a), row-major
  };
  // P' = (sum_j c_j E_{N-j}^N (E_{N-j}^2N)^T) diag(1/C(2N,g)), Np x N2p
  std::vector<double> P((size_t)Np * N2p, 0.0);
  for (int j = 0; j <= N; ++j) {
    const int a = N - j, na = npd(a);
    const std::vector<double> EN = elev(a, N), E2 = elev(a, N2);
    for (int i = 0; i < Np; ++i)
      for (int g = 0; g < N2p; ++g) {
        double acc = 0.0;
        for (int t = 0; t < na; ++t) acc += EN[(size_t)i * na + t] * E2[(size_t)g * na + t];
        P[(size_t)i * N2p + g] += cj[j] * acc;
      }
  }
  std::vector<float> pb, cn(Np);
  lane_blocks(Np, N2p, [&](int i, int g) { return (float)(P[(size_t)i * N2p + g] * invC2Nscale[g]); }, pb);
  for (int i = 0; i < Np; ++i) cn[i] = (float)CNscale[i];
  int rc;
  dfree(c->bb_CN);
  dfree(c->bb_PBlk);
  dfree(c->c2bb);
  if ((rc = upload(c, &c->bb_CN, cn.data(), cn.size()))) return rc;
  if ((rc = upload(c, &c->bb_PBlk, pb.data(), pb.size()))) return rc;
  if ((rc = upload(c, &c->c2bb, c2_bb, (size_t)c->K * Np))) return rc;
  return BBW_OK;
}

int bbw_update_bbwadg(bbw_ctx *c) {
  if (!c) return BBW_ERR_ARG;
  if (!c->bb_PBlk || !c->c2bb) return fail(c, BBW_ERR_STATE, "exact BB-WADG operators not uploaded");
  hipError_t e = launch_update_bbwadg(c);
  return e == hipSuccess ? BBW_OK : hipfail(c, e, "exact BB-WADG launch");
}

int bbw_update(bbw_ctx *c, float fa, float fb, float fdt) {
  if (!c) return BBW_ERR_ARG;
  hipError_t e = launch_update(c, fa, fb, fdt);
  return e == hipSuccess ? BBW_OK : hipfail(c, e, "rk update launch");
}

int bbw_stage_fused(bbw_ctx *c, float fa, float fb, float fdt, float ftime, int projU) {
  if (!c) return BBW_ERR_ARG;
  int rc = check_ready(c);
  if (rc || (rc = check_tiled_order(c, "the fused stage")) || (rc = check_not_curved(c, "the fused stage")))
    return rc;
  if (c->Nf == 9) {
    hipError_t e = launch_fused_elas(c, c->Q[c->cur], c->Q[1 - c->cur], fa, fb, fdt, nullptr, ftime);
    if (e != hipSuccess) return hipfail(c, e, "fused elastic stage launch");
    c->cur = 1 - c->cur;
    return BBW_OK;
  }
  hipError_t e = c->fused_impl == BBW_FUSED_MFMA
                     ? launch_fused_mfma(c, c->Q[c->cur], c->Q[1 - c->cur], fa, fb, fdt, nullptr, ftime, projU)
                     : launch_fused(c, c->Q[c->cur], c->Q[1 - c->cur], fa, fb, fdt, nullptr, ftime, projU);
  if (e != hipSuccess) return hipfail(c, e, "fused stage launch");
  c->cur = 1 - c->cur;
  return BBW_OK;
}

int bbw_step(bbw_ctx *c, int nsteps, double dt, double t0, int mode, int flags) {
  if (!c || nsteps < 0 || !(dt > 0) ||
      (mode != BBW_MODE_PLAIN && mode != BBW_MODE_FUSED && mode != BBW_MODE_CURVED))
    return BBW_ERR_ARG;
  int rc = check_ready(c);
  if (rc) return rc;
  if ((mode == BBW_MODE_CURVED) != c->curved)
    return fail(c, BBW_ERR_ARG, c->curved ? "a curvilinear context steps with BBW_MODE_CURVED"
                                          : "BBW_MODE_CURVED needs bbw_upload_curved");
  if (mode == BBW_MODE_FUSED && (rc = check_tiled_order(c, "the fused stage"))) return rc;
  if (c->nrec && c->rec_steps + nsteps > c->rec_max_steps)
    return fail(c, BBW_ERR_ARG, "receiver trace buffer too short");
  const bool prepare = (flags & 4) != 0;
  flags &= 3;
  const int projU = flags & 1;
  const bool use_graph = (flags & 2) != 0;
  const int G = 10;  // steps per captured graph (even: the fused ping-pong parity is restored)
  if (prepare && !use_graph) return fail(c, BBW_ERR_ARG, "graph preparation needs the graph flag");

  // stage time-function values for all steps (host FP64 -> float)
  const bool has_src = c->nsrc > 0;
  std::vector<float> ft;
  if (has_src && !prepare) {
    ft.resize((size_t)nsteps * 5);
    for (int s = 0; s < nsteps; ++s)
      for (int i = 0; i < 5; ++i)
        ft[(size_t)s * 5 + i] = (float)ricker(stage_time(c, t0 + s * dt, i, dt), c->src_f0);
  }

  int s = 0;
  if (use_graph && (nsteps >= G || prepare)) {
    // window buffer: the graph nodes read their stage values and write receiver rows there.  It
    // is sized for the current receiver count; a larger one (receivers set after a graph step)
    // reallocates it and recaptures the graph, which holds its addresses.
    const size_t win = (size_t)G * 5 + (size_t)G * (c->nrec > 0 ? c->nrec : 1);
    if (c->ftime_win_cap < win) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      reset_graph(c);
      dfree(c->ftime_win);
      c->ftime_win_cap = 0;
      if ((rc = dalloc(c, &c->ftime_win, win))) return rc;
      c->ftime_win_cap = win;
    }
    float *rec_win = c->ftime_win + G * 5;
    const int parity = mode == BBW_MODE_FUSED ? c->cur : 0;
    Graph &g = c->graph;
    if (!g.exec || g.steps != G || g.mode != mode || g.flags != flags || g.parity != parity || g.dt != dt ||
        g.form != elas_form(c)) {
      reset_graph(c);
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int cur0 = c->cur;
      int crc = BBW_OK;
      for (int j = 0; j < G && !crc; ++j) {
        for (int i = 0; i < 5 && !crc; ++i)
          crc = do_stage(c, i, dt, has_src ? c->ftime_win + j * 5 + i : nullptr, 0.f, mode, projU,
                         j + 1 < G);
        if (!crc && c->nrec) {
          hipError_t e = launch_record(c, c->Q[c->cur], rec_win + (size_t)j * c->nrec);
          if (e != hipSuccess) crc = hipfail(c, e, "record launch");
        }
      }
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &graph);
      c->cur = cur0;
      if (crc) {
        if (graph) (void)hipGraphDestroy(graph);
        return crc;
      }
      if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return hipfail(c, e, "hipGraphInstantiate");
      g.steps = G; g.mode = mode; g.flags = flags; g.parity = parity; g.dt = dt; g.form = elas_form(c);
    }
    if (has_src) {
      // the call's stage values go to the device once, through a persistent pinned staging
      // buffer; the buffers grow only when a call has more steps than every earlier one
      const size_t nft = (size_t)nsteps * 5;
      if (!c->ft_evt) HIPCHK(c, hipEventCreateWithFlags(&c->ft_evt, hipEventDisableTiming));
      if (c->ft_evt_pending) {
        HIPCHK(c, hipEventSynchronize(c->ft_evt));  // the last copy out of ft_host has ended
        c->ft_evt_pending = false;
      }
      if (c->ft_all_cap < nft || c->ft_host_cap < nft) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c->ft_all);
        c->ft_all_cap = 0;
        if ((rc = dalloc(c, &c->ft_all, nft))) return rc;
        c->ft_all_cap = nft;
        if (c->ft_host) (void)hipHostFree(c->ft_host);
        c->ft_host = nullptr;
        c->ft_host_cap = 0;
        HIPCHK(c, hipHostMalloc((void **)&c->ft_host, nft * sizeof(float), hipHostMallocDefault));
        c->ft_host_cap = nft;
      }
    }
    if (prepare) {
      // graph and buffers ready for a call of nsteps steps with these arguments; no step is run
      HIPCHK(c, hipGraphUpload(g.exec, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      return BBW_OK;
    }
    if (has_src) {
      const size_t nft = ft.size();
      std::memcpy(c->ft_host, ft.data(), nft * sizeof(float));
      HIPCHK(c, hipMemcpyAsync(c->ft_all, c->ft_host, nft * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipEventRecord(c->ft_evt, c->stream));
      c->ft_evt_pending = true;
    }
    for (; s + G <= nsteps; s += G) {
      if (has_src)
        HIPCHK(c, hipMemcpyAsync(c->ftime_win, c->ft_all + (size_t)s * 5, sizeof(float) * G * 5,
                                 hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
      if (c->nrec) {
        HIPCHK(c, hipMemcpyAsync(c->rec_trace + (size_t)c->rec_steps * c->nrec, rec_win,
                                 sizeof(float) * G * c->nrec, hipMemcpyDeviceToDevice, c->stream));
        c->rec_steps += G;
      }
    }
  }
  for (; s < nsteps; ++s) {
    for (int i = 0; i < 5; ++i)
      if ((rc = do_stage(c, i, dt, nullptr, has_src ? ft[(size_t)s * 5 + i] : 0.f, mode, projU,
                         s + 1 < nsteps)))
        return rc;
    if (c->nrec) {
      hipError_t e = launch_record(c, c->Q[c->cur], c->rec_trace + (size_t)c->rec_steps * c->nrec);
      if (e != hipSuccess) return hipfail(c, e, "record launch");
      c->rec_steps += 1;
    }
  }
  return BBW_OK;
}

static int energy_common(bbw_ctx *c, int kind, double *E) {
  if (!c || !E) return BBW_ERR_ARG;
  int rc = check_ready(c);
  if (rc || (rc = check_tiled_order(c, "the energy reduction")) ||
      (rc = check_not_curved(c, "the energy reduction")))
    return rc;
  if ((rc = dalloc(c, &c->partial, (size_t)BBW_ENERGY_BLOCKS + 1))) return rc;
  hipError_t e = launch_energy(c, kind, c->partial, c->partial + BBW_ENERGY_BLOCKS);
  if (e != hipSuccess) return hipfail(c, e, "energy launch");
  HIPCHK(c, hipMemcpyAsync(E, c->partial + BBW_ENERGY_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BBW_OK;
}

int bbw_energy(bbw_ctx *c, double *E) { return energy_common(c, BBW_ENERGY_QUADRATURE, E); }

int bbw_energy_wadg(bbw_ctx *c, double *E) { return energy_common(c, BBW_ENERGY_WADG, E); }

int bbw_energy_noalloc(bbw_ctx *c, double *E) {
  if (!c || !E) return BBW_ERR_ARG;
  int rc = check_ready(c);
  if (rc || (rc = check_tiled_order(c, "the energy reduction")) ||
      (rc = check_not_curved(c, "the energy reduction")))
    return rc;
  double *scr = energy_scratch(c->dev);
  if (!scr) return fail(c, BBW_ERR_STATE, "energy scratch array not set up on this device");
  // the reduction writes one partial sum per block of a grid capped at BBW_ENERGY_BLOCKS (so any
  // K fits), and the array holds that many and the result
  if (energy_blocks(c, BBW_ENERGY_QUADRATURE) > BBW_ENERGY_BLOCKS)
    return fail(c, BBW_ERR_ARG, "energy reduction needs more partial sums than the scratch array holds");
  hipError_t e = launch_energy(c, BBW_ENERGY_QUADRATURE, scr, scr + BBW_ENERGY_BLOCKS);
  if (e != hipSuccess) return hipfail(c, e, "energy launch");
  HIPCHK(c, hipMemcpyAsync(E, scr + BBW_ENERGY_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BBW_OK;
}

int bbw_upload_adjoint_maps(bbw_ctx *c, const int32_t *mapP) {
  if (!c || !mapP) return BBW_ERR_ARG;
  const size_t n = (size_t)c->K * 4 * c->Nfp;
  for (size_t i = 0; i < n; ++i)
    if (mapP[i] < 0 || (size_t)mapP[i] >= n || (size_t)mapP[mapP[i]] != i)
      return fail(c, BBW_ERR_ARG, "mapP must be an involution on the face nodes");
  return upload(c, &c->mapP, (const int *)mapP, n);
}

namespace {

// Forward nsteps from the current state, recording receivers into the trace buffer.
int forward_record(bbw_ctx *c, int nsteps, double dt, double t0) {
  c->rec_steps = 0;
  return bbw_step(c, nsteps, dt, t0, BBW_MODE_FUSED, 2);
}

// Observations on the device: a caller's device array, or a host array copied into the
// context's buffer.
int device_obs(bbw_ctx *c, int nsteps, const float *d_obs, bool on_device, const float **d_dev) {
  if (on_device) {
    *d_dev = d_obs;
    return BBW_OK;
  }
  const size_t n = (size_t)nsteps * c->nrec;
  if (c->dobs_cap < n) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->dobs);
    c->dobs_cap = 0;
    int rc = dalloc(c, &c->dobs, n);
    if (rc) return rc;
    c->dobs_cap = n;
  }
  HIPCHK(c, hipMemcpyAsync(c->dobs, d_obs, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  *d_dev = c->dobs;
  return BBW_OK;
}

// J = 1/2 sum (trace - d)^2 over nsteps rows of the trace buffer, FP64 on the device; with
// resid, the residuals (FP32) for the adjoint's receiver injection.  *J is written once the
// stream reaches it (the callers synchronize before returning).
int device_misfit(bbw_ctx *c, int nsteps, const float *d_dev, float *resid, double *J) {
  int rc = dalloc(c, &c->partial, (size_t)BBW_ENERGY_BLOCKS + 1);
  if (rc) return rc;
  hipError_t e = launch_misfit(c, c->rec_trace, d_dev, (long long)nsteps * c->nrec, resid, c->partial,
                               c->partial + BBW_ENERGY_BLOCKS);
  if (e != hipSuccess) return hipfail(c, e, "misfit launch");
  HIPCHK(c, hipMemcpyAsync(J, c->partial + BBW_ENERGY_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return BBW_OK;
}

int check_inverse(bbw_ctx *c, int nsteps, const float *d_obs) {
  int rc = check_ready(c);
  if (rc || (rc = check_tiled_order(c, "the misfit / adjoint path")) ||
      (rc = check_not_curved(c, "the misfit / adjoint path")))
    return rc;
  if (!d_obs || nsteps <= 0) return fail(c, BBW_ERR_ARG, "bad observations");
  if (c->nrec == 0 || c->rec_max_steps < nsteps) return fail(c, BBW_ERR_STATE, "receivers not set for nsteps");
  return BBW_OK;
}

int save_state(bbw_ctx *c) {
  const size_t n = nstate(c);
  int rc = dalloc(c, &c->Q0save, n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->Q0save, c->Q[c->cur], n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->res, 0, n * sizeof(float), c->stream));
  return BBW_OK;
}

int restore_state(bbw_ctx *c) {
  const size_t n = nstate(c);
  HIPCHK(c, hipMemcpyAsync(c->Q[c->cur], c->Q0save, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->res, 0, n * sizeof(float), c->stream));
  c->rec_steps = 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BBW_OK;
}

// Marks the adjoint driver's forward passes for the duration of a misfit or gradient call.
struct AdjPass {
  bbw_ctx *c;
  explicit AdjPass(bbw_ctx *ctx) : c(ctx) { c->adj_pass = 1; }
  ~AdjPass() { c->adj_pass = 0; }
};

int misfit_impl(bbw_ctx *c, int nsteps, double dt, double t0, const float *d_obs, bool obs_dev, double *J) {
  if (!c || !J) return BBW_ERR_ARG;
  AdjPass pass(c);
  int rc = check_inverse(c, nsteps, d_obs);
  if (rc) return rc;
  const float *d_dev = nullptr;
  if ((rc = device_obs(c, nsteps, d_obs, obs_dev, &d_dev))) return rc;
  if ((rc = save_state(c))) return rc;
  if ((rc = forward_record(c, nsteps, dt, t0))) return rc;
  if ((rc = device_misfit(c, nsteps, d_dev, nullptr, J))) return rc;
  return restore_state(c);
}

int gradient_impl(bbw_ctx *c, int nsteps, double dt, double t0, const float *d_obs, bool obs_dev,
                  int ckpt_every, float *grad_out, bool grad_dev, double *J) {
  if (!c || !grad_out || !J || ckpt_every <= 0) return BBW_ERR_ARG;
  AdjPass pass(c);
  int rc = check_inverse(c, nsteps, d_obs);
  if (rc) return rc;
  if (!c->mapP) return fail(c, BBW_ERR_STATE, "adjoint maps not uploaded");
  // acoustic: gradient in c2 [K][Nq], one face value per face node, (A Q)_p per stage;
  // elastic: gradient in (1/rho, lambda, mu) [3][K][Nq], six face values per face node in two
  // buffers (the stage before's and this stage's), A Q of all 9 fields
  const bool elas = c->Nf == 9;
  const size_t n = nstate(c), nq = (size_t)c->K * c->Nq * (elas ? 3 : 1);
  const size_t nh1 = (size_t)c->K * 4 * c->Nfp * (elas ? 6 : 1), nh = elas ? 2 * nh1 : nh1;
  const size_t ystage = (size_t)c->K * c->Np * (elas ? 9 : 1);
  const size_t ystep = 5 * ystage;  // floats of the saved stage values per step (5 stages)
  if ((rc = dalloc(c, &c->lQ, n)) || (rc = dalloc(c, &c->lres, n)) || (!elas && (rc = dalloc(c, &c->Zadj, n))) ||
      (rc = dalloc(c, &c->Hadj, nh)) || (rc = dalloc(c, &c->grad, nq)))
    return rc;
  if (c->resid_len < (size_t)nsteps * c->nrec) { dfree(c->resid); c->resid_len = 0; }
  if ((rc = dalloc(c, &c->resid, (size_t)nsteps * c->nrec))) return rc;
  c->resid_len = (size_t)nsteps * c->nrec > c->resid_len ? (size_t)nsteps * c->nrec : c->resid_len;

  // how many final steps keep their Vq (A Q)_p from the first forward pass
  size_t ncache = 0;
  if (c->ycache_mb >= 0) {
    size_t budget;
    if (c->ycache_mb > 0) {
      budget = (size_t)c->ycache_mb << 20;
    } else {
      size_t fr = 0, tot = 0;
      HIPCHK(c, hipMemGetInfo(&fr, &tot));
      budget = (size_t)(0.6 * (double)(fr + c->Ycache_len * sizeof(float)));
    }
    ncache = budget / (ystep * sizeof(float));
    if (ncache > (size_t)nsteps) ncache = nsteps;
  }
  const int s_cache0 = nsteps - (int)ncache;
  if (ncache) {
    if (c->Ycache_len < ncache * ystep) { dfree(c->Ycache); c->Ycache_len = 0; }
    if ((rc = dalloc(c, &c->Ycache, ncache * ystep))) return rc;
    c->Ycache_len = ncache * ystep > c->Ycache_len ? ncache * ystep : c->Ycache_len;
  }
  const int C = ckpt_every < s_cache0 ? ckpt_every : (s_cache0 > 0 ? s_cache0 : 1);
  const int nseg = s_cache0 > 0 ? (s_cache0 + C - 1) / C : 0;
  if (nseg) {
    if (c->ckpt_len < (size_t)nseg * n) { dfree(c->ckpt); c->ckpt_len = 0; }
    if ((rc = dalloc(c, &c->ckpt, (size_t)nseg * n))) return rc;
    c->ckpt_len = (size_t)nseg * n > c->ckpt_len ? (size_t)nseg * n : c->ckpt_len;
    if (c->Ybuf_len < (size_t)C * ystep) { dfree(c->Ybuf); c->Ybuf_len = 0; }
    if ((rc = dalloc(c, &c->Ybuf, (size_t)C * ystep))) return rc;
    c->Ybuf_len = (size_t)C * ystep > c->Ybuf_len ? (size_t)C * ystep : c->Ybuf_len;
  }
  if ((rc = save_state(c))) return rc;

  const bool a0_zero = c->rk4a[0] == 0.0;
  auto stage_value = [&](int step, int i) -> float {
    return c->nsrc ? (float)ricker(stage_time(c, t0 + step * dt, i, dt), c->src_f0) : 0.f;
  };
  auto fwd_step_y = [&](int step, float *Y) -> int {
    for (int i = 0; i < 5; ++i) {
      // the residual left by the last stage is dead when a_0 = 0 (the state is restored and the
      // residual cleared after the gradient)
      hipError_t e =
          elas ? launch_fused_elas(c, c->Q[c->cur], c->Q[1 - c->cur], (float)c->rk4a[i], (float)c->rk4b[i],
                                   (float)dt, nullptr, stage_value(step, i), !(i == 4 && a0_zero),
                                   Y + (size_t)i * ystage)
               : launch_fused_mfma(c, c->Q[c->cur], c->Q[1 - c->cur], (float)c->rk4a[i],
                                   (float)c->rk4b[i], (float)dt, nullptr, stage_value(step, i), 0,
                                   Y + (size_t)i * ystage, !(i == 4 && a0_zero));
      if (e != hipSuccess) return hipfail(c, e, "forward stage with A Q capture");
      c->cur = 1 - c->cur;
    }
    return BBW_OK;
  };
  int hsel = 0;   // elastic: the H buffer the next reverse stage reads
  auto adj_step = [&](int step, const float *Y) -> int {
    hipError_t e = launch_adj_inject(c, c->resid + (size_t)step * c->nrec);
    if (e != hipSuccess) return hipfail(c, e, "receiver injection");
    for (int i = 4; i >= 0; --i) {
      const float *Yi = Y + (size_t)i * ystage;
      // with a_0 = 0 the reverse of stage 0 sets lres to 0 (dead store), so the reverse of stage 4,
      // which always follows it or the zeroed start, reads zeros
      const int lres_io = a0_zero ? (i == 4 ? 0 : 1) | (i == 0 ? 0 : 2) : 3;
      if (elas) {
        e = launch_adj_elas_stage(c, Yi, (float)c->rk4a[i], (float)c->rk4b[i], (float)dt, lres_io,
                                  c->Hadj + (size_t)hsel * nh1, c->Hadj + (size_t)(1 - hsel) * nh1);
        hsel = 1 - hsel;
      } else {
        e = c->adj_impl ? launch_adj_stage_tiled(c, Yi, (float)c->rk4a[i], (float)c->rk4b[i], (float)dt, lres_io)
                        : launch_adj_stage(c, Yi, (float)c->rk4a[i], (float)c->rk4b[i], (float)dt);
      }
      if (e != hipSuccess) return hipfail(c, e, "adjoint stage");
    }
    return BBW_OK;
  };

  // 1. forward: checkpointed segments, then the cached final steps; receiver traces
  c->rec_steps = 0;
  for (int s0 = 0; s0 < s_cache0; s0 += C) {
    const int ns = s_cache0 - s0 < C ? s_cache0 - s0 : C;
    HIPCHK(c, hipMemcpyAsync(c->ckpt + (size_t)(s0 / C) * n, c->Q[c->cur], n * sizeof(float),
                             hipMemcpyDeviceToDevice, c->stream));
    if ((rc = bbw_step(c, ns, dt, t0 + s0 * dt, BBW_MODE_FUSED, 2))) return rc;
  }
  for (int s = s_cache0; s < nsteps; ++s) {
    if ((rc = fwd_step_y(s, c->Ycache + (size_t)(s - s_cache0) * ystep))) return rc;
    hipError_t e = launch_record(c, c->Q[c->cur], c->rec_trace + (size_t)c->rec_steps * c->nrec);
    if (e != hipSuccess) return hipfail(c, e, "record launch");
    c->rec_steps += 1;
  }
  const float *d_dev = nullptr;
  if ((rc = device_obs(c, nsteps, d_obs, obs_dev, &d_dev))) return rc;
  if ((rc = device_misfit(c, nsteps, d_dev, c->resid, J))) return rc;

  // 2. reverse: cached steps first, then checkpointed segments (recomputed with capture)
  HIPCHK(c, hipMemsetAsync(c->lQ, 0, n * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(c->lres, 0, n * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(c->grad, 0, nq * sizeof(float), c->stream));
  // elastic: no neighbour terms are pending before the first reverse stage
  if (elas) HIPCHK(c, hipMemsetAsync(c->Hadj, 0, nh1 * sizeof(float), c->stream));
  for (int s = nsteps - 1; s >= s_cache0; --s)
    if ((rc = adj_step(s, c->Ycache + (size_t)(s - s_cache0) * ystep))) return rc;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * C, ns = s_cache0 - s0 < C ? s_cache0 - s0 : C;
    HIPCHK(c, hipMemcpyAsync(c->Q[c->cur], c->ckpt + (size_t)seg * n, n * sizeof(float),
                             hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->res, 0, n * sizeof(float), c->stream));
    for (int j = 0; j < ns; ++j)
      if ((rc = fwd_step_y(s0 + j, c->Ybuf + (size_t)j * ystep))) return rc;
    for (int j = ns - 1; j >= 0; --j)
      if ((rc = adj_step(s0 + j, c->Ybuf + (size_t)j * ystep))) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(grad_out, c->grad, nq * sizeof(float),
                           grad_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
  return restore_state(c);
}

}  // namespace

int bbw_misfit(bbw_ctx *c, int nsteps, double dt, double t0, const float *d_obs, double *J) {
  return misfit_impl(c, nsteps, dt, t0, d_obs, false, J);
}

int bbw_misfit_device(bbw_ctx *c, int nsteps, double dt, double t0, const float *d_obs_dev, double *J) {
  return misfit_impl(c, nsteps, dt, t0, d_obs_dev, true, J);
}

int bbw_adjoint_gradient(bbw_ctx *c, int nsteps, double dt, double t0, const float *d_obs,
                         int ckpt_every, float *grad_out, double *J) {
  return gradient_impl(c, nsteps, dt, t0, d_obs, false, ckpt_every, grad_out, false, J);
}

int bbw_adjoint_gradient_device(bbw_ctx *c, int nsteps, double dt, double t0, const float *d_obs_dev,
                                int ckpt_every, float *grad_dev, double *J) {
  return gradient_impl(c, nsteps, dt, t0, d_obs_dev, true, ckpt_every, grad_dev, true, J);
}

int bbw_set_option(bbw_ctx *c, int option, int value) {
  if (!c) return BBW_ERR_ARG;
  switch (option) {
    case BBW_OPT_FUSED_IMPL:
      if (value != BBW_FUSED_VECTOR && value != BBW_FUSED_MFMA) return fail(c, BBW_ERR_ARG, "bad fused impl");
      c->fused_impl = value;
      break;
    case BBW_OPT_KBLK_V: c->KblkV = value > 0 ? value : 0; break;
    case BBW_OPT_KBLK_S: c->KblkS = value > 0 ? value : 0; break;
    case BBW_OPT_KBLK_U: c->KblkU = value > 0 ? value : 0; break;
    case BBW_OPT_KBLK_F: c->KblkF = value > 0 ? value : 0; break;
    case BBW_OPT_ADJ_CACHE_MB: c->ycache_mb = value; break;
    case BBW_OPT_ADJ_IMPL: c->adj_impl = value ? 1 : 0; break;
    case BBW_OPT_PLAIN_IMPL:
      if (c->curved) return fail(c, BBW_ERR_UNSUPPORTED, "a curvilinear context keeps its kernels");
      if (value != BBW_PLAIN_FLAT && value != BBW_PLAIN_TILED && value != BBW_PLAIN_SLICE)
        return fail(c, BBW_ERR_ARG, "bad plain impl");
      if (value == BBW_PLAIN_TILED && c->N > BBW_MAX_ORDER_TILED)
        return fail(c, BBW_ERR_UNSUPPORTED, "the tiled plain kernels are compiled for N <= 6");
      if (value == BBW_PLAIN_SLICE && c->Nf != 4)
        return fail(c, BBW_ERR_UNSUPPORTED, "the slice-by-slice lift is acoustic");
      c->plain_impl = value;
      break;
    case BBW_OPT_SRC_STEP_TIME: c->src_step_time = value ? 1 : 0; break;
    case BBW_OPT_REC_FIELD:
      if (value < 0 || value >= c->Nf) return fail(c, BBW_ERR_ARG, "receiver field out of range");
      c->rec_field = value;
      reset_graph(c);
      break;
    case BBW_OPT_PLANAR_SPLIT: c->planar_split = value ? 1 : 0; break;
    case BBW_OPT_ELAS_FUSED:
      if (value != BBW_ELAS_LIFT_NODES && value != BBW_ELAS_LIFT_CUBATURE) return fail(c, BBW_ERR_ARG, "bad elastic fused impl");
      c->elas_fused = value;
      break;
    case BBW_OPT_XCD_CHUNK:
      if (value < 0 || value > 1024 || (value & (value - 1))) return fail(c, BBW_ERR_ARG, "XCD run length must be 0 or a power of 2 up to 1024");
      c->xcd_chunk = value;
      break;
    default: return fail(c, BBW_ERR_ARG, "unknown option");
  }
  reset_graph(c);
  return BBW_OK;
}

int bbw_device_ptr(bbw_ctx *c, int which, void **ptr) {
  if (!c || !ptr) return BBW_ERR_ARG;
  switch (which) {
    case 0: *ptr = c->Q[c->cur]; break;
    case 1: *ptr = c->res; break;
    case 2: *ptr = c->rhs; break;
    case 3: *ptr = c->c2q; break;
    default: return fail(c, BBW_ERR_ARG, "unknown buffer");
  }
  return BBW_OK;
}

}  // extern "C"
