// Fused acoustic stage for gfx950 with the WADG projection on FP32 MFMA.
//
// One workgroup (4 waves) owns a tile of TB = 16 consecutive elements:
//   1  Q_in tile -> LDS (buffer loads, 16 B per lane); prefetched into registers: the thread's
//      res and Q_in row of every tile element, c2 in the MFMA output layout, neighbour traces
//   2  (p, Ur, Us, Ut) as one float4 per node; upwind flux phi at face nodes into registers
//      (the p and U fluxes are phi and -phi)
//   3  phi -> LDS [face node][element]; volume terms (4-nnz barycentric operators, 16 float4
//      LDS reads per node) into registers
//   4  volume terms -> rhs (aliasing the (p, U) tile)
//   5  lift: P_f = LIFT_f . phi_f per face on MFMA 16x16x4 (M = nodes, N = 16 elements,
//      K = face nodes); rhs_p += sum_f P_f, rhs_U -= sum_f n_f P_f.  LIFT = EEL . blkdiag(L0)
//      is applied densely: on the matrix cores the factored sparse form saves nothing
//   6  S = (Vq . rhs_p) * c2      MFMA 16x16x4, M = cubature points, N = 16 elements, K = Np
//   7  rhs_p = Pq . S             MFMA 16x16x4, M = nodes, N = 16 elements, K = Nq
//   8  LSRK: res = fa res + fdt rhs (+ source); Q_out = Q_in + fb res (coalesced rows)
// Thread mappings are fixed per thread (node, face node or state row) and loop over the
// tile's elements, so operator rows are loaded once per tile and no index is divided.
// LDS regions are reused across phases (Q_in tile -> face flux -> S; (p, U) -> rhs) to stay
// at 30 KB per workgroup (N = 5).
// With PROJ_U the velocity fields go through Pq . Vq as well (N = 64 columns).
//
// v_mfma_f32_16x16x4f32 lane layout (checked on gfx950): A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k, D[4*(l/16) + r][l%16] in register r of lane l.
#include <algorithm>

#include "bbw_fused_tile.h"

namespace bbw {

// workgroups per CU the register allocation is sized for (LDS allows 5 at N = 5)
#ifndef BBW_FUSED_WG_PER_CU
#define BBW_FUSED_WG_PER_CU 3
#endif

template <int N, bool PROJ_U>
__global__ __launch_bounds__(256, BBW_FUSED_WG_PER_CU) void k_fused_mfma(FusedMesh m, FusedOps o, FusedShot sh,
                                                                         SrcArgs src, float fa, float fb,
                                                                         const float *ftime_dev) {
  using L = FusedLayout<N, PROJ_U>;
  extern __shared__ float lds[];
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * L::TB, nk = min(L::TB, m.K - k0);
  if (ftime_dev) sh.ft = *ftime_dev;
  TileStage<N, PROJ_U> st;
  stage_load<N, PROJ_U>(st, m, sh.Qin, k0, nk, tid);
  int vm[L::EPT3];
  vm_load<N, PROJ_U>(vm, m, k0, nk, tid);
  stage_store<N, PROJ_U>(st, lds + L::OQ, lds + L::OG, tid);
  __syncthreads();
  tile_shot<N, PROJ_U>(m, o, sh, src, fa, fb, k0, nk, lds, lds + L::OQ, lds + L::OG, vm,
                       face_node<N, PROJ_U>(o, tid), tid, NoHook{}, NoHook{}, NoHook{});
}

FusedOps fused_ops(const bbw_ctx *c) {
  FusedOps o;
  o.Fmask = c->Fmask;
  o.Dids = c->Dids;
  o.Dvals = c->Dvals;
  o.EEL_ids4 = c->EEL_ids4;
  o.EEL_vals = c->EEL_vals;
  o.L0_ids = c->L0_ids;
  o.L0_vals = c->L0_vals;
  o.VqPad = c->VqPad;
  o.PqPad = c->PqPad;
  o.LiftPad = c->LiftPad;
  o.VqBlk = c->VqBlk;
  o.PqBlk = c->PqBlk;
  o.LiftBlk = c->LiftBlk;
  return o;
}

template <int N, bool P>
static hipError_t fused_mfma_n(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                               float fdt, const float *ftd, float ft, const SrcArgs &s,
                               float *yout) {
  using L = FusedLayout<N, P>;
  const size_t sh = sizeof(float) * L::floats;
  const int grid = (c->K + L::TB - 1) / L::TB;
  if ((unsigned long long)c->K * 16 * L::Np >= (1ull << 32)) return hipErrorInvalidValue;
  if (sh > 65536) {
    hipError_t e = hipFuncSetAttribute((const void *)k_fused_mfma<N, P>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  FusedMesh m{c->K, c->vgeo, c->fgeo, c->vmapP};
  FusedShot shot{c->c2q, Qin, Qout, c->res, yout, ft, fdt};
  hipLaunchKernelGGL((k_fused_mfma<N, P>), dim3(grid), dim3(256), sh, c->stream, m, fused_ops(c),
                     shot, s, fa, fb, ftd);
  return hipGetLastError();
}

template <int N>
static hipError_t fused_mfma_sel(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                                 float fdt, const float *ftd, float ft, int projU,
                                 const SrcArgs &s, float *yout) {
  return projU ? fused_mfma_n<N, true>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, yout)
               : fused_mfma_n<N, false>(c, Qin, Qout, fa, fb, fdt, ftd, ft, s, yout);
}

hipError_t launch_fused_mfma(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                             float fdt, const float *ftd, float ft, int projU, float *yout) {
  SrcArgs s;
  s.nsrc = c->nsrc;
  s.first = c->src_first;
  s.last = c->src_last;
  s.elems = c->src_elems;
  s.coef = c->src_coef;
  for (int f = 0; f < 9; ++f) s.w[f] = c->src_w[f];
  switch (c->N) {
    case 1: return fused_mfma_sel<1>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 2: return fused_mfma_sel<2>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 3: return fused_mfma_sel<3>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 4: return fused_mfma_sel<4>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 5: return fused_mfma_sel<5>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
    case 6: return fused_mfma_sel<6>(c, Qin, Qout, fa, fb, fdt, ftd, ft, projU, s, yout);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
