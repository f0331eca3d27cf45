// Per-tile device functions of the fused acoustic stage (included by bbw_fused.hip and
// bbw_batch.hip).  See bbw_fused.hip for the phase structure.
#pragma once

#include "bbw_internal.h"

namespace bbw {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int src_slot2(const SrcArgs &s, int k) {
  if (s.nsrc == 0 || k < s.first || k > s.last) return -1;
  int lo = 0, hi = s.nsrc - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = s.elems[mid];
    if (v == k) return mid;
    if (v < k) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// source weight of acoustic field f (0..3) without indexing the array by a per-lane value
__device__ __forceinline__ float w4(const SrcArgs &s, int f) {
  return f == 0 ? s.w[0] : f == 1 ? s.w[1] : f == 2 ? s.w[2] : s.w[3];
}

template <int N, bool PROJ_U>
struct FusedLayout {
  static constexpr int Np = Dims<N>::Np, Nfp = Dims<N>::Nfp, NfpNf = Dims<N>::NfpNf;
  static constexpr int Nq = Dims<N>::Nq;
  static constexpr int TB = 16;                  // elements per tile = MFMA N
  static constexpr int NQT = (Nq + 15) / 16, NQP = NQT * 16;
  static constexpr int NPT = (Np + 15) / 16, NPP = NPT * 16;
  static constexpr int KTV = (Np + 3) / 4, KTP = NQP / 4;
  static constexpr int NP = PROJ_U ? 4 : 1;
  // rhs row stride: 4 fields plus the K padding the Vq product reads, rounded up to 2 mod 32
  // so that the MFMA operand reads (16 rows x 2 consecutive words per 32 lanes) hit 32 banks
  static constexpr int RS0 = 4 * Np > (NP - 1) * Np + 4 * KTV ? 4 * Np : (NP - 1) * Np + 4 * KTV;
  static constexpr int RS = RS0 + ((34 - RS0 % 32) % 32);
  static constexpr int KF = 4 * ((Nfp + 3) / 4); // face nodes per face, padded for MFMA K
  static constexpr int KB = NfpNf + 4;           // rows of the face-flux tile (zero tail)
  static constexpr int BS = TB + 1;              // its row stride: column writes spread over banks
  static constexpr int SX = NP * TB;             // S row length
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  // W1: face flux [KB][BS] | S [NQP][SX]
  static constexpr int W1 = cmax((KB * BS + TB - 1) / TB, NQP * NP);
  // W2: (p, Ur, Us, Ut) float4 [TB][Np] | rhs [TB][RS]
  static constexpr int W2 = cmax(4 * Np, RS);
  static constexpr int QTW = (NQT + 3) / 4;     // q tiles per wave
  static constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS;        // node groups
  static constexpr int NG3 = 256 / NfpNf, EPT3 = (TB + NG3 - 1) / NG3;   // face-node groups
  static constexpr int EPTV = (TB + NG - 1) / NG;
  static constexpr bool PF = 4 * Np <= 256;    // one state row per thread: prefetch res
  // Q_in tile [TB][4Np] (read in phases 2 and 8) and geometry (vgeo [TB][9], fgeo [TB][16])
  static constexpr int QT = TB * 4 * Np, GT = TB * 25;
  static constexpr unsigned QB = QT * 4u, QIT = (QB + 4095) / 4096;  // 16-byte words per thread
  static constexpr int OQ = 0, OW1 = QT, OW2 = OW1 + TB * W1, OG = OW2 + TB * W2, OSL = OG + GT;
  static constexpr size_t floats = (size_t)OSL + TB;
};

// Operator, mesh and per-shot arguments of one tile.
struct FusedOps {
  const int *Fmask;
  const int4 *Dids;
  const float4 *Dvals;
  const int *EEL_ids4;
  const float *EEL_vals;
  const int *L0_ids;
  const float *L0_vals;
  const float *VqPad;
  const float *PqPad;
  const float *LiftPad;
  const float *VqBlk, *PqBlk, *LiftBlk;   // lane-ordered A operands (bbw_ctx::VqBlk)
};

struct FusedMesh {
  int K;
  const float *vgeo, *fgeo;
  const int *vmapP;
};

struct FusedShot {
  const float *c2q;
  const float *Qin;
  float *Qout, *res;
  float *yout;   // (A Q)_p in BB coefficients [K][Np] for the adjoint, or null
  float ft;      // source time function at this stage
  float fdt;     // time step
  int store_res; // 0: res is dead after this stage and is not stored (the last LSRK45 stage of a
                 // step that another step follows: the next stage multiplies it by a_0 = 0)
};

template <int N, bool PROJ_U>
using FL = FusedLayout<N, PROJ_U>;

// Buffer resources: a scalar base and byte size per array (or per tile), addressed by 32-bit
// per-lane offsets with the constant part in the instruction's offset field.  Reads past the
// size return 0 and writes past it are dropped, which is how partial tiles are handled.
typedef __amdgpu_buffer_rsrc_t brsrc;
__device__ __forceinline__ brsrc mkrsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bld(brsrc r, unsigned off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ int bldi(brsrc r, unsigned off) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
// (the 128-bit result is converted as a whole: extracting its elements one by one before the
// conversion was folded to a single element by this compiler)
__device__ __forceinline__ float4 bld4(brsrc r, unsigned off) {
  const f4 v = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  return make_float4(v[0], v[1], v[2], v[3]);
}
// operator tables: every offset is in range by construction, so the uniform part may go to the
// scalar offset (which the range check does not cover)
__device__ __forceinline__ float bld_s(brsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
// 4 consecutive K-steps of an MFMA A operand (lane-ordered blocks, always in range)
__device__ __forceinline__ f4 bld4_s(brsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(brsrc r, float v, unsigned off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

// A tile's Q_in rows and geometry on their way to LDS through registers (words past the
// tile's end read as 0).
template <int N, bool PROJ_U>
struct TileStage {
  float4 q[FL<N, PROJ_U>::QIT];
  float g[2];
};

template <int N, bool PROJ_U>
__device__ __forceinline__ void stage_load(TileStage<N, PROJ_U> &st, const FusedMesh &m, const float *Qin,
                                           int k0, int nk, const int tid) {
  using L = FL<N, PROJ_U>;
  static_assert(L::TB * 16 <= 256, "one geometry word per thread");
  const brsrc rQ = mkrsrc(Qin + (size_t)k0 * 4 * L::Np, (unsigned)nk * L::Np * 16u);
#pragma unroll
  for (unsigned it = 0; it < L::QIT; ++it) st.q[it] = bld4(rQ, tid * 16u + it * 4096u);
  const brsrc rg = mkrsrc(m.vgeo + (size_t)k0 * 9, (unsigned)nk * 36u);
  const brsrc rf = mkrsrc(m.fgeo + (size_t)k0 * 16, (unsigned)nk * 64u);
  st.g[0] = bld(rg, tid * 4u);
  st.g[1] = bld(rf, tid * 4u);
}

template <int N, bool PROJ_U>
__device__ __forceinline__ void stage_store(const TileStage<N, PROJ_U> &st, float *sQ, float *sG,
                                            const int tid) {
  using L = FL<N, PROJ_U>;
  float4 *sQ4 = (float4 *)sQ;
#pragma unroll
  for (unsigned it = 0; it < L::QIT; ++it) {
    const unsigned off = tid * 16u + it * 4096u;
    if (L::QIT * 4096u == L::QB || off < L::QB) sQ4[off / 16] = st.q[it];
  }
  if (tid < L::TB * 9) sG[tid] = st.g[0];
  sG[L::TB * 9 + tid] = st.g[1];
}

// The thread's vmapP entries of a tile (face-node mapping of the flux phase), vm[EPT3].
template <int N, bool PROJ_U>
__device__ __forceinline__ void vm_load(int *vm, const FusedMesh &m, int k0, int nk, const int tid) {
  using L = FL<N, PROJ_U>;
  constexpr int NfpNf = L::NfpNf;
  const brsrc rv = mkrsrc(m.vmapP + (size_t)k0 * NfpNf, (unsigned)nk * NfpNf * 4u);
  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf;
  const unsigned vo = (unsigned)(g3 * NfpNf + m3) * 4u;
#pragma unroll
  for (int i = 0; i < L::EPT3; ++i) vm[i] = bldi(rv, vo + i * L::NG3 * NfpNf * 4u);
}

// The thread's face node (flux phase): its node id in the element.
template <int N, bool PROJ_U>
__device__ __forceinline__ int face_node(const FusedOps &o, const int tid) {
  using L = FL<N, PROJ_U>;
  const int g3 = tid / L::NfpNf, m3 = tid - g3 * L::NfpNf;
  return g3 < L::NG3 ? o.Fmask[m3] : 0;
}

struct NoHook {
  __device__ void operator()() const {}
};

// The register-resident inputs of a tile's stage (phase 1): res rows, c2 in the MFMA output
// layout and the neighbour traces of the thread's face nodes.  Issued before the tile's Q_in
// rows are stored to LDS, so the vmapP -> trace chain overlaps the tile loads.
template <int N, bool PROJ_U>
struct TilePre {
  float rreg[FL<N, PROJ_U>::PF ? FL<N, PROJ_U>::TB : 1];
  float c2r[FL<N, PROJ_U>::QTW][4];
  float tr[FL<N, PROJ_U>::EPT3][4];
};

// Issues the loads of TilePre.  With fa = 0 (the first LSRK45 stage) res is not read: fa res = 0
// for any finite res.  Requires K * 4 * Np * 4 < 2^32 (32-bit trace offsets).
template <int N, bool PROJ_U>
__device__ __forceinline__ void tile_prefetch(TilePre<N, PROJ_U> &pf, const FusedMesh &m,
                                              const FusedShot &sh, float fa, int k0, int nk,
                                              const int (&vm)[FL<N, PROJ_U>::EPT3], const int tid) {
  using L = FusedLayout<N, PROJ_U>;
  constexpr int Np = L::Np, Nq = L::Nq, TB = L::TB;
  constexpr unsigned ROW = 4 * Np * 4u;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if constexpr (L::PF) {
    const brsrc rR = mkrsrc(sh.res + (size_t)k0 * 4 * Np, (unsigned)nk * ROW);
    if (fa != 0.f) {
#pragma unroll
      for (int e = 0; e < TB; ++e) pf.rreg[e] = bld(rR, tid * 4u + e * ROW);
    } else {
#pragma unroll
      for (int e = 0; e < TB; ++e) pf.rreg[e] = 0.f;
    }
  }
  // rows q >= Nq read a neighbouring value; the padded Vq rows make them 0 in S
  const brsrc rC = mkrsrc(sh.c2q + (size_t)k0 * Nq, (unsigned)nk * Nq * 4u);
  const unsigned co = (unsigned)((lane & 15) * Nq + 4 * (lane >> 4)) * 4u;
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t;
#pragma unroll
    for (int r = 0; r < 4; ++r) pf.c2r[t][r] = qt < L::NQT ? bld(rC, co + (qt * 16 + r) * 4u) : 0.f;
  }
  // the traces last: their addresses wait for vmapP, the loads above do not
  const brsrc rT = mkrsrc(sh.Qin, (unsigned)m.K * ROW);
#pragma unroll
  for (int i = 0; i < L::EPT3; ++i) {
    const unsigned to = (unsigned)vm[i] * 4u;
#pragma unroll
    for (int f = 0; f < 4; ++f) pf.tr[i][f] = bld(rT, to + f * Np * 4u);
  }
}

// One fused LSRK stage of one shot on one tile (phases 1-8 of bbw_fused.hip), with the tile's
// Q_in rows already in LDS at sQ, its geometry at sG and the register inputs in pf
// (tile_prefetch).  at_start runs after phase 1, at_mid at the start of phase 6 and at_end
// before the final barrier.  Ends with a barrier.  Elements past the tile's end are computed on
// zeros and never stored.
template <int N, bool PROJ_U, class HS, class HM, class HE>
__device__ __forceinline__ void tile_shot(const FusedMesh &m, const FusedOps &o, const FusedShot &sh,
                                          const SrcArgs &src, float fa, float fb, int k0, int nk,
                                          float *lds, const float *sQ, const float *sG,
                                          const int (&vm)[FL<N, PROJ_U>::EPT3],
                                          const TilePre<N, PROJ_U> &pf, int fid3,
                                          const int tid, HS &&at_start, HM &&at_mid, HE &&at_end) {
  using L = FusedLayout<N, PROJ_U>;
  constexpr int Np = L::Np, Nfp = L::Nfp, NfpNf = L::NfpNf, Nq = L::Nq, TB = L::TB, RS = L::RS;
  constexpr int NpNf = 4 * Np, KF = L::KF, KB = L::KB, BS = L::BS;
  constexpr int NP = L::NP, SX = L::SX, NQP = L::NQP, NPP = L::NPP;
  constexpr unsigned ROW = NpNf * 4u;    // bytes of one element's state
  const float fdt = sh.fdt;
  const int4 *__restrict__ Dids = o.Dids;
  const float4 *__restrict__ Dvals = o.Dvals;
  float *W1 = lds + L::OW1;              // [TB * W1]
  float *W2 = lds + L::OW2;              // [TB * W2]
  const float *sF = sG + TB * 9;         // [TB][16]
  int *sSlot = (int *)(lds + L::OSL);
  float *sB = W1;                        // phases 3-5: face flux [KB][BS]
  float *S = W1;                         // phases 6-7
  float4 *sV = (float4 *)W2;             // phases 2-3
  float *sR = W2;                        // phases 4-8: rhs [TB][RS]
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const size_t base = (size_t)k0 * NpNf;
  const unsigned tbytes = (unsigned)nk * ROW;
  const brsrc rR = mkrsrc(sh.res + base, tbytes);
  const brsrc rO = mkrsrc(sh.Qout + base, tbytes);
  const float(&rreg)[L::PF ? TB : 1] = pf.rreg;
  const float(&c2r)[L::QTW][4] = pf.c2r;
  const float(&tr)[L::EPT3][4] = pf.tr;
  // face-node mapping (phases 2, 3)
  const int g3 = tid / NfpNf, m3 = tid - g3 * NfpNf;
  const bool on3 = g3 < L::NG3;
  const int f3 = m3 / Nfp;
  // node mapping (phases 2, 3, 4)
  const int gv = tid / L::GS, nv = tid - gv * L::GS;
  const bool onv = gv < L::NG && nv < Np;
  // source: elements are sorted, so a tile outside [first, last] has none (uniform test)
  const bool has_src = src.nsrc > 0 && k0 + nk - 1 >= src.first && k0 <= src.last;

  // ---- 1: source slots of the tile's elements (read in phase 8, after several barriers); the
  // register inputs (res, c2, traces) were issued by tile_prefetch before the Q_in tile's barrier
  if (has_src && tid < TB) sSlot[tid] = tid < nk ? src_slot2(src, k0 + tid) : -1;
  const bool read_res = fa != 0.f;
  at_start();

  // ---- 2: (p, Ur, Us, Ut) per node; upwind flux at face nodes into registers.  The p and U
  // fluxes are (phi, -phi) with phi = (pjump - Unjump) Fscale / 2, so one value per face node.
  if (onv) {
#pragma unroll
    for (int i = 0; i < L::EPTV; ++i) {
      const int e = gv + L::NG * i;
      if (L::NG * L::EPTV == TB || e < TB) {
        const float *s = sQ + e * NpNf;
        const float *G = sG + 9 * e;
        const float uu = s[Np + nv], vv = s[2 * Np + nv], ww = s[3 * Np + nv];
        sV[e * Np + nv] = make_float4(s[nv], uu * G[0] + vv * G[1] + ww * G[2],
                                      uu * G[3] + vv * G[4] + ww * G[5],
                                      uu * G[6] + vv * G[7] + ww * G[8]);
      }
    }
  }
  float phi[L::EPT3];
  if (on3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      phi[i] = 0.f;
      if (L::NG3 * L::EPT3 == TB || e < TB) {
        const float *s = sQ + e * NpNf;
        const float *F = sF + 16 * e + 4 * f3;
        const float pM = s[fid3], uM = s[Np + fid3], vM = s[2 * Np + fid3], wM = s[3 * Np + fid3];
        const int idM = fid3 + (k0 + e) * NpNf;
        float pjump, Unjump;
        if (vm[i] == idM) {
          pjump = -2.f * pM;
          Unjump = 0.f;
        } else {
          pjump = tr[i][0] - pM;
          Unjump = (tr[i][1] - uM) * F[1] + (tr[i][2] - vM) * F[2] + (tr[i][3] - wM) * F[3];
        }
        phi[i] = .5f * (pjump - Unjump) * F[0];
      }
    }
  }
  __syncthreads();

  // ---- 3: face flux -> LDS [face node][element]; volume terms into registers
  if (on3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) sB[m3 * BS + e] = phi[i];
    }
  }
  if (tid < (KB - NfpNf) * BS) sB[NfpNf * BS + tid] = 0.f;   // K padding of the last face
  float vol[L::EPTV][4];
  if (onv) {
    const float4 dv = Dvals[nv];
    const int4 D1 = Dids[nv], D2 = Dids[Np + nv], D3 = Dids[2 * Np + nv], D4 = Dids[3 * Np + nv];
    const float4 *Vb = sV + gv * Np;
    const float4 *pa1 = Vb + D1.x, *pa2 = Vb + D2.x, *pa3 = Vb + D3.x, *pa4 = Vb + D4.x;
    const float4 *pb1 = Vb + D1.y, *pb2 = Vb + D2.y, *pb3 = Vb + D3.y, *pb4 = Vb + D4.y;
    const float4 *pc1 = Vb + D1.z, *pc2 = Vb + D2.z, *pc3 = Vb + D3.z, *pc4 = Vb + D4.z;
    const float4 *pd1 = Vb + D1.w, *pd2 = Vb + D2.w, *pd3 = Vb + D3.w, *pd4 = Vb + D4.w;
    const float *Gb = sG + 9 * gv;
#pragma unroll
    for (int i = 0; i < L::EPTV; ++i) {
      const int o4 = i * L::NG * Np;
      const float4 a1 = pa1[o4], a2 = pa2[o4], a3 = pa3[o4], a4 = pa4[o4];
      const float4 b1 = pb1[o4], b2 = pb2[o4], b3 = pb3[o4], b4 = pb4[o4];
      const float4 c1 = pc1[o4], c2 = pc2[o4], c3 = pc3[o4], c4 = pc4[o4];
      const float4 d1 = pd1[o4], d2 = pd2[o4], d3 = pd3[o4], d4 = pd4[o4];
      const float p1 = dv.x * a1.x + dv.y * a2.x + dv.z * a3.x + dv.w * a4.x;
      const float p2 = dv.x * b1.x + dv.y * b2.x + dv.z * b3.x + dv.w * b4.x;
      const float p3 = dv.x * c1.x + dv.y * c2.x + dv.z * c3.x + dv.w * c4.x;
      const float p4 = dv.x * d1.x + dv.y * d2.x + dv.z * d3.x + dv.w * d4.x;
      float dU1 = dv.x * (b1.y + c1.z + d1.w);
      dU1 += dv.y * (b2.y + c2.z + d2.w);
      dU1 += dv.z * (b3.y + c3.z + d3.w);
      dU1 += dv.w * (b4.y + c4.z + d4.w);
      float dU2 = dv.x * (a1.y + a1.z + a1.w);
      dU2 += dv.y * (a2.y + a2.z + a2.w);
      dU2 += dv.z * (a3.y + a3.z + a3.w);
      dU2 += dv.w * (a4.y + a4.z + a4.w);
      const float dpdr = .5f * (p2 - p1), dpds = .5f * (p3 - p1), dpdt = .5f * (p4 - p1);
      const float *G = Gb + 9 * L::NG * i;
      vol[i][0] = -.5f * (dU1 - dU2);
      vol[i][1] = -(G[0] * dpdr + G[3] * dpds + G[6] * dpdt);
      vol[i][2] = -(G[1] * dpdr + G[4] * dpds + G[7] * dpdt);
      vol[i][3] = -(G[2] * dpdr + G[5] * dpds + G[8] * dpdt);
    }
  }
  __syncthreads();

  // ---- 4: volume terms -> rhs (the (p, U) tile is no longer read)
  if (onv) {
    float *rb = sR + gv * RS + nv;
#pragma unroll
    for (int i = 0; i < L::EPTV; ++i) {
      const int e = gv + L::NG * i;
      if (L::NG * L::EPTV == TB || e < TB) {
        float *r = rb + L::NG * i * RS;
        r[0] = vol[i][0];
        r[Np] = vol[i][1];
        r[2 * Np] = vol[i][2];
        r[3 * Np] = vol[i][3];
      }
    }
  }
  constexpr int KPAD = L::RS0 - NpNf;      // rhs words past the 4 fields that the Vq product reads
  static_assert(TB * KPAD <= 256, "one padding word per thread");
  if (KPAD > 0 && tid < TB * KPAD) sR[(tid / KPAD) * RS + NpNf + tid % KPAD] = 0.f;
  // MFMA A operands are fetched one phase ahead of their products (for the wave's first M tile)
  constexpr int KTL = KF / 4, KBL = (KTL + 3) / 4;
  constexpr int KBV = (L::KTV + 3) / 4, KBP = (L::KTP + 3) / 4;
  const brsrc rAL = mkrsrc(o.LiftBlk, 4u * L::NPT * KBL * 1024u);
  const brsrc rAV = mkrsrc(o.VqBlk, (unsigned)L::NQT * KBV * 1024u);
  const brsrc rAP = mkrsrc(o.PqBlk, (unsigned)L::NPT * KBP * 1024u);
  f4 aL[4][KBL];
  if (wave < L::NPT) {
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int b = 0; b < KBL; ++b)
        aL[f][b] = bld4_s(rAL, lane * 16u, (unsigned)((f * L::NPT + wave) * KBL + b) * 1024u);
  }
  __syncthreads();

  // ---- 5: lift on MFMA, one product per face (M = nodes, N = 16 elements, K = face nodes):
  // P_f = LIFT_f phi_f;  rhs_p += sum_f P_f,  rhs_U -= sum_f n_f P_f
  f4 aV[L::QTW][KBV];
#pragma unroll
  for (int t = 0; t < L::QTW; ++t)
    if (wave + 4 * t < L::NQT) {
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb)
        aV[t][kb] = bld4_s(rAV, lane * 16u, (unsigned)((wave + 4 * t) * KBV + kb) * 1024u);
    }
  {
    const int col = lane & 15, kr = lane >> 4;
    const float *Bp = sB + kr * BS + col;
    const float *Fe = sF + 16 * col;
    for (int mt = wave; mt < L::NPT; mt += 4) {
      if (mt != wave) {
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
          for (int b = 0; b < KBL; ++b)
            aL[f][b] = bld4_s(rAL, lane * 16u, (unsigned)((f * L::NPT + mt) * KBL + b) * 1024u);
      }
      f4 acc[4];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        acc[f] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < KBL; ++b) {
#pragma unroll
          for (int sb = 0; sb < 4; ++sb) {
            const int kt = 4 * b + sb;
            if (kt < KTL)
              acc[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(aL[f][b][sb], Bp[(f * Nfp + 4 * kt) * BS], acc[f], 0, 0, 0);
          }
        }
      }
      float nx[4], ny[4], nz[4];
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        nx[f] = Fe[4 * f + 1];
        ny[f] = Fe[4 * f + 2];
        nz[f] = Fe[4 * f + 3];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
          float dp = 0.f, du = 0.f, dv = 0.f, dw = 0.f;
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            dp += acc[f][r];
            du += nx[f] * acc[f][r];
            dv += ny[f] * acc[f][r];
            dw += nz[f] * acc[f][r];
          }
          float *rr = sR + col * RS + n;
          rr[0] += dp;
          rr[Np] -= du;
          rr[2 * Np] -= dv;
          rr[3 * Np] -= dw;
        }
      }
    }
  }
  __syncthreads();

  // ---- 6: S = (Vq . rhs) x c2 on MFMA (A from the zero-padded Vq, fixed offsets)
  at_mid();
  if (sh.yout && tid < Np) {
    // (A Q)_p before the projection, kept for the adjoint's c2 gradient
    const brsrc rY = mkrsrc(sh.yout + (size_t)k0 * Np, (unsigned)nk * Np * 4u);
#pragma unroll
    for (int e = 0; e < TB; ++e) bst(rY, sR[e * RS + tid], (unsigned)(e * Np + tid) * 4u);
  }
  f4 aP[KBP];
  if (wave < L::NPT) {
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld4_s(rAP, lane * 16u, (unsigned)(wave * KBP + kb) * 1024u);
  }
  {
    const int col = lane & 15, kr = lane >> 4;
    const float *Bp = sR + col * RS + kr;
#pragma unroll
    for (int t = 0; t < L::QTW; ++t) {
      const int qt = wave + 4 * t;
      if (qt >= L::NQT) break;
      f4 acc[NP];
#pragma unroll
      for (int b = 0; b < NP; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBV; ++kb) {
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < L::KTV) {
#pragma unroll
            for (int b = 0; b < NP; ++b)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aV[t][kb][sb], Bp[b * Np + kt * 4], acc[b], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = qt * 16 + 4 * kr + r;
        S[q * SX + col] = acc[0][r] * c2r[t][r];
#pragma unroll
        for (int b = 1; b < NP; ++b) S[q * SX + b * 16 + col] = q < Nq ? acc[b][r] : 0.f;
      }
    }
  }
  __syncthreads();

  // ---- 7: rhs = Pq . S on MFMA
  {
    const int col = lane & 15, kr = lane >> 4;
    for (int mt = wave; mt < L::NPT; mt += 4) {
      if (mt != wave) {
#pragma unroll
        for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld4_s(rAP, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u);
      }
      f4 acc[NP];
#pragma unroll
      for (int b = 0; b < NP; ++b) acc[b] = f4{0.f, 0.f, 0.f, 0.f};
      const float *Bp = S + kr * SX + col;
#pragma unroll
      for (int kb = 0; kb < KBP; ++kb) {
#pragma unroll
        for (int sb = 0; sb < 4; ++sb) {
          const int kt = 4 * kb + sb;
          if (kt < L::KTP) {
#pragma unroll
            for (int b = 0; b < NP; ++b)
              acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[kb][sb], Bp[kt * 4 * SX + b * 16], acc[b], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = mt * 16 + 4 * kr + r;
        if (n < Np) {
#pragma unroll
          for (int b = 0; b < NP; ++b) sR[col * RS + b * Np + n] = acc[b][r];
        }
      }
    }
  }
  __syncthreads();

  // ---- 8: LSRK stage: res = fa res + fdt rhs (+ source); Q_out = Q_in + fb res
  const float ft = sh.ft;
  if constexpr (L::PF) {
    if (tid < NpNf) {
      if (!has_src) {
#pragma unroll
        for (int e = 0; e < TB; ++e) {
          const float x = fa * rreg[e] + fdt * sR[e * RS + tid];
          if (sh.store_res) bst(rR, x, tid * 4u + e * ROW);
          bst(rO, sQ[e * NpNf + tid] + fb * x, tid * 4u + e * ROW);
        }
      } else {
        const int fr = tid / Np, nr = tid - fr * Np;
        const float wsel = w4(src, fr);
#pragma unroll
        for (int e = 0; e < TB; ++e) {
          float v = sR[e * RS + tid];
          const int slot = sSlot[e];
          if (slot >= 0) v += wsel * ft * src.coef[(size_t)slot * Np + nr];
          const float x = fa * rreg[e] + fdt * v;
          if (sh.store_res) bst(rR, x, tid * 4u + e * ROW);
          bst(rO, sQ[e * NpNf + tid] + fb * x, tid * 4u + e * ROW);
        }
      }
    }
  } else {
    for (int rr = tid; rr < NpNf; rr += 256) {
      const int fr = rr / Np, nr = rr - fr * Np;
      for (int e = 0; e < nk; ++e) {
        float v = sR[e * RS + rr];
        if (has_src) {
          const int slot = sSlot[e];
          if (slot >= 0) v += w4(src, fr) * ft * src.coef[(size_t)slot * Np + nr];
        }
        const unsigned g = rr * 4u + e * ROW;
        const float x = (read_res ? fa * bld(rR, g) : 0.f) + fdt * v;
        if (sh.store_res) bst(rR, x, g);
        bst(rO, sQ[e * NpNf + rr] + fb * x, g);
      }
    }
  }
  at_end();
  __syncthreads();
}

FusedOps fused_ops(const bbw_ctx *c);

}  // namespace bbw
