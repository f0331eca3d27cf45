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

// Source elements of the 16-element tile at k0 (nk elements): .y a mask of the elements that
// may carry a source (0: none, uniform over the workgroup), .x the slot of the first.  With the
// per-tile table this is one scalar load and the mask is exact; without it, the elements are
// sorted, so a tile outside [first, last] has none and the others search the list.
__device__ __forceinline__ int2 src_tile(const SrcArgs &s, int k0, int nk) {
  if (s.nsrc == 0) return make_int2(0, 0);
  if (s.tiles) {
    const int t = k0 >> 4;
    return (t < s.tfirst || t > s.tlast) ? make_int2(0, 0) : s.tiles[t - s.tfirst];
  }
  return make_int2(0, (k0 + nk - 1 >= s.first && k0 <= s.last) ? (1 << nk) - 1 : 0);
}

// slot of the tile's element e (-1: no source) from src_tile's value st
__device__ __forceinline__ int src_slot_tile(const SrcArgs &s, int2 st, int k0, int e) {
  if (e >= 16 || !((st.y >> e) & 1)) return -1;
  if (s.tiles) return st.x + __builtin_popcount(st.y & ((1 << e) - 1));
  return src_slot2(s, k0 + e);
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
  // W2: volume step-1 values F float4 [TB][Np1] | rhs [TB][RS]
  static constexpr int W2 = cmax(4 * Np, RS);
  static constexpr int QTW = (NQT + 3) / 4;     // q tiles per wave
  static constexpr int GS = ((Np + 63) / 64) * 64, NG = 256 / GS;        // node groups
  static constexpr int NG3 = 256 / NfpNf, EPT3 = (TB + NG3 - 1) / NG3;   // face-node groups
  static constexpr int EPTV = (TB + NG - 1) / NG;
  // degree N-1 nodes of the factored derivative (volume step 1): groups of 32 or 64 lanes
  static constexpr int Np1 = N * (N + 1) * (N + 2) / 6;
  static constexpr int GS1 = Np1 <= 32 ? 32 : 64, NG1 = 256 / GS1, EPT1 = TB / NG1;
  static_assert(Np1 <= 64 && NG1 * EPT1 == TB, "volume step 1 covers the tile");
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
  const float *VqCBlk;                    // the folded elastic stage's operator (bbw_ctx::VqCBlk)
  const int4 *Sids;     // [Np1]: degree-N nodes b + e_1 .. b + e_4 of degree-(N-1) node b
  const int4 *Eids;     // [Np]: degree-(N-1) nodes a - e_1 .. a - e_4 of degree-N node a
  const float4 *Evals;  // [Np]: (a_1, a_2, a_3, a_4), the elevation weights
};

struct FusedMesh {
  int K;
  const float *vgeo, *fgeo;
  const int *vmapP;
  int xcd_chunk;  // tiles per XCD run (power of 2), 0: the dispatch order (tile = workgroup)
};

// Tile of this workgroup.  Workgroups go to the 8 XCDs round-robin (b mod 8), so in the
// dispatch order consecutive tiles, and the neighbours whose traces a tile gathers, sit in
// different L2s.  With xcd_chunk = C, XCD x takes the runs of C consecutive tiles
// x, x + 8, x + 16, ... (run r = tiles rC .. rC + C - 1): workgroup b = 8i + x takes tile
// (8 (i / C) + x) C + i mod C.  The workgroups past the last whole 8C span keep b.
__device__ __forceinline__ int tile_index(const int C) {
  const int b = blockIdx.x;
  if (C == 0) return b;
  const int lc = __builtin_ctz(C), full = (int)gridDim.x & ~(8 * C - 1);
  if (b >= full) return b;
  const int i = b >> 3;
  return ((((i >> lc) << 3) + (b & 7)) << lc) + (i & (C - 1));
}

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
typedef unsigned u2v __attribute__((ext_vector_type(2)));
typedef unsigned u3v __attribute__((ext_vector_type(3)));
// Block kb of an operand with KT K-steps.  In the last block only the KT mod 4 K-steps that
// the products use are loaded and the others are zero constants: a load into registers that
// are never read lets the compiler reuse them while the load is in flight, and the reuse then
// waits for that load and every load issued before it (the next phase's prefetch included).
// For fully unrolled block loops only (kb a compile-time constant).  In a partly unrolled loop
// the choice becomes a run-time branch; the exact BB-WADG update's product loop (unrolled by 2,
// four accumulator chains, the K-steps past KT skipped) compiled that way at N = 3 with this
// compiler has, on the path that skips accumulator 1's last MFMA, four v_accvgpr_mov that copy
// one component of another accumulator into all four of accumulator 1, which loses its sum (the
// wrong results seen at N = 3).  With the 128-bit load in every block the copy is not emitted,
// so that loop keeps it; tests/test_codegen.py checks the shipped kernels' code for the pattern.
template <int KT>
__device__ __forceinline__ f4 bld_blk(brsrc r, unsigned voff, unsigned soff, int kb) {
  constexpr int KB = (KT + 3) / 4, REM = KT - 4 * (KB - 1);
  if (REM == 4 || kb != KB - 1) return bld4_s(r, voff, soff);
  if constexpr (REM == 1) {
    return f4{bld_s(r, voff, soff), 0.f, 0.f, 0.f};
  } else if constexpr (REM == 2) {
    const u2v v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return f4{__uint_as_float(v[0]), __uint_as_float(v[1]), 0.f, 0.f};
  } else {
    const u3v v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
    return f4{__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), 0.f};
  }
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
  // unconditional (m3 < NfpNf for every thread; threads past the face-node groups do not use
  // it), so that the load is issued with the tile's loads and not behind the first barrier
  const int g3 = tid / L::NfpNf, m3 = tid - g3 * L::NfpNf;
  return o.Fmask[m3];
}

struct NoHook {
  __device__ void operator()() const {}
};

// The register-resident inputs of a tile's stage (phase 1): the thread's rows of the volume
// shift/elevation tables, res rows, c2 in the MFMA output layout and the neighbour traces of
// the thread's face nodes.  Issued before the tile's Q_in rows are stored to LDS, so the
// vmapP -> trace chain overlaps the tile loads and no phase waits on a table load after its
// barrier.
template <int N, bool PROJ_U>
struct TilePre {
  int4 S, E;
  float4 ev;
  float rreg[FL<N, PROJ_U>::PF ? FL<N, PROJ_U>::TB : 1];
  float c2r[FL<N, PROJ_U>::QTW][4];
  float tr[FL<N, PROJ_U>::EPT3][4];
};

// Issues the loads of TilePre.  With fa = 0 (the first LSRK45 stage) res is not read: fa res = 0
// for any finite res.  Requires K * 4 * Np * 4 < 2^32 (32-bit trace offsets).
template <int N, bool PROJ_U>
__device__ __forceinline__ void tile_prefetch(TilePre<N, PROJ_U> &pf, const FusedMesh &m, const FusedOps &o,
                                              const FusedShot &sh, float fa, int k0, int nk,
                                              const int (&vm)[FL<N, PROJ_U>::EPT3], const int tid) {
  using L = FusedLayout<N, PROJ_U>;
  constexpr int Np = L::Np, Nq = L::Nq, TB = L::TB;
  constexpr unsigned ROW = 4 * Np * 4u;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  {
    // rows of the thread's degree-(N-1) node (phase 3) and node (phase 3b); lanes past them
    // read row 0 and do not use it
    const int n1 = tid % L::GS1, nv = tid % L::GS;
    pf.S = o.Sids[n1 < L::Np1 ? n1 : 0];
    pf.E = o.Eids[nv < Np ? nv : 0];
    pf.ev = o.Evals[nv < Np ? nv : 0];
  }
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
  float *W1 = lds + L::OW1;              // [TB * W1]
  float *W2 = lds + L::OW2;              // [TB * W2]
  const float *sF = sG + TB * 9;         // [TB][16]
  int *sSlot = (int *)(lds + L::OSL);
  float *sB = W1;                        // phases 3-5: face flux [KB][BS]
  float *S = W1;                         // phases 6-7
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
  // source elements of the tile (uniform test)
  static_assert(TB == 16, "the per-tile source table is built for 16-element tiles");
  const int2 stile = src_tile(src, k0, nk);
  const bool has_src = stile.y != 0;

  // ---- 1: source slots of the tile's elements (read in phase 8, after several barriers); the
  // register inputs (res, c2, traces) were issued by tile_prefetch before the Q_in tile's barrier
  if (has_src && tid < TB) sSlot[tid] = src_slot_tile(src, stile, k0, tid);
  at_start();

  // ---- 2: upwind flux at face nodes into registers.  The p and U fluxes are (phi, -phi) with
  // phi = (pjump - Unjump) Fscale / 2, so one value per face node.
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

  // ---- 3: face flux -> LDS [face node][element]; volume step 1 from the Q_in tile.
  // The derivative along barycentric m is D_m u [a] = sum_j a_j u[a - e_j + e_m]: the shift to
  // degree N - 1, u[b + e_m], then elevation.  The volume terms are linear in the shifted values,
  // so step 1 forms, at each degree-(N-1) node b, F[b] = (H, X, Y, Z) with
  //   H = -(Ur[b+e2] + Us[b+e3] + Ut[b+e4] - (Ur + Us + Ut)[b+e1]) / 2     (Ur = U . grad r, ...)
  //   (X, Y, Z) = -G^T ((p[b+e2] - p[b+e1]) / 2, (p[b+e3] - p[b+e1]) / 2, (p[b+e4] - p[b+e1]) / 2)
  // and step 2 (phase 3b) elevates: vol[a] = sum_j a_j F[a - e_j].
  if (on3) {
#pragma unroll
    for (int i = 0; i < L::EPT3; ++i) {
      const int e = g3 + L::NG3 * i;
      if (L::NG3 * L::EPT3 == TB || e < TB) sB[m3 * BS + e] = phi[i];
    }
  }
  if (tid < (KB - NfpNf) * BS) sB[NfpNf * BS + tid] = 0.f;   // K padding of the last face
  float4 *sF1 = (float4 *)W2;                                // [TB][Np1]: F
  {
    const int g1 = tid / L::GS1, n1 = tid - g1 * L::GS1;
    if (n1 < L::Np1) {
      const int4 S = pf.S;
#pragma unroll
      for (int i = 0; i < L::EPT1; ++i) {
        const int e = g1 + L::NG1 * i;
        const float *q = sQ + e * NpNf;
        const float *G = sG + 9 * e;
        const float p1 = q[S.x], p2 = q[S.y], p3 = q[S.z], p4 = q[S.w];
        const float ur = q[Np + S.y] * G[0] + q[2 * Np + S.y] * G[1] + q[3 * Np + S.y] * G[2];
        const float us = q[Np + S.z] * G[3] + q[2 * Np + S.z] * G[4] + q[3 * Np + S.z] * G[5];
        const float ut = q[Np + S.w] * G[6] + q[2 * Np + S.w] * G[7] + q[3 * Np + S.w] * G[8];
        const float u1 = q[Np + S.x] * (G[0] + G[3] + G[6]) + q[2 * Np + S.x] * (G[1] + G[4] + G[7]) +
                         q[3 * Np + S.x] * (G[2] + G[5] + G[8]);
        const float a = .5f * (p2 - p1), b = .5f * (p3 - p1), cc = .5f * (p4 - p1);
        sF1[e * L::Np1 + n1] = make_float4(-.5f * ((ur + us + ut) - u1), -(G[0] * a + G[3] * b + G[6] * cc),
                                           -(G[1] * a + G[4] * b + G[7] * cc), -(G[2] * a + G[5] * b + G[8] * cc));
      }
    }
  }
  __syncthreads();

  // ---- 3b: volume step 2 (elevation) into registers: vol[a] = sum_j a_j F[a - e_j]
  float vol[L::EPTV][4];
  if (onv) {
    const int4 E = pf.E;
    const float4 ev = pf.ev;
    const float4 *Fb = sF1 + gv * L::Np1;
#pragma unroll
    for (int i = 0; i < L::EPTV; ++i) {
      const int o4 = i * L::NG * L::Np1;
      const float4 f1 = Fb[E.x + o4], f2 = Fb[E.y + o4], f3 = Fb[E.z + o4], f4 = Fb[E.w + o4];
      vol[i][0] = ev.x * f1.x + ev.y * f2.x + ev.z * f3.x + ev.w * f4.x;
      vol[i][1] = ev.x * f1.y + ev.y * f2.y + ev.z * f3.y + ev.w * f4.y;
      vol[i][2] = ev.x * f1.z + ev.y * f2.z + ev.z * f3.z + ev.w * f4.z;
      vol[i][3] = ev.x * f1.w + ev.y * f2.w + ev.z * f3.w + ev.w * f4.w;
    }
  }
  __syncthreads();

  // ---- 4: volume terms -> rhs (F is no longer read)
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
  // The prefetches are unconditional (a wave past the last tile loads the last one and does not
  // use it): loads under a wave test leave the compiler unsure how many are in flight where the
  // paths join, and its waits then cover the prefetches as well.
  f4 aL[4][KBL];
  {
    const int mt0 = wave < L::NPT ? wave : L::NPT - 1;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int b = 0; b < KBL; ++b)
        aL[f][b] = bld_blk<KTL>(rAL, lane * 16u, (unsigned)((f * L::NPT + mt0) * KBL + b) * 1024u, b);
  }
  __syncthreads();

  // ---- 5: lift on MFMA, one product per face (M = nodes, N = 16 elements, K = face nodes):
  // P_f = LIFT_f phi_f;  rhs_p += sum_f P_f,  rhs_U -= sum_f n_f P_f
  f4 aV[L::QTW][KBV];
#pragma unroll
  for (int t = 0; t < L::QTW; ++t) {
    const int qt = wave + 4 * t < L::NQT ? wave + 4 * t : L::NQT - 1;
#pragma unroll
    for (int kb = 0; kb < KBV; ++kb)
      aV[t][kb] = bld_blk<L::KTV>(rAV, lane * 16u, (unsigned)(qt * KBV + kb) * 1024u, kb);
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
            aL[f][b] = bld_blk<KTL>(rAL, lane * 16u, (unsigned)((f * L::NPT + mt) * KBL + b) * 1024u, b);
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
  {
    const int mt0 = wave < L::NPT ? wave : L::NPT - 1;
#pragma unroll
    for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<L::KTP>(rAP, lane * 16u, (unsigned)(mt0 * KBP + kb) * 1024u, kb);
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
        for (int kb = 0; kb < KBP; ++kb) aP[kb] = bld_blk<L::KTP>(rAP, lane * 16u, (unsigned)(mt * KBP + kb) * 1024u, kb);
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
      float x[TB];
      if (!has_src) {
#pragma unroll
        for (int e = 0; e < TB; ++e) x[e] = fa * rreg[e] + fdt * sR[e * RS + tid];
      } else {
        const int fr = tid / Np, nr = tid - fr * Np;
        const float wsel = w4(src, fr);
#pragma unroll
        for (int e = 0; e < TB; ++e) {
          float v = sR[e * RS + tid];
          const int slot = sSlot[e];
          if (slot >= 0) v += wsel * ft * src.coef[(size_t)slot * Np + nr];
          x[e] = fa * rreg[e] + fdt * v;
        }
      }
#pragma unroll
      for (int e = 0; e < TB; ++e) bst(rO, sQ[e * NpNf + tid] + fb * x[e], tid * 4u + e * ROW);
      if (sh.store_res) {
#pragma unroll
        for (int e = 0; e < TB; ++e) bst(rR, x[e], tid * 4u + e * ROW);
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
        const float x = fa * bld(rR, g) + fdt * v;
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
