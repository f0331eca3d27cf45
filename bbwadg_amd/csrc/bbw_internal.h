// Internal definitions shared by the HIP translation units of libbbwadg_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bbwadg.h"

namespace bbw {

// Compile-time sizes for polynomial order N.
template <int N>
struct Dims {
  static constexpr int Np = (N + 1) * (N + 2) * (N + 3) / 6;
  static constexpr int Nfp = (N + 1) * (N + 2) / 2;
  static constexpr int NfpNf = 4 * Nfp;
  // points of the WADG rule: the degree 2N+1 tetrahedral rule for N <= 7, the (N+1)^3
  // tensor-product set for N = 8, 9
  static constexpr int Nq = N == 1 ? 6 : N == 2 ? 14 : N == 3 ? 31 : N == 4 ? 57 : N == 5 ? 95
                          : N == 6 ? 146 : N == 7 ? 214 : (N + 1) * (N + 1) * (N + 1);
  static constexpr int EEL_nnz = Nfp + 3;
  static constexpr int L0_nnz = Nfp < 7 ? Nfp : 7;
};

inline int np_of(int N) { return (N + 1) * (N + 2) * (N + 3) / 6; }
inline int nfp_of(int N) { return (N + 1) * (N + 2) / 2; }
inline int nq_of(int N) {
  static const int t[10] = {0, 6, 14, 31, 57, 95, 146, 214, 729, 1000};
  return (N >= 1 && N <= 9) ? t[N] : -1;
}
// Orders 7..9 run the acoustic plain stage on the flat kernels with the slice-by-slice lift;
// the tiled, fused, adjoint, batched, elastic and energy kernels are instantiated for N <= 6.
constexpr int BBW_MAX_ORDER_TILED = 6;

// Energy reductions: kinds and the fixed number of partial sums (one per block of the grid).
enum { BBW_ENERGY_QUADRATURE = 0, BBW_ENERGY_WADG = 1 };
constexpr int BBW_ENERGY_BLOCKS = 4096;

// Source description passed to the update kernels.
struct SrcArgs {
  int nsrc;
  int first, last;        // smallest / largest source element id
  const int *elems;       // ascending element ids
  const float *coef;      // [nsrc][Np]
  float w[9];             // per-field weights x amplitude x time function (set per stage)
  // per 16-element tile t = tfirst .. tlast: (slot of the tile's first source element, mask of
  // its source elements); null: the 16-element tile kernels search elems instead
  const int2 *tiles = nullptr;
  int tfirst = 0, tlast = -1;
};

// The per-tile source table of a sorted element list (SrcArgs::tiles), for tiles
// elems[0] / 16 .. elems[n - 1] / 16.
inline std::vector<int2> src_tile_table(const int32_t *elems, int n) {
  std::vector<int2> t;
  if (n <= 0) return t;
  const int t0 = elems[0] >> 4;
  t.assign((size_t)((elems[n - 1] >> 4) - t0 + 1), make_int2(0, 0));
  for (int i = n - 1; i >= 0; --i) {
    int2 &v = t[(size_t)((elems[i] >> 4) - t0)];
    v.x = i;   // descending, so the last write is the tile's first source element
    v.y |= 1 << (elems[i] & 15);
  }
  return t;
}

struct Graph {
  hipGraphExec_t exec = nullptr;
  int steps = 0, mode = -1, flags = -1, parity = -1;
  int form = -1;   // the elastic stage's form at capture (elas_form)
  double dt = 0;
};

}  // namespace bbw

struct bbw_ctx {
  int N = 0, Np = 0, Nfp = 0, Nq = 0, Nf = 0, K = 0, dev = 0;
  int ncu = 0;  // compute units of the device
  int KblkV = 0, KblkS = 0, KblkU = 0, KblkF = 0;
  int fused_impl = 1;  // BBW_FUSED_VECTOR or BBW_FUSED_MFMA
  int elas_fused = 1;  // BBW_ELAS_LIFT_NODES or BBW_ELAS_LIFT_CUBATURE (default by N: bbw_create)
  int xcd_chunk = 16;  // fused and adjoint stages: tiles per XCD run (BBW_OPT_XCD_CHUNK)
  // set inside the adjoint driver (misfit and gradient): its forward passes run the elastic stage
  // in the nodal form, whose A Q capture the reverse pass reads, so J and the states of every
  // forward pass agree bitwise
  int adj_pass = 0;
  int adj_impl = 1;    // 0: per-element adjoint kernels, 1: tiled MFMA adjoint kernels
  int plain_impl = 1;  // plain kernels: 0 flat per-element kernels (Kblk*), 1 16-element tiles,
                       // 2 flat kernels with the slice-by-slice lift
  hipStream_t own_stream = nullptr, stream = nullptr;
  std::string err;

  // operators
  int4 *Dids = nullptr;
  float4 *Dvals = nullptr;
  int *Fmask = nullptr, *EEL_ids = nullptr, *L0_ids = nullptr;
  float *EEL_vals = nullptr, *L0_vals = nullptr, *Vq = nullptr, *Pq = nullptr;
  // zero-padded copies for the MFMA kernels: VqPad [4*ceil(Np/4)][NQP] (j rows, q contiguous),
  // PqPad [NQP][NPP] (q rows, n contiguous); NQP, NPP = Nq, Np rounded up to 16
  float *VqPad = nullptr, *PqPad = nullptr;
  // the same for the adjoint: PqTPad [4*ceil(Np/4)][NQP] = Pq^T, VqTPad [NQP][NPP] = Vq^T
  float *PqTPad = nullptr, *VqTPad = nullptr;
  // EEL column ids as float4 slots and L0 ids with the face offset folded in, for the fused kernel
  int *EEL_ids4 = nullptr;
  // the lift EEL . blkdiag(L0, L0, L0, L0) per face, zero-padded for the MFMA kernels:
  // LiftPad [4 faces][KF = 4*ceil(Nfp/4) face nodes][NPP] (node contiguous)
  float *LiftPad = nullptr;
  // the MFMA A operands of the fused kernel in lane order: for M tile t and K-step block b,
  // 64 lanes x 4 consecutive K-steps, value A[16 t + (lane % 16)][4 (4 b + s) + lane / 16] at
  // float ((t * KB + b) * 64 + lane) * 4 + s (zero outside A).  VqBlk: A = Vq (M = cubature
  // points, K = nodes); PqBlk: A = Pq (M = nodes, K = cubature points); LiftBlk: per face f,
  // A = LIFT_f (M = nodes, K = face nodes), face blocks of NPT * KB tiles
  float *VqBlk = nullptr, *PqBlk = nullptr, *LiftBlk = nullptr;
  // the folded elastic stage's A = Vq [E | LIFT_0 | .. | LIFT_3] (M = cubature points, K = the
  // degree-(N-1) nodes of the factored derivative, then the face nodes of each face)
  float *VqCBlk = nullptr;
  // the same for the adjoint stage: PqTBlk: A = Pq^T (M = cubature points, K = nodes);
  // VqTBlk: A = Vq^T (M = nodes, K = cubature points); LiftTBlk: per face f, A = LIFT_f^T
  // (M = face nodes, K = nodes), face blocks of ceil(Nfp/16) * KB tiles
  float *PqTBlk = nullptr, *VqTBlk = nullptr, *LiftTBlk = nullptr;
  double *Vq64 = nullptr, *wq64 = nullptr;
  double *Pq64 = nullptr;  // (Vq64^T W Vq64)^-1 Vq64^T W, column-major Np x Nq (energy reductions)
  double rk4a[5] = {0}, rk4b[5] = {0}, rk4c[6] = {0};
  bool have_ops = false;
  // slice-by-slice lift (bbw_upload_slice_lift): per node and face, the node's row in that
  // face's concatenated slices plus (slice << 16) [Np] int4; one-step reductions [Np-Nfp][3]
  // (ids into the previous slice, values); cEL [N+1]
  int4 *slice_inv = nullptr;
  int *red_ids = nullptr;
  float *red_vals = nullptr, *cEL = nullptr;
  bool have_slice = false;
  // curvilinear path (bbw_upload_curved): nodal lane-ordered operator blocks, per-element
  // cubature geometry and the element lists
  bool curved = false;
  int Nfq = 0, Kcurved = 0, Kplanar = 0;
  float *cv_Vq = nullptr, *cv_Vr = nullptr, *cv_Vs = nullptr, *cv_Vt = nullptr;
  float *cv_Pq = nullptr, *cv_Pr = nullptr, *cv_Ps = nullptr, *cv_Pt = nullptr;
  float *cv_Vf = nullptr, *cv_Pf = nullptr;
  float *cv_Dr = nullptr, *cv_Ds = nullptr, *cv_Dt = nullptr, *cv_Lift = nullptr;  // planar elements
  int planar_split = 0;   // 1: planar elements take the split form with constant factors
  hipStream_t side_stream = nullptr;            // the curved list's stream (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  float *vgeoq = nullptr, *cv_Jq = nullptr, *fgeoq = nullptr, *Jplanar = nullptr;
  int *Klist_curved = nullptr, *Klist_planar = nullptr;

  // mesh
  float *vgeo = nullptr, *fgeo = nullptr;
  int *vmapP = nullptr;
  double *J = nullptr;
  bool have_mesh = false;

  // media
  float *c2q = nullptr, *Jq = nullptr;
  float *rhoinv = nullptr, *lambda = nullptr, *mu = nullptr;
  float tau_v = 1.f, tau_s = 1.f;
  bool have_media = false;

  // state
  float *Q[2] = {nullptr, nullptr};
  int cur = 0;
  float *res = nullptr, *rhs = nullptr;

  // source
  int nsrc = 0, src_first = 0, src_last = -1;
  int *src_elems = nullptr;
  float *src_coef = nullptr;
  int2 *src_tiles = nullptr;   // src_tile_table of src_elems
  int src_tfirst = 0, src_tlast = -1;
  float src_w[9] = {0};
  double src_f0 = 0, src_amp = 0;
  int src_step_time = 0;   // 1: every stage evaluates the source at its step's start time

  // receivers
  int nrec = 0, rec_max_steps = 0, rec_steps = 0;
  int rec_field = 0;   // the field the receivers sample
  int *rec_elems = nullptr;
  float *rec_basis = nullptr, *rec_trace = nullptr;

  // energy reductions: BBW_ENERGY_BLOCKS partial sums and the result
  double *partial = nullptr;

  // adjoint: transposed operators (D_m^T rows [4][Np] int4/float4, EEL^T [WE][4Nfp], L0^T
  // [WL][Nfp]), node -> face-node slots [Np][3], face-node neighbour map [K][4Nfp], buffers
  // exact BB-WADG update (bbw_upload_bbwadg): C(N,a) [Np], the folded projection
  // P diag(1/C(2N,g)) as lane-ordered MFMA A blocks (M = Np, K = N2p) and c2 in BB
  // coefficients [K][Np]
  float *bb_CN = nullptr, *bb_PBlk = nullptr, *c2bb = nullptr;
  int4 *DTids = nullptr;
  // factored Bernstein derivative (fused stage): shift [Np1] int4, elevation ids [Np] int4 and
  // weights [Np] float4
  int4 *Sids = nullptr, *Eids = nullptr;
  float4 *Evals = nullptr;
  float4 *Swts = nullptr;  // [Np1]: (b_1 + 1, .., b_4 + 1), weights of N E^T (adjoint volume)
  float4 *DTvals = nullptr;
  int *EELT_ids = nullptr, *L0T_ids = nullptr, *FmaskInv = nullptr, *mapP = nullptr;
  float *EELT_vals = nullptr, *L0T_vals = nullptr;
  int WE = 0, WL = 0;
  float *lQ = nullptr, *lres = nullptr, *Zadj = nullptr, *Hadj = nullptr, *grad = nullptr;
  float *Ybuf = nullptr, *ckpt = nullptr, *resid = nullptr, *Q0save = nullptr;
  float *dobs = nullptr;   // observations copied from the host [nsteps][nrec]
  size_t dobs_cap = 0;
  size_t Ybuf_len = 0, ckpt_len = 0, resid_len = 0;
  // Vq (A Q)_p of the final steps kept from the first forward pass (no recompute for them)
  float *Ycache = nullptr;
  size_t Ycache_len = 0;
  int ycache_mb = 0;  // device memory for Ycache in MiB: 0 = automatic, < 0 = none

  // graph cache and per-window stage data: ftime_win holds the window's stage values
  // [G][5] followed by its receiver rows [G][max(nrec, 1)]; ftime_win_cap is its length in floats
  bbw::Graph graph;
  float *ftime_win = nullptr;
  size_t ftime_win_cap = 0;
  // stage values of a whole bbw_step call: device copy, its pinned host staging buffer and the
  // event that marks the end of the last host-to-device copy from it (capacities in floats)
  float *ft_all = nullptr, *ft_host = nullptr;
  size_t ft_all_cap = 0, ft_host_cap = 0;
  hipEvent_t ft_evt = nullptr;
  bool ft_evt_pending = false;
};

namespace bbw {

// Launchers implemented in bbw_acoustic.hip / bbw_elastic.hip; return hipError_t.
hipError_t launch_volume(bbw_ctx *c);
hipError_t launch_surface(bbw_ctx *c);
hipError_t launch_update_wadg(bbw_ctx *c, float fa, float fb, float fdt, const float *ftime_dev,
                              float ftime);
hipError_t launch_update(bbw_ctx *c, float fa, float fb, float fdt);
hipError_t launch_update_wadg_mfma(bbw_ctx *c, float fa, float fb, float fdt, const float *ftime_dev,
                                   float ftime);
hipError_t launch_volume_tiled(bbw_ctx *c);
hipError_t launch_surface_elas_tiled(bbw_ctx *c);
hipError_t launch_volume_elas_tiled(bbw_ctx *c);
hipError_t launch_update_elas_mfma(bbw_ctx *c, float fa, float fb, float fdt, const float *ftime_dev,
                                   float ftime);
hipError_t launch_surface_tiled(bbw_ctx *c);
hipError_t launch_surface_slice(bbw_ctx *c);
hipError_t launch_curved_rhs(bbw_ctx *c);
int curved_nfq(int N);   // face cubature points of the curvilinear path (-1: order not compiled)
hipError_t launch_fused(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb, float fdt,
                        const float *ftime_dev, float ftime, int project_velocity);
hipError_t launch_fused_mfma(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb,
                             float fdt, const float *ftime_dev, float ftime, int project_velocity,
                             float *yout = nullptr, int store_res = 1);
// fused elastic stage: Qin -> Qout, res updated (stored unless store_res = 0); with yout, also
// A Q_in [K][9][Np] (the right-hand side before the WADG projection) for the adjoint
hipError_t launch_fused_elas(bbw_ctx *c, const float *Qin, float *Qout, float fa, float fb, float fdt,
                             const float *ftime_dev, float ftime, int store_res = 1, float *yout = nullptr);
// reverse elastic stage (one launch): the neighbour flux terms of the previous reverse stage come in
// through Hin, this stage's go out through Hout (see bbw_adjoint_elas.hip)
hipError_t launch_adj_elas_stage(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io,
                                 const float *Hin, float *Hout);
hipError_t launch_adj_stage(bbw_ctx *c, const float *Y, float fa, float fb, float fdt);
// lres_io bit 0: read the adjoint residual (else it is zero); bit 1: store it (else it is dead)
hipError_t launch_adj_stage_tiled(bbw_ctx *c, const float *Y, float fa, float fb, float fdt, int lres_io = 3);
hipError_t launch_adj_inject(bbw_ctx *c, const float *resid_row);
// energy of the current state (kind BBW_ENERGY_*): partial sums in partial[BBW_ENERGY_BLOCKS],
// the total in *out (device memory)
hipError_t launch_energy(bbw_ctx *c, int kind, double *partial, double *out);
int energy_blocks(const bbw_ctx *c, int kind);   // partial sums the reduction writes
// receiver misfit 1/2 sum (tr - d)^2 over n values, FP64, partial sums in
// partial[BBW_ENERGY_BLOCKS] and the total in *out (device); resid (or null) gets tr - d
hipError_t launch_misfit(bbw_ctx *c, const float *tr, const float *d, long long n, float *resid,
                         double *partial, double *out);
// the device address of the energy scratch array (BBW_ENERGY_BLOCKS + 1 doubles) on device dev:
// looked up once per device by energy_scratch_init (called by bbw_create), null before
hipError_t energy_scratch_init(int dev);
double *energy_scratch(int dev);
hipError_t launch_record(bbw_ctx *c, const float *Q, float *row);
hipError_t launch_update_bbwadg(bbw_ctx *c);

hipError_t launch_volume_elas(bbw_ctx *c);
hipError_t launch_surface_elas(bbw_ctx *c);
hipError_t launch_update_elas(bbw_ctx *c, float fa, float fb, float fdt, const float *ftime_dev,
                              float ftime);

bool order_supported(int N);

// The context's source as the update kernels take it (weights as set, time function applied
// by the kernels).
inline SrcArgs ctx_src(const bbw_ctx *c) {
  SrcArgs s;
  s.nsrc = c->nsrc;
  s.first = c->src_first;
  s.last = c->src_last;
  s.elems = c->src_elems;
  s.coef = c->src_coef;
  for (int f = 0; f < 9; ++f) s.w[f] = c->src_w[f];
  s.tiles = c->src_tiles;
  s.tfirst = c->src_tfirst;
  s.tlast = c->src_tlast;
  return s;
}

// the form of the fused elastic stage a launch runs (BBW_ELAS_LIFT_*): the nodal form inside
// the adjoint driver, else the context's option
inline int elas_form(const bbw_ctx *c) { return c->adj_pass ? BBW_ELAS_LIFT_NODES : c->elas_fused; }

}  // namespace bbw
