/*
 * bbwadg.h -- C ABI of libbbwadg_hip.so, the MI355X (gfx950) BB-WADG wave solver.
 *
 * The library replaces the device half of the reference's solver: the OCCA
 * device/memory/kernel globals and their launches in cpp/src/WaveOKL3d.cpp
 * (and elas_cpp/src/WaveOKL3d.cpp).  The reference's host setup (mesh,
 * operators, geometric factors, maps: Mesh3d.cpp, StartUp3d.cpp,
 * InitWADG_subelem) keeps producing the arrays; these entry points take them
 * as plain pointers in the reference's layouts:
 *
 *   Q / rhsQ / resQ   float [K][Nfields][Np]
 *   vgeo              float [K][9]     (rx ry rz sx sy sz tx ty tz)
 *   fgeo              float [K][4][4]  (Fscale nx ny nz per face)
 *   vmapP             int32 [K][4*Nfp] flat index into Q of the exterior trace
 *   Fmask             int32 [4][Nfp]
 *   D_ids             int32 [4][Np][4]: D_ids[j][i] = (D1[i][j],D2[i][j],D3[i][j],D4[i][j])
 *   Dvals             float [Np][4]
 *   EEL_ids/vals      [EEL_nnz][Np] (column-major Np x EEL_nnz)
 *   L0_ids/vals       [L0_nnz][Nfp] (column-major Nfp x L0_nnz)
 *   Vq                float, column-major Nq x Np   (entry (q,j) at q + j*Nq)
 *   Pq                float, column-major Np x Nq   (entry (n,q) at n + q*Np)
 *   c2q, Jq           float [K][Nq]
 *
 * Conventions: every function returns 0 (BBW_OK) or a negative error class;
 * bbw_last_error() describes the last failure on a context.  Host pointers are
 * borrowed for the duration of the call.  All device work is enqueued on the
 * context's stream (bbw_set_stream may substitute the caller's); copies back to
 * the host synchronize that stream.  One context per host thread.
 */
#ifndef BBWADG_H
#define BBWADG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BBW_OK 0
#define BBW_ERR_ARG (-1)          /* invalid argument or size */
#define BBW_ERR_HIP (-2)          /* HIP runtime error */
#define BBW_ERR_STATE (-3)        /* required data not uploaded yet */
#define BBW_ERR_UNSUPPORTED (-4)  /* order / field count not compiled */
#define BBW_ERR_NOMEM (-5)        /* device allocation failed */

#define BBW_MODE_PLAIN 0  /* volume, surface, update launches per stage (RK_step_WADG_subelem) */
#define BBW_MODE_FUSED 1  /* one fused stage kernel per stage, ping-pong state */
#define BBW_MODE_CURVED 2 /* curvilinear context (bbw_upload_curved): RK_step_WADG's stage */

typedef struct bbw_ctx bbw_ctx;

typedef struct {
  int N;        /* polynomial order: 1..6 every path; 7..9 acoustic plain stages (flat kernels,
                   slice-by-slice lift), see bbw_upload_slice_lift */
  int Nfields;  /* 4 acoustic (p,u,v,w), 9 elastic (v1,v2,v3,sxx,syy,szz,syz,sxz,sxy) */
  int K;        /* number of elements */
  int Nq;       /* WADG cubature points per element */
  int KblkV, KblkS, KblkU, KblkF; /* elements per block: volume, surface, update, fused (0 = default) */
  int device;   /* HIP device ordinal */
} bbw_config;

typedef struct {
  const int32_t *Fmask;
  const int32_t *D_ids;
  const float *Dvals;
  const int32_t *EEL_ids;
  const float *EEL_vals;
  int EEL_nnz;
  const int32_t *L0_ids;
  const float *L0_vals;
  int L0_nnz;
  const float *Vq;
  const float *Pq;
  const double *wq;   /* cubature weights (Nq), FP64: energy reduction */
  const double *Vq64; /* Vq in FP64 (same layout as Vq) for the energy reduction; NULL: Vq */
  const double *rk4a; /* LSRK45 coefficients (5, 5, 6 values), FP64 */
  const double *rk4b;
  const double *rk4c;
} bbw_ops;

/* ≈ WaveInitOCCA3d, device part (cpp/src/WaveOKL3d.cpp:1400-1680): device,
 * stream and state buffers Q/resQ/rhsQ (zeroed). */
int bbw_create(const bbw_config *cfg, bbw_ctx **out);
void bbw_destroy(bbw_ctx *ctx);
const char *bbw_last_error(const bbw_ctx *ctx);
int bbw_set_stream(bbw_ctx *ctx, void *hip_stream); /* NULL restores the context's own stream */
int bbw_sync(bbw_ctx *ctx);

/* ≈ the operator uploads of WaveInitOCCA3d (:1426-1526) and InitWADG_subelem (:906-915). */
int bbw_upload_operators(bbw_ctx *ctx, const bbw_ops *ops);
/* ≈ the slice-lift upload of WaveInitOCCA3d (slice_ids, EEL_val_vec / EEL_id_vec and cEL,
 * cpp/src/WaveOKL3d.cpp:1515-1526; tables built in cpp/src/StartUp3d.cpp:409-505) for the
 * surface kernel rk_surface_bern_slice (cpp/okl/WaveKernels.okl:652-939).  The lift of face f
 * puts cEL[s] (E_{N-s}^N)^T (L0 flux_f) on slice s of that face (the nodes s layers from it),
 * computed slice by slice with one-step degree reductions:
 *   slice_ids  int32 [Np][4]: row sk of face f (slices 0..N concatenated, slice s holding
 *              (N-s+1)(N-s+2)/2 nodes) -> volume node; slice 0 must be the face in Fmask order
 *   red_ids    int32 [Np-Nfp][3], red_vals float [Np-Nfp][3]: row r of slice s >= 1 (rows of
 *              slices 1..N concatenated) = sum_c red_vals * (slice s-1 value at red_ids), the
 *              rows of (E_{N-s}^{N-s+1})^T; ids index the previous slice, unused entries 0
 *   cEL        float [N+1]
 * Required for N >= 7 (the default surface kernel there) and for BBW_PLAIN_SLICE. */
int bbw_upload_slice_lift(bbw_ctx *ctx, const int32_t *slice_ids, const int32_t *red_ids,
                          const float *red_vals, const float *cEL);
/* ≈ the vgeo/fgeo/vmapP uploads of WaveInitOCCA3d (:1528-1616).  J: FP64 [K] (energy). */
int bbw_upload_mesh(bbw_ctx *ctx, const float *vgeo, const float *fgeo, const int32_t *vmapP,
                    const double *J);
/* ≈ c_c2q / c_Jq_reduced uploads of InitWADG_subelem (:906-909).  Jq may be NULL (1/J = 1). */
int bbw_upload_media(bbw_ctx *ctx, const float *c2q, const float *Jq);
/* The same from device memory (no reference counterpart): c2q_dev [K][Nq] is copied on the
 * context's stream (Jq = 1). */
int bbw_upload_media_device(bbw_ctx *ctx, const float *c2q_dev);
/* Elastic media at cubature (elas_cpp/src/WaveOKL3d.cpp:445-511): rhoinv (the kernel's rhoq
 * multiplies), lambda, mu, each [K][Nq]; penalties tau_v, tau_s (:516-517). */
int bbw_upload_elastic_media(bbw_ctx *ctx, const float *rhoinvq, const float *lambdaq,
                             const float *muq, float tau_v, float tau_s);
/* The same from device memory (no reference counterpart), copied on the context's stream. */
int bbw_upload_elastic_media_device(bbw_ctx *ctx, const float *rhoinvq_dev, const float *lambdaq_dev,
                                    const float *muq_dev, float tau_v, float tau_s);
/* Point-like source: nsrc elements (ascending ids), BB coefficients coef[nsrc][Np] added to
 * the right-hand side of each field f as amp * w[f] * ricker(t, f0) * coef (elas_cpp fsrc/ftime,
 * :487-538, :1677-1680).  weights: Nfields floats.  nsrc = 0 removes the source. */
int bbw_set_source(bbw_ctx *ctx, int nsrc, const int32_t *elems, const float *coef,
                   const float *weights, double f0, double amp);
/* Receivers: element id and BB basis values at the receiver point ([nrec][Np]); field 0 is
 * recorded after every step of bbw_step into a trace [nsteps][nrec]. */
int bbw_set_receivers(bbw_ctx *ctx, int nrec, const int32_t *elems, const float *basis,
                      int max_steps);
int bbw_get_traces(bbw_ctx *ctx, float *out, int nsteps);

/* ≈ WaveSetData3d / WaveGetData3d device copies (:2527-2537).  Q in BB coefficients.
 * bbw_set_state(ctx, NULL) sets the state to zero on the device. */
int bbw_set_state(bbw_ctx *ctx, const float *Q);
int bbw_get_state(bbw_ctx *ctx, float *Q);
int bbw_set_res(bbw_ctx *ctx, const float *res);
int bbw_get_res(bbw_ctx *ctx, float *res);
int bbw_set_rhs(bbw_ctx *ctx, const float *rhs);
int bbw_get_rhs(bbw_ctx *ctx, float *rhs);

/* Single launches on the context's current state, positional equivalents of:
 *   bbw_volume       rk_volume_bern  (cpp/okl/WaveKernels.okl:344)      rhsQ  = vol(Q)
 *   bbw_surface      rk_surface_bern (cpp/okl/WaveKernels.okl:479)      rhsQ += lift(flux(Q))
 *   bbw_update_wadg  rk_update_WADG  (cpp/okl/WaveKernelsWADG.okl:1760) / rk_update_elas
 *                    (elas_cpp/okl/ElasKernelsWADG.okl:286): projection + LSRK stage
 *   bbw_update       rk_update       (cpp/okl/WaveKernels.okl:262)
 * ftime is the source time-function value for this stage (0 without a source). */
int bbw_volume(bbw_ctx *ctx);
int bbw_surface(bbw_ctx *ctx);
int bbw_update_wadg(bbw_ctx *ctx, float fa, float fb, float fdt, float ftime);
int bbw_update(bbw_ctx *ctx, float fa, float fb, float fdt);
/* One fused stage (volume + surface + WADG + LSRK) reading the current state and writing
 * the other state buffer, which becomes current.  project_velocity: 1 projects the
 * velocity fields through Pq*Vq as rk_update_WADG does, 0 uses Pq*Vq = I. */
int bbw_stage_fused(bbw_ctx *ctx, float fa, float fb, float fdt, float ftime, int project_velocity);

/* Curvilinear WADG (the reference's main run: InitQuadratureArrays, InitWADG_curved,
 * GordonHallSphere, RK_step_WADG; cpp/src/StartUp3d.cpp:829-960, cpp/src/WaveOKL3d.cpp:164-417,
 * :929-1297, :2111-2185).  The context holds NODAL coefficients (the reference builds this path
 * with the nodal basis).  Operators, single precision, the reference's column-major layouts:
 *   Vq, Vrq, Vsq, Vtq  Nq x Np   nodal interpolation / r, s, t derivatives at the volume rule
 *                                (the degree 2N+1 rule of the WADG update, Nq = bbw_config.Nq)
 *   Pq, Prq, Psq, Ptq  Np x Nq   M^-1 V*^T W
 *   VfqFace            Nfq x Nfp face nodes (Fmask order) -> the face rule
 *   Pfq                Np x 4Nfq M^-1 Vfq^T Wf (faces in order)
 *   Dr, Ds, Dt         Np x Np   nodal derivatives at the nodes (planar elements)
 *   LIFT               Np x 4Nfp nodal lift (planar elements)
 * Vq and Pq also serve as the WADG update's Vq_reduced / Pq_reduced (the same rule; its 1/J is
 * bbw_upload_media's Jq).  Per element: vgeoq [K][9][Nq] (rx sx tx ry sy ty rz sz tz at the volume
 * points), Jq [K][Nq], fgeoq [K][4][4][Nfq] (per face nx ny nz sJ at the face points), Jplanar [K],
 * and the element lists KlistCurved / KlistPlanar, which must partition 0..K-1.  Planar elements
 * use bbw_upload_mesh's vgeo / fgeo.  Orders 1..6, acoustic. */
typedef struct {
  int Nfq;
  const float *Vq, *Vrq, *Vsq, *Vtq;
  const float *Pq, *Prq, *Psq, *Ptq;
  const float *VfqFace;
  const float *Pfq;
  const float *Dr, *Ds, *Dt;
  const float *LIFT;
  const int32_t *Fmask;
  const double *wq;
  const double *rk4a, *rk4b, *rk4c;
} bbw_curved_ops;
int bbw_upload_curved(bbw_ctx *ctx, const bbw_curved_ops *ops, const float *vgeoq, const float *Jq,
                      const float *fgeoq, const float *Jplanar, int Kcurved,
                      const int32_t *KlistCurved, int Kplanar, const int32_t *KlistPlanar);
/* ≈ rk_volume_WADG_skew_combine + rk_surface_WADG_skew_combine on KlistCurved and
 * rk_volume_planar + rk_surface_planar on KlistPlanar (cpp/okl/WaveKernelsWADG.okl:514-682,
 * :3024-3195, :2203-2441): rhsQ = the split-form right-hand side of the current state times J.
 * bbw_update_wadg then applies rk_update_WADG with 1/J. */
int bbw_curved_rhs(bbw_ctx *ctx);

/* Exact Bernstein-Bezier WADG pressure update, ≈ rk_update_BBWADG
 * (cpp/okl/WaveKernelsBBWADG.okl:14-321, driven by test_RK, cpp/src/WaveOKL3d.cpp:2288-2296):
 * rhs_p <- P_N(c2 rhs_p) in place on the context's rhsQ, with c2 given by its degree-N BB
 * coefficients and P_N the L2 projection of the degree-2N product.  The operators are the
 * reference's (InitWADG_subelem BB branch, cpp/src/WaveOKL3d.cpp:486-690), FP64 on the host:
 * subtet_ids [Np][Np] degree-2N index of a + b; CNscale [Np] = C(N,a); invC2Nscale [N2p] =
 * 1/C(2N,g); Ei_ids4 / Ei_vals4 [2N][N2p][4] one-step elevations E_i (degree i -> i+1) as
 * 4-wide sparse rows; cj [N+1] of P = sum_j c_j E_{N-j}^N (E_{N-j}^{2N})^T; c2_bb [K][Np]
 * (= Pq c2q per element).  N2p = (2N+1)(2N+2)(2N+3)/6.  The projection is folded into one
 * dense Np x N2p matrix at upload (applied on MFMA); subtet_ids is checked against the index
 * formula the kernel evaluates. */
int bbw_upload_bbwadg(bbw_ctx *ctx, const int32_t *subtet_ids, const double *CNscale,
                      const double *invC2Nscale, const int32_t *Ei_ids4, const double *Ei_vals4,
                      const double *cj, const float *c2_bb);
int bbw_update_bbwadg(bbw_ctx *ctx);

/* ≈ Wave_RK inner loop (cpp/src/WaveOKL3d.cpp:2029-2080): nsteps LSRK45 steps of size dt
 * from time t0; stage coefficients cast to float as the reference does.  mode: BBW_MODE_*.
 * flags bit 0: project velocity in the fused path; bit 1: capture steps in a HIP graph;
 * bit 2 (with bit 1, no reference counterpart): prepare only -- capture, instantiate and upload
 * the graph for these arguments and the current state buffer, and size the stage-time buffers
 * for a call of nsteps steps, without running a step (so that a later call's time excludes it). */
#define BBW_STEP_PROJECT_U 1
#define BBW_STEP_GRAPH 2
#define BBW_STEP_PREPARE 4
int bbw_step(bbw_ctx *ctx, int nsteps, double dt, double t0, int mode, int flags);

/* FP64 energies of the current state, reduced on the device (no C++ counterpart in the
 * reference; the weighted-mass form follows matlab/WADG/check_conservation_loss.m:51-54).
 *   bbw_energy       quadrature form 1/2 sum_K J sum_q w_q e_q with e = p^2 / c2 + |u|^2
 *                    (acoustic) or rho |v|^2 + sigma : C^-1 sigma (elastic, isotropic compliance)
 *   bbw_energy_wadg  the energy the WADG scheme dissipates: 1/2 sum_K J (MB x)^T A_m^-1 (MB x)
 *                    per block of fields, A_m = Vq^T W m Vq with the block's media m (c2 on p,
 *                    1 on u; 1/rho on v, the isotropic stiffness on the normal stresses, mu on
 *                    the shear stresses), solved per element by conjugate gradients in FP64 */
int bbw_energy(bbw_ctx *ctx, double *E);
int bbw_energy_wadg(bbw_ctx *ctx, double *E);
/* bbw_energy without allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The reduction's partial sums and result go to a file-scope device
 * array of BBW_ENERGY_BLOCKS + 1 = 4097 doubles (32 KiB), one per GPU, held until the process
 * ends; the grid is capped at that many blocks, so any K fits.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (on different streams or threads) race. */
int bbw_energy_noalloc(bbw_ctx *ctx, double *E);

/* Discrete adjoint (no reference counterpart) of the fused forward stage, both physics.
 * Misfit J = 1/2 sum_n sum_r (u_r(Q_{n+1}) - d_obs[n][r])^2 over nsteps steps from the current
 * state, u_r the receiver sample (bbw_set_receivers) of the field BBW_OPT_REC_FIELD selects
 * (default 0: p, or v1 for the elastic fields).  Gradient by the reverse LSRK45 stages with the
 * state checkpointed every ckpt_every steps and each segment recomputed: acoustic (velocity
 * not projected, Pq Vq = I) dJ/dc2q [K][Nq]; elastic [3][K][Nq], dJ/d(1/rho), dJ/dlambda, dJ/dmu
 * at the cubature points (the media of bbw_upload_elastic_media).  Both calls leave the
 * context's state as it was. */
int bbw_upload_adjoint_maps(bbw_ctx *ctx, const int32_t *mapP); /* [K][4Nfp] face-node map */
int bbw_misfit(bbw_ctx *ctx, int nsteps, double dt, double t0, const float *d_obs, double *J);
int bbw_adjoint_gradient(bbw_ctx *ctx, int nsteps, double dt, double t0, const float *d_obs,
                         int ckpt_every, float *grad, double *J);
/* The same with the observations [nsteps][nrec] and the gradient [K][Nq] in device memory, all
 * work on the context's stream (for callers that keep both on the GPU, e.g. torch tensors); J
 * is reduced in FP64 on the device in all four and returned on the host. */
int bbw_misfit_device(bbw_ctx *ctx, int nsteps, double dt, double t0, const float *d_obs_dev, double *J);
int bbw_adjoint_gradient_device(bbw_ctx *ctx, int nsteps, double dt, double t0, const float *d_obs_dev,
                                int ckpt_every, float *grad_dev, double *J);

/* Batched multi-shot acoustic runs (no reference counterpart): many shots on several meshes,
 * uneven durations, one persistent launch per LSRK stage for all active shots, HIP-graph
 * replay.  Each shot: its mesh, c2q [K][Nq], initial state [K][4][Np], a source as in
 * bbw_set_source, time step dt and step count.  bbw_batch_run advances every shot to its own
 * step count; shots_per_tile shots of one mesh share a tile's geometry; blocks_per_cu sets the
 * persistent grid (0: one workgroup per work item).  bbw_batch_reset restores every shot's initial state. */
typedef struct bbw_batch bbw_batch;
int bbw_batch_create(int N, int device, bbw_batch **out);
void bbw_batch_destroy(bbw_batch *b);
const char *bbw_batch_last_error(const bbw_batch *b);
int bbw_batch_upload_operators(bbw_batch *b, const bbw_ops *ops);
int bbw_batch_add_mesh(bbw_batch *b, int K, const float *vgeo, const float *fgeo,
                       const int32_t *vmapP, int *mesh_id);
int bbw_batch_add_shot(bbw_batch *b, int mesh_id, const float *c2q, const float *Q0, int nsrc,
                       const int32_t *src_elems, const float *src_coef, const float *weights,
                       double f0, double amp, double dt, int nsteps, int *shot_id);
int bbw_batch_run(bbw_batch *b, int shots_per_tile, int blocks_per_cu, int use_graph);
int bbw_batch_get_state(bbw_batch *b, int shot_id, float *Q);
int bbw_batch_reset(bbw_batch *b);
int bbw_batch_set_step_cap(bbw_batch *b, int cap); /* > 0: run at most cap steps per shot (tuning) */
double bbw_batch_last_ms(const bbw_batch *b); /* device time of the last bbw_batch_run, ms */

/* Tuning options (no reference counterpart beyond the Kblk* compile-time defines of
 * WaveInitOCCA3d, cpp/src/WaveOKL3d.cpp:1631-1635): fused kernel implementation and
 * elements per block; used by the autotuner. */
#define BBW_OPT_FUSED_IMPL 0 /* BBW_FUSED_VECTOR or BBW_FUSED_MFMA (default) */
#define BBW_OPT_KBLK_V 1
#define BBW_OPT_KBLK_S 2
#define BBW_OPT_KBLK_U 3
#define BBW_OPT_KBLK_F 4 /* elements per block of the vector fused kernel */
/* adjoint: MiB of device memory that keep Vq (A Q)_p of the final steps from the first forward
 * pass, so those steps are not recomputed (0 = automatic: up to 60% of free memory, < 0 = none;
 * earlier steps are checkpointed and recomputed) */
#define BBW_OPT_ADJ_CACHE_MB 5
#define BBW_OPT_ADJ_IMPL 6 /* acoustic adjoint kernels: 0 per-element, 1 tiled MFMA (default) */
#define BBW_OPT_PLAIN_IMPL 7 /* plain kernels (bbw_volume, bbw_surface, bbw_update_wadg): */
#define BBW_PLAIN_FLAT 0     /* per-element loops over Kblk* elements per block */
#define BBW_PLAIN_TILED 1    /* default for N <= 6: 16-element tiles, WADG projections on MFMA */
#define BBW_PLAIN_SLICE 2    /* flat kernels, surface lifted slice by slice (default for N >= 7) */
#define BBW_FUSED_VECTOR 0
#define BBW_FUSED_MFMA 1
/* source time: 0 (default) evaluates the Ricker pulse at each stage's time t + c_i dt; 1 at the
 * step's start time for all five stages, as the reference's elastic driver does
 * (elas_cpp/src/WaveOKL3d.cpp:1675-1680; first order in time for the source term) */
#define BBW_OPT_SRC_STEP_TIME 8
#define BBW_OPT_PLANAR_SPLIT 9 /* curvilinear context: 1 runs the planar elements through the split
                                 form with constant factors; 0 (default) the strong nodal form of
                                 rk_volume_planar / rk_surface_planar */
#define BBW_OPT_REC_FIELD 10  /* the field the receivers sample (0 .. Nfields-1; default 0) */
/* fused elastic stage: BBW_ELAS_LIFT_CUBATURE (default for N <= 4) forms Vq (E F + sum_f
 * LIFT_f flux_f) in one product at the cubature points (the volume elevation and the lift folded
 * into the Vq operator); BBW_ELAS_LIFT_NODES (default for N >= 5) lifts and elevates at the
 * nodes, then applies Vq.  The adjoint driver's forward passes always run the nodal form. */
#define BBW_OPT_ELAS_FUSED 11
#define BBW_ELAS_LIFT_NODES 0
#define BBW_ELAS_LIFT_CUBATURE 1
/* fused and adjoint stages (acoustic and elastic): 16-element tiles per XCD run.  Workgroups
 * go to the device's 8 XCDs round-robin; with a run length C (a power of 2, at most 1024;
 * default 16) each XCD takes runs of C consecutive tiles, so that neighbouring elements' face
 * traces are read through one L2.  0 keeps the dispatch order (tile = workgroup).  Results do
 * not depend on it. */
#define BBW_OPT_XCD_CHUNK 12
int bbw_set_option(bbw_ctx *ctx, int option, int value);

/* Device addresses of the context's buffers (for copies into caller-owned device memory):
 * which = 0 current Q, 1 resQ, 2 rhsQ, 3 c2q. */
int bbw_device_ptr(bbw_ctx *ctx, int which, void **ptr);

#ifdef __cplusplus
}
#endif

#endif /* BBWADG_H */
