// Batched multi-shot acoustic solver: many shots of uneven mesh size and duration advance in
// lockstep, one persistent launch per LSRK stage (no reference counterpart; the per-shot stage is
// the reference's RK_step_WADG_subelem, cpp/src/WaveOKL3d.cpp:2084-2109, as fused on one tile).
//
// Work item = (mesh, 16-element tile, group of up to S shots on that mesh).  A workgroup loads
// the tile's geometry and neighbour indices once and runs the fused stage for each shot of the
// group (shots-per-tile S); workgroups loop over the item table with a grid stride (persistent).
// The active set of shots changes only when a shot reaches its step count, so the run is cut
// into phases with one item table each; within a phase, steps are replayed from a captured HIP
// graph.  Every shot keeps its own state ping-pong pair, residual and c2; the source time
// function is tabulated per shot and step and read through a device step counter.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <type_traits>

#include "bbw_fused_tile.h"

namespace bbw {

struct BatchItem {
  int mesh, k0, shot0, nshot;  // shots: list[shot0 .. shot0 + nshot)
};

struct BatchShotDev {
  float *Q[2];
  float *res;
  const float *c2q;
  const float *ftab;  // [nsteps][5] source time function values (x amplitude in weights)
  SrcArgs src;
  float dt;
  int mesh;
};

// Descriptor values are the same for the whole workgroup: move them to scalar registers.
__device__ __forceinline__ int su(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float sf(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
// Pointers read from the descriptor tables: uniform (scalar registers) and global memory, so
// accesses through them compile to global loads with a scalar base, as for kernel arguments.
template <typename T>
__device__ __forceinline__ T *sp(T *p) {
  const unsigned long long u = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  typedef __attribute__((address_space(1))) T GT;
  return (T *)(GT *)(((unsigned long long)hi << 32) | lo);
}

#ifndef BBW_FUSED_WG_PER_CU
#define BBW_FUSED_WG_PER_CU 3
#endif

// Work items blockIdx.x, + gridDim.x, ... (PERSIST; otherwise one item per workgroup, which
// keeps the register count of the single-shot kernel, in runs of xcd_chunk items per XCD),
// each with its shots in order on the same tile (the shots after the first read the tile's
// geometry and vmapP entries from cache).
template <int N, bool PERSIST>
__global__ __launch_bounds__(256, PERSIST ? 3 : 4) void k_fused_batch(
    int nitems, const BatchItem *__restrict__ items, const FusedMesh *__restrict__ meshes,
    const BatchShotDev *__restrict__ shots, const int *__restrict__ list, FusedOps o, int stage,
    int par, float fa, float fb, int store_res, const int *__restrict__ step, int xcd_chunk) {
  using L = FusedLayout<N, false>;
  extern __shared__ float lds[];
  const int stp = *step;
  auto item = [&](int w) {
    // thread index re-derived per item and shot: address arithmetic is redone per tile instead
    // of being hoisted out of the loops and kept live across them (it would not fit the
    // register budget)
    int tid;
    asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
    const int mesh = su(items[w].mesh), k0 = su(items[w].k0);
    const int shot0 = su(items[w].shot0), nshot = su(items[w].nshot);
    FusedMesh m;
    m.K = su(meshes[mesh].K);
    m.vgeo = sp(meshes[mesh].vgeo);
    m.fgeo = sp(meshes[mesh].fgeo);
    m.vmapP = sp(meshes[mesh].vmapP);
    const int nk = min(L::TB, m.K - k0);
#pragma unroll 1
    for (int j = 0; j < nshot; ++j) {
      int tj;
      asm volatile("v_mov_b32 %0, %1" : "=v"(tj) : "v"(tid));
      // vmapP entries per shot (the shots of an item read them from cache; keeping them live
      // across the shot loop would not fit the register budget)
      int vm[L::EPT3];
      vm_load<N, false>(vm, m, k0, nk, tj);
      const int fid3 = face_node<N, false>(o, tj);
      const BatchShotDev *sd = shots + su(list[shot0 + j]);
      FusedShot sh;
      sh.c2q = sp(sd->c2q);
      sh.Qin = sp(sd->Q[par]);
      sh.Qout = sp(sd->Q[1 - par]);
      sh.res = sp(sd->res);
      sh.yout = nullptr;
      SrcArgs src;
      src.nsrc = su(sd->src.nsrc);
      src.first = su(sd->src.first);
      src.last = su(sd->src.last);
      src.elems = sp(sd->src.elems);
      src.coef = sp(sd->src.coef);
      src.tiles = sp(sd->src.tiles);
      src.tfirst = su(sd->src.tfirst);
      src.tlast = su(sd->src.tlast);
#pragma unroll
      for (int f = 0; f < 9; ++f) src.w[f] = sf(sd->src.w[f]);
      sh.ft = src.nsrc ? sf(sp(sd->ftab)[stp * 5 + stage]) : 0.f;
      sh.fdt = sf(sd->dt);
      sh.store_res = store_res;
      TileStage<N, false> st;
      stage_load<N, false>(st, m, sh.Qin, k0, nk, tj);
      TilePre<N, false> pf;
      tile_prefetch<N, false>(pf, m, o, sh, fa, k0, nk, vm, tj);
      stage_store<N, false>(st, lds + L::OQ, lds + L::OG, tj);
      __syncthreads();
      tile_shot<N, false>(m, o, sh, src, fa, fb, k0, nk, lds, lds + L::OQ, lds + L::OG, vm, pf, fid3, tj,
                          NoHook{}, NoHook{}, NoHook{});
    }
  };
  if constexpr (PERSIST) {
#pragma unroll 1
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) item(w);
  } else {
    // consecutive items are consecutive tiles of one shot group: runs of them per XCD, as in
    // the single-shot stage
    item(tile_index(xcd_chunk));
  }
}

__global__ void k_step_inc(int *step) { *step += 1; }

}  // namespace bbw

using namespace bbw;

struct bbw_batch {
  bbw_ctx *ops = nullptr;  // operators, stream and device (a context of one tile)
  std::string err;
  std::vector<FusedMesh> meshes;
  std::vector<float *> owned;
  struct Shot {
    int mesh = 0, nsteps = 0;
    double dt = 0, f0 = 0, amp = 0;
    float *Q[2] = {nullptr, nullptr};
    float *res = nullptr, *c2q = nullptr, *ftab = nullptr, *coef = nullptr, *Q0 = nullptr;
    int *elems = nullptr;
    int nsrc = 0, first = 0, last = -1;
    int2 *tiles = nullptr;   // src_tile_table of elems
    int tfirst = 0, tlast = -1;
    float w[9] = {0};
  };
  std::vector<Shot> shots;
  double last_ms = 0;
  int step_cap = 0;  // > 0: every shot runs min(nsteps, step_cap) steps (tuning trials)
  std::vector<int> last_steps;  // steps each shot ran in the last run
};

namespace {

int bfail(bbw_batch *b, int code, const std::string &m) {
  b->err = m;
  return code;
}

template <typename T>
int balloc(bbw_batch *b, T **p, size_t n) {
  if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return bfail(b, BBW_ERR_NOMEM, "hipMalloc failed");
  }
  b->owned.push_back((float *)*p);
  return BBW_OK;
}

template <typename T>
int bup(bbw_batch *b, T **p, const T *src, size_t n) {
  int rc = balloc(b, p, n);
  if (rc) return rc;
  if (hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, b->ops->stream) != hipSuccess ||
      hipStreamSynchronize(b->ops->stream) != hipSuccess)
    return bfail(b, BBW_ERR_HIP, "upload failed");
  return BBW_OK;
}

double ricker_b(double t, double f0) {
  const double a = M_PI * f0 * (t - 1.0 / f0);
  return (1.0 - 2.0 * a * a) * std::exp(-a * a);
}

template <int N>
hipError_t launch_batch_n(bbw_batch *b, int nitems, const BatchItem *items, const FusedMesh *meshes,
                          const BatchShotDev *shots, const int *list, int stage, int par, float fa,
                          float fb, int store_res, const int *step, int grid) {
  using L = FusedLayout<N, false>;
  const size_t sh = sizeof(float) * L::floats;
  if (grid < nitems) {
    if (sh > 65536)
      hipFuncSetAttribute((const void *)k_fused_batch<N, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL((k_fused_batch<N, true>), dim3(grid), dim3(256), sh, b->ops->stream, nitems, items,
                       meshes, shots, list, fused_ops(b->ops), stage, par, fa, fb, store_res, step,
                       b->ops->xcd_chunk);
  } else {
    if (sh > 65536)
      hipFuncSetAttribute((const void *)k_fused_batch<N, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL((k_fused_batch<N, false>), dim3(nitems), dim3(256), sh, b->ops->stream, nitems, items,
                       meshes, shots, list, fused_ops(b->ops), stage, par, fa, fb, store_res, step,
                       b->ops->xcd_chunk);
  }
  return hipGetLastError();
}

hipError_t launch_batch(bbw_batch *b, int nitems, const BatchItem *items, const FusedMesh *meshes,
                        const BatchShotDev *shots, const int *list, int stage, int par, float fa,
                        float fb, int store_res, const int *step, int grid) {
  switch (b->ops->N) {
    case 1: return launch_batch_n<1>(b, nitems, items, meshes, shots, list, stage, par, fa, fb, store_res, step, grid);
    case 2: return launch_batch_n<2>(b, nitems, items, meshes, shots, list, stage, par, fa, fb, store_res, step, grid);
    case 3: return launch_batch_n<3>(b, nitems, items, meshes, shots, list, stage, par, fa, fb, store_res, step, grid);
    case 4: return launch_batch_n<4>(b, nitems, items, meshes, shots, list, stage, par, fa, fb, store_res, step, grid);
    case 5: return launch_batch_n<5>(b, nitems, items, meshes, shots, list, stage, par, fa, fb, store_res, step, grid);
    case 6: return launch_batch_n<6>(b, nitems, items, meshes, shots, list, stage, par, fa, fb, store_res, step, grid);
  }
  return hipErrorInvalidValue;
}

}  // namespace

extern "C" {

int bbw_batch_create(int N, int device, bbw_batch **out) {
  if (!out) return BBW_ERR_ARG;
  *out = nullptr;
  if (N < 1 || N > BBW_MAX_ORDER_TILED) return BBW_ERR_UNSUPPORTED;
  bbw_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.N = N;
  cfg.Nfields = 4;
  cfg.K = 16;
  cfg.Nq = nq_of(N);
  cfg.device = device;
  bbw_ctx *c = nullptr;
  int rc = bbw_create(&cfg, &c);
  if (rc) return rc;
  bbw_batch *b = new bbw_batch();
  b->ops = c;
  *out = b;
  return BBW_OK;
}

void bbw_batch_destroy(bbw_batch *b) {
  if (!b) return;
  hipStreamSynchronize(b->ops->stream);
  for (float *p : b->owned) hipFree(p);
  bbw_destroy(b->ops);
  delete b;
}

const char *bbw_batch_last_error(const bbw_batch *b) {
  if (!b) return "null batch";
  return b->err.empty() ? bbw_last_error(b->ops) : b->err.c_str();
}

int bbw_batch_upload_operators(bbw_batch *b, const bbw_ops *o) {
  if (!b) return BBW_ERR_ARG;
  int rc = bbw_upload_operators(b->ops, o);
  if (rc) b->err = bbw_last_error(b->ops);
  return rc;
}

int bbw_batch_add_mesh(bbw_batch *b, int K, const float *vgeo, const float *fgeo,
                       const int32_t *vmapP, int *mesh_id) {
  if (!b || K <= 0 || !vgeo || !fgeo || !vmapP || !mesh_id) return BBW_ERR_ARG;
  const int Np = b->ops->Np, NfpNf = 4 * b->ops->Nfp;
  const long long lim = (long long)K * 4 * Np;
  // the stage kernel addresses a shot's state (K * 4 * Np floats) through 32-bit byte offsets
  if ((unsigned long long)K * 16 * Np >= (1ull << 32))
    return bfail(b, BBW_ERR_ARG, "mesh too large: K * 4 * Np * 4 bytes must stay below 2^32");
  for (size_t i = 0; i < (size_t)K * NfpNf; ++i)
    if (vmapP[i] < 0 || vmapP[i] >= lim || (vmapP[i] % (4LL * Np)) >= Np)
      return bfail(b, BBW_ERR_ARG, "vmapP entry out of range");
  FusedMesh m;
  m.K = K;
  m.xcd_chunk = 0;
  float *dv = nullptr, *df = nullptr;
  int *dm = nullptr;
  int rc;
  if ((rc = bup(b, &dv, vgeo, (size_t)K * 9)) || (rc = bup(b, &df, fgeo, (size_t)K * 16)) ||
      (rc = bup(b, &dm, (const int *)vmapP, (size_t)K * NfpNf)))
    return rc;
  m.vgeo = dv;
  m.fgeo = df;
  m.vmapP = dm;
  *mesh_id = (int)b->meshes.size();
  b->meshes.push_back(m);
  return BBW_OK;
}

int bbw_batch_add_shot(bbw_batch *b, int mesh_id, const float *c2q, const float *Q0, int nsrc,
                       const int32_t *src_elems, const float *src_coef, const float *weights,
                       double f0, double amp, double dt, int nsteps, int *shot_id) {
  if (!b || !c2q || !Q0 || !shot_id || mesh_id < 0 || mesh_id >= (int)b->meshes.size() ||
      nsteps < 0 || !(dt > 0) || nsrc < 0)
    return BBW_ERR_ARG;
  const int K = b->meshes[mesh_id].K, Np = b->ops->Np, Nq = b->ops->Nq;
  bbw_batch::Shot s;
  s.mesh = mesh_id;
  s.nsteps = nsteps;
  s.dt = dt;
  s.f0 = f0;
  s.amp = amp;
  const size_t n = (size_t)K * 4 * Np;
  int rc;
  if ((rc = bup(b, &s.Q[0], Q0, n)) || (rc = bup(b, &s.Q0, Q0, n)) || (rc = balloc(b, &s.Q[1], n)) ||
      (rc = balloc(b, &s.res, n)) ||
      (rc = bup(b, &s.c2q, c2q, (size_t)K * Nq)))
    return rc;
  if (hipMemsetAsync(s.res, 0, n * sizeof(float), b->ops->stream) != hipSuccess ||
      hipMemsetAsync(s.Q[1], 0, n * sizeof(float), b->ops->stream) != hipSuccess)
    return bfail(b, BBW_ERR_HIP, "memset failed");
  if (nsrc > 0) {
    if (!src_elems || !src_coef || !weights || !(f0 > 0)) return bfail(b, BBW_ERR_ARG, "bad source");
    for (int i = 0; i < nsrc; ++i)
      if (src_elems[i] < 0 || src_elems[i] >= K || (i && src_elems[i] <= src_elems[i - 1]))
        return bfail(b, BBW_ERR_ARG, "source elements must ascend within the mesh");
    if ((rc = bup(b, &s.elems, (const int *)src_elems, (size_t)nsrc)) ||
        (rc = bup(b, &s.coef, src_coef, (size_t)nsrc * Np)))
      return rc;
    const std::vector<int2> tiles = src_tile_table(src_elems, nsrc);
    if ((rc = bup(b, &s.tiles, tiles.data(), tiles.size()))) return rc;
    s.nsrc = nsrc;
    s.first = src_elems[0];
    s.last = src_elems[nsrc - 1];
    s.tfirst = src_elems[0] >> 4;
    s.tlast = src_elems[nsrc - 1] >> 4;
    for (int f = 0; f < 4; ++f) s.w[f] = (float)(weights[f] * amp);
    std::vector<float> ft((size_t)std::max(nsteps, 1) * 5);
    for (int st = 0; st < nsteps; ++st)
      for (int i = 0; i < 5; ++i) ft[(size_t)st * 5 + i] = (float)ricker_b(st * dt + b->ops->rk4c[i] * dt, f0);
    if ((rc = bup(b, &s.ftab, ft.data(), ft.size()))) return rc;
  }
  *shot_id = (int)b->shots.size();
  b->shots.push_back(s);
  return BBW_OK;
}

int bbw_batch_run(bbw_batch *b, int shots_per_tile, int blocks_per_cu, int use_graph) {
  if (!b || shots_per_tile <= 0 || blocks_per_cu < 0) return BBW_ERR_ARG;
  bbw_ctx *c = b->ops;
  if (!c->have_ops) return bfail(b, BBW_ERR_STATE, "operators not uploaded");
  const int S = shots_per_tile, TB = 16;
  const int ns = (int)b->shots.size();
  if (ns == 0) return BBW_OK;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->dev) != hipSuccess) return bfail(b, BBW_ERR_HIP, "device query");
  // blocks_per_cu = 0: one workgroup per item (the hardware dispatcher balances the tiles)
  const int max_grid = blocks_per_cu > 0 ? prop.multiProcessorCount * blocks_per_cu : (1 << 30);

  std::vector<int> nst(ns);
  for (int i = 0; i < ns; ++i)
    nst[i] = b->step_cap > 0 ? std::min(b->shots[i].nsteps, b->step_cap) : b->shots[i].nsteps;
  b->last_steps = nst;
  // shots by duration (longest first): the active set at any step is a prefix
  std::vector<int> order(ns);
  for (int i = 0; i < ns; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int d) { return nst[a] > nst[d]; });
  std::vector<int> bounds;  // distinct durations ascending, with 0
  bounds.push_back(0);
  for (int v : nst) bounds.push_back(v);
  std::sort(bounds.begin(), bounds.end());
  bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());

  // device tables: meshes, shots, per-phase items and shot lists
  std::vector<BatchShotDev> sdev(ns);
  for (int i = 0; i < ns; ++i) {
    auto &s = b->shots[i];
    BatchShotDev d;
    d.Q[0] = s.Q[0];
    d.Q[1] = s.Q[1];
    d.res = s.res;
    d.c2q = s.c2q;
    d.ftab = s.ftab;
    d.src.nsrc = s.nsrc;
    d.src.first = s.first;
    d.src.last = s.last;
    d.src.elems = s.elems;
    d.src.coef = s.coef;
    d.src.tiles = s.tiles;
    d.src.tfirst = s.tfirst;
    d.src.tlast = s.tlast;
    for (int f = 0; f < 9; ++f) d.src.w[f] = s.w[f];
    d.dt = (float)s.dt;
    d.mesh = s.mesh;
    sdev[i] = d;
  }
  struct Phase {
    int s0, s1, item0, nitems;
  };
  std::vector<Phase> phases;
  std::vector<BatchItem> items;
  std::vector<int> list;
  for (size_t p = 0; p + 1 < bounds.size(); ++p) {
    const int s0 = bounds[p], s1 = bounds[p + 1];
    Phase ph{s0, s1, (int)items.size(), 0};
    for (int m = 0; m < (int)b->meshes.size(); ++m) {
      std::vector<int> act;
      for (int i : order)
        if (nst[i] > s0 && b->shots[i].mesh == m) act.push_back(i);
      for (size_t g = 0; g < act.size(); g += S) {
        const int l0 = (int)list.size();
        const int cnt = (int)std::min<size_t>(S, act.size() - g);
        for (int j = 0; j < cnt; ++j) list.push_back(act[g + j]);
        for (int k0 = 0; k0 < b->meshes[m].K; k0 += TB) items.push_back({m, k0, l0, cnt});
      }
    }
    ph.nitems = (int)items.size() - ph.item0;
    phases.push_back(ph);
  }
  FusedMesh *dmesh = nullptr;
  BatchShotDev *dshot = nullptr;
  BatchItem *ditems = nullptr;
  int *dlist = nullptr, *dstep = nullptr;
  int rc;
  std::vector<float *> tmp;
  auto tmp_up = [&](auto **p, const auto *src, size_t n) -> int {
    using T = std::remove_reference_t<decltype(**p)>;
    if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
      return bfail(b, BBW_ERR_NOMEM, "hipMalloc failed");
    tmp.push_back((float *)*p);
    if (n && hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream) != hipSuccess)
      return bfail(b, BBW_ERR_HIP, "table upload failed");
    return BBW_OK;
  };
  const int zero = 0;
  rc = tmp_up(&dmesh, b->meshes.data(), b->meshes.size());
  if (!rc) rc = tmp_up(&dshot, sdev.data(), sdev.size());
  if (!rc) rc = tmp_up(&ditems, items.data(), items.size());
  if (!rc) rc = tmp_up(&dlist, list.data(), list.size());
  if (!rc) rc = tmp_up(&dstep, &zero, 1);
  auto cleanup = [&]() {
    hipStreamSynchronize(c->stream);
    for (float *p : tmp) hipFree(p);
  };
  if (rc) {
    cleanup();
    return rc;
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, c->stream);
  int par = 0;
  const int G = 10;
  for (const Phase &ph : phases) {
    const int grid = std::min(ph.nitems, max_grid);
    auto one_step = [&](int parity) -> hipError_t {
      int p = parity;
      for (int i = 0; i < 5; ++i) {
        // a shot's residual is not part of its readable state: after the last stage it is dead
        // whenever a_0 = 0 (the next step's first stage scales it by 0)
        const int store_res = !(i == 4 && c->rk4a[0] == 0.0);
        hipError_t e = launch_batch(b, ph.nitems, ditems + ph.item0, dmesh, dshot, dlist, i, p,
                                    (float)c->rk4a[i], (float)c->rk4b[i], store_res, dstep, grid);
        if (e != hipSuccess) return e;
        p ^= 1;
      }
      hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(1), 0, c->stream, dstep);
      return hipGetLastError();
    };
    int st = ph.s0;
    if (use_graph && ph.s1 - ph.s0 >= G) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
      int p = par;
      for (int j = 0; j < G && e == hipSuccess; ++j) {
        e = one_step(p);
        p ^= 1;  // five stages flip the parity once per step
      }
      hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
      if (e == hipSuccess) e = e2;
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph) hipGraphDestroy(graph);
      if (e != hipSuccess) {
        cleanup();
        return bfail(b, BBW_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
      }
      for (; st + G <= ph.s1; st += G) {
        if ((e = hipGraphLaunch(exec, c->stream)) != hipSuccess) break;
      }
      hipGraphExecDestroy(exec);
      if (e != hipSuccess) {
        cleanup();
        return bfail(b, BBW_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(e));
      }
      // G is even, so the parity at st equals the parity at the phase start
    }
    for (; st < ph.s1; ++st) {
      hipError_t e = one_step(par);
      if (e != hipSuccess) {
        cleanup();
        return bfail(b, BBW_ERR_HIP, std::string("batch stage: ") + hipGetErrorString(e));
      }
      par ^= 1;
    }
  }
  hipEventRecord(e1, c->stream);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  b->last_ms = ms;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  cleanup();
  if (e != hipSuccess) return bfail(b, BBW_ERR_HIP, std::string("batch run: ") + hipGetErrorString(e));
  return BBW_OK;
}

int bbw_batch_get_state(bbw_batch *b, int shot_id, float *Q) {
  if (!b || !Q || shot_id < 0 || shot_id >= (int)b->shots.size()) return BBW_ERR_ARG;
  auto &s = b->shots[shot_id];
  const size_t n = (size_t)b->meshes[s.mesh].K * 4 * b->ops->Np;
  // five stage flips per step: after n steps the state is in buffer n % 2
  const int ran = (size_t)shot_id < b->last_steps.size() ? b->last_steps[shot_id] : 0;
  if (hipMemcpyAsync(Q, s.Q[ran & 1], n * sizeof(float), hipMemcpyDeviceToHost, b->ops->stream) != hipSuccess ||
      hipStreamSynchronize(b->ops->stream) != hipSuccess)
    return bfail(b, BBW_ERR_HIP, "state copy failed");
  return BBW_OK;
}

double bbw_batch_last_ms(const bbw_batch *b) { return b ? b->last_ms : -1.0; }

int bbw_batch_set_step_cap(bbw_batch *b, int cap) {
  if (!b || cap < 0) return BBW_ERR_ARG;
  b->step_cap = cap;
  return BBW_OK;
}

int bbw_batch_reset(bbw_batch *b) {
  if (!b) return BBW_ERR_ARG;
  b->last_steps.assign(b->shots.size(), 0);
  for (auto &s : b->shots) {
    const size_t n = (size_t)b->meshes[s.mesh].K * 4 * b->ops->Np;
    if (hipMemcpyAsync(s.Q[0], s.Q0, n * sizeof(float), hipMemcpyDeviceToDevice, b->ops->stream) != hipSuccess ||
        hipMemsetAsync(s.res, 0, n * sizeof(float), b->ops->stream) != hipSuccess)
      return bfail(b, BBW_ERR_HIP, "reset failed");
  }
  if (hipStreamSynchronize(b->ops->stream) != hipSuccess) return bfail(b, BBW_ERR_HIP, "reset failed");
  return BBW_OK;
}

}  // extern "C"
