// FP64 energy reductions of the state on the device (no reference counterpart in C++; the
// weighted-mass form is that of matlab/WADG/check_conservation_loss.m:51-54).
//
//   quadrature  E = 1/2 sum_K J sum_q w_q e(u_q)
//               acoustic e = p^2 / c2 + |u|^2; elastic e = rho |v|^2 + sigma : C^-1 sigma
//               (isotropic compliance, shear stresses with engineering strains)
//   WADG        the energy the WADG scheme dissipates.  Its update applies T = Pq diag(m) Vq to
//               a block of fields (m = c2 on p; 1/rho on v; the isotropic stiffness on the normal
//               stresses; mu on the shear stresses; 1 on the acoustic velocity), so its mass
//               matrix is M A_m^-1 M with A_m = Vq^T W m Vq, and
//                 E = 1/2 sum_K J sum_blocks (MB x)^T A_m^-1 (MB x).
//               A_m^-1 (MB x) is found per element by conjugate gradients preconditioned with
//               MB^-1, whose spectrum lies in [min m, max m]: with r = Vq^T W g the residual,
//               MB^-1 r = Pq g, so every product is a Vq or Pq product at the cubature points.
//
// Both reductions write one partial sum per block of a fixed grid (at most BBW_ENERGY_BLOCKS),
// which one block then adds in a fixed order: the result does not depend on the launch, and the
// partial-sum buffer has a fixed size whatever K is.
#include <mutex>

#include "bbw_internal.h"

namespace bbw {

// Sum over the 64 lanes of a wave, equal in every lane: the butterfly adds in a different order
// in each lane, so lane 0's sum is broadcast (the conjugate-gradient loop's trip count and its
// step sizes must be the same in all lanes).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return __shfl(v, 0, 64);
}

__device__ __forceinline__ double elastic_density(const double (&u)[9], double rhoinv, double lam, double mu) {
  const double kin = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / rhoinv;
  const double tr = u[3] + u[4] + u[5];
  const double nrm = (u[3] * u[3] + u[4] * u[4] + u[5] * u[5] - lam / (2.0 * mu + 3.0 * lam) * tr * tr) / (2.0 * mu);
  const double shr = (u[6] * u[6] + u[7] * u[7] + u[8] * u[8]) / mu;
  return kin + nrm + shr;
}

// One partial sum per block: threads over (element, cubature point) pairs, grid-stride.
template <int N, int NF>
__global__ __launch_bounds__(256) void k_energy_quad(int K, const double *__restrict__ Vq64,
                                                     const double *__restrict__ wq,
                                                     const double *__restrict__ J,
                                                     const float *__restrict__ m0,   // c2 | 1/rho
                                                     const float *__restrict__ m1,   // - | lambda
                                                     const float *__restrict__ m2,   // - | mu
                                                     const float *__restrict__ Q,
                                                     double *__restrict__ partial) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq;
  __shared__ double red[256];
  double acc = 0.0;
  const long long total = (long long)K * Nq;
  for (long long w = (long long)blockIdx.x * 256 + threadIdx.x; w < total; w += (long long)gridDim.x * 256) {
    const int k = (int)(w / Nq), q = (int)(w - (long long)k * Nq);
    const float *Qk = Q + (size_t)k * NF * Np;
    double u[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) u[f] = 0.0;
    for (int j = 0; j < Np; ++j) {
      const double v = Vq64[q + j * Nq];
#pragma unroll
      for (int f = 0; f < NF; ++f) u[f] += v * (double)Qk[f * Np + j];
    }
    const size_t mq = (size_t)k * Nq + q;
    double e;
    if constexpr (NF == 4) {
      e = u[0] * u[0] / (double)m0[mq] + u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    } else {
      e = elastic_density(u, (double)m0[mq], (double)m1[mq], (double)m2[mq]);
    }
    acc += J[k] * wq[q] * e;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = 0.5 * red[0];
}

// media of one block of NC fields at a cubature point: scalar m0 (UNIT: m = 1), or the
// isotropic normal-stress stiffness C = lambda 1 1^T + 2 mu I (ISO, m0 = lambda, m1 = mu)
enum { MED_SCALAR = 0, MED_UNIT = 1, MED_ISO = 2 };

template <int NC, int KIND>
__device__ __forceinline__ void apply_media(double m0, double m1, const double (&v)[NC], double (&out)[NC]) {
  if constexpr (KIND == MED_ISO) {
    double tr = 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) tr += v[c];
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = 2.0 * m1 * v[c] + m0 * tr;
  } else if constexpr (KIND == MED_UNIT) {
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = v[c];
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = m0 * v[c];
  }
}

// (MB x)^T A_m^-1 (MB x) for the NC fields f0 .. f0 + NC - 1 of element row Qk, by preconditioned
// conjugate gradients on one wave (64 threads, the whole block): lanes over cubature points for
// Vq products and over nodes for Pq products, vectors exchanged through LDS (sg [NC][Nq],
// sz [NC][Np]).  Returns the value in every lane.
template <int N, int NC, int KIND>
__device__ double wadg_block(const float *Qk, int f0, const float *ma, const float *mb,
                             const double *__restrict__ Vq, const double *__restrict__ Pq,
                             const double *__restrict__ wq, double *sg, double *sz, int lane) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq;
  constexpr int QPL = (Nq + 63) / 64, NPL = (Np + 63) / 64;
  constexpr int MAXIT = 200;
  constexpr double TOL = 1e-26;   // on r^T MB^-1 r relative to its start: residual / 1e13
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int n = lane + 64 * i;
    if (n < Np) {
#pragma unroll
      for (int c = 0; c < NC; ++c) sz[c * Np + n] = (double)Qk[(f0 + c) * Np + n];
    }
  }
  __syncthreads();
  double w[QPL], a0[QPL], a1[QPL];
  double g0[QPL][NC], g[QPL][NC], Vd[QPL][NC], Vy[QPL][NC];
#pragma unroll
  for (int i = 0; i < QPL; ++i) {
    const int q = lane + 64 * i;
    const bool on = q < Nq;
    w[i] = on ? wq[q] : 0.0;
    a0[i] = (on && KIND != MED_UNIT) ? (double)ma[q] : 0.0;
    a1[i] = (on && KIND == MED_ISO) ? (double)mb[q] : 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double acc = 0.0;
      if (on)
        for (int j = 0; j < Np; ++j) acc += Vq[q + j * Nq] * sz[c * Np + j];
      g0[i][c] = acc;
      g[i][c] = acc;
      Vd[i][c] = acc;
      Vy[i][c] = 0.0;
    }
  }
  // the iterate y and the search direction d enter only through Vq y and Vq d (kept at the
  // lane's cubature points): y_0 = 0, d_0 = z_0 = Pq Vq x = x, so Vq d_0 = g_0
  double rz = 0.0;
#pragma unroll
  for (int i = 0; i < QPL; ++i)
#pragma unroll
    for (int c = 0; c < NC; ++c) rz += w[i] * g0[i][c] * g0[i][c];
  rz = wave_sum(rz);
  const double rz0 = rz;
  for (int it = 0; it < MAXIT && rz > TOL * rz0; ++it) {
    // alpha = r^T z / d^T A d, d^T A d = sum_q w_q (Vq d)_q . m_q (Vq d)_q
    double Cd[QPL][NC];
    double dAd = 0.0;
#pragma unroll
    for (int i = 0; i < QPL; ++i) {
      apply_media<NC, KIND>(a0[i], a1[i], Vd[i], Cd[i]);
#pragma unroll
      for (int c = 0; c < NC; ++c) dAd += w[i] * Vd[i][c] * Cd[i][c];
    }
    dAd = wave_sum(dAd);
    const double alpha = rz / dAd;
    // Vq y += alpha Vq d; g -= alpha m Vq d (r -= alpha A d)
#pragma unroll
    for (int i = 0; i < QPL; ++i) {
      const int q = lane + 64 * i;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        Vy[i][c] += alpha * Vd[i][c];
        g[i][c] -= alpha * Cd[i][c];
        if (q < Nq) sg[c * Nq + q] = g[i][c];
      }
    }
    __syncthreads();
    // z = MB^-1 r = Pq g (lanes over nodes; Pq column-major Np x Nq)
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int n = lane + 64 * i;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double acc = 0.0;
        if (n < Np) {
          for (int q = 0; q < Nq; ++q) acc += Pq[n + q * Np] * sg[c * Nq + q];
          sz[c * Np + n] = acc;
        }
      }
    }
    __syncthreads();
    // Vq z (lanes over cubature points) and r^T z = sum_q w_q g_q . (Vq z)_q
    double Vz[QPL][NC];
    double rzn = 0.0;
#pragma unroll
    for (int i = 0; i < QPL; ++i) {
      const int q = lane + 64 * i;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double acc = 0.0;
        if (q < Nq)
          for (int j = 0; j < Np; ++j) acc += Vq[q + j * Nq] * sz[c * Np + j];
        Vz[i][c] = acc;
        rzn += w[i] * g[i][c] * acc;
      }
    }
    rzn = wave_sum(rzn);
    const double beta = rzn / rz;
    rz = rzn;
    // d = z + beta d
#pragma unroll
    for (int i = 0; i < QPL; ++i)
#pragma unroll
      for (int c = 0; c < NC; ++c) Vd[i][c] = Vz[i][c] + beta * Vd[i][c];
    __syncthreads();
  }
  // (MB x)^T y = sum_q w_q (Vq x)_q . (Vq y)_q
  double e = 0.0;
#pragma unroll
  for (int i = 0; i < QPL; ++i)
#pragma unroll
    for (int c = 0; c < NC; ++c) e += w[i] * g0[i][c] * Vy[i][c];
  return wave_sum(e);
}

template <int N, int NF>
__global__ __launch_bounds__(64) void k_energy_wadg(int K, const double *__restrict__ Vq64,
                                                    const double *__restrict__ Pq64,
                                                    const double *__restrict__ wq,
                                                    const double *__restrict__ J,
                                                    const float *__restrict__ m0,
                                                    const float *__restrict__ m1,
                                                    const float *__restrict__ m2,
                                                    const float *__restrict__ Q,
                                                    double *__restrict__ partial) {
  constexpr int Np = Dims<N>::Np, Nq = Dims<N>::Nq;
  constexpr int NC = NF == 4 ? 1 : 3;
  __shared__ double sg[3 * Nq], sz[3 * Np];
  const int lane = threadIdx.x;
  double acc = 0.0;
  for (int k = blockIdx.x; k < K; k += gridDim.x) {
    const float *Qk = Q + (size_t)k * NF * Np;
    const size_t mo = (size_t)k * Nq;
    double e;
    if constexpr (NF == 4) {
      e = wadg_block<N, NC, MED_SCALAR>(Qk, 0, m0 + mo, nullptr, Vq64, Pq64, wq, sg, sz, lane) +
          wadg_block<N, 3, MED_UNIT>(Qk, 1, nullptr, nullptr, Vq64, Pq64, wq, sg, sz, lane);
    } else {
      e = wadg_block<N, 3, MED_SCALAR>(Qk, 0, m0 + mo, nullptr, Vq64, Pq64, wq, sg, sz, lane) +
          wadg_block<N, 3, MED_ISO>(Qk, 3, m1 + mo, m2 + mo, Vq64, Pq64, wq, sg, sz, lane) +
          wadg_block<N, 3, MED_SCALAR>(Qk, 6, m2 + mo, nullptr, Vq64, Pq64, wq, sg, sz, lane);
    }
    acc += J[k] * e;
  }
  if (lane == 0) partial[blockIdx.x] = 0.5 * acc;
}

__global__ __launch_bounds__(256) void k_misfit(const float *__restrict__ tr, const float *__restrict__ d,
                                                long long n, float *__restrict__ resid,
                                                double *__restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double r = (double)tr[i] - (double)d[i];
    if (resid) resid[i] = (float)r;
    acc += 0.5 * r * r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ partial, int n,
                                                double *__restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

int energy_blocks(const bbw_ctx *c, int kind) {
  const long long want = kind == BBW_ENERGY_QUADRATURE ? ((long long)c->K * c->Nq + 255) / 256 : c->K;
  return (int)(want < BBW_ENERGY_BLOCKS ? want : BBW_ENERGY_BLOCKS);
}

template <int N>
static hipError_t energy_n(bbw_ctx *c, int kind, double *partial, double *out) {
  const float *m0 = c->Nf == 4 ? c->c2q : c->rhoinv;
  const float *m1 = c->Nf == 4 ? nullptr : c->lambda;
  const float *m2 = c->Nf == 4 ? nullptr : c->mu;
  const float *Q = c->Q[c->cur];
  const int grid = energy_blocks(c, kind);
  if (grid > BBW_ENERGY_BLOCKS) return hipErrorInvalidValue;
  if (kind == BBW_ENERGY_QUADRATURE) {
    if (c->Nf == 4)
      hipLaunchKernelGGL((k_energy_quad<N, 4>), dim3(grid), dim3(256), 0, c->stream, c->K, c->Vq64, c->wq64,
                         c->J, m0, m1, m2, Q, partial);
    else
      hipLaunchKernelGGL((k_energy_quad<N, 9>), dim3(grid), dim3(256), 0, c->stream, c->K, c->Vq64, c->wq64,
                         c->J, m0, m1, m2, Q, partial);
  } else {
    if (c->Nf == 4)
      hipLaunchKernelGGL((k_energy_wadg<N, 4>), dim3(grid), dim3(64), 0, c->stream, c->K, c->Vq64, c->Pq64,
                         c->wq64, c->J, m0, m1, m2, Q, partial);
    else
      hipLaunchKernelGGL((k_energy_wadg<N, 9>), dim3(grid), dim3(64), 0, c->stream, c->K, c->Vq64, c->Pq64,
                         c->wq64, c->J, m0, m1, m2, Q, partial);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, c->stream, partial, grid, out);
  return hipGetLastError();
}

// Partial sums and result of bbw_energy_noalloc: a file-scope device array, one per GPU, held
// until the process ends.  Host code never passes its name to a kernel (on the host it is a
// stand-in, not a device address); it uses the address hipGetSymbolAddress returned for each
// device, recorded at context creation.
static __device__ __align__(256) double scratch[BBW_ENERGY_BLOCKS + 1];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[MAX_DEVICES] = {};
}  // namespace

hipError_t energy_scratch_init(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[dev]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) hipSetDevice(cur);
  if (e != hipSuccess) return e;
  scratch_ptr[dev] = (double *)p;
  return hipSuccess;
}

double *energy_scratch(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_ptr[dev];
}

hipError_t launch_misfit(bbw_ctx *c, const float *tr, const float *d, long long n, float *resid,
                         double *partial, double *out) {
  const long long want = (n + 255) / 256;
  const int grid = (int)(want < 1 ? 1 : want < BBW_ENERGY_BLOCKS ? want : BBW_ENERGY_BLOCKS);
  hipLaunchKernelGGL(k_misfit, dim3(grid), dim3(256), 0, c->stream, tr, d, n, resid, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, c->stream, partial, grid, out);
  return hipGetLastError();
}

hipError_t launch_energy(bbw_ctx *c, int kind, double *partial, double *out) {
  switch (c->N) {
    case 1: return energy_n<1>(c, kind, partial, out);
    case 2: return energy_n<2>(c, kind, partial, out);
    case 3: return energy_n<3>(c, kind, partial, out);
    case 4: return energy_n<4>(c, kind, partial, out);
    case 5: return energy_n<5>(c, kind, partial, out);
    case 6: return energy_n<6>(c, kind, partial, out);
  }
  return hipErrorInvalidValue;
}

}  // namespace bbw
