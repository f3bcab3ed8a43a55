// Multi-chain effective sample size and potential scale reduction of a
// [T, C, D] sample array on the device (float64 arithmetic).
//
// Reference: geomjax/diagnostics.py:25-75 (potential_scale_reduction) and
// :78-209 (effective_sample_size: autocovariance of every chain, the
// mean autocovariance over chains, Geyer's initial positive and initial
// monotone sequences, tau floored at 1 / log10(M N)).  The reference forms
// the autocovariances by FFT with zero padding to at least 2N; that is the
// same quantity as the direct sums taken here,
//   acov_c(k) = (1/N) sum_{t < N-k} (x_{t,c} - m_c)(x_{t+k,c} - m_c),
// which for the sample windows of this path (N up to a few thousand) cost
// less than a transform.
//
// Pass 1, one workgroup per (coordinate, tile of up to 16 chains; fewer for
// series longer than 1280 so that the tile fits LDS): the tile's
// centred series go to LDS in float64; thread k sums lag k over the tile's
// chains.  Outputs per tile the lag sums and per chain the mean and the
// variance.  Pass 2, one workgroup per coordinate: the tiles' sums are added,
// and one lane runs the sequential Geyer steps (at most N/2 of them).
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/geomjax_amd.h"
#include "common.h"

namespace gjx {
namespace diag {

constexpr int kMaxTile = 16;  // chains per pass-1 workgroup at most
constexpr int kThreads = 256;
constexpr size_t kLdsBytes = 160 * 1024 - 1024;  // dynamic LDS of pass 1 at most

// chains per pass-1 workgroup for series of length N (0: too long)
__host__ __device__ inline int tile_chains(int N) {
  const size_t t = kLdsBytes / ((size_t)N * sizeof(double));
  return (int)(t < (size_t)kMaxTile ? t : (size_t)kMaxTile);
}

template <class T>
__global__ void __launch_bounds__(kThreads) lag_sums_kernel(const T* __restrict__ x, int N, int C, int tc,
                                                            int64_t st, int64_t sc, int64_t sd,
                                                            double* __restrict__ lag, double* __restrict__ cmean,
                                                            double* __restrict__ cvar) {
  extern __shared__ double xs[];  // [tc][N] centred series
  const int d = blockIdx.y, tile = blockIdx.x, D = gridDim.y;
  const int c0 = tile * tc;
  const int nc = min(tc, C - c0);
  const int tid = threadIdx.x;
  __shared__ double red[kMaxTile][kThreads / 64];
  // load the tile's series and their means
  for (int j = 0; j < nc; ++j) {
    double s = 0.0;
    for (int t = tid; t < N; t += kThreads) {
      const double v = (double)x[(int64_t)t * st + (int64_t)(c0 + j) * sc + (int64_t)d * sd];
      xs[j * N + t] = v;
      s += v;
    }
    s = group_sum<64>(s);
    if (tid % 64 == 0) red[j][tid / 64] = s;
  }
  __syncthreads();
  for (int j = 0; j < nc; ++j) {
    const double m = (red[j][0] + red[j][1] + red[j][2] + red[j][3]) / (double)N;
    for (int t = tid; t < N; t += kThreads) xs[j * N + t] -= m;
    if (tid == 0) cmean[(int64_t)d * C + c0 + j] = m;
  }
  __syncthreads();
  // lag sums over the tile's chains
  for (int k = tid; k < N; k += kThreads) {
    double s = 0.0;
    for (int j = 0; j < nc; ++j) {
      const double* a = xs + j * N;
      for (int t = 0; t + k < N; ++t) s = fma(a[t], a[t + k], s);
    }
    lag[((int64_t)tile * D + d) * N + k] = s;
  }
  // per-chain variance (unbiased), for R-hat
  for (int j = 0; j < nc; ++j) {
    double s = 0.0;
    for (int t = tid; t < N; t += kThreads) s = fma(xs[j * N + t], xs[j * N + t], s);
    s = group_sum<64>(s);
    __syncthreads();
    if (tid % 64 == 0) red[j][tid / 64] = s;
    __syncthreads();
    if (tid == 0) cvar[(int64_t)d * C + c0 + j] = (red[j][0] + red[j][1] + red[j][2] + red[j][3]) / (double)(N - 1);
  }
}

// Sum of v over the workgroup (kThreads lanes), identical in every lane.
__device__ double block_sum(double v, double* sh) {
  v = group_sum<64>(v);
  __syncthreads();
  if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kThreads / 64; ++w) t += sh[w];
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(kThreads) finalize_kernel(const double* __restrict__ lag,
                                                            const double* __restrict__ cmean,
                                                            const double* __restrict__ cvar, int N, int C,
                                                            int ntiles, double* __restrict__ rho,
                                                            double* __restrict__ ess, double* __restrict__ rhat) {
  const int d = blockIdx.x, D = gridDim.x, tid = threadIdx.x;
  __shared__ double sh[kThreads / 64];
  double* r = rho + (int64_t)d * N;  // mean autocovariance, then rho
  for (int k = tid; k < N; k += kThreads) {
    double s = 0.0;
    for (int tl = 0; tl < ntiles; ++tl) s += lag[((int64_t)tl * D + d) * N + k];
    r[k] = s / ((double)N * (double)C);
  }
  // variance of the chain means (unbiased) and the mean within-chain variance
  double sm = 0.0, sv = 0.0;
  for (int c = tid; c < C; c += kThreads) {
    sm += cmean[(int64_t)d * C + c];
    sv += cvar[(int64_t)d * C + c];
  }
  const double mbar = block_sum(sm, sh) / (double)C;
  const double within = block_sum(sv, sh) / (double)C;
  double sq = 0.0;
  for (int c = tid; c < C; c += kThreads) {
    const double e = cmean[(int64_t)d * C + c] - mbar;
    sq = fma(e, e, sq);
  }
  const double var_means = block_sum(sq, sh) / (double)(C - 1);
  __syncthreads();
  if (tid != 0) return;
  if (rhat) rhat[d] = sqrt(((double)N * var_means / within + (double)N - 1.0) / (double)N);
  if (!ess) return;
  // Geyer's initial positive and initial monotone sequences (diagnostics.py:150-209)
  const double var0 = r[0] * (double)N / ((double)N - 1.0);
  const double weighted_var = var0 * ((double)N - 1.0) / (double)N + var_means;
  const int n_even = N - N % 2, K = n_even / 2;
  for (int k = 1; k < n_even; ++k) r[k] = 1.0 - (var0 - r[k]) / weighted_var;
  r[0] = 1.0;
  // rho_even[i] = r[2i], rho_odd[i] = r[2i + 1]; mask = positive pair sums so far
  int count = 0;
  bool all_pos = true;
  for (int i = 0; i < K; ++i) {
    all_pos = all_pos && (r[2 * i] + r[2 * i + 1] > 0.0);
    if (all_pos) ++count;
  }
  const int max_t = count > 0 ? count - 1 : 0;
  const int idx = max_t + 1;
  const bool in_range = idx < K;
  const int idc = idx < K ? idx : K - 1;
  const double gathered = r[2 * idc];
  double sum = 0.0, running = 0.0, last_even_f = 0.0;
  for (int i = 0; i < K; ++i) {
    const bool m = i < count;  // the mask is a prefix
    const bool me = (in_range && i == idx) ? (gathered > 0.0) : m;
    const double re = me ? r[2 * i] : 0.0;
    const double ro = m ? r[2 * i + 1] : 0.0;
    const double s = re + ro;
    const double prev = i == 0 ? s : running;
    running = i == 0 ? s : fmin(running, s);
    const bool update = s > prev;
    const double ref = update ? running / 2.0 : re;
    const double rof = update ? running / 2.0 : ro;
    sum += ref + rof;
    if (i == idc) last_even_f = ref;
  }
  double tau = -1.0 + 2.0 * sum - last_even_f;
  tau = fmax(tau, 1.0 / log10((double)C * (double)N));
  ess[d] = (double)C * (double)N / tau;
}

}  // namespace diag

int diagnostics_max_samples() { return (int)(diag::kLdsBytes / sizeof(double)); }

uint64_t diagnostics_workspace_bytes(int N, int C, int D) {
  const int tc = diag::tile_chains(N);
  const int64_t ntiles = (C + tc - 1) / tc;
  return (uint64_t)(ntiles * D * N + 2 * (int64_t)D * C + (int64_t)D * N) * sizeof(double);
}

hipError_t launch_diagnostics(const void* x, int dtype, int N, int C, int D, int64_t st, int64_t sc, int64_t sd,
                              void* workspace, double* ess, double* rhat, hipStream_t stream) {
  const int tc = diag::tile_chains(N);
  const int ntiles = (C + tc - 1) / tc;
  double* lag = static_cast<double*>(workspace);
  double* cmean = lag + (int64_t)ntiles * D * N;
  double* cvar = cmean + (int64_t)D * C;
  double* rho = cvar + (int64_t)D * C;
  const size_t lds = (size_t)tc * N * sizeof(double);
  const dim3 g1((unsigned)ntiles, (unsigned)D);
  if (dtype == GJX_F32) {
    auto k = diag::lag_sums_kernel<float>;
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, g1, dim3(diag::kThreads), lds, stream, static_cast<const float*>(x), N, C, tc, st, sc, sd,
                       lag, cmean, cvar);
  } else {
    auto k = diag::lag_sums_kernel<double>;
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, g1, dim3(diag::kThreads), lds, stream, static_cast<const double*>(x), N, C, tc, st, sc,
                       sd, lag, cmean, cvar);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(diag::finalize_kernel, dim3((unsigned)D), dim3(diag::kThreads), 0, stream, lag, cmean, cvar, N,
                     C, ntiles, rho, ess, rhat);
  return hipGetLastError();
}

}  // namespace gjx
