// General dense metrics for LMC (lmcmc): the per-chain linear algebra of the
// Lagrangian integrator for a metric the caller evaluates (G(q) and its
// Jacobian dG_ab/dq_i, e.g. by automatic differentiation of a user metric_fn),
// float64, D <= 64, one 256-lane workgroup per chain.
//
// Reference: geomjax/lmcmc/metrics.py:75-112 (velocity draw L^-T z, kinetic
// energy), :158-189 (Omega~ from the Jacobian, grad log det G),
// geomjax/lmcmc/integrators.py:61-91 (the half step: LU factorisations of
// G +- (eps/2) Omega~, the solve, the volume change).
//
// Omega~(v)_lj = 1/2 (T_jl + U_lj - U_jl) with T_ab = sum_i v_i dG_ab/dq_i and
// U_bj = sum_i v_i dG_ib/dq_j: two contractions of the D^3 Jacobian, read once
// each (coalesced along its fastest index).  G + (eps/2) Omega~ is not
// symmetric for a general metric, so it is factored by LU with partial
// pivoting in LDS (log |det| from the pivots, the forward elimination applied
// to the right side as the factorisation proceeds, then a back substitution).
// The position quantities use the Cholesky factor L of (G + G^T)/2: log det,
// v0 = L^-T z, G^-1 = L^-T L^-1 and grad log det G = tr(G^-1 dG/dq_k).
#include "common.h"
#include "dense_metric.h"

namespace gjx {
namespace {

constexpr int NT = 256;

// LDS of one chain for a capacity MD (32 or 64): a D <= 32 problem takes a
// quarter of the LDS, so four times as many chains are resident per CU.
template <int MD>
struct SmemT {
  static constexpr int LDA = MD + 1;
  double A[MD * LDA];  // the matrix being factored
  double W[MD * LDA];  // G (half step) / L^-1, then G^-1 (position)
  double T[MD * LDA];  // the contraction T (half step)
  double v[MD], rhs[MD], x[MD];
  double part[4][64];
  double red[8];
  int piv;
};

template <class Smem>
__device__ __forceinline__ double block_sum(Smem& S, double v) {
  v = group_sum<64>(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) S.red[w] = v;
  __syncthreads();
  return S.red[0] + S.red[1] + S.red[2] + S.red[3];
}

// A = G + coef (T^T + U - U^T) for the velocity in S.v, with G in S.W and
// coef = step / 4 (the reference's metric + 0.5 step * 0.5 (P1 + P2 - P3)).
// U is formed in S.A and overwritten in place pair by pair: the lane that owns
// {(l, j), (j, l)} reads both U entries before writing either.
template <class Smem>
__device__ void omega_into_A(Smem& S, const double* __restrict__ dG, int D, double coef) {
  constexpr int LDA = Smem::LDA;
  const int t = threadIdx.x;
  const int DD = D * D;
  for (int ab = t; ab < DD; ab += NT) {  // T_ab = sum_i v_i dG[a][b][i]
    const double* p = dG + (size_t)ab * D;
    double s = 0.0;
    for (int i = 0; i < D; ++i) s = fma(S.v[i], p[i], s);
    S.T[(ab / D) * LDA + ab % D] = s;
  }
  for (int bj = t; bj < DD; bj += NT) {  // U_bj = sum_i v_i dG[i][b][j]
    double s = 0.0;
    for (int i = 0; i < D; ++i) s = fma(S.v[i], dG[(size_t)i * DD + bj], s);
    S.A[(bj / D) * LDA + bj % D] = s;
  }
  __syncthreads();
  for (int lj = t; lj < DD; lj += NT) {
    const int l = lj / D, j = lj % D;
    if (l > j) continue;
    const double ulj = S.A[l * LDA + j], ujl = S.A[j * LDA + l];
    const double tjl = S.T[j * LDA + l], tlj = S.T[l * LDA + j];
    S.A[l * LDA + j] = S.W[l * LDA + j] + coef * (tjl + ulj - ujl);
    if (l != j) S.A[j * LDA + l] = S.W[j * LDA + l] + coef * (tlj + ujl - ulj);
  }
  __syncthreads();
}

// LU with partial pivoting of the leading D x D block of S.A, the forward
// elimination applied to S.rhs when solve is set; then S.x = A^-1 rhs.
// Returns log |det A|.  Every lane leaves after the same D steps.
template <class Smem>
__device__ double lu_logdet_solve(Smem& S, int D, bool solve) {
  constexpr int LDA = Smem::LDA;
  const int t = threadIdx.x;
  double logdet = 0.0;
  for (int k = 0; k < D; ++k) {
    if (t < 64) {  // pivot row: largest |A_ik|, i >= k (ties and NaN keep the lower index)
      double best = (t >= k && t < D) ? fabs(S.A[t * LDA + k]) : -1.0;
      if (best != best) best = -1.0;
      int bi = t;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (t == 0) S.piv = (bi >= k && bi < D) ? bi : k;
    }
    __syncthreads();
    const int p = S.piv;
    if (p != k) {
      for (int j = t; j < D; j += NT) {
        const double a = S.A[k * LDA + j];
        S.A[k * LDA + j] = S.A[p * LDA + j];
        S.A[p * LDA + j] = a;
      }
      if (solve && t == 0) {
        const double r = S.rhs[k];
        S.rhs[k] = S.rhs[p];
        S.rhs[p] = r;
      }
    }
    __syncthreads();
    const double d = S.A[k * LDA + k];
    logdet += log(fabs(d));
    const double rd = 1.0 / d;
    for (int i = k + 1 + t; i < D; i += NT) {
      const double m = S.A[i * LDA + k] * rd;
      S.A[i * LDA + k] = m;
      if (solve) S.rhs[i] = fma(-m, S.rhs[k], S.rhs[i]);
    }
    __syncthreads();
    const int w = D - k - 1;
    for (int idx = t; idx < w * w; idx += NT) {
      const int i = k + 1 + idx / w, j = k + 1 + idx % w;
      S.A[i * LDA + j] = fma(-S.A[i * LDA + k], S.A[k * LDA + j], S.A[i * LDA + j]);
    }
    __syncthreads();
  }
  if (solve) {
    for (int k = D - 1; k >= 0; --k) {
      if (t == 0) S.x[k] = S.rhs[k] / S.A[k * LDA + k];
      __syncthreads();
      const double xk = S.x[k];
      for (int i = t; i < k; i += NT) S.rhs[i] = fma(-S.A[i * LDA + k], xk, S.rhs[i]);
      __syncthreads();
    }
  }
  return logdet;
}

// Right-looking Cholesky of the symmetric part of S.A (lower, in place).
// Returns log det.  A non-positive pivot gives NaN (as the reference's failed
// factorisation does), never a hang.
template <class Smem>
__device__ double cholesky(Smem& S, int D) {
  constexpr int LDA = Smem::LDA;
  const int t = threadIdx.x;
  double logdet = 0.0;
  for (int k = 0; k < D; ++k) {
    const double d = sqrt(S.A[k * LDA + k]);
    logdet += 2.0 * log(d);
    __syncthreads();  // every lane has read the pivot before column k is scaled
    for (int i = k + 1 + t; i < D; i += NT) S.A[i * LDA + k] /= d;
    if (t == 0) S.A[k * LDA + k] = d;
    __syncthreads();
    const int w = D - k - 1;
    for (int idx = t; idx < w * w; idx += NT) {
      const int i = k + 1 + idx / w, j = k + 1 + idx % w;
      if (j <= i) S.A[i * LDA + j] = fma(-S.A[i * LDA + k], S.A[j * LDA + k], S.A[i * LDA + j]);
    }
    __syncthreads();
  }
  return logdet;
}

template <int MD>
__global__ void __launch_bounds__(NT) dense_metric_position_kernel(const DenseMetricArgs a) {
  using Smem = SmemT<MD>;
  constexpr int LDA = Smem::LDA;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& S = *reinterpret_cast<Smem*>(smem_raw);
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int t = threadIdx.x, D = a.D, DD = D * D;
  const double* G = a.metric + (size_t)c * DD;
  const double* dG = a.metric_jac + (size_t)c * DD * D;
  for (int ij = t; ij < DD; ij += NT) {
    const int i = ij / D, j = ij % D;
    S.A[i * LDA + j] = 0.5 * (G[i * D + j] + G[j * D + i]);
  }
  if (t < D) S.v[t] = a.velocity ? a.velocity[(size_t)c * D + t] : 0.0;
  __syncthreads();
  // v^T G v of the given velocity with the symmetric part (kinetic energy, metrics.py:93-112)
  double vgv = 0.0;
  if (a.velocity != nullptr) {
    double s = 0.0;
    for (int ij = t; ij < DD; ij += NT) s += S.v[ij / D] * S.A[(ij / D) * LDA + ij % D] * S.v[ij % D];
    vgv = block_sum(S, s);
  }
  const double logdet = cholesky(S, D);
  // W = L^-1, one column per lane (forward substitution on e_j)
  if (t < D) {
    const int j = t;
    for (int i = 0; i < D; ++i) {
      double s = i == j ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) s = fma(-S.A[i * LDA + k], S.W[k * LDA + j], s);
      S.W[i * LDA + j] = i < j ? 0.0 : s / S.A[i * LDA + i];
    }
  }
  // v0 = L^-T z (metrics.py:75-91): back substitution on L^T
  if (a.z != nullptr) {
    if (t < D) S.rhs[t] = a.z[(size_t)c * D + t];
    __syncthreads();
    for (int k = D - 1; k >= 0; --k) {
      if (t == 0) S.x[k] = S.rhs[k] / S.A[k * LDA + k];
      __syncthreads();
      const double xk = S.x[k];
      for (int i = t; i < k; i += NT) S.rhs[i] = fma(-S.A[k * LDA + i], xk, S.rhs[i]);
      __syncthreads();
    }
  }
  __syncthreads();
  // G^-1 = W^T W into S.T
  for (int ab = t; ab < DD; ab += NT) {
    const int ai = ab / D, bi = ab % D;
    double s = 0.0;
    for (int k = (ai > bi ? ai : bi); k < D; ++k) s = fma(S.W[k * LDA + ai], S.W[k * LDA + bi], s);
    S.T[ai * LDA + bi] = s;
  }
  __syncthreads();
  // grad log det G_k = sum_ab (G^-1)_ba dG[a][b][k]: lanes along k, four parts over a
  {
    const int k = t % 64, part = t / 64;
    double s = 0.0;
    if (k < D)
      for (int ai = part; ai < D; ai += 4)
        for (int bi = 0; bi < D; ++bi) s = fma(S.T[bi * LDA + ai], dG[((size_t)ai * D + bi) * D + k], s);
    S.part[part][k] = s;
  }
  __syncthreads();
  if (t < D) {
    const size_t r = (size_t)c * D + t;
    if (a.grad_logdet) a.grad_logdet[r] = S.part[0][t] + S.part[1][t] + S.part[2][t] + S.part[3][t];
    if (a.velocity_out && a.z) a.velocity_out[r] = S.x[t];
  }
  if (t == 0) {
    if (a.logdet) a.logdet[c] = logdet;
    if (a.vgv) a.vgv[c] = vgv;
  }
}

template <int MD>
__global__ void __launch_bounds__(NT) dense_lmc_half_step_kernel(const DenseMetricArgs a) {
  using Smem = SmemT<MD>;
  constexpr int LDA = Smem::LDA;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& S = *reinterpret_cast<Smem*>(smem_raw);
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int t = threadIdx.x, D = a.D, DD = D * D;
  const double eps = a.step_sizes ? a.step_sizes[c] : a.step_size;
  const double* G = a.metric + (size_t)c * DD;
  const double* dG = a.metric_jac + (size_t)c * DD * D;
  for (int ij = t; ij < DD; ij += NT) S.W[(ij / D) * LDA + ij % D] = G[ij];
  if (t < D) S.v[t] = a.velocity[(size_t)c * D + t];
  __syncthreads();
  // right side G v - (eps/2)(-grad l + grad log det G / 2) (integrators.py:61-91)
  if (t < D) {
    double gv = 0.0;
    for (int j = 0; j < D; ++j) gv = fma(S.W[t * LDA + j], S.v[j], gv);
    const size_t r = (size_t)c * D + t;
    S.rhs[t] = gv - 0.5 * eps * (-a.logdensity_grad[r] + 0.5 * a.grad_logdet_in[r]);
  }
  omega_into_A(S, dG, D, 0.25 * eps);
  const double ld_plus = lu_logdet_solve(S, D, true);
  // the new velocity, then G - (eps/2) Omega~(v')
  if (t < D) S.v[t] = S.x[t];
  __syncthreads();
  omega_into_A(S, dG, D, -0.25 * eps);
  const double ld_minus = lu_logdet_solve(S, D, false);
  if (t < D) a.velocity_out[(size_t)c * D + t] = S.v[t];
  if (t == 0) a.volume_adjustment[c] = a.volume_adjustment[c] - ld_plus + ld_minus;
}

// Manifold MALA quantities of a state (geomjax/lmcmc/mmala.py:91-113,
// geomjax/lmcmc/diffusions.py:55-94): with G^-1 from the Cholesky factor of
// (G + G^T)/2, Gamma = -G^-1 h with h_a = sum_bj dG_ab/dq_j (G^-1)_bj (so that
// Gamma_i = sum_j d(G^-1)_ij/dq_j), the drift G^-1 grad l + Gamma, log det G and
// the draw chol(G^-1) z (the lower factor of the inverse, which the reference
// factors as the proposal covariance 2 eps G^-1).
template <int MD>
__global__ void __launch_bounds__(NT) dense_mmala_drift_kernel(const DenseMetricArgs a) {
  using Smem = SmemT<MD>;
  constexpr int LDA = Smem::LDA;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& S = *reinterpret_cast<Smem*>(smem_raw);
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int t = threadIdx.x, D = a.D, DD = D * D;
  const double* G = a.metric + (size_t)c * DD;
  for (int ij = t; ij < DD; ij += NT) {
    const int i = ij / D, j = ij % D;
    S.A[i * LDA + j] = 0.5 * (G[i * D + j] + G[j * D + i]);
  }
  if (t < D) S.v[t] = a.logdensity_grad ? a.logdensity_grad[(size_t)c * D + t] : 0.0;
  __syncthreads();
  const double logdet = cholesky(S, D);
  if (t < D) {  // W = L^-1, one column per lane
    const int j = t;
    for (int i = 0; i < D; ++i) {
      double s = i == j ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) s = fma(-S.A[i * LDA + k], S.W[k * LDA + j], s);
      S.W[i * LDA + j] = i < j ? 0.0 : s / S.A[i * LDA + i];
    }
  }
  __syncthreads();
  for (int ab = t; ab < DD; ab += NT) {  // G^-1 = W^T W into S.T
    const int ai = ab / D, bi = ab % D;
    double s = 0.0;
    for (int k = (ai > bi ? ai : bi); k < D; ++k) s = fma(S.W[k * LDA + ai], S.W[k * LDA + bi], s);
    S.T[ai * LDA + bi] = s;
  }
  __syncthreads();
  // h_a = sum_b sum_j dG[a][b][j] (G^-1)_bj: one (a, b) per lane pass (contiguous j) into S.A
  if (a.metric_jac != nullptr) {
    const double* dG = a.metric_jac + (size_t)c * DD * D;
    for (int ab = t; ab < DD; ab += NT) {
      const int ai = ab / D, bi = ab % D;
      const double* p = dG + (size_t)ab * D;
      double s = 0.0;
      for (int j = 0; j < D; ++j) s = fma(p[j], S.T[bi * LDA + j], s);
      S.A[ai * LDA + bi] = s;
    }
    __syncthreads();
    if (t < D) {
      double h = 0.0;
      for (int bi = 0; bi < D; ++bi) h += S.A[t * LDA + bi];
      S.rhs[t] = h;
    }
  } else if (t < D) {
    S.rhs[t] = 0.0;
  }
  __syncthreads();
  // Gamma = -G^-1 h, drift = G^-1 grad l + Gamma
  if (t < D) {
    double gh = 0.0, gg = 0.0;
    for (int j = 0; j < D; ++j) {
      gh = fma(S.T[t * LDA + j], S.rhs[j], gh);
      gg = fma(S.T[t * LDA + j], S.v[j], gg);
    }
    const size_t r = (size_t)c * D + t;
    if (a.gamma) a.gamma[r] = -gh;
    if (a.drift) a.drift[r] = gg - gh;
  }
  // chol(G^-1) z: the lower Cholesky factor of the inverse, formed in S.A
  if (a.z != nullptr && a.chol_inv_z != nullptr) {
    __syncthreads();
    for (int ij = t; ij < DD; ij += NT) {
      const int i = ij / D, j = ij % D;
      S.A[i * LDA + j] = 0.5 * (S.T[i * LDA + j] + S.T[j * LDA + i]);
    }
    if (t < D) S.x[t] = a.z[(size_t)c * D + t];
    __syncthreads();
    (void)cholesky(S, D);
    if (t < D) {
      double s = 0.0;
      for (int k = 0; k <= t; ++k) s = fma(S.A[t * LDA + k], S.x[k], s);
      a.chol_inv_z[(size_t)c * D + t] = s;
    }
  }
  if (t == 0 && a.logdet) a.logdet[c] = logdet;
}

// The implicit-midpoint vector field of RMHMC for a dense metric
// (geomjax/rmhmc/metrics.py:42-129, geomjax/rmhmc/integrators.py:125-135):
// with T(q, p) = p^T G^-1 p / 2 + log det G / 2 + (D/2) log 2 pi and L the
// Cholesky factor of (G + G^T)/2, dT/dp = u = G^-1 p and
// dT/dq_k = 1/2 sum_ab ((G^-1)_ba - u_a u_b) dG_ab/dq_k; the momentum draw is
// p = L z.
template <int MD>
__global__ void __launch_bounds__(NT) dense_rmhmc_field_kernel(const DenseMetricArgs a) {
  using Smem = SmemT<MD>;
  constexpr int LDA = Smem::LDA;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& S = *reinterpret_cast<Smem*>(smem_raw);
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int t = threadIdx.x, D = a.D, DD = D * D;
  const double* G = a.metric + (size_t)c * DD;
  for (int ij = t; ij < DD; ij += NT) {
    const int i = ij / D, j = ij % D;
    S.A[i * LDA + j] = 0.5 * (G[i * D + j] + G[j * D + i]);
  }
  if (t < D) S.rhs[t] = a.z ? a.z[(size_t)c * D + t] : 0.0;
  __syncthreads();
  const double logdet = cholesky(S, D);
  // the momentum: given, or p = L z
  if (t < D) {
    double pv;
    if (a.momentum != nullptr) {
      pv = a.momentum[(size_t)c * D + t];
    } else {
      pv = 0.0;
      for (int k = 0; k <= t; ++k) pv = fma(S.A[t * LDA + k], S.rhs[k], pv);
    }
    S.v[t] = pv;
    const int j = t;  // W = L^-1, one column per lane
    for (int i = 0; i < D; ++i) {
      double s = i == j ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) s = fma(-S.A[i * LDA + k], S.W[k * LDA + j], s);
      S.W[i * LDA + j] = i < j ? 0.0 : s / S.A[i * LDA + i];
    }
  }
  __syncthreads();
  for (int ab = t; ab < DD; ab += NT) {  // G^-1 = W^T W into S.T
    const int ai = ab / D, bi = ab % D;
    double s = 0.0;
    for (int k = (ai > bi ? ai : bi); k < D; ++k) s = fma(S.W[k * LDA + ai], S.W[k * LDA + bi], s);
    S.T[ai * LDA + bi] = s;
  }
  __syncthreads();
  if (t < D) {  // u = G^-1 p
    double s = 0.0;
    for (int j = 0; j < D; ++j) s = fma(S.T[t * LDA + j], S.v[j], s);
    S.x[t] = s;
  }
  __syncthreads();
  const double pu = block_sum(S, t < D ? S.v[t] * S.x[t] : 0.0);
  if (a.metric_jac != nullptr) {
    const double* dG = a.metric_jac + (size_t)c * DD * D;
    const int k = t % 64, part = t / 64;
    double s = 0.0;
    if (k < D)
      for (int ai = part; ai < D; ai += 4) {
        const double ua = S.x[ai];
        for (int bi = 0; bi < D; ++bi)
          s = fma(S.T[bi * LDA + ai] - ua * S.x[bi], dG[((size_t)ai * D + bi) * D + k], s);
      }
    S.part[part][k] = s;
    __syncthreads();
  }
  if (t < D) {
    const size_t r = (size_t)c * D + t;
    if (a.velocity_out) a.velocity_out[r] = S.x[t];
    if (a.momentum_out) a.momentum_out[r] = S.v[t];
    if (a.dtdq && a.metric_jac) a.dtdq[r] = 0.5 * (S.part[0][t] + S.part[1][t] + S.part[2][t] + S.part[3][t]);
  }
  if (t == 0) {
    if (a.logdet) a.logdet[c] = logdet;
    if (a.pu) a.pu[c] = pu;
  }
}

}  // namespace

size_t dense_metric_smem_bytes() { return sizeof(SmemT<64>); }

namespace {

// the capacity-32 instantiation for D <= 32 (a quarter of the LDS), else capacity 64
hipError_t launch_dense(const DenseMetricArgs& a, hipStream_t stream, void (*k32)(DenseMetricArgs),
                        void (*k64)(DenseMetricArgs)) {
  const bool small = a.D <= 32;
  const size_t smem = small ? sizeof(SmemT<32>) : sizeof(SmemT<64>);
  void (*k)(DenseMetricArgs) = small ? k32 : k64;
  (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL(k, dim3(a.C), dim3(NT), smem, stream, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_dense_rmhmc_field(const DenseMetricArgs& a, hipStream_t stream) {
  return launch_dense(a, stream, dense_rmhmc_field_kernel<32>, dense_rmhmc_field_kernel<64>);
}

hipError_t launch_dense_mmala_drift(const DenseMetricArgs& a, hipStream_t stream) {
  return launch_dense(a, stream, dense_mmala_drift_kernel<32>, dense_mmala_drift_kernel<64>);
}

hipError_t launch_dense_metric_position(const DenseMetricArgs& a, hipStream_t stream) {
  return launch_dense(a, stream, dense_metric_position_kernel<32>, dense_metric_position_kernel<64>);
}

hipError_t launch_dense_lmc_half_step(const DenseMetricArgs& a, hipStream_t stream) {
  return launch_dense(a, stream, dense_lmc_half_step_kernel<32>, dense_lmc_half_step_kernel<64>);
}

}  // namespace gjx
