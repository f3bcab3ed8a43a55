// Manifold MALA (the Euler step of the Riemannian overdamped Langevin
// diffusion with a Metropolis test) for Bayesian logistic regression with the
// Fisher metric, float64 (MFMA Gram passes), D <= 64.
//
// Reference: geomjax/lmcmc/mmala.py:53-148 (init, transition energy, kernel),
// geomjax/lmcmc/diffusions.py:35-94 (the step and Gamma),
// geomjax/mcmc/proposal.py:87-121,124-160,168-185 (asymmetric proposal, accept).
//
// One transition from x:
//   mu(x) = x + eps G^-1 grad l + eps Gamma,  Gamma_i = sum_j d(G^-1)_ij / dx_j,
//   x'    = mu(x) + chol(2 eps G^-1) z,
// and the test on the transition energies -l(x) - log N(x'; mu(x), 2 eps G(x)^-1)
// and -l(x') - log N(x; mu(x'), 2 eps G(x')^-1).
//
// For G = X^T diag(lambda) X + I/prior_var, d(G^-1)/dx_j = -G^-1 (dG/dx_j) G^-1
// with dG/dx_j = X^T diag(w X_{:,j}) X, w = s(1-s)(1-2s), so
//   Gamma = -G^-1 X^T (w a) = -G^-1 grad log det G,   a_n = x_n^T G^-1 x_n,
// and mu(x) = x + eps G^-1 (grad l - grad log det G): one position pass (Gram,
// sweep, quadratic forms) and two products with G^-1 per state.  The draw uses
// the lower Cholesky factor of G^-1 (the reference factors the covariance
// 2 eps G^-1), formed by a second sweep over G^-1 that accumulates chol(G^-1) z
// on the fly.  log N(x'; mu, 2 eps G^-1) = -|z|^2/2 - (D log 2 eps - log det G)/2
// - (D/2) log 2 pi; the reverse density needs r^T G(x') r with r = x - mu(x'),
// one Gram pass with r.
//
// The reference pairs the energies so that the weight is -log alpha (the
// Metropolis-Hastings log ratio with its sign flipped); mmala_exact = 1 uses
// +log alpha.
#include "common.h"
#include "logistic.h"
#include "logistic_common.h"

namespace gjx {

using namespace logistic;

namespace {

using C64 = Cfg<64>;
constexpr int MAXD = C64::MAXD;

// Write G (in S.Gi after a Gram pass) as the [Dmax, Dmax] block of one chain,
// zeros past D.  Every call maps element (i, j) to the same lane, so a later
// call overwrites an earlier one in program order.
__device__ __forceinline__ void write_metric(const Smem<C64>& S, double* dst, int D, int Dp) {
  for (int idx = threadIdx.x; idx < Dp * Dp; idx += NT) {
    const int i = idx / Dp, j = idx % Dp;
    dst[idx] = (i < D && j < D) ? S.Gi[i * C64::LDX + j] : 0.0;
  }
}

// Quantities at S.q: S.Gi = G^-1, S.g = grad l, S.dH = grad log det G; returns
// log det G and (through loglik) the log-likelihood sum.  With metric_dst the
// metric G itself is written there before it is inverted.
__device__ __forceinline__ double mmala_position(Smem<C64>& S, double (&A)[4][4], const double* X, const double* y,
                                                 int N, int D, int Dp, double prior, double* metric_dst,
                                                 double& loglik) {
  pass_gram(S, X, y, N, D, prior, nullptr, 0.0, true, false, loglik);
  if (metric_dst != nullptr) write_metric(S, metric_dst, D, Dp);
  load_block(S, A);
  __syncthreads();
  const double ld = sweep(S, A, D);
  __syncthreads();
  store_neg_block(S, A);
  __syncthreads();
  pass_quad(S, X, y, N, D, prior, false);
  return ld;
}

// S.u = G^-1 (grad l - grad log det G) (the drift direction) and, per lane,
// Gamma = -G^-1 grad log det G; G^-1 in S.Gi.
__device__ __forceinline__ double drift(Smem<C64>& S, int D) {
  const int t = threadIdx.x;
  if (t < MAXD) S.vt[t] = t < D ? S.g[t] - S.dH[t] : 0.0;
  __syncthreads();
  ginv_times(S, S.vt, S.u, D);
  ginv_times(S, S.dH, S.qs, D);
  return t < D ? -S.qs[t] : 0.0;
}

struct Chain {
  int N, D;
  const double* X;
  const double* y;
};

__device__ __forceinline__ bool chain_of(const LogisticArgs& a, int c, Chain& ch) {
  if (c >= a.C) return false;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return false;
  ch.N = a.post_N[post];
  ch.D = a.post_D[post];
  if (ch.D > MAXD) return false;  // rejected by the entry points
  ch.X = a.X + a.post_xoff[post];
  ch.y = a.y + a.post_yoff[post];
  return true;
}

// Start of a chain: its position in S.q0 / S.q and every LDS vector that a
// pass multiplies by whole zero-padded rows of X zeroed over its capacity.
__device__ __forceinline__ void load_position(const LogisticArgs& a, Smem<C64>& S, int c, int D) {
  const int t = threadIdx.x;
  if (t < MAXD) {
    const double qv = t < D ? static_cast<const double*>(a.q_in)[(size_t)c * a.Dmax + t] : 0.0;
    S.q0[t] = qv;
    S.q[t] = qv;
    S.p[t] = 0.0;
    S.u[t] = 0.0;
    S.g[t] = 0.0;
    S.dH[t] = 0.0;
    S.vt[t] = 0.0;
    S.qs[t] = 0.0;
    S.gv[t] = 0.0;
    S.z[t] = 0.0;
  }
  __syncthreads();
}

}  // namespace

template <bool STEP>
__global__ void __launch_bounds__(NT, 2) logistic_mmala_f64_kernel(const LogisticArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C64>& S = *reinterpret_cast<Smem<C64>*>(smem_raw);
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  Chain ch;
  if (!chain_of(a, c, ch)) return;
  const int N = ch.N, D = ch.D, Dp = a.Dmax;
  const double* X = ch.X;
  const double* y = ch.y;
  const bool own = t < D;
  const double eps = a.eps;
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  // log N(.; mu, 2 eps G^-1) = -quad/2 + log det G / 2 + norm_const
  const double norm_const = -0.5 * (double)D * (log(2.0 * eps) + log(2.0 * M_PI));
  const double sq2e = sqrt(2.0 * eps);
  const int T_steps = STEP ? 1 : a.T;
  const size_t row = (size_t)c * Dp;
  double* const metric_row = (STEP && a.metric_out) ? a.metric_out + (size_t)c * Dp * Dp : nullptr;
  double* const prop_metric_row = (STEP && a.prop_metric) ? a.prop_metric + (size_t)c * Dp * Dp : nullptr;

  load_position(a, S, c, D);
  double logp = a.logp_in[c];
  int n_acc = 0;
  double A[4][4];

  for (int tr = 0; tr < T_steps; ++tr) {
    const uint64_t tt = a.counter + (uint64_t)tr;
    double uacc;
    if (a.z != nullptr) {
      if (t < MAXD) S.z[t] = t < D ? static_cast<const double*>(a.z)[((size_t)tr * a.C + c) * Dp + t] : 0.0;
      uacc = a.u[(size_t)tr * a.C + c];
    } else {
      if (t < MAXD / 2) {
        u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * t),
                                      (uint32_t)(tt >> 32)}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        double n0, n1;
        normal2(r, n0, n1);
        S.z[2 * t] = 2 * t < D ? n0 : 0.0;
        S.z[2 * t + 1] = 2 * t + 1 < D ? n1 : 0.0;
      }
      u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamAccept, (uint32_t)(tt >> 32)},
                              (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
      uacc = u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
    }
    if (t < MAXD) S.q[t] = S.q0[t];
    __syncthreads();

    // ---- the current state x ----
    double ll0 = 0.0;
    const double logdet0 = mmala_position(S, A, X, y, N, D, Dp, prior, metric_row, ll0);
    const double g0 = own ? S.g[t] : 0.0;
    const double gamma0 = drift(S, D);
    const double mu = own ? S.q0[t] + eps * S.u[t] : 0.0;
    // chol(G^-1) z by a sweep over G^-1 (the inverse itself stays in S.Gi)
    load_block(S, A);
    __syncthreads();
    (void)sweep(S, A, D, S.z, S.p);
    __syncthreads();
    const double x1 = own ? mu + sq2e * S.p[t] : 0.0;
    const double zz = block_sum(S, own ? S.z[t] * S.z[t] : 0.0);
    const double logq_fwd = -0.5 * zz + 0.5 * logdet0 + norm_const;
    if (t < MAXD) S.q[t] = x1;
    __syncthreads();

    // ---- the proposal x' ----
    double ll1 = 0.0;
    const double logdet1 = mmala_position(S, A, X, y, N, D, Dp, prior, prop_metric_row, ll1);
    const double g1 = own ? S.g[t] : 0.0;
    const double qq1 = block_sum(S, own ? x1 * x1 : 0.0);
    const double logp1 = ll1 - 0.5 * qq1 / prior + lp_const;
    const double gamma1 = drift(S, D);
    const double r = own ? S.q0[t] - (x1 + eps * S.u[t]) : 0.0;
    if (t < MAXD) S.vt[t] = r;
    __syncthreads();
    double dd = 0.0;
    pass_gram(S, X, y, N, D, prior, S.vt, 0.0, false, true, dd);  // S.gv = G(x') r, S.Gi = G(x')
    const double rGr = block_sum(S, own ? r * S.gv[t] : 0.0);
    const double logq_bwd = -rGr / (4.0 * eps) + 0.5 * logdet1 + norm_const;

    // ---- the test (proposal.py:87-121,168-185) ----
    const double new_energy = -logp - logq_fwd;
    const double prev_energy = -logp1 - logq_bwd;
    double delta = a.mmala_exact ? new_energy - prev_energy : prev_energy - new_energy;
    if (isnan(delta)) delta = -INFINITY;
    const double p_acc = fmin(exp(delta), 1.0);
    const bool accept = uacc < p_acc;

    if (STEP) {
      if (t == 0) {
        if (a.acc_rate) a.acc_rate[c] = p_acc;
        if (a.accepted) a.accepted[c] = (uint8_t)accept;
        if (a.divergent) a.divergent[c] = (uint8_t)(-delta > a.div_threshold);
        if (a.prop_weight) a.prop_weight[c] = delta;
        if (a.prop_logp) a.prop_logp[c] = logp1;
        a.logp_out[c] = accept ? logp1 : logp;
      }
      if (own) {
        if (a.prop_q) static_cast<double*>(a.prop_q)[row + t] = x1;
        if (a.prop_g) static_cast<double*>(a.prop_g)[row + t] = g1;
        if (a.prop_gamma) a.prop_gamma[row + t] = gamma1;
        static_cast<double*>(a.q_out)[row + t] = accept ? x1 : S.q0[t];
        if (a.g_out) {
          const double gin = a.g_in ? static_cast<const double*>(a.g_in)[row + t] : g0;
          static_cast<double*>(a.g_out)[row + t] = accept ? g1 : gin;
        }
        if (a.gamma_out) a.gamma_out[row + t] = accept ? gamma1 : gamma0;
      }
      if (accept && metric_row != nullptr) write_metric(S, metric_row, D, Dp);
    }
    __syncthreads();  // every lane has read S.q0 before it moves
    if (accept) {
      if (t < MAXD) S.q0[t] = S.q[t];
      logp = logp1;
      ++n_acc;
    }
    __syncthreads();
    if (!STEP) {
      if (a.samples != nullptr && own) static_cast<double*>(a.samples)[((size_t)tr * a.C + c) * Dp + t] = S.q0[t];
      if (a.sample_acc != nullptr && t == 0) a.sample_acc[(size_t)tr * a.C + c] = (float)p_acc;
    }
  }

  if (!STEP) {
    if (own) static_cast<double*>(a.final_q)[row + t] = S.q0[t];
    if (t == 0) {
      a.final_logp[c] = logp;
      if (a.num_accepted) a.num_accepted[c] = n_acc;
    }
  }
  if (STEP) {
    for (int j = D + t; j < Dp; j += NT) {
      if (a.gamma_out) a.gamma_out[row + j] = 0.0;
      if (a.prop_gamma) a.prop_gamma[row + j] = 0.0;
    }
  }
  zero_row_tails<double, STEP>(a, c, D, t, NT);
}

// init (mmala.py:53-60): log density, gradient, metric and Gamma at q_in.
__global__ void __launch_bounds__(NT, 2) logistic_mmala_init_kernel(const LogisticArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C64>& S = *reinterpret_cast<Smem<C64>*>(smem_raw);
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  Chain ch;
  if (!chain_of(a, c, ch)) return;
  const int D = ch.D, Dp = a.Dmax;
  const double prior = a.prior_var;
  const size_t row = (size_t)c * Dp;
  load_position(a, S, c, D);
  double A[4][4];
  double ll = 0.0;
  (void)mmala_position(S, A, ch.X, ch.y, ch.N, D, Dp, prior, a.metric_out ? a.metric_out + (size_t)c * Dp * Dp : nullptr,
                       ll);
  const double qq = block_sum(S, t < D ? S.q[t] * S.q[t] : 0.0);
  (void)drift(S, D);  // S.qs = G^-1 grad log det G = -Gamma
  if (t == 0 && a.logp_out) a.logp_out[c] = ll - 0.5 * qq / prior - 0.5 * (double)D * log(2.0 * M_PI * prior);
  for (int j = t; j < Dp; j += NT) {
    if (a.g_out) static_cast<double*>(a.g_out)[row + j] = j < D ? S.g[j] : 0.0;
    if (a.gamma_out) a.gamma_out[row + j] = j < D ? -S.qs[j] : 0.0;
  }
}

namespace {

template <class K>
void launch_one(K kern, const LogisticArgs& a, hipStream_t stream) {
  const size_t smem = sizeof(Smem<C64>);
  (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL(kern, dim3(a.C), dim3(NT), smem, stream, a);
}

}  // namespace

hipError_t launch_logistic_mmala_init(const LogisticArgs& a, hipStream_t stream) {
  launch_one(logistic_mmala_init_kernel, a, stream);
  return hipGetLastError();
}

hipError_t launch_logistic_mmala_f64(const LogisticArgs& a, hipStream_t stream) {
  if (a.q_out != nullptr) {
    launch_one(logistic_mmala_f64_kernel<true>, a, stream);
    return hipGetLastError();
  }
  launch_one(logistic_mmala_f64_kernel<false>, a, stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.final_g != nullptr || a.metric_out != nullptr || a.gamma_out != nullptr) {
    // the final state's gradient, metric and Gamma, evaluated at final_q
    LogisticArgs b = a;
    b.q_in = a.final_q;
    b.logp_out = nullptr;
    b.g_out = a.final_g;
    launch_one(logistic_mmala_init_kernel, b, stream);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace gjx
