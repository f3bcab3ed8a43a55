// Lagrangian Monte Carlo with a general metric (lmcmc) for Bayesian logistic
// regression with the Fisher metric, float64 (MFMA Gram passes).
//
// Reference: geomjax/lmcmc/lmc.py:114-182,411-499 (transition, proposal,
// flip), geomjax/lmcmc/integrators.py:51-144 (lan_integrator: half steps with
// LU factorisations of G +- (eps/2) Omega~), geomjax/lmcmc/metrics.py:42-221
// (velocity draw L^-T z, kinetic energy -1/2 log det G + 1/2 v^T G v,
// Omega~ from jacfwd(metric_fn), grad log det G).
//
// For G(q) = X^T diag(lambda) X + I/prior_var, lambda = s(1-s), every
// derivative of G is a weighted Gram: dG/dq_i = X^T diag(w X_{:,i}) X with
// w = s(1-s)(1-2s).  The Christoffel contraction of the reference
// (metrics.py:158-179) is therefore
//   Omega~(v) = 1/2 X^T diag(w (X v)) X,
// so G +- (eps/2) Omega~(v) = X^T diag(lambda +- (eps/4) w (X v)) X + I/prior_var
// is one more Gram pass, and grad log det G = X^T (w a), a_n = x_n^T G^-1 x_n.
// Determinants come from the pivots of the symmetric sweep (log |det|); the
// solve (G + (eps/2) Omega~)^-1 (G v - (eps/2) grad phi) uses the swept
// inverse.  Per integrator step: one position pass set (Gram, sweep,
// quadratic forms) and two half steps of two Gram passes and two sweeps.
#include "common.h"
#include "logistic.h"
#include "logistic_common.h"
#include "logistic_lmc_ops.h"

namespace gjx {

using namespace logistic;

constexpr int MAXD = C64::MAXD;  // the LMC kernel serves D <= 64

template <bool STEP>
__global__ void __launch_bounds__(NT, 2) logistic_lmc_f64_kernel(const LogisticArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C64>& S = *reinterpret_cast<Smem<C64>*>(smem_raw);
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  const int N = a.post_N[post], D = a.post_D[post];
  const double* X = a.X + a.post_xoff[post];
  const double* y = a.y + a.post_yoff[post];
  const int Dp = a.Dmax;
  const double eps = a.eps;
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const int T_steps = STEP ? 1 : a.T;

  if (t < MAXD) {
    const double qv = t < D ? static_cast<const double*>(a.q_in)[(size_t)c * Dp + t] : 0.0;
    S.q0[t] = qv;
    S.q[t] = qv;
    S.p[t] = 0.0;
    S.vt[t] = 0.0;
  }
  double logp = a.logp_in[c];
  double J = a.J_in[c];
  int n_acc = 0;
  double A[4][4];
  __syncthreads();

  auto position_pass = [&](bool chol_draw, double& loglik) __attribute__((always_inline)) -> double {
    return lmc_position_pass(S, A, X, y, N, D, prior, chol_draw, loglik);
  };
  auto half_step = [&](double& dJ) __attribute__((always_inline)) {
    lmc_half_step(S, A, X, y, N, D, prior, eps, dJ);
  };

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

    // ---- velocity draw and initial kinetic energy ----
    double ll0 = 0.0;
    double logdetG = position_pass(true, ll0);
    if (STEP && a.momentum != nullptr && t < D) static_cast<double*>(a.momentum)[(size_t)c * Dp + t] = S.p[t];
    // v0^T G v0 = |L^T v0|^2 = |z|^2
    const double zz = block_sum(S, t < D ? S.z[t] * S.z[t] : 0.0);
    const double K0 = -0.5 * logdetG + 0.5 * zz;

    // ---- trajectory ----
    double dJ = 0.0, ll = ll0;
    const int Ls = chain_steps(a.Lc, a.L, a.C, STEP ? 0 : tr, c);
    for (int s = 0; s < Ls; ++s) {
      half_step(dJ);
      if (t < D) S.q[t] = S.q[t] + eps * S.p[t];
      __syncthreads();
      logdetG = position_pass(false, ll);
      half_step(dJ);
    }

    // ---- end: G v (momentum) and the energy (lmc_energy, metrics.py:209-221) ----
    double dd = 0.0;
    if (t < D) S.u[t] = S.g[t];
    __syncthreads();
    pass_gram(S, X, y, N, D, prior, S.p, 0.0, false, true, dd);  // S.gv = G v
    if (t < D) S.g[t] = S.u[t];
    const double vGv = block_sum(S, t < D ? S.p[t] * S.gv[t] : 0.0);
    const double qq = block_sum(S, t < D ? S.q[t] * S.q[t] : 0.0);
    const double logp1 = ll - 0.5 * qq / prior + lp_const;
    const double K1 = -0.5 * logdetG + 0.5 * vGv;
    const double J1 = J + dJ;
    const double H1 = -logp1 + K1 - J1;
    double dHe = (K0 - logp) - (K1 - logp1) + dJ;  // H0 - H1 with the carried J cancelled exactly
    if (isnan(dHe)) dHe = -INFINITY;
    const double p_acc = fmin(exp(dHe), 1.0);
    const bool accept = uacc < p_acc;

    if (STEP) {
      if (t == 0) {
        if (a.acc_rate) a.acc_rate[c] = p_acc;
        if (a.accepted) a.accepted[c] = (uint8_t)accept;
        if (a.divergent) a.divergent[c] = (uint8_t)(-dHe > a.div_threshold);
        if (a.energy) a.energy[c] = H1;
        if (a.prop_logp) a.prop_logp[c] = logp1;
        if (a.prop_J) a.prop_J[c] = J1;
        if (a.prop_weight) a.prop_weight[c] = dHe;
        if (a.prop_slpa) a.prop_slpa[c] = fmin(dHe, 0.0);
        a.logp_out[c] = accept ? logp1 : logp;
        if (a.J_out) a.J_out[c] = accept ? J1 : J;
      }
      if (t < D) {
        const size_t i = (size_t)c * Dp + t;
        if (a.prop_q) static_cast<double*>(a.prop_q)[i] = S.q[t];
        if (a.prop_p) static_cast<double*>(a.prop_p)[i] = -S.gv[t];
        if (a.prop_v) static_cast<double*>(a.prop_v)[i] = -S.p[t];
        if (a.prop_g) static_cast<double*>(a.prop_g)[i] = S.g[t];
        static_cast<double*>(a.q_out)[i] = accept ? S.q[t] : S.q0[t];
        if (a.g_out)
          static_cast<double*>(a.g_out)[i] = accept ? S.g[t] : static_cast<const double*>(a.g_in)[i];
      }
    }
    if (accept) {
      if (t < MAXD) S.q0[t] = S.q[t];
      logp = logp1;
      J = J1;
      ++n_acc;
    }
    __syncthreads();
    if (!STEP) {
      if (a.samples != nullptr && t < D)
        static_cast<double*>(a.samples)[((size_t)tr * a.C + c) * Dp + t] = S.q0[t];
      if (a.sample_acc != nullptr && t == 0) a.sample_acc[(size_t)tr * a.C + c] = (float)p_acc;
    }
  }

  if (!STEP) {
    if (t < D) static_cast<double*>(a.final_q)[(size_t)c * Dp + t] = S.q0[t];
    if (t == 0) {
      a.final_logp[c] = logp;
      if (a.final_J) a.final_J[c] = J;
      if (a.num_accepted) a.num_accepted[c] = n_acc;
    }
  }
  zero_row_tails<double, STEP>(a, c, D, t, NT);
}

hipError_t launch_logistic_lmc_f64(const LogisticArgs& a, hipStream_t stream) {
  const size_t smem = sizeof(Smem<C64>);
  if (a.q_out != nullptr) {
    (void)hipFuncSetAttribute((const void*)logistic_lmc_f64_kernel<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipLaunchKernelGGL(logistic_lmc_f64_kernel<true>, dim3(a.C), dim3(NT), smem, stream, a);
  } else {
    (void)hipFuncSetAttribute((const void*)logistic_lmc_f64_kernel<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipLaunchKernelGGL(logistic_lmc_f64_kernel<false>, dim3(a.C), dim3(NT), smem, stream, a);
    if (a.final_g != nullptr) {
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      LogisticArgs b = a;
      b.q_in = a.final_q;
      b.logp_out = nullptr;
      b.g_out = a.final_g;
      return launch_logistic_value_and_grad(b, stream);
    }
  }
  return hipGetLastError();
}

}  // namespace gjx
