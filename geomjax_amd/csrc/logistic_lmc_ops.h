// Device building blocks of the general-metric LMC integrator (lmcmc) for
// Bayesian logistic regression with the Fisher metric, float64, capacity 64:
// the position pass and the half step of lan_integrator, shared by the LMC
// transition kernel (logistic_lmc.hip) and NUTS (logistic_nuts.hip).  One
// 256-lane workgroup per chain on the Smem layout of logistic_common.h; the
// algebra is derived at the top of logistic_lmc.hip.
//
// Reference: geomjax/lmcmc/integrators.py:51-144, geomjax/lmcmc/metrics.py:42-221.
#pragma once

#include "logistic_common.h"

namespace gjx {
namespace logistic {

using C64 = Cfg<64>;

// Position quantities at S.q: S.Gi = G^-1, S.g = grad l, S.dH = grad log det G;
// returns log det G and (through loglik) the log-likelihood sum.  With chol_draw
// the velocity v0 = L^-T S.z (L = chol(G), metrics.py:75-91) goes to S.p first.
__device__ __forceinline__ double lmc_position_pass(Smem<C64>& S, double (&A)[4][4], const double* X, const double* y,
                                                    int N, int D, double prior, bool chol_draw, double& loglik) {
  constexpr int MAXD = C64::MAXD, LDX = C64::LDX;
  const int t = threadIdx.x;
  pass_gram(S, X, y, N, D, prior, nullptr, 0.0, true, false, loglik);
  load_block(S, A);
  __syncthreads();
  if (chol_draw) {
    cholesky(S, D);
    if (t < MAXD) S.vt[t] = t < D ? S.z[t] : 0.0;
    __syncthreads();
    for (int i = D - 1; i >= 0; --i) {
      const double vi = S.vt[i] / S.Gi[i * LDX + i];
      if (t < i) S.vt[t] -= S.Gi[i * LDX + t] * vi;
      if (t == i) S.p[i] = vi;
      __syncthreads();
    }
  }
  const double ld = sweep(S, A, D);
  __syncthreads();
  store_neg_block(S, A);
  __syncthreads();
  pass_quad(S, X, y, N, D, prior, false);  // S.dH = X^T (w a)
  return ld;
}

// One half step of lan_integrator (integrators.py:61-91) of (signed) size eps / 2
// with the velocity in S.p; accumulates the volume change in dJ.  S.g and S.dH
// (grad l and grad log det G at S.q) are kept; S.gv ends as G v of the input
// velocity and S.p as the new velocity.
__device__ __forceinline__ void lmc_half_step(Smem<C64>& S, double (&A)[4][4], const double* X, const double* y, int N,
                                              int D, double prior, double eps, double& dJ) {
  const int t = threadIdx.x;
  double dd = 0.0;
  // S.vt = grad log det G and S.u = grad l survive the two Gram passes below
  if (t < D) {
    S.u[t] = S.g[t];
    S.vt[t] = S.dH[t];
  }
  __syncthreads();
  // A+ = G + (eps/2) Omega~(v) and G v
  pass_gram(S, X, y, N, D, prior, S.p, 0.25 * eps, false, true, dd);
  load_block(S, A);
  __syncthreads();
  dJ -= sweep(S, A, D);
  __syncthreads();
  store_neg_block(S, A);  // S.Gi = A+^-1
  // v_temp = G v - (eps/2) (-g + grad log det G / 2)
  if (t < D) S.qs[t] = S.gv[t] - 0.5 * eps * (-S.u[t] + 0.5 * S.vt[t]);
  __syncthreads();
  ginv_times(S, S.qs, S.p, D);  // v' = A+^-1 v_temp
  // A- = G - (eps/2) Omega~(v')
  pass_gram(S, X, y, N, D, prior, S.p, -0.25 * eps, false, false, dd);
  load_block(S, A);
  __syncthreads();
  dJ += sweep(S, A, D);
  __syncthreads();
  if (t < D) {
    S.g[t] = S.u[t];
    S.dH[t] = S.vt[t];
  }
  __syncthreads();
}

}  // namespace logistic
}  // namespace gjx
