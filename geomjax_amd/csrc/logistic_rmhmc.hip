// RMHMC (implicit midpoint, fixed-point solve) for Bayesian logistic
// regression with the Fisher metric G(q) = X^T diag(s(1-s)) X + I / prior_var,
// s = sigmoid(X q).  Float64 path: the two dense contractions per evaluation
// of the vector field run on v_mfma_f64_16x16x4f64.
//
// Reference: geomjax/rmhmc/rmhmc.py:131-174,376-462,
// geomjax/rmhmc/integrators.py:53-156, geomjax/rmhmc/metrics.py:42-129
// (kinetic energy -log N(p; 0, G) through a Cholesky, differentiated by
// autodiff), geomjax/mcmc/proposal.py:87-185.  The model and metric are the
// user-supplied logdensity_fn / metric_fn of the logistic configurations.
//
// Closed forms (w_n = s(1-s)(1-2s) = d lambda / d eta, r_n = y_n - s_n):
//   dT/dp = u = G^-1 p
//   dT/dq = 1/2 X^T [ w (a - b^2) ],  a_n = x_n^T G^-1 x_n,  b_n = x_n . u
//   grad l = X^T r - q / prior_var
// so dH/dq = X^T [ w (a - b^2) / 2 - r ] + q / prior_var.
//
// One workgroup (256 lanes, 4 waves) owns one chain for the whole launch, so
// every chain runs its own fixed-point iteration count.  Two instantiations by
// feature capacity: 64 (X in LDS chunks of 32 rows, 10 lower 16x16 Gram tiles,
// 4x4 inverse entries per lane) and 128 (chunks of 16 rows, 36 tiles, 8x8
// entries per lane; G^-1 alone is 129 KiB of LDS).  Per evaluation:
//   pass 1: eta = X q, lambda, and the Gram X^T diag(lambda) X on MFMA
//           (lower tiles owned by waves);
//   sweep : G^-1 by the symmetric sweep operator with G held in registers
//           (one barrier per pivot), log det G = sum of the pivots' logs;
//   pass 2: eta, s, w, r, b = X u on VALU, V = X_chunk G^-1 on MFMA,
//           a_n = x_n . V_n, h_n, and X^T h.
// The momentum draw p0 = chol(G) z comes out of the first sweep of each
// transition: before pivot k is eliminated, row k of the partly swept matrix
// holds column k of the Schur complement, i.e. sqrt(d_k) L[:, k].
#include "common.h"
#include "logistic.h"
#include "logistic_common.h"

namespace gjx {

using namespace logistic;

// MD: feature capacity (64 or 128); chains whose posterior has D outside
// (MD/2, MD] (or D <= 64 for MD = 64) belong to the other instantiation.
template <int MD, bool STEP>
__global__ void __launch_bounds__(NT, MD == 64 ? 2 : 1) logistic_rmhmc_f64_kernel(const LogisticArgs a) {
  using C = Cfg<MD>;
  constexpr int MAXD = C::MAXD, BS = C::BS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C>& S = *reinterpret_cast<Smem<C>*>(smem_raw);
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  if (c >= a.C) return;  // whole workgroups leave together
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;  // invalid posterior index: chain left untouched
  const int N = a.post_N[post], D = a.post_D[post];
  if (D > MD || (MD == 128 && D <= 64)) return;  // served by the other capacity
  const double* X = a.X + a.post_xoff[post];
  const double* y = a.y + a.post_yoff[post];
  const int Dp = a.Dmax;  // row stride of the chain-major state arrays
  const double h = 0.5 * a.eps;
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;
  const int T_steps = STEP ? 1 : a.T;

  // chain state
  if (t < MAXD) {
    const double qv = t < D ? static_cast<const double*>(a.q_in)[(size_t)c * Dp + t] : 0.0;
    S.q0[t] = qv;
    S.q[t] = qv;
    S.qs[t] = 0.0;
  }
  double logp = a.logp_in[c];
  int n_acc = 0;
  int64_t n_fev = 0;
  __syncthreads();

  for (int tr = 0; tr < T_steps; ++tr) {
    // ---- noise ----
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
    // the current position S.q0 is the chain state; S.q = q0
    if (t < MAXD) S.q[t] = S.q0[t];
    __syncthreads();

    // ---- momentum p0 = chol(G(q0)) z and H0 ----
    double dummy = 0.0;
    pass1(S, X, y, N, D, prior, false, dummy);
    double A[BS][BS];
    load_block(S, A);
    __syncthreads();
    // the sweep also draws p0 = chol(G(q0)) z from its Schur-complement columns
    double logdet = sweep(S, A, D, S.z, S.p0);
    __syncthreads();
    if (t < MAXD) S.p[t] = S.p0[t];
    if (STEP && a.momentum != nullptr && t < D) static_cast<double*>(a.momentum)[(size_t)c * Dp + t] = S.p0[t];
    store_neg_block(S, A);
    __syncthreads();
    ginv_times(S, S.p, S.u, D);
    double H0;
    {
      const double pu = block_sum(S, t < D ? S.p[t] * S.u[t] : 0.0);
      H0 = -logp + 0.5 * pu + 0.5 * logdet + ke_const;
    }
    bool have_ginv = true;  // S.Gi = G(S.q)^-1 and S.u = G^-1 S.p

    // one evaluation of the vector field at (S.q, S.p): S.u = dH/dp, S.dH = dH/dq
    auto feval = [&]() __attribute__((always_inline)) {
      if (!have_ginv) {
        double dd = 0.0;
        pass1(S, X, y, N, D, prior, false, dd);
        load_block(S, A);
        __syncthreads();
        sweep(S, A, D);
        __syncthreads();
        store_neg_block(S, A);
        __syncthreads();
        ginv_times(S, S.p, S.u, D);
      }
      pass2(S, X, y, N, D, prior);
      have_ginv = false;
      ++n_fev;
    };

    // ---- trajectory ----
    const int Ls = chain_steps(a.Lc, a.L, a.C, STEP ? 0 : tr, c);
    for (int s = 0; s < Ls; ++s) {
      // z0 of this step is (S.q, S.p), kept in registers of lanes t < D
      double q0v = 0.0, p0v = 0.0;
      if (t < D) {
        q0v = S.q[t];
        p0v = S.p[t];
      }
      // first evaluation outside the loop (integrators.py:84)
      feval();
      double norm;
      {
        double dq = 0.0;
        if (t < D) {
          const double qn = q0v + h * S.u[t];
          const double pn = p0v - h * S.dH[t];
          dq = nan_max(fabs(qn - S.q[t]), fabs(pn - S.p[t]));
          S.q[t] = qn;
          S.p[t] = pn;
        }
        norm = block_max(S, dq);
      }
      int n = 0;
      while (n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol) {
        feval();
        double dq = 0.0;
        if (t < D) {
          const double qn = q0v + h * S.u[t];
          const double pn = p0v - h * S.dH[t];
          dq = nan_max(fabs(qn - S.q[t]), fabs(pn - S.p[t]));
          S.q[t] = qn;
          S.p[t] = pn;
        }
        norm = block_max(S, dq);
        ++n;
      }
      if (STEP && t == 0) {
        const size_t idx = (size_t)c * a.L + s;
        if (a.fp_iters) a.fp_iters[idx] = n;
        if (a.fp_norm) a.fp_norm[idx] = norm;
        if (a.fp_success) a.fp_success[idx] = (uint8_t)(isfinite(norm) && norm <= a.conv_tol);
      }
      // explicit half step from the last iterate (integrators.py:146-148)
      feval();
      if (t < D) {
        S.q[t] = S.q[t] + h * S.u[t];
        S.p[t] = S.p[t] - h * S.dH[t];
      }
      __syncthreads();
    }

    // ---- end: l, grad l, G^-1, velocity and energy at (q_L, p_L) ----
    double loglik = 0.0;
    pass1(S, X, y, N, D, prior, true, loglik);  // S.g = grad l(q_L)
    load_block(S, A);
    __syncthreads();
    logdet = sweep(S, A, D);
    __syncthreads();
    store_neg_block(S, A);
    __syncthreads();
    ginv_times(S, S.p, S.u, D);
    const double qq = block_sum(S, t < D ? S.q[t] * S.q[t] : 0.0);
    const double pu = block_sum(S, t < D ? S.p[t] * S.u[t] : 0.0);
    const double logp1 = loglik - 0.5 * qq / prior + lp_const;
    const double H1 = -logp1 + 0.5 * pu + 0.5 * logdet + ke_const;

    double dHe = H0 - H1;
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
        if (a.prop_weight) a.prop_weight[c] = dHe;
        if (a.prop_slpa) a.prop_slpa[c] = fmin(dHe, 0.0);
        a.logp_out[c] = accept ? logp1 : logp;
      }
      if (t < D) {
        const size_t i = (size_t)c * Dp + t;
        double* qo = static_cast<double*>(a.q_out);
        if (a.prop_q) static_cast<double*>(a.prop_q)[i] = S.q[t];
        if (a.prop_p) static_cast<double*>(a.prop_p)[i] = -S.p[t];
        if (a.prop_v) static_cast<double*>(a.prop_v)[i] = -S.u[t];
        if (a.prop_g) static_cast<double*>(a.prop_g)[i] = S.g[t];
        qo[i] = accept ? S.q[t] : S.q0[t];
        if (a.g_out)
          static_cast<double*>(a.g_out)[i] = accept ? S.g[t] : static_cast<const double*>(a.g_in)[i];
      }
    }
    if (accept) {
      if (t < MAXD) S.q0[t] = S.q[t];
      logp = logp1;
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
      if (a.num_accepted) a.num_accepted[c] = n_acc;
      if (a.num_fevals) a.num_fevals[c] = n_fev;
    }
  }
  zero_row_tails<double, STEP>(a, c, D, t, NT);
}

// value_and_grad of the logistic log density (the reference's init,
// geomjax/rmhmc/rmhmc.py:96-98): one workgroup per chain, pass 1 with the
// likelihood and gradient accumulations.
template <int MD>
__global__ void __launch_bounds__(NT) logistic_value_and_grad_kernel(const LogisticArgs a) {
  using C = Cfg<MD>;
  constexpr int MAXD = C::MAXD;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C>& S = *reinterpret_cast<Smem<C>*>(smem_raw);
  const int t = threadIdx.x, c = blockIdx.x;
  if (c >= a.C) return;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  const int N = a.post_N[post], D = a.post_D[post];
  if (D > MD || (MD == 128 && D <= 64)) return;
  const double* X = a.X + a.post_xoff[post];
  const double* y = a.y + a.post_yoff[post];
  if (t < MAXD) S.q[t] = t < D ? static_cast<const double*>(a.q_in)[(size_t)c * a.Dmax + t] : 0.0;
  __syncthreads();
  double ll = 0.0;
  pass1(S, X, y, N, D, a.prior_var, true, ll);
  const double qq = block_sum(S, t < D ? S.q[t] * S.q[t] : 0.0);
  if (t == 0 && a.logp_out != nullptr)
    a.logp_out[c] = ll - 0.5 * qq / a.prior_var - 0.5 * (double)D * log(2.0 * M_PI * a.prior_var);
  if (t < D) static_cast<double*>(a.g_out)[(size_t)c * a.Dmax + t] = S.g[t];
  for (int j = D + t; j < a.Dmax; j += NT) static_cast<double*>(a.g_out)[(size_t)c * a.Dmax + j] = 0.0;
}

template <int MD>
static void launch_vg(const LogisticArgs& a, hipStream_t stream) {
  const size_t smem = sizeof(Smem<Cfg<MD>>);
  (void)hipFuncSetAttribute((const void*)logistic_value_and_grad_kernel<MD>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL(logistic_value_and_grad_kernel<MD>, dim3(a.C), dim3(NT), smem, stream, a);
}

// max_features <= 64: the 64-capacity kernels only; otherwise both (each
// chain runs in the one whose capacity fits its posterior).
hipError_t launch_logistic_value_and_grad(const LogisticArgs& a, hipStream_t stream) {
  launch_vg<64>(a, stream);
  if (a.Dmax > 64 && a.max_features > 64) launch_vg<128>(a, stream);
  return hipGetLastError();
}

template <int MD, bool STEP>
static void launch_rmhmc(const LogisticArgs& a, hipStream_t stream) {
  const size_t smem = sizeof(Smem<Cfg<MD>>);
  (void)hipFuncSetAttribute((const void*)logistic_rmhmc_f64_kernel<MD, STEP>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL((logistic_rmhmc_f64_kernel<MD, STEP>), dim3(a.C), dim3(NT), smem, stream, a);
}

hipError_t launch_logistic_rmhmc_f64(const LogisticArgs& a, hipStream_t stream) {
  const bool wide = a.max_features > 64;
  const bool narrow = a.P > 1 || !wide;
  if (a.q_out != nullptr) {
    if (narrow) launch_rmhmc<64, true>(a, stream);
    if (wide) launch_rmhmc<128, true>(a, stream);
  } else {
    if (narrow) launch_rmhmc<64, false>(a, stream);
    if (wide) launch_rmhmc<128, false>(a, stream);
    if (a.final_g != nullptr) {
      // gradient at the final positions: value_and_grad on them (log density already carried)
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
