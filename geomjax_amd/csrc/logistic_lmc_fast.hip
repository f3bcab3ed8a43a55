// Lagrangian Monte Carlo with a general metric (lmcmc) for Bayesian logistic
// regression with the Fisher metric, tuned form: float32 state, every Gram
// matrix on split-bf16 MFMA with float32 accumulation, float32 sweeps with one
// refinement step per solve, float64 log-determinant sums, volume change,
// energies and accept test.  D <= 64, one wave per chain, the LDS layout and
// passes of the tuned RMHMC kernel (logistic_fast_ops.h).
//
// Reference: geomjax/lmcmc/lmc.py:114-182,411-499 (transition, proposal,
// flip), geomjax/lmcmc/integrators.py:51-144 (lan_integrator: half steps with
// factorisations of G +- (eps/2) Omega~), geomjax/lmcmc/metrics.py:42-221
// (velocity draw L^-T z, kinetic energy -1/2 log det G + 1/2 v^T G v, Omega~,
// grad log det G).
//
// The algebra is that of the float64 kernel (logistic_lmc.hip): with
// w = lambda (1 - 2 s),
//   G +- (eps/2) Omega~(v) = X^T diag(lambda +- (eps/4) w (X v)) X + I/prior_var,
//   v_temp = G v - (eps/2) grad phi,  phi = -l + 1/2 log det G,
//   grad phi = X^T (w a / 2 - (y - s)) + q / prior_var,  a_n = x_n^T G^-1 x_n,
// where grad phi is the tuned RMHMC vector field's dH/dq at u = 0.  The
// weights lambda +- (eps/4) w (X v) can be negative: their Gram passes use the
// signed-weight form.  A transition runs a flat loop of phases, each one Gram
// pass and one sweep (one inlined copy of each in the kernel):
//   draw:  G(q0), its sweep with the rows of L = chol(G) kept, v0 = L^-T z,
//          G^-1 and grad phi (field pass);
//   per integration step s:
//     half A:  row pass (X v, weights, G v), A+ = G + (eps/2) Omega~(v), sweep
//              (-log|det A+| into the volume change), v' = A+^-1 v_temp;
//     half B:  row pass (X v', weights; on the first half step also the
//              position update eta += eps X v' and sqrt(lambda) there),
//              A- = G - (eps/2) Omega~(v'), sweep (+log|det A-|);
//     position (after the first half step): q += eps v', G(q), sweep, G^-1,
//              grad phi;  then the second half step (A, B);
//   end:   eta, log likelihood and gradient from X in float64, G v, energy.
// Per-chain scratch rows (float32, global): eta at the current position,
// sqrt(lambda) there, and the signed square roots of the half steps' weights.
#include <type_traits>

#include "logistic_fast_ops.h"

namespace gjx {
namespace lfast {

enum LmcPhase { kLmcDraw, kLmcHalfA, kLmcHalfB, kLmcPos, kLmcEnd };

template <int NB, bool STEP>
__device__ __forceinline__ void run_lmc_chain(const LogisticArgs& a, Smem<1>& S, int c, int post) {
  const int l = threadIdx.x % 64;
  const PackedPost pp = a.packed_hdr[post];
  const int N = a.post_N[post], D = a.post_D[post];
  const float* Xg = a.packed + pp.xg;
  const float* Xr = a.packed + pp.xr;
  const float* Xs = a.packed + pp.xs;
  const float* yf = a.packed + pp.yf;
  const int npad = pp.npad;
  const int64_t ns = a.work_rows;
  float* const W = a.work + (size_t)c * 5 * ns;
  float* const ETA = W;
  float* const SL = W + ns;
  float* const SS = W + 2 * ns;
  const int Dp = a.Dmax;
  const float eps = (float)a.eps;
  const float inv_prior = (float)(1.0 / a.prior_var);
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const int T_steps = STEP ? 1 : a.T;

  const float q0v = l < D ? static_cast<const float*>(a.q_in)[(size_t)c * Dp + l] : 0.0f;
  S.q0[l] = q0v;
  double logp = a.logp_in[c];
  double J = a.J_in[c];
  int n_acc = 0;
  __syncthreads();

  for (int tr = 0; tr < T_steps; ++tr) {
    const uint64_t tt = a.counter + (uint64_t)tr;
    // ---- chain state: eta, sqrt(lambda) at q0 (and its gradient, for final_g) ----
    S.q[l] = S.q0[l];
    __syncthreads();
    const bool grad0 = !STEP && tr == 0 && a.final_g != nullptr;
    (void)fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, ETA, SL, grad0, inv_prior);
    if (grad0 && l < D) static_cast<float*>(a.final_g)[(size_t)c * Dp + l] = S.g[l];
    // ---- noise (the float64 kernel's streams) ----
    double uacc;
    if (a.z != nullptr) {
      S.z[l] = l < D ? static_cast<const float*>(a.z)[((size_t)tr * a.C + c) * Dp + l] : 0.0f;
      uacc = a.u[(size_t)tr * a.C + c];
    } else {
      if (l < MAXD / 2) {
        u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * l),
                                      (uint32_t)(tt >> 32)}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        double n0, n1;
        normal2(r, n0, n1);
        S.z[2 * l] = 2 * l < D ? (float)n0 : 0.0f;
        S.z[2 * l + 1] = 2 * l + 1 < D ? (float)n1 : 0.0f;
      }
      u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamAccept, (uint32_t)(tt >> 32)},
                              (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
      uacc = u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
    }
    __syncthreads();

    const int Ls = chain_steps(a.Lc, a.L, a.C, STEP ? 0 : tr, c);
    double logdetG = 0.0, K0 = 0.0, dJ = 0.0;
    int phase = kLmcDraw, s = 0;
    bool second = false;  // the half steps after the position update
    while (phase != kLmcEnd) {
      // ---- the row pass before this phase's matrix ----
      const float* wts = SL;
      if (phase == kLmcHalfA) {
        lmc_row_pass<NB>(S, Xr, ETA, SL, SS, npad, D, 0.25f * eps, eps, true, false, inv_prior);
        wts = SS;
      } else if (phase == kLmcHalfB) {
        lmc_row_pass<NB>(S, Xr, ETA, SL, SS, npad, D, -0.25f * eps, eps, false, !second, inv_prior);
        wts = SS;
      } else if (phase == kLmcPos) {
        S.q[l] = l < D ? fmaf(eps, S.p[l], S.q[l]) : 0.0f;
        __syncthreads();
      }
      // ---- the phase's matrix: Gram pass and sweep ----
      gram_pass<NB, 1, true>(S, Xg, wts, npad, D, inv_prior);
      f32x32 A[NB];
      float p0j;
      sweep<NB, 1>(S, A, D, false, p0j, phase == kLmcDraw);
      const double lad = logabsdet_from_pivots(S, D);
      if (phase == kLmcHalfA) {
        // v_temp = G v - (eps/2) grad phi;  v' = A+^-1 v_temp
        dJ -= lad;
        S.p0[l] = l < D ? fmaf(-0.5f * eps, S.dH[l], S.u[l]) : 0.0f;
        __syncthreads();
        solve_refined<NB, 1>(S, A, S.p0, S.p, S.u);
        phase = kLmcHalfB;
        continue;
      }
      if (phase == kLmcHalfB) {
        dJ += lad;
        if (!second) {
          phase = kLmcPos;
        } else {
          second = false;
          phase = ++s < Ls ? kLmcHalfA : kLmcEnd;
        }
        continue;
      }
      // ---- G at the position: log det G, G^-1, grad phi ----
      logdetG = lad;
      if (phase == kLmcDraw) {
        // v0 = L^-T z (metrics.py:75-91): back substitution over the rows of L
        float r = S.z[l], vj = 0.0f;
        for (int i = D - 1; i >= 0; --i) {
          const float vi = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, r), i)) /
                           S.G[i * LDG + i];
          if (l == i) vj = vi;
          if (l < i) r = fmaf(-S.G[i * LDG + l], vi, r);
        }
        S.p[l] = vj;
        const double zz = group_sum<64>(l < D ? (double)S.z[l] * (double)S.z[l] : 0.0);
        K0 = -0.5 * logdetG + 0.5 * zz;  // v0^T G v0 = |L^T v0|^2 = |z|^2
        if (STEP && a.momentum != nullptr && l < D) static_cast<float*>(a.momentum)[(size_t)c * Dp + l] = vj;
      }
      store_ginv<NB, 1>(S, A);
      S.u[l] = 0.0f;
      __syncthreads();
      field_pass<NB, 1>(S, Xr, Xs, yf, ETA, ETA, SL, npad, D, 0.0f, inv_prior, true);
      if (phase == kLmcDraw) {
        phase = Ls > 0 ? kLmcHalfA : kLmcEnd;
      } else {
        second = true;
        phase = kLmcHalfA;
      }
    }

    // ---- end: eta, log likelihood and gradient in float64, G v, energy (metrics.py:209-221) ----
    const double loglik = fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, ETA, SL, true, inv_prior);
    lmc_row_pass<NB>(S, Xr, ETA, SL, SS, npad, D, 0.0f, eps, true, false, inv_prior);  // S.u = G v
    const double vGv = group_sum<64>(l < D ? (double)S.p[l] * (double)S.u[l] : 0.0);
    const double qq = group_sum<64>(l < D ? (double)S.q[l] * (double)S.q[l] : 0.0);
    const double logp1 = loglik - 0.5 * qq / prior + lp_const;
    const double K1 = -0.5 * logdetG + 0.5 * vGv;
    const double J1 = J + dJ;
    const double H1 = -logp1 + K1 - J1;
    double dHe = (K0 - logp) - (K1 - logp1) + dJ;  // H0 - H1 with the carried J cancelled exactly
    if (isnan(dHe)) dHe = -INFINITY;
    const double p_acc = fmin(exp(dHe), 1.0);
    const bool accept = uacc < p_acc;

    if (STEP) {
      if (l == 0) {
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
      if (l < D) {
        const size_t i = (size_t)c * Dp + l;
        if (a.prop_q) static_cast<float*>(a.prop_q)[i] = S.q[l];
        if (a.prop_p) static_cast<float*>(a.prop_p)[i] = -S.u[l];
        if (a.prop_v) static_cast<float*>(a.prop_v)[i] = -S.p[l];
        if (a.prop_g) static_cast<float*>(a.prop_g)[i] = S.g[l];
        static_cast<float*>(a.q_out)[i] = accept ? S.q[l] : S.q0[l];
        if (a.g_out) static_cast<float*>(a.g_out)[i] = accept ? S.g[l] : static_cast<const float*>(a.g_in)[i];
      }
    }
    if (accept) {
      S.q0[l] = S.q[l];
      if (!STEP && a.final_g != nullptr && l < D) static_cast<float*>(a.final_g)[(size_t)c * Dp + l] = S.g[l];
      logp = logp1;
      J = J1;
      ++n_acc;
    }
    __syncthreads();
    if (!STEP) {
      if (a.samples != nullptr && l < D) static_cast<float*>(a.samples)[((size_t)tr * a.C + c) * Dp + l] = S.q0[l];
      if (a.sample_acc != nullptr && l == 0) a.sample_acc[(size_t)tr * a.C + c] = (float)p_acc;
    }
  }

  if (!STEP) {
    if (l < D) static_cast<float*>(a.final_q)[(size_t)c * Dp + l] = S.q0[l];
    if (l == 0) {
      a.final_logp[c] = logp;
      if (a.final_J) a.final_J[c] = J;
      if (a.num_accepted) a.num_accepted[c] = n_acc;
    }
  }
  zero_row_tails<float, STEP>(a, c, D, l, 64);
}

// One wave per chain, chains dealt to the XCDs in contiguous ranges (as the
// tuned RMHMC kernel); chains of the other tile width leave at once.
template <int NB, bool STEP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
logistic_lmc_f32_kernel(const LogisticArgs a) {
  __shared__ Smem<1> S;
  const int per_xcd = gridDim.x / 8;
  const int slot = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (slot >= a.C) return;
  const int c = a.dispatch_order ? a.dispatch_order[slot] : slot;
  if (c < 0 || c >= a.C) return;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  if (a.packed_hdr[post].nb != NB) return;
  run_lmc_chain<NB, STEP>(a, S, c, post);
}

}  // namespace lfast

hipError_t launch_logistic_lmc_f32(const LogisticArgs& a, int max_features, hipStream_t stream) {
  const int blocks = (a.C + 7) / 8 * 8;
  const bool nb1 = max_features <= 32 || a.P > 1;
  const bool nb2 = max_features > 32;
  const bool step = a.q_out != nullptr;
  if (nb2) {
    if (step)
      hipLaunchKernelGGL((lfast::logistic_lmc_f32_kernel<2, true>), dim3(blocks), dim3(64), 0, stream, a);
    else
      hipLaunchKernelGGL((lfast::logistic_lmc_f32_kernel<2, false>), dim3(blocks), dim3(64), 0, stream, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (nb1) {
    if (step)
      hipLaunchKernelGGL((lfast::logistic_lmc_f32_kernel<1, true>), dim3(blocks), dim3(64), 0, stream, a);
    else
      hipLaunchKernelGGL((lfast::logistic_lmc_f32_kernel<1, false>), dim3(blocks), dim3(64), 0, stream, a);
  }
  return hipGetLastError();
}

}  // namespace gjx
