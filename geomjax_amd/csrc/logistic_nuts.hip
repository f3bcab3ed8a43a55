// NUTS with the RMHMC implicit-midpoint integrator for Bayesian logistic
// regression with the Fisher metric, float64, D <= 64.
//
// Reference: geomjax/rmhmc/nuts.py:78-161,253-348 (kernel, iterative proposal),
// geomjax/mcmc/trajectory.py:149-317 (a subtree of up to 2^e steps, progressive
// uniform sampling, stop at divergence or an iterative U-turn), :517-668
// (random direction, biased merge, U-turn of the merged trajectory),
// geomjax/mcmc/termination.py:30-119 (numpyro checkpoints),
// geomjax/rmhmc/metrics.py:100-117 (U-turn criteria), and the integrator and
// energy of logistic_rmhmc.hip (geomjax/rmhmc/integrators.py:53-156,
// geomjax/rmhmc/metrics.py:42-129).
//
// One workgroup (256 lanes) per chain, the float64 RMHMC kernel's LDS layout
// (capacity 64) and passes.  Lane t < D holds coordinate t of the trajectory's
// two ends (position, momentum, velocity), of the momentum and velocity sums of
// the trajectory and of the subtree being built, and of the two proposals.  The
// state being integrated lives in LDS (S.q, S.p); after each step its log
// density, gradient, G^-1, velocity and energy are formed once, and they also
// serve as the first evaluation of the next step from that end.  U-turn
// checkpoints ([NA][M][64] doubles) follow the Smem block in dynamic LDS.
#include "common.h"
#include "logistic.h"
#include "logistic_common.h"
#include "nuts.h"

namespace gjx {

using namespace logistic;
using nuts::logaddexp;

__global__ void __launch_bounds__(NT, 1) logistic_nuts_f64_kernel(const LogisticArgs a, const NutsArgs nu) {
  using C = Cfg<64>;
  constexpr int MAXD = C::MAXD, BS = C::BS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C>& S = *reinterpret_cast<Smem<C>*>(smem_raw);
  double* const ck = reinterpret_cast<double*>(smem_raw + sizeof(Smem<C>));
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  if (c >= a.C) return;  // whole workgroups leave together
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  const int N = a.post_N[post], D = a.post_D[post];
  if (D > MAXD) return;  // rejected by the entry point
  const double* X = a.X + a.post_xoff[post];
  const double* y = a.y + a.post_yoff[post];
  const int Dp = a.Dmax;
  const bool own = t < D;  // this lane holds coordinate t
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;
  const int M = nu.max_doublings;
  const int NS = (1 << M) - 1;
  const int crit = nu.criterion;
  const int NA = nuts_checkpoint_arrays(crit);
  const uint64_t tt = a.counter;
  const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
  auto uniform = [&](const double* host, size_t idx, uint32_t stream, uint32_t i) -> double {
    if (host != nullptr) return host[idx];
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, stream + i, (uint32_t)(tt >> 32)}, k0, k1);
    return u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
  };

  for (int i = t; i < NA * M * MAXD; i += NT) ck[i] = 0.0;
  // ---- state and noise ----
  if (t < MAXD) {
    const double qv = own ? static_cast<const double*>(a.q_in)[(size_t)c * Dp + t] : 0.0;
    S.q[t] = qv;
    S.q0[t] = qv;
  }
  if (a.z != nullptr) {
    if (t < MAXD) S.z[t] = own ? static_cast<const double*>(a.z)[(size_t)c * Dp + t] : 0.0;
  } else if (t < MAXD / 2) {
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * t),
                                  (uint32_t)(tt >> 32)}, k0, k1);
    double n0, n1;
    normal2(r, n0, n1);
    S.z[2 * t] = 2 * t < D ? n0 : 0.0;
    S.z[2 * t + 1] = 2 * t + 1 < D ? n1 : 0.0;
  }
  const double logp0 = a.logp_in[c];
  __syncthreads();

  // ---- momentum p0 = chol(G(q0)) z, velocity and H0 (as logistic_rmhmc.hip) ----
  double dummy = 0.0;
  pass1(S, X, y, N, D, prior, false, dummy);
  double A[BS][BS];
  load_block(S, A);
  __syncthreads();
  double logdet = sweep(S, A, D, S.z, S.p0);
  __syncthreads();
  if (t < MAXD) S.p[t] = S.p0[t];
  if (a.momentum != nullptr && own) static_cast<double*>(a.momentum)[(size_t)c * Dp + t] = S.p0[t];
  store_neg_block(S, A);
  __syncthreads();
  ginv_times(S, S.p, S.u, D);
  const double H0 = -logp0 + 0.5 * block_sum(S, own ? S.p[t] * S.u[t] : 0.0) + 0.5 * logdet + ke_const;
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
  };
  // one implicit-midpoint step of (signed) size hstep from (S.q, S.p), then the new
  // state's log density, gradient (S.g), G^-1, velocity (S.u) and energy
  auto integrate = [&](double hstep, double& logp1) -> double {
    const double h = 0.5 * hstep;
    const double q0v = t < MAXD ? S.q[t] : 0.0, p0v = t < MAXD ? S.p[t] : 0.0;
    feval();  // first evaluation outside the loop (integrators.py:84)
    double norm;
    {
      double dq = 0.0;
      if (own) {
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
      if (own) {
        const double qn = q0v + h * S.u[t];
        const double pn = p0v - h * S.dH[t];
        dq = nan_max(fabs(qn - S.q[t]), fabs(pn - S.p[t]));
        S.q[t] = qn;
        S.p[t] = pn;
      }
      norm = block_max(S, dq);
      ++n;
    }
    // explicit half step from the last iterate (integrators.py:146-148)
    feval();
    if (own) {
      S.q[t] = S.q[t] + h * S.u[t];
      S.p[t] = S.p[t] - h * S.dH[t];
    }
    __syncthreads();
    double loglik = 0.0;
    pass1(S, X, y, N, D, prior, true, loglik);
    load_block(S, A);
    __syncthreads();
    const double ld = sweep(S, A, D);
    __syncthreads();
    store_neg_block(S, A);
    __syncthreads();
    ginv_times(S, S.p, S.u, D);
    have_ginv = true;
    const double qq = block_sum(S, own ? S.q[t] * S.q[t] : 0.0);
    const double pu = block_sum(S, own ? S.p[t] * S.u[t] : 0.0);
    logp1 = loglik - 0.5 * qq / prior + lp_const;
    return -logp1 + 0.5 * pu + 0.5 * ld + ke_const;
  };
  // U-turn criteria (geomjax/rmhmc/metrics.py:100-117) on the two ends' momenta and
  // velocities and the sums between them (one coordinate per lane)
  auto is_turning = [&](double ml, double mr, double ms, double vl, double vr, double vs) -> bool {
    if (crit == 1) return block_sum(S, vl * ms) <= 0.0 || block_sum(S, vr * ms) <= 0.0;  // "bet"
    if (crit == 2) return block_sum(S, ml * vs) <= 0.0 || block_sum(S, mr * vs) <= 0.0;  // "riem"
    return block_sum(S, vl * vs) <= 0.0 || block_sum(S, vr * vs) <= 0.0;                 // "euc"
  };
  auto ck_at = [&](int arr, int idx) -> double& { return ck[(arr * M + idx) * MAXD + t]; };

  // ---- the trajectory: two ends, sums (the start state included), length ----
  const double q_start = t < MAXD ? S.q[t] : 0.0;
  double lq = q_start, lp = t < MAXD ? S.p[t] : 0.0, lu = t < MAXD ? S.u[t] : 0.0;
  double rq = lq, rp = lp, ru = lu;
  double psum = lp, vsum = lu;
  int ntraj = 0;
  // proposal of the transition: position, log density, energy, weight, accumulated acceptance
  double prop_q = q_start, prop_logp = logp0, prop_energy = H0, prop_w = 0.0, prop_slpa = -INFINITY;
  int at_end = 0;  // which end (S.q, S.p) hold: 0 the start (both ends), +1 right, -1 left
  int e = 0, j = 0;
  bool diverging = false, turning = false;
  while (e < M && !diverging && !turning) {
    const double ud = uniform(nu.u_dir, (size_t)c * M + e, kStreamNutsDirection, (uint32_t)e);
    const int dir = ud < 0.5 ? 1 : -1;
    if (at_end != 0 && at_end != dir) {  // LDS holds the other end: reload this one
      if (t < MAXD) {
        S.q[t] = dir > 0 ? rq : lq;
        S.p[t] = dir > 0 ? rp : lp;
      }
      have_ginv = false;
      __syncthreads();
    }
    // ---- subtree of up to 2^e steps from that end (trajectory.py:149-317) ----
    double sp_q = 0.0, sp_logp = 0.0, sp_energy = 0.0, sp_w = 0.0, sp_slpa = 0.0;
    double ssp = 0.0, ssv = 0.0;
    int nsub = 0, step = 0;
    const int max_steps = 1 << e;
    bool div_s = false, term_s = false;
    while (step < max_steps && !term_s && !div_s) {
      double logp1;
      const double en = integrate(dir * a.eps, logp1);
      double w = H0 - en;
      if (isnan(w)) w = -INFINITY;
      div_s = -w > a.div_threshold;
      const double us = uniform(nu.u_step, (size_t)c * NS + j, kStreamNutsStep, (uint32_t)j);
      const double qv = own ? S.q[t] : 0.0, pv = own ? S.p[t] : 0.0, uv = own ? S.u[t] : 0.0;
      if (step == 0) {
        ssp = pv;
        ssv = uv;
        nsub = 1;
        sp_q = qv;
        sp_logp = logp1;
        sp_energy = en;
        sp_w = w;
        sp_slpa = fmin(w, 0.0);
      } else {
        ssp += pv;
        ssv += uv;
        ++nsub;
        // progressive uniform sampling (proposal.py:198-222)
        const double pacc = 1.0 / (1.0 + exp(-(w - sp_w)));
        const double nw = logaddexp(sp_w, w), ns = logaddexp(sp_slpa, fmin(w, 0.0));
        if (us < pacc) {
          sp_q = qv;
          sp_logp = logp1;
          sp_energy = en;
        }
        sp_w = nw;
        sp_slpa = ns;
      }
      ++j;
      // iterative U-turn check (termination.py:38-117)
      const int idx_max = nuts::ckpt_idx_max(step), idx_min = nuts::ckpt_idx_min(step);
      if ((step & 1) == 0 && t < MAXD) {
        ck_at(0, idx_max) = uv;
        ck_at(1, idx_max) = ssv;
        if (NA == 4) {
          ck_at(2, idx_max) = pv;
          ck_at(3, idx_max) = ssp;
        }
      }
      __syncthreads();
      for (int i = idx_max; i >= idx_min && !term_s; --i) {
        const double vl = t < MAXD ? ck_at(0, i) : 0.0;
        const double vs = t < MAXD ? ssv - ck_at(1, i) + vl : 0.0;
        const double ml = (NA == 4 && t < MAXD) ? ck_at(2, i) : 0.0;
        const double ms = (NA == 4 && t < MAXD) ? ssp - ck_at(3, i) + ml : 0.0;
        term_s = is_turning(ml, pv, ms, vl, uv, vs);
      }
      ++step;
    }
    // ---- merge (trajectory.py:580-660) ----
    if (div_s || term_s) {
      prop_slpa = logaddexp(prop_slpa, sp_slpa);
    } else {
      // progressive biased sampling (proposal.py:225-254)
      const double um = uniform(nu.u_merge, (size_t)c * M + e, kStreamNutsMerge, (uint32_t)e);
      const double pacc = fmin(exp(sp_w - prop_w), 1.0);
      const double nw = logaddexp(prop_w, sp_w), ns = logaddexp(prop_slpa, sp_slpa);
      if (um < pacc) {
        prop_q = sp_q;
        prop_logp = sp_logp;
        prop_energy = sp_energy;
      }
      prop_w = nw;
      prop_slpa = ns;
    }
    const double eq = own ? S.q[t] : 0.0, ep = own ? S.p[t] : 0.0, eu = own ? S.u[t] : 0.0;
    if (dir > 0) {
      rq = eq;
      rp = ep;
      ru = eu;
    } else {
      lq = eq;
      lp = ep;
      lu = eu;
    }
    at_end = dir;
    psum += ssp;
    vsum += ssv;
    ntraj += nsub;
    const bool turn_full = is_turning(lp, rp, psum, lu, ru, vsum);
    ++e;
    diverging = div_s;
    turning = term_s || turn_full;
  }

  // ---- outputs: the proposal's state (its gradient recomputed) and NUTSInfo ----
  __syncthreads();
  if (t < MAXD) S.q[t] = own ? prop_q : 0.0;
  __syncthreads();
  double ll_end = 0.0;
  pass1(S, X, y, N, D, prior, true, ll_end);  // S.g = grad l(prop_q)
  const size_t row = (size_t)c * Dp;
  if (own) {
    static_cast<double*>(a.q_out)[row + t] = prop_q;
    if (a.g_out) static_cast<double*>(a.g_out)[row + t] = S.g[t];
    if (nu.left_position) static_cast<double*>(nu.left_position)[row + t] = lq;
    if (nu.right_position) static_cast<double*>(nu.right_position)[row + t] = rq;
  }
  for (int k = D + t; k < Dp; k += NT) {  // row tails past this posterior's D
    static_cast<double*>(a.q_out)[row + k] = 0.0;
    if (a.g_out) static_cast<double*>(a.g_out)[row + k] = 0.0;
    if (a.momentum) static_cast<double*>(a.momentum)[row + k] = 0.0;
    if (nu.left_position) static_cast<double*>(nu.left_position)[row + k] = 0.0;
    if (nu.right_position) static_cast<double*>(nu.right_position)[row + k] = 0.0;
  }
  if (t == 0) {
    a.logp_out[c] = prop_logp;
    if (nu.energy) nu.energy[c] = prop_energy;
    if (nu.acceptance_rate) nu.acceptance_rate[c] = exp(prop_slpa) / (double)ntraj;
    if (nu.is_divergent) nu.is_divergent[c] = (uint8_t)diverging;
    if (nu.is_turning) nu.is_turning[c] = (uint8_t)turning;
    if (nu.num_expansions) nu.num_expansions[c] = e;
    if (nu.num_steps) nu.num_steps[c] = ntraj;
  }
}

size_t logistic_nuts_smem_bytes(int max_doublings, int criterion) {
  return sizeof(Smem<Cfg<64>>) + (size_t)nuts_checkpoint_arrays(criterion) * max_doublings * 64 * sizeof(double);
}

hipError_t launch_logistic_nuts_f64(const LogisticArgs& a, const NutsArgs& n, hipStream_t stream) {
  const size_t smem = logistic_nuts_smem_bytes(n.max_doublings, n.criterion);
  hipError_t e = hipFuncSetAttribute((const void*)logistic_nuts_f64_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(logistic_nuts_f64_kernel, dim3(a.C), dim3(NT), smem, stream, a, n);
  return hipGetLastError();
}

}  // namespace gjx
