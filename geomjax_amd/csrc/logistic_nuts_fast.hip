// NUTS with the RMHMC implicit-midpoint integrator for Bayesian logistic
// regression with the Fisher metric, tuned form: float32 state and trajectory,
// Gram and quadratic-form contractions on split-bf16 MFMA, float32 sweeps with
// refined solves; log densities, log-determinant sums, energies, tree weights
// and the U-turn dot products and sums in float64.  D <= 64, one wave per
// chain.
//
// Reference: geomjax/rmhmc/nuts.py:78-161,253-348 (kernel, iterative
// proposal), geomjax/mcmc/trajectory.py:149-317 (a subtree of up to 2^e steps,
// progressive uniform sampling, stop at divergence or an iterative U-turn),
// :517-668 (random direction, biased merge, U-turn of the merged trajectory),
// geomjax/mcmc/termination.py:30-119 (numpyro checkpoints),
// geomjax/rmhmc/metrics.py:100-117 (U-turn criteria); the integrator and
// energy of the tuned RMHMC kernel (geomjax/rmhmc/integrators.py:53-156,
// geomjax/rmhmc/metrics.py:42-129).
//
// The tree building is that of the float64 logistic NUTS kernel
// (logistic_nuts.hip) on one coordinate per lane; the step is the tuned
// RMHMC kernel's implicit midpoint (logistic_fast_ops.h): a flat loop over
// the metric (Gram pass + sweep), the vector field (field pass, which also
// advances the per-row eta scratch of the next iterate) and, at the new
// state, eta, the log likelihood (float64, from X) and the metric again.
// What a step leaves (G^-1 and the velocity at the new state) serves the next
// step from that end; a switch of direction reloads the other end and
// recomputes its eta rows.  Checkpoints ([NA][M][64] doubles) sit in dynamic
// LDS after the chain's 20 KiB block.
#include "logistic_fast_ops.h"
#include "nuts.h"

namespace gjx {
namespace lfast {

using nuts::logaddexp;

// coordinate l of a trajectory state (zero past D): position, momentum, velocity
// and (LMC) grad(-l + log det G / 2); its log density and (LMC) volume adjustment
struct NPt {
  float q, p, u, dphi;
  double logp, J;
};

struct NProp {
  float q;
  double logp, J, energy, weight, slpa;
};

struct NTree {
  int expansions, num_states;
  bool diverging, turning;
};

__device__ __forceinline__ NProp nprop(const NPt& s, double en, double w) {
  return NProp{s.q, s.logp, s.J, en, w, fmin(w, 0.0)};
}

// The transition's uniforms: host-fed or the device Philox stream of the
// purpose tag (the float64 NUTS kernels' streams).
struct NUniform {
  int c;
  uint64_t tt;
  uint32_t k0, k1;
  __device__ double operator()(const double* host, size_t idx, uint32_t stream, uint32_t i) const {
    if (host != nullptr) return host[idx];
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, stream + i, (uint32_t)(tt >> 32)}, k0, k1);
    return u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
  }
};

// The iterative NUTS expansion (trajectory.py:149-317,517-668, termination.py:
// 30-119) of one chain from s0 (energy H0, integrator state in LDS).  load(end)
// makes a trajectory end the state to integrate; step(h, out) takes one step of
// signed size h, leaves the new state in LDS and in out, and returns its
// energy.  Decisions use wave-reduced values, the same in every lane.
template <class Load, class Step>
__device__ NTree tuned_nuts_tree(const NPt& s0, double H0, Load load, Step step, const NUniform& uniform,
                                 const NutsArgs& nu, float eps, double div_threshold, double* ck, int c, NPt& left,
                                 NPt& right, NProp& prop) {
  const int l = threadIdx.x % 64;
  const int M = nu.max_doublings;
  const int NS = (1 << M) - 1;
  const int crit = nu.criterion;
  const int NA = nuts_checkpoint_arrays(crit);
  auto dot = [&](double a, double b) { return group_sum<64>(a * b); };
  // U-turn criteria (geomjax/rmhmc/metrics.py:100-117)
  auto is_turning = [&](double ml, double mr, double ms, double vl, double vr, double vs) -> bool {
    if (crit == 1) return dot(vl, ms) <= 0.0 || dot(vr, ms) <= 0.0;  // "bet"
    if (crit == 2) return dot(ml, vs) <= 0.0 || dot(mr, vs) <= 0.0;  // "riem"
    return dot(vl, vs) <= 0.0 || dot(vr, vs) <= 0.0;                 // "euc"
  };
  // checkpoints [NA][M][64]: velocity, velocity sum (0, 1), momentum, momentum sum (2, 3)
  auto ck_at = [&](int arr, int idx) -> double& { return ck[(arr * M + idx) * 64 + l]; };

  for (int i = l; i < NA * M * 64; i += 64) ck[i] = 0.0;
  __syncthreads();
  left = s0;
  right = s0;
  double psum = s0.p, vsum = s0.u;  // the start state is in the trajectory's sums
  int ntraj = 0;
  prop = nprop(s0, H0, 0.0);
  prop.slpa = -INFINITY;
  int at_end = 0;  // the end the integrator state holds: 0 the start (both ends), +1 right, -1 left
  int e = 0, j = 0;
  bool diverging = false, turning = false;
  while (e < M && !diverging && !turning) {
    const double ud = uniform(nu.u_dir, (size_t)c * M + e, kStreamNutsDirection, (uint32_t)e);
    const int dir = ud < 0.5 ? 1 : -1;
    if (at_end != 0 && at_end != dir) load(dir > 0 ? right : left);
    // ---- subtree of up to 2^e steps from that end (trajectory.py:149-317) ----
    NProp sp{};
    NPt cur{};
    double ssp = 0.0, ssv = 0.0;
    int nsub = 0, st = 0;
    const int max_steps = 1 << e;
    bool div_s = false, term_s = false;
    while (st < max_steps && !term_s && !div_s) {
      const double en = step((float)dir * eps, cur);
      double w = H0 - en;
      if (isnan(w)) w = -INFINITY;
      div_s = -w > div_threshold;
      const double us = uniform(nu.u_step, (size_t)c * NS + j, kStreamNutsStep, (uint32_t)j);
      if (st == 0) {
        ssp = cur.p;
        ssv = cur.u;
        nsub = 1;
        sp = nprop(cur, en, w);
      } else {
        ssp += cur.p;
        ssv += cur.u;
        ++nsub;
        // progressive uniform sampling (proposal.py:198-222)
        const double pacc = 1.0 / (1.0 + exp(-(w - sp.weight)));
        const double nw = logaddexp(sp.weight, w), ns = logaddexp(sp.slpa, fmin(w, 0.0));
        if (us < pacc) sp = nprop(cur, en, w);
        sp.weight = nw;
        sp.slpa = ns;
      }
      ++j;
      // iterative U-turn check (termination.py:38-117)
      const int idx_max = nuts::ckpt_idx_max(st), idx_min = nuts::ckpt_idx_min(st);
      if ((st & 1) == 0) {
        ck_at(0, idx_max) = cur.u;
        ck_at(1, idx_max) = ssv;
        if (NA == 4) {
          ck_at(2, idx_max) = cur.p;
          ck_at(3, idx_max) = ssp;
        }
      }
      __syncthreads();
      for (int i = idx_max; i >= idx_min && !term_s; --i) {
        const double vl = ck_at(0, i);
        const double vs = ssv - ck_at(1, i) + vl;
        const double ml = NA == 4 ? ck_at(2, i) : 0.0;
        const double ms = NA == 4 ? ssp - ck_at(3, i) + ml : 0.0;
        term_s = is_turning(ml, cur.p, ms, vl, cur.u, vs);
      }
      ++st;
    }
    // ---- merge (trajectory.py:580-660) ----
    if (div_s || term_s) {
      prop.slpa = logaddexp(prop.slpa, sp.slpa);
    } else {
      // progressive biased sampling (proposal.py:225-254)
      const double um = uniform(nu.u_merge, (size_t)c * M + e, kStreamNutsMerge, (uint32_t)e);
      const double pacc = fmin(exp(sp.weight - prop.weight), 1.0);
      const double nw = logaddexp(prop.weight, sp.weight), ns = logaddexp(prop.slpa, sp.slpa);
      if (um < pacc) prop = sp;
      prop.weight = nw;
      prop.slpa = ns;
    }
    if (dir > 0)
      right = cur;
    else
      left = cur;
    at_end = dir;
    psum += ssp;
    vsum += ssv;
    ntraj += nsub;
    const bool turn_full = is_turning(left.p, right.p, psum, left.u, right.u, vsum);
    ++e;
    diverging = div_s;
    turning = term_s || turn_full;
  }
  return NTree{e, ntraj, diverging, turning};
}

template <int NB>
__device__ __forceinline__ void run_nuts_chain(const LogisticArgs& a, const NutsArgs& nu, Smem<1>& S, double* ck,
                                               int c, int post) {
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
  float* const E0 = W;
  float* const E1 = W + ns;
  float* const SL = W + 2 * ns;
  const int Dp = a.Dmax;
  const float inv_prior = (float)(1.0 / a.prior_var);
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;
  const NUniform uniform{c, a.counter, (uint32_t)a.seed, (uint32_t)(a.seed >> 32)};
  const bool own = l < D;

  // ---- chain state, normals, eta rows ----
  S.q[l] = own ? static_cast<const float*>(a.q_in)[(size_t)c * Dp + l] : 0.0f;
  if (a.z != nullptr) {
    S.z[l] = own ? static_cast<const float*>(a.z)[(size_t)c * Dp + l] : 0.0f;
  } else if (l < MAXD / 2) {
    const uint64_t tt = a.counter;
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * l),
                                  (uint32_t)(tt >> 32)}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
    double n0, n1;
    normal2(r, n0, n1);
    S.z[2 * l] = 2 * l < D ? (float)n0 : 0.0f;
    S.z[2 * l + 1] = 2 * l + 1 < D ? (float)n1 : 0.0f;
  }
  __syncthreads();
  const double logp0 = a.logp_in[c];
  // eta of the position in LDS from X in float64 (E0 = E1, sqrt(lambda) in SL);
  // with want_grad also the log likelihood (returned) and the gradient (S.g)
  auto fresh_eta = [&](bool want_grad) -> double {
    const double ll = fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, E0, SL, want_grad, inv_prior);
    prep_pass<1>(E0, E0, E1, SL, npad);
    return ll;
  };
  (void)fresh_eta(false);
  // ---- momentum p0 = chol(G(q0)) z, velocity and H0 (as the tuned RMHMC kernel) ----
  gram_pass<NB, 1>(S, Xg, SL, npad, D, inv_prior);
  invert<NB, 1>(S, D, true, true);
  if (a.momentum != nullptr && own) static_cast<float*>(a.momentum)[(size_t)c * Dp + l] = S.p[l];
  const double H0 = -logp0 + 0.5 * group_sum<64>(own ? (double)S.p[l] * (double)S.u[l] : 0.0) +
                    0.5 * logdet_from_pivots(S, D) + ke_const;
  bool have_ginv = true;  // S.gi() = G(S.q)^-1 and S.u = G^-1 S.p

  // one implicit-midpoint step (integrators.py:92-156) of signed size hstep, then
  // the new state's eta, log density, metric, velocity and energy
  auto step = [&](float hstep, NPt& out) -> double {
    const float h = 0.5f * hstep;
    const float q0v = S.q[l], p0v = S.p[l];
    int mode = 0;  // 0 first evaluation, 1 fixed-point iterations, 2 explicit half step, 3 the new state
    int n = 0;
    double loglik = 0.0;
    for (;;) {
      if (!have_ginv) {
        gram_pass<NB, 1>(S, Xg, SL, npad, D, inv_prior);
        invert<NB, 1>(S, D, false, true);
        have_ginv = true;
      }
      if (mode == 3) break;
      field_pass<NB, 1>(S, Xr, Xs, yf, E0, E1, SL, npad, D, h, inv_prior, mode == 2);
      have_ginv = false;
      if (mode == 2) {  // explicit half step from the last iterate (integrators.py:146-148)
        const float qn = own ? fmaf(h, S.u[l], S.q[l]) : 0.0f;
        const float pn = own ? fmaf(-h, S.dH[l], S.p[l]) : 0.0f;
        S.q[l] = qn;
        S.p[l] = pn;
        __syncthreads();
        loglik = fresh_eta(true);  // (the log likelihood comes with the gradient pass)
        mode = 3;
        continue;
      }
      float dq = 0.0f, qn = 0.0f, pn = 0.0f;
      if (own) {
        qn = fmaf(h, S.u[l], q0v);
        pn = fmaf(-h, S.dH[l], p0v);
        dq = nan_max(fabsf(qn - S.q[l]), fabsf(pn - S.p[l]));
      }
      S.q[l] = qn;
      S.p[l] = pn;
      const double norm = (double)group_reduce<64, NanMax>(dq);
      __syncthreads();
      if (mode == 0) {
        mode = 1;
        n = 0;
      } else {
        ++n;
      }
      if (!(n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol)) mode = 2;
    }
    const double qq = group_sum<64>(own ? (double)S.q[l] * (double)S.q[l] : 0.0);
    const double pu = group_sum<64>(own ? (double)S.p[l] * (double)S.u[l] : 0.0);
    const double logp1 = loglik - 0.5 * qq / prior + lp_const;
    // lanes past D hold zeros (S.u past the tile width is never written)
    out = NPt{own ? S.q[l] : 0.0f, own ? S.p[l] : 0.0f, own ? S.u[l] : 0.0f, 0.0f, logp1, 0.0};
    return -logp1 + 0.5 * pu + 0.5 * logdet_from_pivots(S, D) + ke_const;
  };
  auto load = [&](const NPt& e) {
    S.q[l] = e.q;
    S.p[l] = e.p;
    __syncthreads();
    (void)fresh_eta(false);
    have_ginv = false;
  };

  const NPt s0{own ? S.q[l] : 0.0f, own ? S.p[l] : 0.0f, own ? S.u[l] : 0.0f, 0.0f, logp0, 0.0};
  NPt left, right;
  NProp prop;
  const NTree res = tuned_nuts_tree(s0, H0, load, step, uniform, nu, (float)a.eps, a.div_threshold, ck, c, left,
                                    right, prop);
  // the proposal's gradient from X at its position
  __syncthreads();
  S.q[l] = own ? prop.q : 0.0f;
  __syncthreads();
  (void)fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, E0, SL, true, inv_prior);
  const size_t row = (size_t)c * Dp;
  if (own) {
    static_cast<float*>(a.q_out)[row + l] = prop.q;
    if (a.g_out) static_cast<float*>(a.g_out)[row + l] = S.g[l];
    if (nu.left_position) static_cast<float*>(nu.left_position)[row + l] = left.q;
    if (nu.right_position) static_cast<float*>(nu.right_position)[row + l] = right.q;
  }
  for (int k = D + l; k < Dp; k += 64) {  // row tails past this posterior's D
    static_cast<float*>(a.q_out)[row + k] = 0.0f;
    if (a.g_out) static_cast<float*>(a.g_out)[row + k] = 0.0f;
    if (a.momentum) static_cast<float*>(a.momentum)[row + k] = 0.0f;
    if (nu.left_position) static_cast<float*>(nu.left_position)[row + k] = 0.0f;
    if (nu.right_position) static_cast<float*>(nu.right_position)[row + k] = 0.0f;
  }
  if (l == 0) {
    a.logp_out[c] = prop.logp;
    if (nu.energy) nu.energy[c] = prop.energy;
    if (nu.acceptance_rate) nu.acceptance_rate[c] = exp(prop.slpa) / (double)res.num_states;
    if (nu.is_divergent) nu.is_divergent[c] = (uint8_t)res.diverging;
    if (nu.is_turning) nu.is_turning[c] = (uint8_t)res.turning;
    if (nu.num_expansions) nu.num_expansions[c] = res.expansions;
    if (nu.num_steps) nu.num_steps[c] = res.num_states;
  }
}

// NUTS over the general-metric LMC integrator (geomjax/lmcmc/nuts.py:78-161),
// tuned form: the tree above over lan_integrator steps made of the tuned LMC
// kernel's phases (logistic_lmc_fast.hip: signed-weight Gram passes of
// G +- (h/2) Omega~, sweeps, refined solve, the position pass with grad phi), with
// the energy -l - log det G / 2 + v^T G v / 2 - J (geomjax/lmcmc/metrics.py:
// 209-221) in float64.  A trajectory end is (q, G v, v, grad phi, log density, J):
// reloading it recomputes only its eta rows.
template <int NB>
__device__ __forceinline__ void run_nuts_lmc_chain(const LogisticArgs& a, const NutsArgs& nu, Smem<1>& S, double* ck,
                                                   int c, int post) {
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
  const float inv_prior = (float)(1.0 / a.prior_var);
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const NUniform uniform{c, a.counter, (uint32_t)a.seed, (uint32_t)(a.seed >> 32)};
  const bool own = l < D;

  S.q[l] = own ? static_cast<const float*>(a.q_in)[(size_t)c * Dp + l] : 0.0f;
  if (a.z != nullptr) {
    S.z[l] = own ? static_cast<const float*>(a.z)[(size_t)c * Dp + l] : 0.0f;
  } else if (l < MAXD / 2) {
    const uint64_t tt = a.counter;
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * l),
                                  (uint32_t)(tt >> 32)}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
    double n0, n1;
    normal2(r, n0, n1);
    S.z[2 * l] = 2 * l < D ? (float)n0 : 0.0f;
    S.z[2 * l + 1] = 2 * l + 1 < D ? (float)n1 : 0.0f;
  }
  S.p[l] = 0.0f;
  __syncthreads();
  const double logp0 = a.logp_in[c];
  double Jc = a.J_in[c];  // volume adjustment of the state in LDS
  (void)fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, ETA, SL, false, inv_prior);

  // The phases of one lan_integrator step of signed size h (integrators.py:61-144)
  // from (S.q, S.p = velocity) with grad phi in S.dH, or with draw the velocity
  // draw v0 = L^-T z at S.q; each phase is one Gram pass and one sweep.  Leaves
  // log det G at the final position in logdetG; returns the volume change.
  enum { kDraw, kHalfA, kHalfB, kPos, kDone };
  double logdetG = 0.0;
  auto phases = [&](bool draw, float h) -> double {
    double dJ = 0.0;
    int phase = draw ? kDraw : kHalfA;
    bool second = false;  // the half step after the position update
    while (phase != kDone) {
      const float* wts = SL;
      if (phase == kHalfA) {
        lmc_row_pass<NB>(S, Xr, ETA, SL, SS, npad, D, 0.25f * h, h, true, false, inv_prior);
        wts = SS;
      } else if (phase == kHalfB) {
        lmc_row_pass<NB>(S, Xr, ETA, SL, SS, npad, D, -0.25f * h, h, false, !second, inv_prior);
        wts = SS;
      } else if (phase == kPos) {
        S.q[l] = own ? fmaf(h, S.p[l], S.q[l]) : 0.0f;
        __syncthreads();
      }
      gram_pass<NB, 1, true>(S, Xg, wts, npad, D, inv_prior);
      f32x32 A[NB];
      float p0j;
      sweep<NB, 1>(S, A, D, false, p0j, phase == kDraw);
      const double lad = logabsdet_from_pivots(S, D);
      if (phase == kHalfA) {  // v' = A+^-1 (G v - (h/2) grad phi)
        dJ -= lad;
        S.p0[l] = own ? fmaf(-0.5f * h, S.dH[l], S.u[l]) : 0.0f;
        __syncthreads();
        solve_refined<NB, 1>(S, A, S.p0, S.p, S.u);
        phase = kHalfB;
        continue;
      }
      if (phase == kHalfB) {
        dJ += lad;
        phase = second ? kDone : kPos;
        continue;
      }
      logdetG = lad;
      if (phase == kDraw) {  // v0 = L^-T z by back substitution over the rows of L
        float r = S.z[l], vj = 0.0f;
        for (int i = D - 1; i >= 0; --i) {
          const float vi = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, r), i)) /
                           S.G[i * LDG + i];
          if (l == i) vj = vi;
          if (l < i) r = fmaf(-S.G[i * LDG + l], vi, r);
        }
        S.p[l] = vj;
        if (a.momentum != nullptr && own) static_cast<float*>(a.momentum)[(size_t)c * Dp + l] = vj;
      }
      store_ginv<NB, 1>(S, A);
      S.u[l] = 0.0f;
      __syncthreads();
      field_pass<NB, 1>(S, Xr, Xs, yf, ETA, ETA, SL, npad, D, 0.0f, inv_prior, true);  // S.dH = grad phi
      if (phase == kDraw) {
        phase = kDone;
      } else {
        second = true;
        phase = kHalfA;
      }
    }
    return dJ;
  };
  // S.u = G v at the position; returns v^T G v
  auto vgv = [&]() -> double {
    lmc_row_pass<NB>(S, Xr, ETA, SL, SS, npad, D, 0.0f, 0.0f, true, false, inv_prior);
    return group_sum<64>(own ? (double)S.p[l] * (double)S.u[l] : 0.0);
  };

  // ---- velocity v0 = L^-T z, momentum G v0 and H0 (lmcmc/nuts.py:146-158) ----
  (void)phases(true, 0.0f);
  const double H0 = -logp0 + (-0.5 * logdetG + 0.5 * vgv()) - Jc;

  auto step = [&](float h, NPt& out) -> double {
    Jc += phases(false, h);
    const double ll = fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, ETA, SL, true, inv_prior);
    const double vGv = vgv();
    const double qq = group_sum<64>(own ? (double)S.q[l] * (double)S.q[l] : 0.0);
    const double logp1 = ll - 0.5 * qq / prior + lp_const;
    out = NPt{own ? S.q[l] : 0.0f, own ? S.u[l] : 0.0f, own ? S.p[l] : 0.0f, own ? S.dH[l] : 0.0f, logp1, Jc};
    return -logp1 + (-0.5 * logdetG + 0.5 * vGv) - Jc;
  };
  auto load = [&](const NPt& e) {
    S.q[l] = e.q;
    S.p[l] = e.u;
    S.dH[l] = e.dphi;
    Jc = e.J;
    __syncthreads();
    (void)fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, ETA, SL, false, inv_prior);
  };

  const NPt s0{own ? S.q[l] : 0.0f, own ? S.u[l] : 0.0f, own ? S.p[l] : 0.0f, own ? S.dH[l] : 0.0f, logp0, Jc};
  NPt left, right;
  NProp prop;
  const NTree res = tuned_nuts_tree(s0, H0, load, step, uniform, nu, (float)a.eps, a.div_threshold, ck, c, left,
                                    right, prop);
  // the proposal's gradient from X at its position
  __syncthreads();
  S.q[l] = own ? prop.q : 0.0f;
  __syncthreads();
  (void)fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, ETA, SL, true, inv_prior);
  const size_t row = (size_t)c * Dp;
  if (own) {
    static_cast<float*>(a.q_out)[row + l] = prop.q;
    if (a.g_out) static_cast<float*>(a.g_out)[row + l] = S.g[l];
    if (nu.left_position) static_cast<float*>(nu.left_position)[row + l] = left.q;
    if (nu.right_position) static_cast<float*>(nu.right_position)[row + l] = right.q;
  }
  for (int k = D + l; k < Dp; k += 64) {  // row tails past this posterior's D
    static_cast<float*>(a.q_out)[row + k] = 0.0f;
    if (a.g_out) static_cast<float*>(a.g_out)[row + k] = 0.0f;
    if (a.momentum) static_cast<float*>(a.momentum)[row + k] = 0.0f;
    if (nu.left_position) static_cast<float*>(nu.left_position)[row + k] = 0.0f;
    if (nu.right_position) static_cast<float*>(nu.right_position)[row + k] = 0.0f;
  }
  if (l == 0) {
    a.logp_out[c] = prop.logp;
    if (a.J_out) a.J_out[c] = prop.J;
    if (nu.energy) nu.energy[c] = prop.energy;
    if (nu.acceptance_rate) nu.acceptance_rate[c] = exp(prop.slpa) / (double)res.num_states;
    if (nu.is_divergent) nu.is_divergent[c] = (uint8_t)res.diverging;
    if (nu.is_turning) nu.is_turning[c] = (uint8_t)res.turning;
    if (nu.num_expansions) nu.num_expansions[c] = res.expansions;
    if (nu.num_steps) nu.num_steps[c] = res.num_states;
  }
}

template <int NB>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
logistic_nuts_lmc_f32_kernel(const LogisticArgs a, const NutsArgs nu) {
  __shared__ Smem<1> S;
  extern __shared__ __align__(16) double ck_dyn[];
  const int per_xcd = gridDim.x / 8;
  const int slot = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (slot >= a.C) return;
  const int c = a.dispatch_order ? a.dispatch_order[slot] : slot;
  if (c < 0 || c >= a.C) return;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  if (a.packed_hdr[post].nb != NB) return;
  run_nuts_lmc_chain<NB>(a, nu, S, ck_dyn, c, post);
}

// One wave per chain, chains dealt to the XCDs in contiguous ranges; chains of
// the other tile width leave at once.
template <int NB>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
logistic_nuts_f32_kernel(const LogisticArgs a, const NutsArgs nu) {
  __shared__ Smem<1> S;
  extern __shared__ __align__(16) double ck_dyn[];
  const int per_xcd = gridDim.x / 8;
  const int slot = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (slot >= a.C) return;
  const int c = a.dispatch_order ? a.dispatch_order[slot] : slot;
  if (c < 0 || c >= a.C) return;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  if (a.packed_hdr[post].nb != NB) return;
  run_nuts_chain<NB>(a, nu, S, ck_dyn, c, post);
}

}  // namespace lfast

hipError_t launch_logistic_nuts_lmc_f32(const LogisticArgs& a, const NutsArgs& n, int max_features,
                                        hipStream_t stream) {
  const int blocks = (a.C + 7) / 8 * 8;
  const size_t smem = (size_t)nuts_checkpoint_arrays(n.criterion) * n.max_doublings * 64 * sizeof(double);
  const bool nb1 = max_features <= 32 || a.P > 1;
  const bool nb2 = max_features > 32;
  if (nb2) {
    hipLaunchKernelGGL((lfast::logistic_nuts_lmc_f32_kernel<2>), dim3(blocks), dim3(64), smem, stream, a, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (nb1) hipLaunchKernelGGL((lfast::logistic_nuts_lmc_f32_kernel<1>), dim3(blocks), dim3(64), smem, stream, a, n);
  return hipGetLastError();
}

hipError_t launch_logistic_nuts_f32(const LogisticArgs& a, const NutsArgs& n, int max_features,
                                    hipStream_t stream) {
  const int blocks = (a.C + 7) / 8 * 8;
  const size_t smem = (size_t)nuts_checkpoint_arrays(n.criterion) * n.max_doublings * 64 * sizeof(double);
  const bool nb1 = max_features <= 32 || a.P > 1;
  const bool nb2 = max_features > 32;
  if (nb2) {
    hipLaunchKernelGGL((lfast::logistic_nuts_f32_kernel<2>), dim3(blocks), dim3(64), smem, stream, a, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (nb1) hipLaunchKernelGGL((lfast::logistic_nuts_f32_kernel<1>), dim3(blocks), dim3(64), smem, stream, a, n);
  return hipGetLastError();
}

}  // namespace gjx
