// NUTS for Bayesian logistic regression with the Fisher metric, float64,
// D <= 64, with the RMHMC implicit-midpoint integrator or the general-metric
// LMC integrator (lmcmc).
//
// Reference: geomjax/rmhmc/nuts.py:78-161,253-348 and
// geomjax/lmcmc/nuts.py:78-161,253-351 (kernels, iterative proposal),
// geomjax/mcmc/trajectory.py:149-317 (a subtree of up to 2^e steps, progressive
// uniform sampling, stop at divergence or an iterative U-turn), :517-668
// (random direction, biased merge, U-turn of the merged trajectory),
// geomjax/mcmc/termination.py:30-119 (numpyro checkpoints),
// geomjax/rmhmc/metrics.py:100-117 (U-turn criteria); the integrators and
// energies of logistic_rmhmc.hip (geomjax/rmhmc/integrators.py:53-156,
// geomjax/rmhmc/metrics.py:42-129) and logistic_lmc.hip
// (geomjax/lmcmc/integrators.py:51-144, energy -l - log det G / 2 + v^T G v / 2
// - J, geomjax/lmcmc/metrics.py:209-221).
//
// One workgroup (256 lanes) per chain, the float64 kernels' LDS layout
// (capacity 64) and passes.  Lane t < D holds coordinate t of the trajectory's
// two ends, of the momentum and velocity sums of the trajectory and of the
// subtree being built, and of the two proposals (struct Pt / Prop).  The state
// being integrated lives in LDS; after each step the new state's log density,
// gradient and energy are formed once, and what the integrator needs at a state
// (RMHMC: G^-1 and the velocity; LMC: grad l and grad log det G) also serves the
// next step from that end.  The tree building (lnuts_tree) is written once over
// a load(end) / step(h) pair.  U-turn checkpoints ([NA][M][64] doubles) follow
// the Smem block in dynamic LDS.
#include "common.h"
#include "logistic.h"
#include "logistic_common.h"
#include "logistic_lmc_ops.h"
#include "nuts.h"

namespace gjx {

using namespace logistic;
using nuts::logaddexp;

namespace lnuts {

using C = Cfg<64>;
constexpr int MAXD = C::MAXD, BS = C::BS;

// A state on the trajectory, coordinate t of lane t < D (zeros elsewhere): position,
// momentum, velocity, grad l, grad log det G (LMC); its volume adjustment (LMC) and
// log density (the same in every lane)
struct Pt {
  double q, p, u, g, gl;
  double J, logp;
};

struct Prop {
  double q, g, logp, J, energy, weight, slpa;
};

struct TreeOut {
  int expansions, num_states;
  bool diverging, turning;
};

__device__ __forceinline__ Prop make_prop(const Pt& s, double en, double w) {
  return Prop{s.q, s.g, s.logp, s.J, en, w, fmin(w, 0.0)};
}

// The transition's uniforms: host-fed or the device Philox stream of the purpose tag.
struct Uniform {
  int c;
  uint64_t tt;
  uint32_t k0, k1;
  __device__ double operator()(const double* host, size_t idx, uint32_t stream, uint32_t i) const {
    if (host != nullptr) return host[idx];
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, stream + i, (uint32_t)(tt >> 32)}, k0, k1);
    return u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
  }
};

// The iterative NUTS expansion of one chain's trajectory from s0 (energy H0, held
// in LDS by the caller): geomjax/mcmc/trajectory.py:149-317,517-668 with the
// numpyro U-turn check of geomjax/mcmc/termination.py:30-119.  load(end) puts a
// trajectory end into LDS as the state to integrate; step(h, out) takes one step
// of signed size h from the LDS state, leaves the new state there and in out,
// and returns its energy.  Every decision is taken on block-reduced values, the
// same in all lanes, so the workgroup stays uniform.
template <class Load, class Step>
__device__ TreeOut lnuts_tree(const Pt& s0, double H0, Load load, Step step, const Uniform& uniform,
                              const NutsArgs& nu, double eps, double div_threshold, Smem<C>& S, double* ck, int c,
                              Pt& left, Pt& right, Prop& prop) {
  const int t = threadIdx.x;
  const int M = nu.max_doublings;
  const int NS = (1 << M) - 1;
  const int crit = nu.criterion;
  const int NA = nuts_checkpoint_arrays(crit);
  // U-turn criteria (geomjax/rmhmc/metrics.py:100-117) on the two ends' momenta and
  // velocities and the sums between them (one coordinate per lane)
  auto is_turning = [&](double ml, double mr, double ms, double vl, double vr, double vs) -> bool {
    if (crit == 1) return block_sum(S, vl * ms) <= 0.0 || block_sum(S, vr * ms) <= 0.0;  // "bet"
    if (crit == 2) return block_sum(S, ml * vs) <= 0.0 || block_sum(S, mr * vs) <= 0.0;  // "riem"
    return block_sum(S, vl * vs) <= 0.0 || block_sum(S, vr * vs) <= 0.0;                 // "euc"
  };
  // checkpoints [NA][M][64]: velocity, velocity sum (0, 1), momentum, momentum sum (2, 3)
  auto ck_at = [&](int arr, int idx) -> double& { return ck[(arr * M + idx) * MAXD + t]; };

  for (int i = t; i < NA * M * MAXD; i += NT) ck[i] = 0.0;
  __syncthreads();
  left = s0;
  right = s0;
  double psum = s0.p, vsum = s0.u;  // the start state is in the trajectory's sums
  int ntraj = 0;
  prop = make_prop(s0, H0, 0.0);
  prop.slpa = -INFINITY;
  int at_end = 0;  // which end LDS holds: 0 the start (both ends), +1 right, -1 left
  int e = 0, j = 0;
  bool diverging = false, turning = false;
  while (e < M && !diverging && !turning) {
    const double ud = uniform(nu.u_dir, (size_t)c * M + e, kStreamNutsDirection, (uint32_t)e);
    const int dir = ud < 0.5 ? 1 : -1;
    if (at_end != 0 && at_end != dir) load(dir > 0 ? right : left);
    // ---- subtree of up to 2^e steps from that end (trajectory.py:149-317) ----
    Prop sp{};
    Pt cur{};
    double ssp = 0.0, ssv = 0.0;
    int nsub = 0, st = 0;
    const int max_steps = 1 << e;
    bool div_s = false, term_s = false;
    while (st < max_steps && !term_s && !div_s) {
      const double en = step(dir * eps, cur);
      double w = H0 - en;
      if (isnan(w)) w = -INFINITY;
      div_s = -w > div_threshold;
      const double us = uniform(nu.u_step, (size_t)c * NS + j, kStreamNutsStep, (uint32_t)j);
      if (st == 0) {
        ssp = cur.p;
        ssv = cur.u;
        nsub = 1;
        sp = make_prop(cur, en, w);
      } else {
        ssp += cur.p;
        ssv += cur.u;
        ++nsub;
        // progressive uniform sampling (proposal.py:198-222)
        const double pacc = 1.0 / (1.0 + exp(-(w - sp.weight)));
        const double nw = logaddexp(sp.weight, w), ns = logaddexp(sp.slpa, fmin(w, 0.0));
        if (us < pacc) sp = make_prop(cur, en, w);
        sp.weight = nw;
        sp.slpa = ns;
      }
      ++j;
      // iterative U-turn check (termination.py:38-117)
      const int idx_max = nuts::ckpt_idx_max(st), idx_min = nuts::ckpt_idx_min(st);
      if ((st & 1) == 0 && t < MAXD) {
        ck_at(0, idx_max) = cur.u;
        ck_at(1, idx_max) = ssv;
        if (NA == 4) {
          ck_at(2, idx_max) = cur.p;
          ck_at(3, idx_max) = ssp;
        }
      }
      __syncthreads();
      for (int i = idx_max; i >= idx_min && !term_s; --i) {
        const double vl = t < MAXD ? ck_at(0, i) : 0.0;
        const double vs = t < MAXD ? ssv - ck_at(1, i) + vl : 0.0;
        const double ml = (NA == 4 && t < MAXD) ? ck_at(2, i) : 0.0;
        const double ms = (NA == 4 && t < MAXD) ? ssp - ck_at(3, i) + ml : 0.0;
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
  return TreeOut{e, ntraj, diverging, turning};
}

// Per-chain set-up shared by both kernels: the posterior of chain c, or false if
// the chain is not served (whole workgroups leave together).
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

// Position from q_in and the normals into S.z (host-fed or device Philox).
__device__ __forceinline__ void load_state_and_noise(const LogisticArgs& a, Smem<C>& S, int c, int D) {
  const int t = threadIdx.x, Dp = a.Dmax;
  const uint64_t tt = a.counter;
  const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
  if (t < MAXD) {
    const double qv = t < D ? static_cast<const double*>(a.q_in)[(size_t)c * Dp + t] : 0.0;
    S.q[t] = qv;
    S.q0[t] = qv;
    // every vector a pass multiplies by whole zero-padded rows of X starts at zero
    // over its capacity (later writes touch only the first D entries)
    S.p[t] = 0.0;
    S.u[t] = 0.0;
    S.g[t] = 0.0;
    S.dH[t] = 0.0;
    S.vt[t] = 0.0;
  }
  if (a.z != nullptr) {
    if (t < MAXD) S.z[t] = t < D ? static_cast<const double*>(a.z)[(size_t)c * Dp + t] : 0.0;
  } else if (t < MAXD / 2) {
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * t),
                                  (uint32_t)(tt >> 32)}, k0, k1);
    double n0, n1;
    normal2(r, n0, n1);
    S.z[2 * t] = 2 * t < D ? n0 : 0.0;
    S.z[2 * t + 1] = 2 * t + 1 < D ? n1 : 0.0;
  }
  __syncthreads();
}

// The proposal's state (gradient g given per lane), the trajectory's ends, row
// tails past D and NUTSInfo.
__device__ void write_outputs(const LogisticArgs& a, const NutsArgs& nu, int c, int D, const Prop& prop, double g,
                              const Pt& left, const Pt& right, const TreeOut& t_out) {
  const int t = threadIdx.x, Dp = a.Dmax;
  const size_t row = (size_t)c * Dp;
  if (t < D) {
    static_cast<double*>(a.q_out)[row + t] = prop.q;
    if (a.g_out) static_cast<double*>(a.g_out)[row + t] = g;
    if (nu.left_position) static_cast<double*>(nu.left_position)[row + t] = left.q;
    if (nu.right_position) static_cast<double*>(nu.right_position)[row + t] = right.q;
  }
  for (int k = D + t; k < Dp; k += NT) {  // row tails past this posterior's D
    static_cast<double*>(a.q_out)[row + k] = 0.0;
    if (a.g_out) static_cast<double*>(a.g_out)[row + k] = 0.0;
    if (a.momentum) static_cast<double*>(a.momentum)[row + k] = 0.0;
    if (nu.left_position) static_cast<double*>(nu.left_position)[row + k] = 0.0;
    if (nu.right_position) static_cast<double*>(nu.right_position)[row + k] = 0.0;
  }
  if (t == 0) {
    a.logp_out[c] = prop.logp;
    if (a.J_out) a.J_out[c] = prop.J;
    if (nu.energy) nu.energy[c] = prop.energy;
    if (nu.acceptance_rate) nu.acceptance_rate[c] = exp(prop.slpa) / (double)t_out.num_states;
    if (nu.is_divergent) nu.is_divergent[c] = (uint8_t)t_out.diverging;
    if (nu.is_turning) nu.is_turning[c] = (uint8_t)t_out.turning;
    if (nu.num_expansions) nu.num_expansions[c] = t_out.expansions;
    if (nu.num_steps) nu.num_steps[c] = t_out.num_states;
  }
}

}  // namespace lnuts

using namespace lnuts;

// ---------------------------------------------------------------------------
// NUTS with the RMHMC implicit midpoint (geomjax/rmhmc/nuts.py)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT, 1) logistic_nuts_f64_kernel(const LogisticArgs a, const NutsArgs nu) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C>& S = *reinterpret_cast<Smem<C>*>(smem_raw);
  double* const ck = reinterpret_cast<double*>(smem_raw + sizeof(Smem<C>));
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  Chain ch;
  if (!chain_of(a, c, ch)) return;
  const int N = ch.N, D = ch.D;
  const double* X = ch.X;
  const double* y = ch.y;
  const bool own = t < D;  // this lane holds coordinate t
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;
  const Uniform uniform{c, a.counter, (uint32_t)a.seed, (uint32_t)(a.seed >> 32)};
  load_state_and_noise(a, S, c, D);
  const double logp0 = a.logp_in[c];

  // ---- momentum p0 = chol(G(q0)) z, velocity and H0 (as logistic_rmhmc.hip) ----
  double dummy = 0.0;
  pass1(S, X, y, N, D, prior, false, dummy);
  double A[BS][BS];
  load_block(S, A);
  __syncthreads();
  const double logdet = sweep(S, A, D, S.z, S.p0);
  __syncthreads();
  if (t < MAXD) S.p[t] = S.p0[t];
  if (a.momentum != nullptr && own) static_cast<double*>(a.momentum)[(size_t)c * a.Dmax + t] = S.p0[t];
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
  // one implicit-midpoint step (integrators.py:92-156), then the new state's log
  // density, gradient (S.g), G^-1, velocity (S.u) and energy
  auto step = [&](double hstep, Pt& out) -> double {
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
    const double logp1 = loglik - 0.5 * qq / prior + lp_const;
    out = Pt{own ? S.q[t] : 0.0, own ? S.p[t] : 0.0, own ? S.u[t] : 0.0, own ? S.g[t] : 0.0, 0.0, 0.0, logp1};
    return -logp1 + 0.5 * pu + 0.5 * ld + ke_const;
  };
  auto load = [&](const Pt& e) {
    if (t < MAXD) {
      S.q[t] = e.q;
      S.p[t] = e.p;
    }
    have_ginv = false;
    __syncthreads();
  };

  const Pt s0{own ? S.q[t] : 0.0, own ? S.p[t] : 0.0, own ? S.u[t] : 0.0, 0.0, 0.0, 0.0, logp0};
  Pt left, right;
  Prop prop;
  const TreeOut res = lnuts_tree(s0, H0, load, step, uniform, nu, a.eps, a.div_threshold, S, ck, c, left, right,
                                 prop);
  // the proposal's gradient: one Gram pass at its position
  __syncthreads();
  if (t < MAXD) S.q[t] = own ? prop.q : 0.0;
  __syncthreads();
  double ll_end = 0.0;
  pass1(S, X, y, N, D, prior, true, ll_end);
  write_outputs(a, nu, c, D, prop, own ? S.g[t] : 0.0, left, right, res);
}

// ---------------------------------------------------------------------------
// NUTS with the general-metric LMC integrator (geomjax/lmcmc/nuts.py)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT, 1) logistic_nuts_lmc_f64_kernel(const LogisticArgs a, const NutsArgs nu) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<C>& S = *reinterpret_cast<Smem<C>*>(smem_raw);
  double* const ck = reinterpret_cast<double*>(smem_raw + sizeof(Smem<C>));
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  Chain ch;
  if (!chain_of(a, c, ch)) return;
  const int N = ch.N, D = ch.D;
  const double* X = ch.X;
  const double* y = ch.y;
  const bool own = t < D;
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const Uniform uniform{c, a.counter, (uint32_t)a.seed, (uint32_t)(a.seed >> 32)};
  load_state_and_noise(a, S, c, D);
  const double logp0 = a.logp_in[c];
  double Jc = a.J_in[c];  // volume adjustment of the state in LDS
  double A[BS][BS];

  // ---- velocity v0 = L^-T z, momentum G v0 and H0 (lmcmc/nuts.py:146-158) ----
  double ll0 = 0.0, dd = 0.0;
  const double logdet0 = lmc_position_pass(S, A, X, y, N, D, prior, true, ll0);
  pass_gram(S, X, y, N, D, prior, S.p, 0.0, false, true, dd);  // S.gv = G v0
  if (a.momentum != nullptr && own) static_cast<double*>(a.momentum)[(size_t)c * a.Dmax + t] = S.p[t];
  const double vGv0 = block_sum(S, own ? S.p[t] * S.gv[t] : 0.0);
  const double H0 = -logp0 + (-0.5 * logdet0 + 0.5 * vGv0) - Jc;

  // one lan_integrator step (integrators.py:51-144) from (S.q, S.p = velocity) with
  // S.g, S.dH current, then the new state's log density, momentum G v and energy
  auto step = [&](double hstep, Pt& out) -> double {
    double dJ = 0.0;
    lmc_half_step(S, A, X, y, N, D, prior, hstep, dJ);
    if (own) S.q[t] = S.q[t] + hstep * S.p[t];
    __syncthreads();
    double ll = 0.0;
    const double ld = lmc_position_pass(S, A, X, y, N, D, prior, false, ll);
    lmc_half_step(S, A, X, y, N, D, prior, hstep, dJ);
    double d2 = 0.0;
    pass_gram(S, X, y, N, D, prior, S.p, 0.0, false, true, d2);  // S.gv = G v
    const double vGv = block_sum(S, own ? S.p[t] * S.gv[t] : 0.0);
    const double qq = block_sum(S, own ? S.q[t] * S.q[t] : 0.0);
    const double logp1 = ll - 0.5 * qq / prior + lp_const;
    Jc += dJ;
    out = Pt{own ? S.q[t] : 0.0, own ? S.gv[t] : 0.0, own ? S.p[t] : 0.0, own ? S.g[t] : 0.0,
             own ? S.dH[t] : 0.0, Jc, logp1};
    return -logp1 + (-0.5 * ld + 0.5 * vGv) - Jc;
  };
  auto load = [&](const Pt& e) {
    if (t < MAXD) {
      S.q[t] = e.q;
      S.p[t] = e.u;
      S.g[t] = e.g;
      S.dH[t] = e.gl;
    }
    Jc = e.J;
    __syncthreads();
  };

  const Pt s0{own ? S.q[t] : 0.0, own ? S.gv[t] : 0.0, own ? S.p[t] : 0.0, own ? S.g[t] : 0.0,
              own ? S.dH[t] : 0.0, Jc, logp0};
  Pt left, right;
  Prop prop;
  const TreeOut res = lnuts_tree(s0, H0, load, step, uniform, nu, a.eps, a.div_threshold, S, ck, c, left, right,
                                 prop);
  write_outputs(a, nu, c, D, prop, prop.g, left, right, res);
}

size_t logistic_nuts_smem_bytes(int max_doublings, int criterion) {
  return sizeof(Smem<Cfg<64>>) + (size_t)nuts_checkpoint_arrays(criterion) * max_doublings * 64 * sizeof(double);
}

namespace {

template <class K>
hipError_t launch_lnuts(K kern, const LogisticArgs& a, const NutsArgs& n, hipStream_t stream) {
  const size_t smem = logistic_nuts_smem_bytes(n.max_doublings, n.criterion);
  const hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(a.C), dim3(NT), smem, stream, a, n);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_logistic_nuts_f64(const LogisticArgs& a, const NutsArgs& n, hipStream_t stream) {
  return launch_lnuts(logistic_nuts_f64_kernel, a, n, stream);
}

hipError_t launch_logistic_nuts_lmc_f64(const LogisticArgs& a, const NutsArgs& n, hipStream_t stream) {
  return launch_lnuts(logistic_nuts_lmc_f64_kernel, a, n, stream);
}

}  // namespace gjx
