// NUTS on Neal's funnel with the RMHMC implicit-midpoint integrator (Monge
// metric; float64, or the tuned float32 form: float32 state, trajectory sums
// and checkpoints, the fixed-point iterates carried in span{x0, p0} as in the
// tuned sampler, float64 energies, weights and U-turn dot products) or the
// LMC-Monge Lagrangian integrator (float64).
//
// Reference: geomjax/rmhmc/nuts.py:78-161,253-348 and
// geomjax/lmcmonge/nuts.py:80-207,300-404 (kernels, iterative proposal),
// geomjax/mcmc/trajectory.py:149-317 (dynamic_progressive_integration: a
// subtree of up to 2^e steps in one direction, progressive uniform sampling,
// stopping at divergence or an iterative U-turn), :517-668
// (dynamic_multiplicative_expansion: random direction, biased merge, U-turn
// of the merged trajectory), geomjax/mcmc/termination.py:30-119 (numpyro
// checkpoints), geomjax/rmhmc/metrics.py:100-117 (the U-turn criteria: "euc"
// the velocity sum against the end velocities, "bet" the momentum sum against
// the end velocities, "riem" the velocity sum against the end momenta).
//
// One wave per chain: every branch of the tree building is wave-uniform.  Lane
// l holds x coordinates [l EPL, (l+1) EPL) of the D-1 "x" part; the log-scale
// coordinate v is replicated.  The tree building (nuts_tree) is written once
// over a chain state with position (x, v), momentum (px, pv) and velocity
// (ux, uv); each kernel supplies its integrator step and energy:
//   RMHMC: the closed forms of funnel_rmhmc.hip (energy -l + T(q, p));
//   LMC  : the lmc_ops of funnel_lmc_ops.h (energy -l + K(q, v) - J with the
//          volume adjustment J carried along the trajectory,
//          geomjax/lmcmonge/metrics.py:267-284).
// The U-turn checkpoints (velocity and velocity sum at up to M leaves, and
// momentum and momentum sum for "bet" and "riem") live in LDS; the
// trajectory's ends, sums and the proposals live in registers.
#include "../../include/geomjax_amd.h"
#include "common.h"
#include "funnel_field.h"
#include "funnel_lmc.h"
#include "funnel_lmc_ops.h"
#include "funnel_rmhmc.h"
#include "nuts.h"

#include <math.h>

namespace gjx {
namespace nuts {

// RMHMC chain state: |x|^2 and x . p_x are kept as group totals
template <int EPL, class T = double>
struct RState {
  T x[EPL], px[EPL], ux[EPL];
  T v, pv, uv;
  T sx, xpx;
  double J;  // no volume adjustment in RMHMC (kept at 0 for the shared proposal code)
};

// LMC-Monge chain state: the position's scalars and the carried volume adjustment
template <int EPL, class T = double>
struct LState {
  T x[EPL], px[EPL], ux[EPL];
  T v, pv, uv;
  lmc_ops::PosScalars<T> P;
  double J;
};

template <int EPL, class T = double>
struct Prop {
  T x[EPL];
  T v;
  double J;
  double energy, weight, slpa;
};

struct TreeOut {
  int expansions, num_states;
  bool diverging, turning;
};

// dot products of the U-turn criteria, accumulated in float64
template <int EPL, class T>
__device__ __forceinline__ double dot64(const T (&ux)[EPL], T uv, const T (&wx)[EPL], T wv) {
  double d = 0.0;
#pragma unroll
  for (int k = 0; k < EPL; ++k) d += (double)ux[k] * (double)wx[k];
  return group_sum<64>(d) + (double)uv * (double)wv;
}

// The transition's uniforms: host-fed or the device Philox stream of the purpose tag.
struct NutsUniform {
  int c;
  uint64_t tt;
  uint32_t k0, k1;
  __device__ double operator()(const double* host, size_t idx, uint32_t stream, uint32_t i) const {
    if (host != nullptr) return host[idx];
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, stream + i, (uint32_t)(tt >> 32)}, k0, k1);
    return u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
  }
};

// The iterative NUTS expansion of one chain's trajectory from s0 (energy H0):
// geomjax/mcmc/trajectory.py:149-317,517-668 with the numpyro U-turn check of
// geomjax/mcmc/termination.py:30-119.  step(s, signed_eps) advances a state by
// one integrator step (its velocity and momentum current), energy(s) is its
// energy.  On return left / right are the trajectory's ends and prop the
// sampled proposal.
template <int EPL, class S, class Step, class Energy, class T>
__device__ TreeOut nuts_tree(const S& s0, double H0, Step step, Energy energy, const NutsUniform& uniform,
                             const NutsArgs& nu, double eps, double div_threshold, T* ck, int c, S& left,
                             S& right, Prop<EPL, T>& prop) {
  const int l = threadIdx.x;
  const int M = nu.max_doublings;
  const int NS = (1 << M) - 1;
  const int crit = nu.criterion;
  const int NA = nuts_checkpoint_arrays(crit);
  const int XS = M * 64 * EPL;  // one checkpoint array's x part
  // U-turn criteria (geomjax/rmhmc/metrics.py:100-117) on the two ends' momenta ml, mr and
  // velocities vl, vr and the sums ms, vs over the states between them; the dot products
  // are group totals, so the answer is the same in every lane
  auto is_turning = [&](const T (&mlx)[EPL], T mlv, const T (&mrx)[EPL], T mrv, const T (&msx)[EPL], T msv,
                        const T (&vlx)[EPL], T vlv, const T (&vrx)[EPL], T vrv, const T (&vsx)[EPL],
                        T vsv) -> bool {
    if (crit == 1) return dot64(vlx, vlv, msx, msv) <= 0.0 || dot64(vrx, vrv, msx, msv) <= 0.0;  // "bet"
    if (crit == 2) return dot64(mlx, mlv, vsx, vsv) <= 0.0 || dot64(mrx, mrv, vsx, vsv) <= 0.0;  // "riem"
    return dot64(vlx, vlv, vsx, vsv) <= 0.0 || dot64(vrx, vrv, vsx, vsv) <= 0.0;                 // "euc"
  };
  // checkpoints: [NA][M][64 EPL] coordinates x, then [NA][M] for v; arrays 0, 1 =
  // velocity, velocity sum; 2, 3 = momentum, momentum sum
  auto ck_put = [&](int arr, int idx, const T (&x)[EPL], T v) {
#pragma unroll
    for (int k = 0; k < EPL; ++k) ck[arr * XS + idx * 64 * EPL + l * EPL + k] = x[k];
    if (l == 0) ck[NA * XS + arr * M + idx] = v;  // v is replicated: lane 0 stores it
  };
  auto ck_get = [&](int arr, int idx, T (&x)[EPL]) -> T {
#pragma unroll
    for (int k = 0; k < EPL; ++k) x[k] = ck[arr * XS + idx * 64 * EPL + l * EPL + k];
    return ck[NA * XS + arr * M + idx];
  };
  auto set_prop = [&](Prop<EPL, T>& pr, const S& s, double en, double w) {
#pragma unroll
    for (int k = 0; k < EPL; ++k) pr.x[k] = s.x[k];
    pr.v = s.v;
    pr.J = s.J;
    pr.energy = en;
    pr.weight = w;
    pr.slpa = fmin(w, 0.0);
  };

  for (int i = l; i < NA * (XS + M); i += 64) ck[i] = T(0);
  __syncthreads();
  // the trajectory: its two ends, sums (the start state included), length
  left = s0;
  right = s0;
  T psx[EPL], vsx[EPL], psv = s0.pv, vsv = s0.uv;
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    psx[k] = s0.px[k];
    vsx[k] = s0.ux[k];
  }
  int ntraj = 0;
  set_prop(prop, s0, H0, 0.0);
  prop.slpa = -INFINITY;
  int e = 0, j = 0;
  bool diverging = false, turning = false;
  while (e < M && !diverging && !turning) {
    const double ud = uniform(nu.u_dir, (size_t)c * M + e, kStreamNutsDirection, (uint32_t)e);
    const int dir = ud < 0.5 ? 1 : -1;
    // ---- subtree of up to 2^e steps from the trajectory's end (trajectory.py:149-317) ----
    S cur = dir > 0 ? right : left;
    Prop<EPL, T> sp;
    T ssx[EPL], ssv_x[EPL], ssp_v = T(0), ssv_v = T(0);
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      ssx[k] = T(0);
      ssv_x[k] = T(0);
    }
    int nsub = 0, st = 0;
    const int max_steps = 1 << e;
    bool div_s = false, term_s = false;
    while (st < max_steps && !term_s && !div_s) {
      step(cur, dir * eps);
      const double en = energy(cur);
      double w = H0 - en;
      if (isnan(w)) w = -INFINITY;
      div_s = -w > div_threshold;
      const double us = uniform(nu.u_step, (size_t)c * NS + j, kStreamNutsStep, (uint32_t)j);
      if (st == 0) {
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
          ssx[k] = cur.px[k];
          ssv_x[k] = cur.ux[k];
        }
        ssp_v = cur.pv;
        ssv_v = cur.uv;
        nsub = 1;
        set_prop(sp, cur, en, w);
      } else {
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
          ssx[k] += cur.px[k];
          ssv_x[k] += cur.ux[k];
        }
        ssp_v += cur.pv;
        ssv_v += cur.uv;
        ++nsub;
        // progressive uniform sampling (proposal.py:198-222)
        const double dw = w - sp.weight;
        const double pacc = 1.0 / (1.0 + exp(-dw));
        const double nw = logaddexp(sp.weight, w), ns = logaddexp(sp.slpa, fmin(w, 0.0));
        if (us < pacc) set_prop(sp, cur, en, w);
        sp.weight = nw;
        sp.slpa = ns;
      }
      ++j;
      // iterative U-turn check (termination.py:38-117)
      const int idx_max = ckpt_idx_max(st), idx_min = ckpt_idx_min(st);
      if ((st & 1) == 0) {
        ck_put(0, idx_max, cur.ux, cur.uv);
        ck_put(1, idx_max, ssv_x, ssv_v);
        if (NA == 4) {
          ck_put(2, idx_max, cur.px, cur.pv);
          ck_put(3, idx_max, ssx, ssp_v);
        }
      }
      __syncthreads();  // one wave: orders the checkpoint writes before the reads below
      for (int i = idx_max; i >= idx_min && !term_s; --i) {
        // the sub-subtree from checkpoint i to the current state (termination.py:89-117)
        T vl[EPL], tmp[EPL], vs[EPL], ml[EPL], ms[EPL];
        const T vlv = ck_get(0, i, vl);
        const T tv = ck_get(1, i, tmp);
#pragma unroll
        for (int k = 0; k < EPL; ++k) vs[k] = ssv_x[k] - tmp[k] + vl[k];
        const T vsv_i = ssv_v - tv + vlv;
        T mlv = T(0), msv = T(0);
        if (NA == 4) {
          mlv = ck_get(2, i, ml);
          const T tm = ck_get(3, i, tmp);
#pragma unroll
          for (int k = 0; k < EPL; ++k) ms[k] = ssx[k] - tmp[k] + ml[k];
          msv = ssp_v - tm + mlv;
        } else {
#pragma unroll
          for (int k = 0; k < EPL; ++k) ml[k] = ms[k] = T(0);
        }
        term_s = is_turning(ml, mlv, cur.px, cur.pv, ms, msv, vl, vlv, cur.ux, cur.uv, vs, vsv_i);
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
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      psx[k] += ssx[k];
      vsx[k] += ssv_x[k];
    }
    psv += ssp_v;
    vsv += ssv_v;
    ntraj += nsub;
    const bool turn_full = is_turning(left.px, left.pv, right.px, right.pv, psx, psv, left.ux, left.uv, right.ux,
                                      right.uv, vsx, vsv);
    ++e;
    diverging = div_s;
    turning = term_s || turn_full;
  }
  return TreeOut{e, ntraj, diverging, turning};
}

// Funnel log density (examples/funnel/main.py:34-39) from v and |x|^2.
__device__ __forceinline__ double funnel_logp(double v, double sx, double sigma, int Dx) {
  return -0.5 * v * v / (sigma * sigma) - log(sigma) - 0.5 * kLog2Pi - 0.5 * exp(-v) * sx - 0.5 * (double)Dx * v -
         0.5 * (double)Dx * kLog2Pi;
}

// Outputs shared by both kernels: the proposal's position, log density and
// gradient (closed forms at its position), optionally its volume adjustment,
// the trajectory's ends and the rest of NUTSInfo.
template <int EPL, class T>
__device__ void write_outputs(const Prop<EPL, T>& prop, const T (&lx)[EPL], T lv, const T (&rx)[EPL], T rv,
                              const TreeOut& t, int c, int D, double sigma, void* q_out, double* logp_out,
                              void* g_out, double* J_out, const NutsArgs& nu) {
  const int l = threadIdx.x, Dx = D - 1, j0 = l * EPL;
  const size_t base = (size_t)c * D;
  double sxp = 0.0;
#pragma unroll
  for (int k = 0; k < EPL; ++k) sxp += (double)prop.x[k] * (double)prop.x[k];
  sxp = group_sum<64>(sxp);
  const double pv = (double)prop.v;
  const double ep = exp(-pv);
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    if (j0 + k >= Dx) continue;
    static_cast<T*>(q_out)[base + j0 + k] = prop.x[k];
    if (g_out) static_cast<T*>(g_out)[base + j0 + k] = (T)(-ep * (double)prop.x[k]);
    if (nu.left_position) static_cast<T*>(nu.left_position)[base + j0 + k] = lx[k];
    if (nu.right_position) static_cast<T*>(nu.right_position)[base + j0 + k] = rx[k];
  }
  if (l == 0) {
    static_cast<T*>(q_out)[base + Dx] = prop.v;
    if (g_out) static_cast<T*>(g_out)[base + Dx] = (T)(-pv / (sigma * sigma) - 0.5 * Dx + 0.5 * ep * sxp);
    if (nu.left_position) static_cast<T*>(nu.left_position)[base + Dx] = lv;
    if (nu.right_position) static_cast<T*>(nu.right_position)[base + Dx] = rv;
    logp_out[c] = funnel_logp(pv, sxp, sigma, Dx);
    if (J_out) J_out[c] = prop.J;
    if (nu.energy) nu.energy[c] = prop.energy;
    if (nu.acceptance_rate) nu.acceptance_rate[c] = exp(prop.slpa) / (double)t.num_states;
    if (nu.is_divergent) nu.is_divergent[c] = (uint8_t)t.diverging;
    if (nu.is_turning) nu.is_turning[c] = (uint8_t)t.turning;
    if (nu.num_expansions) nu.num_expansions[c] = t.expansions;
    if (nu.num_steps) nu.num_steps[c] = t.num_states;
  }
}

}  // namespace nuts

using namespace nuts;

// ---------------------------------------------------------------------------
// NUTS with the RMHMC implicit midpoint, Monge metric (geomjax/rmhmc/nuts.py)
// ---------------------------------------------------------------------------
template <int EPL, class T>
__global__ void __launch_bounds__(64) funnel_nuts_kernel(const FunnelArgs a, const NutsArgs nu) {
  extern __shared__ __align__(16) double ck_raw[];
  T* const ck = reinterpret_cast<T*>(ck_raw);
  using RS = RState<EPL, T>;
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int l = threadIdx.x;
  const int D = a.D, Dx = a.D - 1;
  const int j0 = l * EPL;
  // the state's arithmetic in T, energies in float64
  const T a2 = (T)a.alpha2;
  const T alpha = (T)sqrt(a.alpha2);
  const T inv_s2 = (T)(1.0 / (a.sigma * a.sigma));
  const T halfDx = (T)(0.5 * Dx);
  const double a2d = a.alpha2, inv_s2d = 1.0 / (a.sigma * a.sigma), halfDxd = 0.5 * Dx;
  const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
  const size_t base = (size_t)c * D;
  auto on = [&](int k) { return j0 + k < Dx; };

  // With a.final_q set the launch runs a.T transitions of the chain (the
  // state stays in registers; the launch's time is the chains' summed
  // trajectory lengths, not one transition's longest); else one transition
  // with every NUTSInfo output.  Host-fed draws are then [T, C, ...].
  const bool multi = a.final_q != nullptr;
  const int ntrans = multi ? a.T : 1;
  T qx[EPL], qv;
  double logp_c = a.logp_in[c];
#pragma unroll
  for (int k = 0; k < EPL; ++k) qx[k] = on(k) ? static_cast<const T*>(a.q_in)[base + j0 + k] : T(0);
  qv = static_cast<const T*>(a.q_in)[base + Dx];
  int64_t n_steps = 0;

  // velocity u = G^-1 p at a state whose sums sx, xpx are current
  auto set_velocity = [&](RS& s) {
    const FieldCoef<T> f = field_coef<T>(s.v, s.pv, s.sx, s.xpx, a2, inv_s2, halfDx);
#pragma unroll
    for (int k = 0; k < EPL; ++k) s.ux[k] = s.px[k] + f.ce * s.x[k];
    s.uv = f.uv;
  };
  // kinetic energy T(q, p) = -log N(p; 0, G(q)) (geomjax/rmhmc/metrics.py:66-69), float64
  // (for a float32 state: sums of the widened coordinates); returns |x|^2 too
  auto kinetic = [&](const RS& s, double& sxd) -> double {
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const double xk = (double)s.x[k], pk = (double)s.px[k];
      d1 += xk * xk;
      d2 += xk * pk;
      d3 += pk * pk;
    }
    const double v = (double)s.v, pv = (double)s.pv;
    sxd = group_sum<64>(d1);
    const double xp = group_sum<64>(d2);
    const double pp = group_sum<64>(d3) + pv * pv;
    const double e = exp(-v);
    const double gv = -v * inv_s2d - halfDxd + 0.5 * e * sxd;
    const double gp = -e * xp + gv * pv;
    const double gg = e * e * sxd + gv * gv;
    const double cp = monge_perp_coef(gp, gg);
    double d4 = 0.0;
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const double r = (double)s.px[k] + cp * e * (double)s.x[k];
      d4 += r * r;
    }
    const double rv = pv - cp * gv;
    const double perp2 = group_sum<64>(d4) + rv * rv;
    return 0.5 * monge_quad(pp, perp2, gg, a2d) + 0.5 * log(1.0 + a2d * gg) + 0.5 * (double)D * kLog2Pi;
  };
  auto energy = [&](const RS& s) -> double {
    double sxd;
    const double kin = kinetic(s, sxd);
    return -funnel_logp((double)s.v, sxd, a.sigma, Dx) + kin;
  };
  // one implicit-midpoint step of size hstep = direction * eps (integrators.py:92-156)
  auto one_step = [&](RS& s, double hstep) {
    const T h = (T)(0.5 * hstep);
    T x0[EPL], p0[EPL];
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      x0[k] = s.x[k];
      p0[k] = s.px[k];
    }
    const T v0 = s.v, pv0 = s.pv;
    int n = -1;
    double norm = 0.0;
    if constexpr (sizeof(T) == sizeof(float)) {
      // the iterates in span{x0, p0} (funnel_rmhmc.hip): x = ca x0 + cb p0, p = cg x0 + cd p0
      T spp = T(0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) spp += p0[k] * p0[k];
      const T Sxx = s.sx, Sxp = s.xpx, Spp = group_sum<64>(spp);
      T ca = T(1), cb = T(0), cg = T(0), cd = T(1);
      while (true) {
        const FieldCoef<T> f = field_coef<T>(s.v, s.pv, s.sx, s.xpx, a2, inv_s2, halfDx);
        const T na = T(1) + h * (cg + f.ce * ca), nb = h * (cd + f.ce * cb);
        const T ng = -h * (f.A * ca + f.B * cg), nd = T(1) - h * (f.A * cb + f.B * cd);
        const T da = na - ca, db = nb - cb, dg = ng - cg, dd = nd - cd;
        T mx = T(0);
#pragma unroll
        for (int k = 0; k < EPL; ++k)
          mx = fmaxf(mx, fmaxf(fabsf(da * x0[k] + db * p0[k]), fabsf(dg * x0[k] + dd * p0[k])));
        mx = group_reduce<64, OpMax>(mx);
        const T vn = v0 + h * f.uv;
        const T pvn = pv0 - h * f.dHv;
        mx = nan_max(mx, nan_max(fabsf(vn - s.v), fabsf(pvn - s.pv)));
        if (!isfinite(na + nb + ng + nd)) mx = T(NAN);
        ca = na;
        cb = nb;
        cg = ng;
        cd = nd;
        s.v = vn;
        s.pv = pvn;
        s.sx = (ca * ca) * Sxx + (T(2) * ca * cb) * Sxp + (cb * cb) * Spp;
        s.xpx = (ca * cg) * Sxx + (ca * cd + cb * cg) * Sxp + (cb * cd) * Spp;
        norm = (double)mx;
        ++n;
        if (!(n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol)) break;
      }
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        s.x[k] = ca * x0[k] + cb * p0[k];
        s.px[k] = cg * x0[k] + cd * p0[k];
      }
    } else while (true) {
      const FieldCoef<T> f = field_coef<T>(s.v, s.pv, s.sx, s.xpx, a2, inv_s2, halfDx);
      T s1 = T(0), s2 = T(0), mx = T(0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const T xn = x0[k] + h * (s.px[k] + f.ce * s.x[k]);
        const T pn = p0[k] - h * (f.A * s.x[k] + f.B * s.px[k]);
        mx = nan_max(mx, nan_max(fabs(xn - s.x[k]), fabs(pn - s.px[k])));
        s1 += xn * xn;
        s2 += xn * pn;
        s.x[k] = xn;
        s.px[k] = pn;
      }
      const T vn = v0 + h * f.uv;
      const T pvn = pv0 - h * f.dHv;
      mx = group_reduce<64, NanMax>(mx);
      mx = nan_max(mx, nan_max(fabs(vn - s.v), fabs(pvn - s.pv)));
      s.v = vn;
      s.pv = pvn;
      s.sx = group_sum<64>(s1);
      s.xpx = group_sum<64>(s2);
      norm = mx;
      ++n;
      if (!(n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol)) break;
    }
    const FieldCoef<T> f = field_coef<T>(s.v, s.pv, s.sx, s.xpx, a2, inv_s2, halfDx);
    T s1 = T(0), s2 = T(0);
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const T xo = s.x[k], po = s.px[k];
      s.x[k] = xo + h * (po + f.ce * xo);
      s.px[k] = po - h * (f.A * xo + f.B * po);
      s1 += s.x[k] * s.x[k];
      s2 += s.x[k] * s.px[k];
    }
    s.v = s.v + h * f.uv;
    s.pv = s.pv - h * f.dHv;
    s.sx = group_sum<64>(s1);
    s.xpx = group_sum<64>(s2);
    set_velocity(s);
  };

  for (int tr = 0; tr < ntrans; ++tr) {
    const uint64_t tt = a.counter + (uint64_t)tr;
    const NutsUniform uniform{c, tt, k0, k1};
    NutsArgs ntr = nu;
    const int M = nu.max_doublings;
    if (nu.u_dir) {
      ntr.u_dir = nu.u_dir + (size_t)tr * a.C * M;
      ntr.u_merge = nu.u_merge + (size_t)tr * a.C * M;
      ntr.u_step = nu.u_step + (size_t)tr * a.C * ((1 << M) - 1);
    }
    // ---- chain state and momentum p0 = chol(G) z (funnel_rmhmc.hip) ----
    RS s0;
    s0.J = 0.0;
    T z[EPL], zv;
  #pragma unroll
    for (int k = 0; k < EPL; ++k) s0.x[k] = qx[k];
    s0.v = qv;
    const double logp0 = logp_c;
    if (a.z != nullptr) {
      const size_t zb = ((size_t)tr * a.C + c) * D;
  #pragma unroll
      for (int k = 0; k < EPL; ++k) z[k] = on(k) ? static_cast<const T*>(a.z)[zb + j0 + k] : T(0);
      zv = static_cast<const T*>(a.z)[zb + Dx];
    } else {
  #pragma unroll
      for (int k = 0; k < EPL; k += 2) {
        u32x4 r = philox4x32_10(
            u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(j0 + k), (uint32_t)(tt >> 32)}, k0, k1);
        double n0, n1;
        normal2(r, n0, n1);
        z[k] = on(k) ? (T)n0 : T(0);
        if (k + 1 < EPL) z[k + 1] = on(k + 1) ? (T)n1 : T(0);
      }
      u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentumV, (uint32_t)(tt >> 32)}, k0, k1);
      double n0, n1;
      normal2(r, n0, n1);
      zv = (T)n0;
    }
    {
      T loc_inc[EPL], acc = T(0);
  #pragma unroll
      for (int k = 0; k < EPL; ++k) {
        acc += s0.x[k] * s0.x[k];
        loc_inc[k] = acc;
      }
      s0.sx = group_sum<64>(acc);
      const T off1 = group_exclusive_scan<64>(acc, l);
      const T e = exp(-s0.v);
      const T a2e2 = a2 * e * e;
      const T gv = -s0.v * inv_s2 - halfDx + T(0.5) * e * s0.sx;
      T cj[EPL], acc2 = T(0);
  #pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const T tj = T(1) + a2e2 * (off1 + loc_inc[k]);
        const T tjm = T(1) + a2e2 * (off1 + loc_inc[k] - s0.x[k] * s0.x[k]);
        const T aj = -alpha * e * s0.x[k];
        cj[k] = aj * z[k] / sqrt(tj * tjm);
        s0.px[k] = sqrt(tj / tjm) * z[k];
        acc2 += cj[k];
      }
      const T off2 = group_exclusive_scan<64>(acc2, l);
      const T tot2 = group_sum<64>(acc2);
      T run = off2;
  #pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const T aj = -alpha * e * s0.x[k];
        s0.px[k] = on(k) ? s0.px[k] + aj * run : T(0);
        run += cj[k];
      }
      const T txe = T(1) + a2e2 * s0.sx;
      const T tv = txe + a2 * gv * gv;
      s0.pv = sqrt(tv / txe) * zv + alpha * gv * tot2;
    }
    if (!multi && a.momentum != nullptr) {
  #pragma unroll
      for (int k = 0; k < EPL; ++k)
        if (on(k)) static_cast<T*>(a.momentum)[base + j0 + k] = s0.px[k];
      if (l == 0) static_cast<T*>(a.momentum)[base + Dx] = s0.pv;
    }

    {
      T d2 = T(0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) d2 += s0.x[k] * s0.px[k];
      s0.xpx = group_sum<64>(d2);
    }
    set_velocity(s0);
    // H0 with the chain's log density as given (the initial IntegratorState carries it)
    double sx0;
    const double H0 = -logp0 + kinetic(s0, sx0);

    RS left, right;
    Prop<EPL, T> prop;
    const TreeOut t = nuts_tree<EPL>(s0, H0, one_step, energy, uniform, ntr, a.eps, a.div_threshold, ck, c, left,
                                     right, prop);
    if (!multi) {
      write_outputs<EPL>(prop, left.x, left.v, right.x, right.v, t, c, D, a.sigma, a.q_out, a.logp_out, a.g_out,
                         nullptr, nu);
      return;
    }
    // the proposal is the chain's next state (geomjax/rmhmc/nuts.py:150-159)
    double sxp = 0.0;
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      qx[k] = prop.x[k];
      sxp += (double)prop.x[k] * (double)prop.x[k];
    }
    qv = prop.v;
    sxp = group_sum<64>(sxp);
    logp_c = funnel_logp((double)prop.v, sxp, a.sigma, Dx);
    n_steps += t.num_states;
    if (a.samples != nullptr) {
      const size_t sb = ((size_t)tr * a.C + c) * D;
#pragma unroll
      for (int k = 0; k < EPL; ++k)
        if (on(k)) static_cast<T*>(a.samples)[sb + j0 + k] = qx[k];
      if (l == 0) static_cast<T*>(a.samples)[sb + Dx] = qv;
    }
    if (l == 0) {
      if (a.sample_acc) a.sample_acc[(size_t)tr * a.C + c] = (float)(exp(prop.slpa) / (double)t.num_states);
      if (nu.num_steps) nu.num_steps[(size_t)tr * a.C + c] = t.num_states;
    }
  }
  // final state of a multi-transition launch
  double sx = 0.0;
#pragma unroll
  for (int k = 0; k < EPL; ++k) sx += (double)qx[k] * (double)qx[k];
  sx = group_sum<64>(sx);
  const double e = exp(-(double)qv);
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    if (!on(k)) continue;
    static_cast<T*>(a.final_q)[base + j0 + k] = qx[k];
    if (a.final_g) static_cast<T*>(a.final_g)[base + j0 + k] = (T)(-e * (double)qx[k]);
  }
  if (l == 0) {
    static_cast<T*>(a.final_q)[base + Dx] = qv;
    if (a.final_g) static_cast<T*>(a.final_g)[base + Dx] = (T)(-(double)qv * inv_s2d - halfDxd + 0.5 * e * sx);
    a.final_logp[c] = logp_c;
    if (a.num_fevals) a.num_fevals[c] = n_steps;
  }
}

// ---------------------------------------------------------------------------
// NUTS with the LMC-Monge integrator (geomjax/lmcmonge/nuts.py:80-207)
// ---------------------------------------------------------------------------
template <int EPL, class T>
__global__ void __launch_bounds__(64) funnel_nuts_lmc_kernel(const FunnelLmcArgs a, const NutsArgs nu) {
  using namespace lmc_ops;
  using LS = LState<EPL, T>;
  extern __shared__ __align__(16) double ck_raw[];
  T* const ck = reinterpret_cast<T*>(ck_raw);
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int l = threadIdx.x;
  const int D = a.D, Dx = a.D - 1;
  const int j0 = l * EPL;
  // the state's arithmetic in T (float32: the tuned form), energies in float64
  const T a2 = (T)a.alpha2;
  const T alpha = (T)sqrt(a.alpha2);
  const T inv_s2 = (T)(1.0 / (a.sigma * a.sigma));
  const T halfDx = (T)(0.5 * Dx);
  const bool exact_sm = a.exact_sm != 0;
  const NutsUniform uniform{c, a.counter, (uint32_t)a.seed, (uint32_t)(a.seed >> 32)};
  const size_t base = (size_t)c * D;

  // diagonal inverse mass (the reference's inverse_mass_matrix)
  T mi[EPL];
#pragma unroll
  for (int k = 0; k < EPL; ++k) mi[k] = j0 + k < Dx ? ld<T>(a.inv_mass, j0 + k) : T(1);
  const T miv = ld<T>(a.inv_mass, Dx);
  double sum_log_mi;
  {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < EPL; ++k)
      if (j0 + k < Dx) s += log((double)mi[k]);
    sum_log_mi = group_sum<64>(s) + log((double)miv);
  }

  // momentum = M v + alpha2 L (v . g) g (the reference's metric_vector_product with its
  // extra L, geomjax/lmcmonge/metrics.py:243-255; exact: alpha2 (v . g) g) from the
  // velocity and the position's scalars
  auto set_momentum = [&](LS& s) {
    T xv = T(0);
#pragma unroll
    for (int k = 0; k < EPL; ++k) xv += s.x[k] * s.ux[k];
    xv = group_sum<64>(xv);
    const T vg = -s.P.e * xv + s.uv * s.P.gv;
    const T coef = a.exact_momentum ? a2 * vg : a2 * s.P.L * vg;
#pragma unroll
    for (int k = 0; k < EPL; ++k) s.px[k] = j0 + k < Dx ? s.ux[k] / mi[k] + coef * (-s.P.e * s.x[k]) : T(0);
    s.pv = s.uv / miv + coef * s.P.gv;
  };
  // energy -l + K - J (lmcmonge_energy, geomjax/lmcmonge/metrics.py:267-284); the log
  // density from the widened coordinates, the kinetic energy's sums in float64
  auto energy = [&](const LS& s) -> double {
    double sxd;
    if constexpr (sizeof(T) == sizeof(double)) {
      sxd = s.P.Sx;
    } else {
      double d1 = 0.0;
#pragma unroll
      for (int k = 0; k < EPL; ++k) d1 += (double)s.x[k] * (double)s.x[k];
      sxd = group_sum<64>(d1);
    }
    return -funnel_logp((double)s.v, sxd, a.sigma, Dx) +
           lmc_kinetic<T, 64, EPL>(s.P, s.x, s.ux, s.uv, mi, miv, j0, Dx, sum_log_mi, a.alpha2) - s.J;
  };
  // one lan_integrator step of size hstep = direction * eps (integrators.py:52-156)
  auto one_step = [&](LS& s, double hstep) {
    double dJ = 0.0;
    lmc_step<T, 64, EPL>(s.P, s.x, s.v, mi, s.ux, s.uv, dJ, (T)hstep, (T)(0.5 * hstep), a2, inv_s2, halfDx, miv,
                         exact_sm);
    s.J += dJ;
    set_momentum(s);
  };

  // ---- state, velocity draw v0 = chol(M^-1 - alpha2 dl dl^T) z, momentum, H0 (lmc.py:151-210) ----
  LS s0;
#pragma unroll
  for (int k = 0; k < EPL; ++k) s0.x[k] = j0 + k < Dx ? ld<T>(a.q_in, base + j0 + k) : T(0);
  s0.v = ld<T>(a.q_in, base + Dx);
  s0.J = a.J_in[c];
  T loc_inc[EPL];
  s0.P = position_scalars<T, 64, EPL>(s0.x, s0.v, mi, loc_inc, l, a2, inv_s2, halfDx, miv);
  T z[EPL], zv;
  double uacc;
  transition_noise<T, EPL>(a, (size_t)c, 0, j0, Dx, true, z, zv, uacc);
  draw_velocity<T, 64, EPL>(s0.P, s0.x, mi, loc_inc, z, zv, a2, alpha, miv, l, j0, Dx, s0.ux, s0.uv);
  set_momentum(s0);
  if (a.velocity != nullptr) {  // NUTSInfo.momentum holds the initial velocity (lmcmonge/nuts.py:386-396)
#pragma unroll
    for (int k = 0; k < EPL; ++k)
      if (j0 + k < Dx) st<T>(a.velocity, base + j0 + k, s0.ux[k]);
    if (l == 0) st<T>(a.velocity, base + Dx, s0.uv);
  }
  // H0 with the chain's log density and volume adjustment as given
  const double H0 = -a.logp_in[c] +
                    lmc_kinetic<T, 64, EPL>(s0.P, s0.x, s0.ux, s0.uv, mi, miv, j0, Dx, sum_log_mi, a.alpha2) - s0.J;

  LS left, right;
  Prop<EPL, T> prop;
  const TreeOut t = nuts_tree<EPL>(s0, H0, one_step, energy, uniform, nu, a.eps, a.div_threshold, ck, c, left, right,
                                   prop);
  write_outputs<EPL>(prop, left.x, left.v, right.x, right.v, t, c, D, a.sigma, a.q_out, a.logp_out, a.g_out, a.J_out,
                     nu);
}

namespace {

// coordinates per lane of a 64-lane chain: 1, 2 or 4 (D - 1 <= 256)
int nuts_epl(int D) {
  const int epl = (D - 1 + 63) / 64;
  return epl <= 1 ? 1 : (epl <= 2 ? 2 : 4);
}

// [NA][M][64 EPL] + [NA][M] checkpoint values of the state's type
size_t nuts_smem(const NutsArgs& n, int epl, size_t elem = sizeof(double)) {
  const int na = nuts_checkpoint_arrays(n.criterion);
  return ((size_t)na * n.max_doublings * 64 * epl + (size_t)na * n.max_doublings) * elem;
}

// up to 4 x 12 x 256 x 8 B = 96 KiB of checkpoints: past the 64 KiB default of dynamic LDS
template <class K, class A>
hipError_t launch_one_nuts(K kern, const A& a, const NutsArgs& n, size_t smem, hipStream_t stream) {
  if (smem > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(a.C), dim3(64), smem, stream, a, n);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_funnel_nuts(const FunnelArgs& a, const NutsArgs& n, int dtype, hipStream_t stream) {
  const int epl = nuts_epl(a.D);
  if (dtype == GJX_F32) {
    const size_t smem = nuts_smem(n, epl, sizeof(float));
    if (epl == 1) return launch_one_nuts(funnel_nuts_kernel<1, float>, a, n, smem, stream);
    if (epl == 2) return launch_one_nuts(funnel_nuts_kernel<2, float>, a, n, smem, stream);
    return launch_one_nuts(funnel_nuts_kernel<4, float>, a, n, smem, stream);
  }
  const size_t smem = nuts_smem(n, epl);
  if (epl == 1) return launch_one_nuts(funnel_nuts_kernel<1, double>, a, n, smem, stream);
  if (epl == 2) return launch_one_nuts(funnel_nuts_kernel<2, double>, a, n, smem, stream);
  return launch_one_nuts(funnel_nuts_kernel<4, double>, a, n, smem, stream);
}

hipError_t launch_funnel_nuts_lmc(const FunnelLmcArgs& a, const NutsArgs& n, int dtype, hipStream_t stream) {
  const int epl = nuts_epl(a.D);
  if (dtype == GJX_F32) {
    const size_t smem = nuts_smem(n, epl, sizeof(float));
    if (epl == 1) return launch_one_nuts(funnel_nuts_lmc_kernel<1, float>, a, n, smem, stream);
    if (epl == 2) return launch_one_nuts(funnel_nuts_lmc_kernel<2, float>, a, n, smem, stream);
    return launch_one_nuts(funnel_nuts_lmc_kernel<4, float>, a, n, smem, stream);
  }
  const size_t smem = nuts_smem(n, epl);
  if (epl == 1) return launch_one_nuts(funnel_nuts_lmc_kernel<1, double>, a, n, smem, stream);
  if (epl == 2) return launch_one_nuts(funnel_nuts_lmc_kernel<2, double>, a, n, smem, stream);
  return launch_one_nuts(funnel_nuts_lmc_kernel<4, double>, a, n, smem, stream);
}

}  // namespace gjx
