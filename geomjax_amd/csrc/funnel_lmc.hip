// Lagrangian Monte Carlo with the Monge metric (lmcmonge) on Neal's funnel.
//
// Reference: geomjax/lmcmonge/lmc.py:130-237,472-565 (transition, proposal,
// flip), geomjax/lmcmonge/integrators.py:52-194 (lan_integrator,
// half_step_omega), geomjax/lmcmonge/metrics.py:104-284 (velocity draw,
// kinetic energy, weighted gradients, momentum), examples/funnel/main.py:34-39.
//
// Metric G = M + alpha2 g g^T with diagonal M = diag(1/mi) (mi = the
// reference's inverse_mass_matrix), g = grad l, L = 1 + alpha2 g^T M^-1 g.
// With x = q[0:D-1], v = q[D-1], e = exp(-v) every vector of the half step
// is a per-coordinate expression in (x_k, vel_k, mi_k) and per-chain scalars:
//   ghat_x = -e x / sqrt(L)              (logdensity_grad_norm)
//   dl_x = mi ghat_x                     (dl_ig)
//   Hdl_x = e (-dl_x + x dl_v) / sqrt(L) (Hdl_ig: funnel Hessian times dl_ig)
//   igHdl_x = mi Hdl_x                   (ig_Hdl_ig)
//   hv_x = e (-vel_x + x vel_v) / sqrt(L)  (logdensity_hvp_velocity_norm)
// and the dot products half_step_omega needs are linear in the velocity, so
// one group reduction of eight sums per half step gives all of them:
//   v_t.hv  = v.hv - (eps/2) sqrt(L) (dphi_ig.hv - alpha2 (dphi.dl) (dl.hv))
//   Hdl.v'  = Hdl.v - (eps/2) sqrt(L) (Hdl.dphi_ig - alpha2 (dphi.dl)(Hdl.dl)) - c (Hdl.dl).
// The velocity draw v0 = chol(M^-1 - alpha2 dl dl^T) z uses the exact lower
// factor of a diagonal-minus-rank-one matrix:
//   a_j = sqrt(alpha2) dl_j / sqrt(mi_j),  t_j = 1 - sum_{k<=j} a_k^2,
//   v_j = sqrt(mi_j) (sqrt(t_j/t_{j-1}) z_j - a_j sum_{k<j} a_k z_k / sqrt(t_k t_{k-1})).
#include "common.h"
#include "dual.h"
#include "funnel_lmc.h"
#include "funnel_lmc_ops.h"

namespace gjx {

using namespace lmc_ops;

template <class T, int LPC, int EPL, bool STEP>
__global__ void __launch_bounds__(256) funnel_lmc_kernel(const FunnelLmcArgs a) {
  using R = T;
  const int tid = blockIdx.x * 256 + threadIdx.x;
  const int chain = tid / LPC;
  const int lg = tid % LPC;
  const bool live = chain < a.C;
  const size_t c = live ? (size_t)chain : 0;
  const int D = a.D, Dx = a.D - 1;
  const int j0 = lg * EPL;
  const R eps = R(a.eps);
  const R h = R(0.5 * a.eps);
  const R a2 = R(a.alpha2);
  const R alpha = R(sqrt(a.alpha2));
  const R inv_s2 = R(1.0 / (a.sigma * a.sigma));
  const R halfDx = R(0.5 * Dx);
  const bool exact_sm = a.exact_sm != 0;
  const int T_steps = STEP ? 1 : a.T;
#define GJX_ON(k) (j0 + (k) < Dx)

  // diagonal inverse mass (the reference's inverse_mass_matrix, "fixed" power)
  R mi[EPL];
#pragma unroll
  for (int k = 0; k < EPL; ++k) mi[k] = GJX_ON(k) ? ld<T>(a.inv_mass, j0 + k) : R(1);
  const R miv = ld<T>(a.inv_mass, Dx);
  double sum_log_mi;  // sum_j log(inverse mass_j) = -sum_j log(mass_j)
  {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < EPL; ++k)
      if (GJX_ON(k)) s += log((double)mi[k]);
    sum_log_mi = group_sum<LPC>(s) + log((double)miv);
  }

  // chain state: LMCState(position, logdensity, logdensity_grad, volume_adjustment)
  R x[EPL];
  R v;
  {
    const size_t base = c * D;
#pragma unroll
    for (int k = 0; k < EPL; ++k) x[k] = (live && GJX_ON(k)) ? ld<T>(a.q_in, base + j0 + k) : R(0);
    v = live ? ld<T>(a.q_in, base + Dx) : R(0);
  }
  double logp = live ? a.logp_in[c] : 0.0;
  double J = live ? a.J_in[c] : 0.0;
  int n_acc = 0;

  for (int t = 0; t < T_steps; ++t) {
    // ---- position quantities at q0 (lmc.py:177-200) ----
    R loc_inc[EPL];
    PosScalars<R> P = position_scalars<R, LPC, EPL>(x, v, mi, loc_inc, lg, a2, inv_s2, halfDx, miv);

    // ---- noise and velocity v0 = chol(M^-1 - alpha2 dl dl^T) z  (metrics.py:155-166) ----
    R z[EPL];
    R zv;
    double uacc;
    transition_noise<T, EPL>(a, c, t, j0, Dx, live, z, zv, uacc);
    R vx[EPL];
    R vv;
    draw_velocity<R, LPC, EPL>(P, x, mi, loc_inc, z, zv, a2, alpha, miv, lg, j0, Dx, vx, vv);
    if (STEP && a.velocity != nullptr && live) {
#pragma unroll
      for (int k = 0; k < EPL; ++k)
        if (GJX_ON(k)) st<T>(a.velocity, c * D + j0 + k, vx[k]);
      if (lg == 0) st<T>(a.velocity, c * D + Dx, vv);
    }

    // ---- kinetic energy (metrics.py:168-186), float64 sums ----
    auto kinetic = [&](const PosScalars<R>& Q, const R (&xx)[EPL], const R (&vel)[EPL], R velv) -> double {
      return lmc_kinetic<R, LPC, EPL>(Q, xx, vel, velv, mi, miv, j0, Dx, sum_log_mi, a.alpha2);
    };
    const double K0 = kinetic(P, x, vx, vv);

    // ---- trajectory (trajectory.py:121-139 over lan_integrator.one_step) ----
    R xc[EPL];
    R vc = v;
#pragma unroll
    for (int k = 0; k < EPL; ++k) xc[k] = x[k];
    double dJ = 0.0;
    // chains of the dynamic kernel stop after their own number of steps; every
    // cross-lane sum stays inside a chain's aligned lane group, so a group whose
    // chain has stopped can sit out the step while the others reduce
    const int Ls = live ? chain_steps(a.Lc, a.L, a.C, STEP ? 0 : t, (int)c) : 0;
    for (int s = 0; s < a.L; ++s) {
      const bool on = s < Ls;
      if (!__any(on)) break;
      if (on) lmc_step<R, LPC, EPL>(P, xc, vc, mi, vx, vv, dJ, eps, h, a2, inv_s2, halfDx, miv, exact_sm);
    }

    // ---- energy at the end (lmcmonge_energy, metrics.py:267-284) ----
    const double K1 = kinetic(P, xc, vx, vv);
    double logp1;
    {
      double d1 = 0.0;
#pragma unroll
      for (int k = 0; k < EPL; ++k) d1 += (double)xc[k] * (double)xc[k];
      const double Sx = group_sum<LPC>(d1);
      const double vd = (double)vc, s2 = a.sigma * a.sigma;
      logp1 = -0.5 * vd * vd / s2 - log(a.sigma) - 0.5 * kLog2Pi - 0.5 * exp(-vd) * Sx - 0.5 * (double)Dx * vd -
              0.5 * (double)Dx * kLog2Pi;
    }
    const double J1 = J + dJ;
    const double H1 = -logp1 + K1 - J1;
    double dH = (K0 - logp) - (K1 - logp1) + dJ;  // H0 - H1 with the carried J cancelled exactly
    if (isnan(dH)) dH = -INFINITY;
    const double p_acc = fmin(exp(dH), 1.0);
    const bool accept = uacc < p_acc;

    if (STEP && live) {
      if (lg == 0) {
        if (a.acc_rate) a.acc_rate[c] = p_acc;
        if (a.accepted) a.accepted[c] = (uint8_t)accept;
        if (a.divergent) a.divergent[c] = (uint8_t)(-dH > a.div_threshold);
        if (a.energy) a.energy[c] = H1;
        if (a.prop_logp) a.prop_logp[c] = logp1;
        if (a.prop_L) a.prop_L[c] = (double)P.L;
        if (a.prop_J) a.prop_J[c] = J1;
        if (a.prop_weight) a.prop_weight[c] = dH;
        if (a.prop_slpa) a.prop_slpa[c] = fmin(dH, 0.0);
      }
      // v . g and x . v at the end for the momentum and hv
      R xv = R(0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) xv += xc[k] * vx[k];
      xv = group_sum<LPC>(xv);
      const R vg = -P.e * xv + vv * P.gv;
      const R mom_coef = a.exact_momentum ? a2 * vg : a2 * P.L * vg;  // metrics.py:243-255 (extra L)
      const R hv_v = (P.e * xv - (inv_s2 + R(0.5) * P.e * P.Sx) * vv) * P.isL;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        if (!GJX_ON(k)) continue;
        const size_t i = c * D + j0 + k;
        const R gx = -P.e * xc[k];
        const R gn = gx * P.isL;
        const R dl = mi[k] * gn;
        const R Hdl = P.e * (-dl + xc[k] * P.dlv) * P.isL;
        if (a.prop_q) st<T>(a.prop_q, i, xc[k]);
        if (a.prop_p) st<T>(a.prop_p, i, -(vx[k] / mi[k] + mom_coef * gx));
        if (a.prop_v) st<T>(a.prop_v, i, -vx[k]);
        if (a.prop_gnorm) st<T>(a.prop_gnorm, i, gn);
        if (a.prop_dl_ig) st<T>(a.prop_dl_ig, i, dl);
        if (a.prop_Hdl_ig) st<T>(a.prop_Hdl_ig, i, Hdl);
        if (a.prop_ig_Hdl_ig) st<T>(a.prop_ig_Hdl_ig, i, mi[k] * Hdl);
        if (a.prop_hvn) st<T>(a.prop_hvn, i, P.e * (-vx[k] + xc[k] * vv) * P.isL);
      }
      if (lg == 0) {
        const size_t i = c * D + Dx;
        if (a.prop_q) st<T>(a.prop_q, i, vc);
        if (a.prop_p) st<T>(a.prop_p, i, -(vv / miv + mom_coef * P.gv));
        if (a.prop_v) st<T>(a.prop_v, i, -vv);
        if (a.prop_gnorm) st<T>(a.prop_gnorm, i, P.gv * P.isL);
        if (a.prop_dl_ig) st<T>(a.prop_dl_ig, i, P.dlv);
        if (a.prop_Hdl_ig) st<T>(a.prop_Hdl_ig, i, P.Hdlv);
        if (a.prop_ig_Hdl_ig) st<T>(a.prop_ig_Hdl_ig, i, P.igHdlv);
        if (a.prop_hvn) st<T>(a.prop_hvn, i, hv_v);
      }
      // new LMCState (lmc.py:226-233): logdensity_grad = ghat * sqrt(L) = g
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        if (!GJX_ON(k)) continue;
        const size_t i = c * D + j0 + k;
        st<T>(a.q_out, i, accept ? xc[k] : x[k]);
        if (a.g_out) st<T>(a.g_out, i, accept ? (-P.e * xc[k] * P.isL) * P.sL : ld<T>(a.g_in, i));
      }
      if (lg == 0) {
        const size_t i = c * D + Dx;
        st<T>(a.q_out, i, accept ? vc : v);
        if (a.g_out) st<T>(a.g_out, i, accept ? (P.gv * P.isL) * P.sL : ld<T>(a.g_in, i));
        a.logp_out[c] = accept ? logp1 : logp;
        if (a.J_out) a.J_out[c] = accept ? J1 : J;
      }
    }

    if (accept) {
#pragma unroll
      for (int k = 0; k < EPL; ++k) x[k] = xc[k];
      v = vc;
      logp = logp1;
      J = J1;
      ++n_acc;
    }
    if (!STEP) {
      if (a.samples != nullptr && live) {
        const size_t base = ((size_t)t * a.C + c) * D;
#pragma unroll
        for (int k = 0; k < EPL; ++k)
          if (GJX_ON(k)) st<T>(a.samples, base + j0 + k, x[k]);
        if (lg == 0) st<T>(a.samples, base + Dx, v);
      }
      if (a.sample_acc != nullptr && live && lg == 0) a.sample_acc[(size_t)t * a.C + c] = (float)p_acc;
    }
  }

  if (!STEP) {
    if (live) {
#pragma unroll
      for (int k = 0; k < EPL; ++k)
        if (GJX_ON(k)) st<T>(a.final_q, c * D + j0 + k, x[k]);
      if (lg == 0) {
        st<T>(a.final_q, c * D + Dx, v);
        a.final_logp[c] = logp;
        if (a.final_J) a.final_J[c] = J;
        if (a.num_accepted) a.num_accepted[c] = n_acc;
      }
    }
    if (a.final_g != nullptr) {
      R s1 = R(0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) s1 += x[k] * x[k];
      const R sxf = group_sum<LPC>(s1);
      const R e = exp(-v);
      const R gv = -v * inv_s2 - halfDx + R(0.5) * e * sxf;
      if (live) {
#pragma unroll
        for (int k = 0; k < EPL; ++k)
          if (GJX_ON(k)) st<T>(a.final_g, c * D + j0 + k, -e * x[k]);
        if (lg == 0) st<T>(a.final_g, c * D + Dx, gv);
      }
    }
  }
#undef GJX_ON
}

// Forward-mode tangent of one LMC-Monge transition in alpha2: the derivative of
// the proposal position that geomjax/adaptation/chees_adaptation_lmcmonge.py:
// 475-519 takes with jax.jacfwd.  The float64 transition runs on
// Dual<double, 1> seeded d alpha2 = 1, through the velocity draw and every
// integrator step (the helpers above, as in funnel_lmc_kernel); the noise is
// the primal kernel's at the same (seed, counter), so value and tangent belong
// to the trajectory the primal launch integrates.  Outputs the proposal position
// (prop_q, optional) and its derivative (d_prop_q).
template <int LPC, int EPL>
__global__ void __launch_bounds__(256) funnel_lmc_alpha2_tangent_kernel(const FunnelLmcArgs a) {
  using R = Dual<double, 1>;
  const int tid = blockIdx.x * 256 + threadIdx.x;
  const int chain = tid / LPC;
  const int lg = tid % LPC;
  const bool live = chain < a.C;
  const size_t c = live ? (size_t)chain : 0;
  const int D = a.D, Dx = a.D - 1;
  const int j0 = lg * EPL;
  const R eps = R(a.eps);
  const R h = R(0.5 * a.eps);
  R a2 = R(a.alpha2);
  a2.d[0] = 1.0;
  const R alpha = sqrt(a2);
  const R inv_s2 = R(1.0 / (a.sigma * a.sigma));
  const R halfDx = R(0.5 * Dx);
  const bool exact_sm = a.exact_sm != 0;
  R mi[EPL];
#pragma unroll
  for (int k = 0; k < EPL; ++k) mi[k] = R(j0 + k < Dx ? ld<double>(a.inv_mass, j0 + k) : 1.0);
  const R miv = R(ld<double>(a.inv_mass, Dx));
  R x[EPL];
#pragma unroll
  for (int k = 0; k < EPL; ++k) x[k] = R((live && j0 + k < Dx) ? ld<double>(a.q_in, c * D + j0 + k) : 0.0);
  const R v = R(live ? ld<double>(a.q_in, c * D + Dx) : 0.0);
  R loc_inc[EPL];
  PosScalars<R> P = position_scalars<R, LPC, EPL>(x, v, mi, loc_inc, lg, a2, inv_s2, halfDx, miv);
  double zd[EPL], zvd, uacc;
  transition_noise<double, EPL>(a, c, 0, j0, Dx, live, zd, zvd, uacc);
  R z[EPL];
#pragma unroll
  for (int k = 0; k < EPL; ++k) z[k] = R(zd[k]);
  R vx[EPL];
  R vv;
  draw_velocity<R, LPC, EPL>(P, x, mi, loc_inc, z, R(zvd), a2, alpha, miv, lg, j0, Dx, vx, vv);
  R xc[EPL];
  R vc = v;
#pragma unroll
  for (int k = 0; k < EPL; ++k) xc[k] = x[k];
  double dJ = 0.0;
  const int Ls = live ? chain_steps(a.Lc, a.L, a.C, 0, (int)c) : 0;
  for (int s = 0; s < a.L; ++s) {
    const bool on = s < Ls;
    if (!__any(on)) break;
    if (on) lmc_step<R, LPC, EPL>(P, xc, vc, mi, vx, vv, dJ, eps, h, a2, inv_s2, halfDx, miv, exact_sm);
  }
  if (!live) return;
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    if (j0 + k >= Dx) continue;
    if (a.prop_q) st<double>(a.prop_q, c * D + j0 + k, xc[k].v);
    a.d_prop_q[c * D + j0 + k] = xc[k].d[0];
  }
  if (lg == 0) {
    if (a.prop_q) st<double>(a.prop_q, c * D + Dx, vc.v);
    a.d_prop_q[c * D + Dx] = vc.d[0];
  }
}

namespace {

template <class T, int LPC, int EPL>
hipError_t launch_one(const FunnelLmcArgs& a, hipStream_t stream) {
  const long long lanes = (long long)a.C * LPC;
  const unsigned blocks = (unsigned)((lanes + 255) / 256);
  if (a.q_out != nullptr)
    hipLaunchKernelGGL((funnel_lmc_kernel<T, LPC, EPL, true>), dim3(blocks), dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL((funnel_lmc_kernel<T, LPC, EPL, false>), dim3(blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <class T, int LPC>
hipError_t launch_epl(const FunnelLmcArgs& a, int epl, hipStream_t stream) {
  switch (epl) {
    case 1: return launch_one<T, LPC, 1>(a, stream);
    case 2: return launch_one<T, LPC, 2>(a, stream);
    case 4: return launch_one<T, LPC, 4>(a, stream);
    case 8: return launch_one<T, LPC, 8>(a, stream);
    case 16: return launch_one<T, LPC, 16>(a, stream);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
hipError_t launch_lpc(const FunnelLmcArgs& a, int lpc, int epl, hipStream_t stream) {
  switch (lpc) {
    case 4: return launch_epl<T, 4>(a, epl, stream);
    case 8: return launch_epl<T, 8>(a, epl, stream);
    case 16: return launch_epl<T, 16>(a, epl, stream);
    case 32: return launch_epl<T, 32>(a, epl, stream);
    case 64: return launch_epl<T, 64>(a, epl, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

namespace {

template <int LPC>
hipError_t launch_tangent_epl(const FunnelLmcArgs& a, int epl, hipStream_t stream) {
  const long long lanes = (long long)a.C * LPC;
  const unsigned blocks = (unsigned)((lanes + 255) / 256);
  switch (epl) {
    case 1: hipLaunchKernelGGL((funnel_lmc_alpha2_tangent_kernel<LPC, 1>), dim3(blocks), dim3(256), 0, stream, a); break;
    case 2: hipLaunchKernelGGL((funnel_lmc_alpha2_tangent_kernel<LPC, 2>), dim3(blocks), dim3(256), 0, stream, a); break;
    case 4: hipLaunchKernelGGL((funnel_lmc_alpha2_tangent_kernel<LPC, 4>), dim3(blocks), dim3(256), 0, stream, a); break;
    case 8: hipLaunchKernelGGL((funnel_lmc_alpha2_tangent_kernel<LPC, 8>), dim3(blocks), dim3(256), 0, stream, a); break;
    case 16: hipLaunchKernelGGL((funnel_lmc_alpha2_tangent_kernel<LPC, 16>), dim3(blocks), dim3(256), 0, stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_funnel_lmc_alpha2_tangent(const FunnelLmcArgs& a, int lpc, int epl, hipStream_t stream) {
  switch (lpc) {
    case 4: return launch_tangent_epl<4>(a, epl, stream);
    case 8: return launch_tangent_epl<8>(a, epl, stream);
    case 16: return launch_tangent_epl<16>(a, epl, stream);
    case 32: return launch_tangent_epl<32>(a, epl, stream);
    case 64: return launch_tangent_epl<64>(a, epl, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_funnel_lmc(const FunnelLmcArgs& a, int dtype, int lpc, int epl, hipStream_t stream) {
  if (dtype == 0) return launch_lpc<double>(a, lpc, epl, stream);
  return launch_lpc<float>(a, lpc, epl, stream);
}

}  // namespace gjx
