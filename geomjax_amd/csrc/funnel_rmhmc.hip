// RMHMC (implicit midpoint, fixed-point solve) on Neal's funnel with the Monge
// metric G(q) = I + alpha2 g g^T, g = grad l(q).
//
// Reference: geomjax/rmhmc/rmhmc.py:131-174, 376-462 (transition),
// geomjax/rmhmc/integrators.py:53-156 (implicit midpoint + solver),
// geomjax/rmhmc/metrics.py:42-129 (kinetic energy -log N(p; 0, G)),
// geomjax/mcmc/proposal.py:87-121,168-185 (energy difference, accept),
// examples/funnel/main.py:34-39 (funnel, log-scale v = q[D-1]).
//
// Closed forms used instead of the reference's dense Cholesky + autodiff:
//   x = q[0:D-1], v = q[D-1], e = exp(-v), Sx = |x|^2, Xp = x.p_x,
//   g_x = -e x,  g_v = -v/sigma^2 - (D-1)/2 + e Sx/2,
//   L = 1 + alpha2 |g|^2 = det G,  u = G^-1 p = p - (alpha2 (g.p)/L) g,
//   dT/dp = u,  dT/dq = (alpha2/L) H (g - (g.p) u)        (H = Hessian of l)
//   p^T G^-1 p = |p|^2 - alpha2 (g.p)^2 / L,
// and |g|^2 = e^2 Sx + g_v^2, g.p = -e Xp + g_v p_v, so one group reduction
// of (Sx, Xp) per evaluation gives every per-chain scalar.  The update of the
// D-1 "x" coordinates is then linear in (x, p_x):
//   u_x    = p_x + (c e) x,                       c = alpha2 (g.p) / L
//   dH/dq_x = A x + B p_x,   A = e (1 + k (e + w_v + (g.p) c e)),  B = e k (g.p)
// with k = alpha2 / L and w_v = g_v - (g.p) u_v.  The log-scale coordinate v
// is replicated in every lane of the chain's group.
//
// Momentum draw p0 = chol(G) z uses the exact factor of I + a a^T (a = sqrt(alpha2) g):
//   t_j = 1 + sum_{k<=j} a_k^2,  L_jj = sqrt(t_j / t_{j-1}),
//   L_ij = a_i a_j / sqrt(t_j t_{j-1})  (i > j),
// the same lower factor a dense Cholesky of G returns, in O(D) with two prefix
// scans.
#include "common.h"
#include "funnel_field.h"
#include "funnel_rmhmc.h"

#include <math.h>

namespace gjx {

// One workgroup = 256 lanes = 256/LPC chains.  Each chain's group of LPC lanes
// holds EPL of the D-1 "x" coordinates per lane (blocked layout: lane l holds
// x[l*EPL .. l*EPL+EPL-1]); coordinates past D-1 are zero and stay zero.
// STEP = true is the single-transition entry point with every RMHMCInfo
// output; STEP = false runs a.T transitions with the state kept in registers.
template <class T, int LPC, int EPL, bool STEP>
__global__ void __launch_bounds__(256) funnel_rmhmc_kernel(const FunnelArgs a) {
  using R = T;
  // float32 (the tuned form) iterates in span{x0, p0}; float64 updates every
  // coordinate, the reference's own operation order
  constexpr bool kSpan = sizeof(T) == sizeof(float);
  const int tid = blockIdx.x * 256 + threadIdx.x;
  const int chain = tid / LPC;
  const int lg = tid % LPC;
  const bool live = chain < a.C;
  const size_t c = live ? (size_t)chain : 0;
  const int D = a.D, Dx = a.D - 1;
  const int j0 = lg * EPL;  // first x coordinate of this lane
  const R h = R(0.5 * a.eps);
  const R a2 = R(a.alpha2);
  const R alpha = R(sqrt(a.alpha2));
  const R inv_s2 = R(1.0 / (a.sigma * a.sigma));
  const R halfDx = R(0.5 * Dx);
  const int T_steps = STEP ? 1 : a.T;
#define GJX_ON(k) (live && j0 + (k) < Dx)

  // ---- chain state: position (x, v) and its log density ----
  R x[EPL];
  R v;
  {
    const size_t base = c * D;
#pragma unroll
    for (int k = 0; k < EPL; ++k) x[k] = GJX_ON(k) ? ld<T>(a.q_in, base + j0 + k) : R(0);
    v = live ? ld<T>(a.q_in, base + Dx) : R(0);
  }
  double logp = live ? a.logp_in[c] : 0.0;
  int n_acc = 0;
  int64_t n_fev = 0;

  for (int t = 0; t < T_steps; ++t) {
    // ---- momentum p0 = chol(G(q0)) z  (geomjax/rmhmc/metrics.py:45-58) ----
    R px[EPL];
    R pv;
    R sx;
    double uacc;
    {
      // noise: z for the momentum, u for the accept test
      R z[EPL];
      R zv;
      const uint64_t tt = a.counter + (uint64_t)t;
      const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
      if (a.z != nullptr) {
        const size_t base = ((size_t)t * a.C + c) * D;
#pragma unroll
        for (int k = 0; k < EPL; ++k) z[k] = GJX_ON(k) ? ld<T>(a.z, base + j0 + k) : R(0);
        zv = live ? ld<T>(a.z, base + Dx) : R(0);
        uacc = live ? a.u[(size_t)t * a.C + c] : 1.0;
      } else {
#pragma unroll
        for (int k = 0; k < EPL; k += 2) {
          u32x4 r = philox4x32_10(
              u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(j0 + k), (uint32_t)(tt >> 32)}, k0, k1);
          R n0, n1;
          normal2(r, n0, n1);
          z[k] = GJX_ON(k) ? n0 : R(0);
          if (k + 1 < EPL) z[k + 1] = GJX_ON(k + 1) ? n1 : R(0);
        }
        {
          u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentumV, (uint32_t)(tt >> 32)}, k0, k1);
          R n0, n1;
          normal2(r, n0, n1);
          zv = n0;
        }
        u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamAccept, (uint32_t)(tt >> 32)}, k0, k1);
        uacc = u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);  // in [0, 1)
      }

      R loc_inc[EPL];
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        acc += x[k] * x[k];
        loc_inc[k] = acc;
      }
      sx = group_sum<LPC>(acc);
      const R off1 = group_exclusive_scan<LPC>(acc, lg);
      const R e = exp(-v);
      const R a2e2 = a2 * e * e;
      const R gv = -v * inv_s2 - halfDx + R(0.5) * e * sx;
      // c_j = a_j z_j / sqrt(t_j t_{j-1}),  a_j = -alpha e x_j
      R cj[EPL];
      R acc2 = R(0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const R tj = R(1) + a2e2 * (off1 + loc_inc[k]);
        const R tjm = R(1) + a2e2 * (off1 + loc_inc[k] - x[k] * x[k]);
        const R aj = -alpha * e * x[k];
        cj[k] = aj * z[k] / sqrt(tj * tjm);
        px[k] = sqrt(tj / tjm) * z[k];
        acc2 += cj[k];
      }
      const R off2 = group_exclusive_scan<LPC>(acc2, lg);
      const R tot2 = group_sum<LPC>(acc2);
      R run = off2;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const R aj = -alpha * e * x[k];
        px[k] = GJX_ON(k) ? px[k] + aj * run : R(0);
        run += cj[k];
      }
      const R txe = R(1) + a2e2 * sx;
      const R tv = txe + a2 * gv * gv;
      pv = sqrt(tv / txe) * zv + alpha * gv * tot2;
    }
    if (STEP && a.momentum != nullptr && live) {
#pragma unroll
      for (int k = 0; k < EPL; ++k)
        if (GJX_ON(k)) st<T>(a.momentum, c * D + j0 + k, px[k]);
      if (lg == 0) st<T>(a.momentum, c * D + Dx, pv);
    }

    // ---- initial energy H0 = -l0 + T(q0, p0)  (geomjax/mcmc/metrics.py:160-166) ----
    R xpx;
    double H0;
    {
      R a1 = R(0);
      double a3 = 0.0;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        a1 += x[k] * px[k];
        a3 += (double)px[k] * (double)px[k];
      }
      xpx = group_sum<LPC>(a1);
      const double pp = group_sum<LPC>(a3) + (double)pv * pv;
      const double e = exp(-(double)v);
      const double gv = -(double)v / (a.sigma * a.sigma) - 0.5 * Dx + 0.5 * e * (double)sx;
      const double gp = -e * (double)xpx + gv * (double)pv;
      const double gg = e * e * (double)sx + gv * gv;
      const double cp = monge_perp_coef(gp, gg);
      double a4 = 0.0;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const double r = (double)px[k] + cp * e * (double)x[k];
        a4 += r * r;
      }
      const double rv = (double)pv - cp * gv;
      const double perp2 = group_sum<LPC>(a4) + rv * rv;
      const double Tk = 0.5 * monge_quad(pp, perp2, gg, a.alpha2) + 0.5 * log(1.0 + a.alpha2 * gg) +
                        0.5 * (double)D * kLog2Pi;
      H0 = -logp + Tk;
    }

    // ---- trajectory: L implicit-midpoint steps (geomjax/mcmc/trajectory.py:121-139) ----
    R xc[EPL], pc[EPL];
    R vc = v, pvc = pv;
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      xc[k] = x[k];
      pc[k] = px[k];
    }
    // chains of the dynamic kernel stop after their own number of steps
    const int Ls = live ? chain_steps(a.Lc, a.L, a.C, STEP ? 0 : t, (int)c) : 0;
    for (int s = 0; s < a.L; ++s) {
      const bool on = s < Ls;
      if (!__any(on)) break;
      // z0 = (x0, p0) of this step; every iterate is z0 + h F(previous iterate)
      R x0[EPL], p0[EPL];
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        x0[k] = xc[k];
        p0[k] = pc[k];
      }
      const R v0 = vc, pv0 = pvc;
      double norm = 0.0;
      int n = -1;  // the first evaluation (integrators.py:84) is outside the loop count
      bool active = on;
      if constexpr (kSpan) {
        // Every iterate is an affine map of the previous one with per-chain
        // coefficients, so (x, p) stays in span{x0, p0}:
        //   x = ca x0 + cb p0,  p = cg x0 + cd p0,
        // |x|^2 and x.p follow from the three sums of the step start, and only
        // the convergence norm max |x' - x|, |p' - p| is formed per coordinate.
        R spp = R(0);
#pragma unroll
        for (int k = 0; k < EPL; ++k) spp += p0[k] * p0[k];
        const R Sxx = sx, Sxp = xpx, Spp = group_sum<LPC>(spp);
        R ca = R(1), cb = R(0), cg = R(0), cd = R(1);
        while (true) {
          const FieldCoef<R> f = field_coef<R>(vc, pvc, sx, xpx, a2, inv_s2, halfDx);
          // x' = x0 + h (p + ce x),  p' = p0 - h (A x + B p)
          const R na = R(1) + h * (cg + f.ce * ca), nb = h * (cd + f.ce * cb);
          const R ng = -h * (f.A * ca + f.B * cg), nd = R(1) - h * (f.A * cb + f.B * cd);
          const R da = na - ca, db = nb - cb, dg = ng - cg, dd = nd - cd;
          R mx = R(0);
#pragma unroll
          for (int k = 0; k < EPL; ++k)
            mx = fmaxf(mx, fmaxf(fabsf(da * x0[k] + db * p0[k]), fabsf(dg * x0[k] + dd * p0[k])));
          mx = group_reduce<LPC, OpMax>(mx);
          const R vn = v0 + h * f.uv;
          const R pvn = pv0 - h * f.dHv;
          mx = nan_max(mx, nan_max(fabs(vn - vc), fabs(pvn - pvc)));
          // the per-coordinate max drops NaN; a non-finite coefficient (or step
          // start sum, through the field) makes the norm NaN instead
          if (!isfinite(na + nb + ng + nd)) mx = R(NAN);
          if (active) {
            ca = na;
            cb = nb;
            cg = ng;
            cd = nd;
            vc = vn;
            pvc = pvn;
            sx = (ca * ca) * Sxx + (R(2) * ca * cb) * Sxp + (cb * cb) * Spp;
            xpx = (ca * cg) * Sxx + (ca * cd + cb * cg) * Sxp + (cb * cd) * Spp;
            norm = (double)mx;
            ++n;
            ++n_fev;
          }
          // cond_fn (integrators.py:67-74)
          active = on && n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol;
          if (!__any(active)) break;
        }
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
          xc[k] = ca * x0[k] + cb * p0[k];
          pc[k] = cg * x0[k] + cd * p0[k];
        }
      } else while (true) {
        const FieldCoef<R> f = field_coef<R>(vc, pvc, sx, xpx, a2, inv_s2, halfDx);
        R s1 = R(0), s2 = R(0), mx = R(0);
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
          const R xn = x0[k] + h * (pc[k] + f.ce * xc[k]);
          const R pn = p0[k] - h * (f.A * xc[k] + f.B * pc[k]);
          mx = nan_max(mx, nan_max(fabs(xn - xc[k]), fabs(pn - pc[k])));
          s1 += xn * xn;
          s2 += xn * pn;
          xc[k] = active ? xn : xc[k];
          pc[k] = active ? pn : pc[k];
        }
        const R vn = v0 + h * f.uv;
        const R pvn = pv0 - h * f.dHv;
        mx = group_reduce<LPC, NanMax>(mx);
        mx = nan_max(mx, nan_max(fabs(vn - vc), fabs(pvn - pvc)));
        s1 = group_sum<LPC>(s1);
        s2 = group_sum<LPC>(s2);
        if (active) {
          vc = vn;
          pvc = pvn;
          sx = s1;
          xpx = s2;
          norm = (double)mx;
          ++n;
          ++n_fev;
        }
        // cond_fn (integrators.py:67-74)
        active = on && n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol;
        if (!__any(active)) break;
      }
      if (STEP && on && lg == 0) {
        const size_t idx = c * a.L + s;
        if (a.fp_iters) a.fp_iters[idx] = n;
        if (a.fp_norm) a.fp_norm[idx] = norm;
        if (a.fp_success) a.fp_success[idx] = (uint8_t)(isfinite(norm) && norm <= a.conv_tol);
      }
      // explicit half step from the last iterate (integrators.py:146-148)
      {
        const FieldCoef<R> f = field_coef<R>(vc, pvc, sx, xpx, a2, inv_s2, halfDx);
        R s1 = R(0), s2 = R(0);
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
          const R xo = xc[k], po = pc[k];
          const R xn = xo + h * (po + f.ce * xo);
          const R pn = po - h * (f.A * xo + f.B * po);
          xc[k] = on ? xn : xo;
          pc[k] = on ? pn : po;
          s1 += xc[k] * xc[k];
          s2 += xc[k] * pc[k];
        }
        vc = on ? vc + h * f.uv : vc;
        pvc = on ? pvc - h * f.dHv : pvc;
        sx = group_sum<LPC>(s1);
        xpx = group_sum<LPC>(s2);
        n_fev += on ? 1 : 0;
      }
    }

    // ---- end of trajectory: l, grad l, velocity, energy at (q_L, p_L) ----
    double H1, logp1;
    R e1, gv1, c1;
    {
      // sums in float64 for the energy (geomjax/rmhmc/metrics.py:66-69)
      double d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        d1 += (double)xc[k] * (double)xc[k];
        d2 += (double)xc[k] * (double)pc[k];
        d3 += (double)pc[k] * (double)pc[k];
      }
      const double Sx = group_sum<LPC>(d1);
      const double Xp = group_sum<LPC>(d2);
      const double pp = group_sum<LPC>(d3) + (double)pvc * pvc;
      const double vd = (double)vc;
      const double s2 = a.sigma * a.sigma;
      const double e = exp(-vd);
      const double gv = -vd / s2 - 0.5 * Dx + 0.5 * e * Sx;
      const double gp = -e * Xp + gv * (double)pvc;
      const double gg = e * e * Sx + gv * gv;
      const double L = 1.0 + a.alpha2 * gg;
      const double cp = monge_perp_coef(gp, gg);
      double d4 = 0.0;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const double r = (double)pc[k] + cp * e * (double)xc[k];
        d4 += r * r;
      }
      const double rv = (double)pvc - cp * gv;
      const double perp2 = group_sum<LPC>(d4) + rv * rv;
      // funnel log density (examples/funnel/main.py:34-39)
      logp1 = -0.5 * vd * vd / s2 - log(a.sigma) - 0.5 * kLog2Pi - 0.5 * e * Sx - 0.5 * (double)Dx * vd -
              0.5 * (double)Dx * kLog2Pi;
      const double Tk = 0.5 * monge_quad(pp, perp2, gg, a.alpha2) + 0.5 * log(L) + 0.5 * (double)D * kLog2Pi;
      H1 = -logp1 + Tk;
      e1 = R(e);
      gv1 = R(gv);
      c1 = R(a.alpha2 * gp / L);
    }

    // ---- accept / reject (geomjax/mcmc/proposal.py:87-121,168-185) ----
    double dH = H0 - H1;
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
        if (a.prop_weight) a.prop_weight[c] = dH;
        if (a.prop_slpa) a.prop_slpa[c] = fmin(dH, 0.0);
      }
      // proposal: flipped momentum and velocity (rmhmc.py:443-462)
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        if (!GJX_ON(k)) continue;
        const size_t i = c * D + j0 + k;
        if (a.prop_q) st<T>(a.prop_q, i, xc[k]);
        if (a.prop_p) st<T>(a.prop_p, i, -pc[k]);
        if (a.prop_v) st<T>(a.prop_v, i, -(pc[k] + c1 * e1 * xc[k]));
        if (a.prop_g) st<T>(a.prop_g, i, -e1 * xc[k]);
      }
      if (lg == 0) {
        const size_t i = c * D + Dx;
        if (a.prop_q) st<T>(a.prop_q, i, vc);
        if (a.prop_p) st<T>(a.prop_p, i, -pvc);
        if (a.prop_v) st<T>(a.prop_v, i, -(pvc - c1 * gv1));
        if (a.prop_g) st<T>(a.prop_g, i, gv1);
      }
      // new state: RMHMCState of the sampled proposal (rmhmc.py:168-172)
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        if (!GJX_ON(k)) continue;
        const size_t i = c * D + j0 + k;
        st<T>(a.q_out, i, accept ? xc[k] : x[k]);
        if (a.g_out) st<T>(a.g_out, i, accept ? -e1 * xc[k] : ld<T>(a.g_in, i));
      }
      if (lg == 0) {
        const size_t i = c * D + Dx;
        st<T>(a.q_out, i, accept ? vc : v);
        if (a.g_out) st<T>(a.g_out, i, accept ? gv1 : ld<T>(a.g_in, i));
        a.logp_out[c] = accept ? logp1 : logp;
      }
    }

    // carry the chain to the next transition
    if (accept) {
#pragma unroll
      for (int k = 0; k < EPL; ++k) x[k] = xc[k];
      v = vc;
      logp = logp1;
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
    // ---- final state of the sample entry point ----
    if (live) {
#pragma unroll
      for (int k = 0; k < EPL; ++k)
        if (GJX_ON(k)) st<T>(a.final_q, c * D + j0 + k, x[k]);
      if (lg == 0) {
        st<T>(a.final_q, c * D + Dx, v);
        a.final_logp[c] = logp;
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
    if (live && lg == 0) {
      if (a.num_accepted) a.num_accepted[c] = n_acc;
      if (a.num_fevals) a.num_fevals[c] = n_fev;
    }
  }
#undef GJX_ON
}

// ---------------------------------------------------------------------------
// Host-side dispatch over the compiled (lanes per chain, elements per lane).
// ---------------------------------------------------------------------------
namespace {

template <class T, int LPC, int EPL>
hipError_t launch_one(const FunnelArgs& a, hipStream_t stream) {
  const long long lanes = (long long)a.C * LPC;
  const unsigned blocks = (unsigned)((lanes + 255) / 256);
  if (a.q_out != nullptr)
    hipLaunchKernelGGL((funnel_rmhmc_kernel<T, LPC, EPL, true>), dim3(blocks), dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL((funnel_rmhmc_kernel<T, LPC, EPL, false>), dim3(blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <class T, int LPC>
hipError_t launch_epl(const FunnelArgs& a, int epl, hipStream_t stream) {
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
hipError_t launch_lpc(const FunnelArgs& a, int lpc, int epl, hipStream_t stream) {
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

int funnel_pick_epl(int dim, int lpc) {
  const int need = (dim - 1 + lpc - 1) / lpc;
  for (int e : {1, 2, 4, 8, 16})
    if (e >= need) return e;
  return -1;
}

hipError_t launch_funnel_rmhmc(const FunnelArgs& a, int dtype, int lpc, int epl, hipStream_t stream) {
  if (dtype == 0) return launch_lpc<double>(a, lpc, epl, stream);
  return launch_lpc<float>(a, lpc, epl, stream);
}

}  // namespace gjx

// ---------------------------------------------------------------------------
// value_and_grad of the funnel log density (the reference's init,
// geomjax/rmhmc/rmhmc.py:96-98): one wave per chain.
// ---------------------------------------------------------------------------
namespace gjx {

template <class T>
__global__ void __launch_bounds__(256) funnel_value_and_grad_kernel(int C, int D, double sigma, const void* q,
                                                                    double* logp, void* grad) {
  const int chain = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (chain >= C) return;  // whole waves leave together
  const int Dx = D - 1;
  const size_t base = (size_t)chain * D;
  const double v = (double)ld<T>(q, base + Dx);
  double acc = 0.0;
  for (int j = lane; j < Dx; j += 64) {
    const double xj = (double)ld<T>(q, base + j);
    acc += xj * xj;
  }
  const double sx = group_sum<64>(acc);
  const double e = exp(-v);
  const double s2 = sigma * sigma;
  for (int j = lane; j < Dx; j += 64) st<T>(grad, base + j, T(-e * (double)ld<T>(q, base + j)));
  if (lane == 0) {
    st<T>(grad, base + Dx, T(-v / s2 - 0.5 * Dx + 0.5 * e * sx));
    logp[chain] = -0.5 * v * v / s2 - log(sigma) - 0.5 * kLog2Pi - 0.5 * e * sx - 0.5 * (double)Dx * v -
                  0.5 * (double)Dx * kLog2Pi;
  }
}

hipError_t launch_funnel_value_and_grad(int C, int D, double sigma, int dtype, const void* q, double* logp,
                                        void* grad, hipStream_t stream) {
  const unsigned blocks = (unsigned)((C + 3) / 4);
  if (dtype == 0)
    hipLaunchKernelGGL(funnel_value_and_grad_kernel<double>, dim3(blocks), dim3(256), 0, stream, C, D, sigma, q,
                       logp, grad);
  else
    hipLaunchKernelGGL(funnel_value_and_grad_kernel<float>, dim3(blocks), dim3(256), 0, stream, C, D, sigma, q,
                       logp, grad);
  return hipGetLastError();
}

}  // namespace gjx
