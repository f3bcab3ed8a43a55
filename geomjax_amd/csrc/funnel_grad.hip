// Tangent (forward-mode) derivatives of one RMHMC-Monge transition on the
// funnel with respect to (log step_size, log alpha2).
//
// The same closed forms as funnel_rmhmc.hip, evaluated on dual numbers
// Dual<T, 2> seeded with d eps / d log eps = eps and
// d alpha2 / d log alpha2 = alpha2: T = double is the check path, T = float
// the fast path (float32 state and trajectory as in the tuned sampler; the
// energies, their difference and the acceptance on Dual<double, 2> from the
// widened float32 values, as the tuned sampler keeps its energies in float64).  Every executed operation, including each
// fixed-point iterate that the solver actually ran, carries its derivative;
// the solver's stop decisions use values only, so the result is the exact
// derivative of the unrolled computation at the executed iteration counts
// (what jax.jacfwd returns through the reference integrator,
// geomjax/adaptation/chees_adaptation_lmcmonge.py:497-519).  The momentum is
// reparameterised p0 = chol(G(q0; alpha2)) z with z fixed.
//
// Per chain it returns the acceptance probability p = min(1, e^dH), the
// squared jump |q_L - q_0|^2 and their product (the expected squared jump
// distance times acceptance that step-size / alpha2 tuning maximises), each
// with its two derivatives.
#include "../../include/geomjax_amd.h"
#include "common.h"
#include "dual.h"
#include "funnel_field.h"
#include "funnel_rmhmc.h"

namespace gjx {

template <class T, int LPC, int EPL>
__global__ void __launch_bounds__(256) funnel_rmhmc_tangent_kernel(const FunnelArgs a, const FunnelGradOut o) {
  using RD = Dual<T, 2>;       // state and trajectory
  using RE = Dual<double, 2>;  // energies
  const int tid = blockIdx.x * 256 + threadIdx.x;
  const int chain = tid / LPC;
  const int lg = tid % LPC;
  const bool live = chain < a.C;
  const size_t c = live ? (size_t)chain : 0;
  const int D = a.D, Dx = a.D - 1;
  const int j0 = lg * EPL;
#define GJX_ON(k) (live && j0 + (k) < Dx)
  RD eps((T)a.eps);
  eps.d[0] = (T)a.eps;
  RD a2((T)a.alpha2);
  a2.d[1] = (T)a.alpha2;
  const RD h = RD(T(0.5)) * eps;
  const RD alpha = sqrt(a2);
  const RD inv_s2((T)(1.0 / (a.sigma * a.sigma)));
  const RD halfDx((T)(0.5 * Dx));
  RE a2e(a.alpha2);
  a2e.d[1] = a.alpha2;
  const RE inv_s2e(1.0 / (a.sigma * a.sigma));
  const RE halfDxe(0.5 * Dx);

  // ---- inputs (constants: zero tangents) ----
  RD x[EPL];
  RD v;
  {
    const size_t base = c * D;
#pragma unroll
    for (int k = 0; k < EPL; ++k) x[k] = RD(GJX_ON(k) ? static_cast<const T*>(a.q_in)[base + j0 + k] : T(0));
    v = RD(live ? static_cast<const T*>(a.q_in)[base + Dx] : T(0));
  }
  const double logp = live ? a.logp_in[c] : 0.0;
  T z[EPL];
  T zv;
  if (a.z != nullptr) {
    const size_t base = c * D;
#pragma unroll
    for (int k = 0; k < EPL; ++k) z[k] = GJX_ON(k) ? static_cast<const T*>(a.z)[base + j0 + k] : T(0);
    zv = live ? static_cast<const T*>(a.z)[base + Dx] : T(0);
  } else {
    const uint64_t tt = a.counter;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
#pragma unroll
    for (int k = 0; k < EPL; k += 2) {
      u32x4 r = philox4x32_10(
          u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(j0 + k), (uint32_t)(tt >> 32)}, k0, k1);
      double n0, n1;
      normal2(r, n0, n1);
      z[k] = GJX_ON(k) ? (T)n0 : T(0);
      if (k + 1 < EPL) z[k + 1] = GJX_ON(k + 1) ? (T)n1 : T(0);
    }
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentumV, (uint32_t)(tt >> 32)}, k0, k1);
    double n0, n1;
    normal2(r, n0, n1);
    zv = (T)n0;
  }

  // ---- momentum p0 = chol(G) z (as funnel_rmhmc_kernel) ----
  RD px[EPL];
  RD pv;
  RD sx;
  {
    RD loc_inc[EPL];
    RD acc(T(0));
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      acc += x[k] * x[k];
      loc_inc[k] = acc;
    }
    sx = group_sum<LPC>(acc);
    const RD off1 = group_exclusive_scan<LPC>(acc, lg);
    const RD e = exp(-v);
    const RD a2e2 = a2 * e * e;
    const RD gv = -v * inv_s2 - halfDx + RD(T(0.5)) * e * sx;
    RD cj[EPL];
    RD acc2(T(0));
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const RD tj = RD(T(1)) + a2e2 * (off1 + loc_inc[k]);
      const RD tjm = RD(T(1)) + a2e2 * (off1 + loc_inc[k] - x[k] * x[k]);
      const RD aj = -alpha * e * x[k];
      cj[k] = aj * z[k] / sqrt(tj * tjm);
      px[k] = sqrt(tj / tjm) * z[k];
      acc2 += cj[k];
    }
    const RD off2 = group_exclusive_scan<LPC>(acc2, lg);
    const RD tot2 = group_sum<LPC>(acc2);
    RD run = off2;
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const RD aj = -alpha * e * x[k];
      px[k] = GJX_ON(k) ? px[k] + aj * run : RD(T(0));
      run += cj[k];
    }
    const RD txe = RD(T(1)) + a2e2 * sx;
    const RD tv = txe + a2 * gv * gv;
    pv = sqrt(tv / txe) * zv + alpha * gv * tot2;
  }

  // energy from the state (widened to float64 duals)
  auto energy = [&](const RD (&xx)[EPL], const RD (&pp)[EPL], RD vs, RD pvs, RE& logp_out) -> RE {
    const RE vv = widen(vs), pvv = widen(pvs);
    RE d1(0.0), d2(0.0), d3(0.0);
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const RE xk = widen(xx[k]), pk = widen(pp[k]);
      d1 += xk * xk;
      d2 += xk * pk;
      d3 += pk * pk;
    }
    const RE Sx = group_sum<LPC>(d1);
    const RE Xp = group_sum<LPC>(d2);
    const RE P2 = group_sum<LPC>(d3) + pvv * pvv;
    const RE e = exp(-vv);
    const RE gv = -vv * inv_s2e - halfDxe + RE(0.5) * e * Sx;
    const RE gp = -e * Xp + gv * pvv;
    const RE gg = e * e * Sx + gv * gv;
    const RE cp = monge_perp_coef(gp, gg);
    RE d4(0.0);
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const RE r = widen(pp[k]) + cp * e * widen(xx[k]);
      d4 += r * r;
    }
    const RE rv = pvv - cp * gv;
    const RE perp2 = group_sum<LPC>(d4) + rv * rv;
    logp_out = RE(-0.5) * vv * vv * inv_s2e - RE(0.5) * e * Sx - RE(0.5 * Dx) * vv +
               RE(-log(a.sigma) - 0.5 * kLog2Pi - 0.5 * (double)Dx * kLog2Pi);
    return RE(0.5) * monge_quad(P2, perp2, gg, a2e) + RE(0.5) * log(RE(1.0) + a2e * gg) +
           RE(0.5 * (double)D * kLog2Pi);
  };
  RE lp0;
  const RE T0 = energy(x, px, v, pv, lp0);
  const RE H0 = RE(-logp) + T0;  // the state's log density is an input constant

  // ---- trajectory ----
  RD xc[EPL], pc[EPL];
  RD vc = v, pvc = pv;
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    xc[k] = x[k];
    pc[k] = px[k];
  }
  RD xpx;
  {
    RD a1(T(0));
#pragma unroll
    for (int k = 0; k < EPL; ++k) a1 += xc[k] * pc[k];
    xpx = group_sum<LPC>(a1);
  }
  int64_t n_fev = 0;
  for (int s = 0; s < a.L; ++s) {
    RD x0[EPL], p0[EPL];
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      x0[k] = xc[k];
      p0[k] = pc[k];
    }
    const RD v0 = vc, pv0 = pvc;
    double norm = 0.0;
    int n = -1;
    bool active = true;
    while (true) {
      const FieldCoef<RD> f = field_coef<RD>(vc, pvc, sx, xpx, a2, inv_s2, halfDx);
      RD s1(T(0)), s2(T(0));
      double mx = 0.0;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const RD xn = x0[k] + h * (pc[k] + f.ce * xc[k]);
        const RD pn = p0[k] - h * (f.A * xc[k] + f.B * pc[k]);
        mx = nan_max(mx, (double)nan_max(fabs(xn.v - xc[k].v), fabs(pn.v - pc[k].v)));
        s1 += xn * xn;
        s2 += xn * pn;
        if (active) {
          xc[k] = xn;
          pc[k] = pn;
        }
      }
      const RD vn = v0 + h * f.uv;
      const RD pvn = pv0 - h * f.dHv;
      mx = group_reduce<LPC, NanMax>(mx);
      mx = nan_max(mx, (double)nan_max(fabs(vn.v - vc.v), fabs(pvn.v - pvc.v)));
      s1 = group_sum<LPC>(s1);
      s2 = group_sum<LPC>(s2);
      if (active) {
        vc = vn;
        pvc = pvn;
        sx = s1;
        xpx = s2;
        norm = mx;
        ++n;
        ++n_fev;
      }
      active = live && n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol;
      if (!__any(active)) break;
    }
    {
      const FieldCoef<RD> f = field_coef<RD>(vc, pvc, sx, xpx, a2, inv_s2, halfDx);
      RD s1(T(0)), s2(T(0));
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const RD xo = xc[k], po = pc[k];
        xc[k] = xo + h * (po + f.ce * xo);
        pc[k] = po - h * (f.A * xo + f.B * po);
        s1 += xc[k] * xc[k];
        s2 += xc[k] * pc[k];
      }
      vc = vc + h * f.uv;
      pvc = pvc - h * f.dHv;
      sx = group_sum<LPC>(s1);
      xpx = group_sum<LPC>(s2);
      ++n_fev;
    }
  }

  // ---- energy at the end, acceptance probability and squared jump ----
  RE lp1;
  const RE T1 = energy(xc, pc, vc, pvc, lp1);
  const RE H1 = -lp1 + T1;
  RE dH = H0 - H1;
  RE pacc;
  if (isnan(dH.v)) {
    pacc = RE(0.0);  // weight -inf: rejected, derivative 0
  } else {
    const RE ed = exp(dH);
    pacc = ed.v < 1.0 ? ed : RE(1.0);  // min(1, e^dH): derivative 0 where clipped
  }
  RE jump(0.0);
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    const RE dx = widen(xc[k]) - widen(x[k]);
    jump += dx * dx;
  }
  const RE dv = widen(vc) - widen(v);
  jump = group_sum<LPC>(jump) + dv * dv;
  RE obj = pacc * jump;
  if (!isfinite(jump.v) || isnan(dH.v)) {
    // a divergent trajectory contributes nothing (and no derivative)
    obj = RE(0.0);
    pacc = RE(0.0);
  }
  if (live && lg == 0) {
    o.objective[c] = obj.v;
    o.acceptance[c] = pacc.v;
    o.sq_jump[c] = jump.v;
    for (int i = 0; i < 2; ++i) {
      o.d_objective[2 * c + i] = obj.d[i];
      o.d_acceptance[2 * c + i] = pacc.d[i];
      o.d_sq_jump[2 * c + i] = jump.d[i];
    }
    if (o.num_fevals) o.num_fevals[c] = n_fev;
  }
#undef GJX_ON
}

namespace {

template <class T, int LPC>
hipError_t launch_tangent_epl(const FunnelArgs& a, const FunnelGradOut& o, int epl, hipStream_t s) {
  const unsigned blocks = (unsigned)(((long long)a.C * LPC + 255) / 256);
  switch (epl) {
    case 1: hipLaunchKernelGGL((funnel_rmhmc_tangent_kernel<T, LPC, 1>), dim3(blocks), dim3(256), 0, s, a, o); break;
    case 2: hipLaunchKernelGGL((funnel_rmhmc_tangent_kernel<T, LPC, 2>), dim3(blocks), dim3(256), 0, s, a, o); break;
    case 4: hipLaunchKernelGGL((funnel_rmhmc_tangent_kernel<T, LPC, 4>), dim3(blocks), dim3(256), 0, s, a, o); break;
    case 8: hipLaunchKernelGGL((funnel_rmhmc_tangent_kernel<T, LPC, 8>), dim3(blocks), dim3(256), 0, s, a, o); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_tangent_lpc(const FunnelArgs& a, const FunnelGradOut& o, int lpc, int epl, hipStream_t stream) {
  switch (lpc) {
    case 16: return launch_tangent_epl<T, 16>(a, o, epl, stream);
    case 32: return launch_tangent_epl<T, 32>(a, o, epl, stream);
    case 64: return launch_tangent_epl<T, 64>(a, o, epl, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_funnel_rmhmc_tangent(const FunnelArgs& a, const FunnelGradOut& o, int lpc, int epl, int dtype,
                                       hipStream_t stream) {
  return dtype == GJX_F32 ? launch_tangent_lpc<float>(a, o, lpc, epl, stream)
                          : launch_tangent_lpc<double>(a, o, lpc, epl, stream);
}

}  // namespace gjx
