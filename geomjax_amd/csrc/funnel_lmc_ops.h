// Device building blocks of the LMC-Monge integrator on Neal's funnel, shared by
// the transition kernels (funnel_lmc.hip) and NUTS (funnel_nuts.hip).  R is
// float, double or a dual number; LPC lanes per chain hold EPL coordinates each.
//
// Reference: geomjax/lmcmonge/integrators.py:52-194 (lan_integrator,
// half_step_omega), geomjax/lmcmonge/metrics.py:104-284 (velocity draw,
// kinetic energy, weighted gradients), examples/funnel/main.py:34-39.
// The closed forms are derived at the top of funnel_lmc.hip.
#pragma once

#include "common.h"
#include "dual.h"
#include "funnel_lmc.h"

namespace gjx {
namespace lmc_ops {

template <class R>
struct PosScalars {
  R e, Sx, Smx, gv, L, sL, isL, dlv, Hdlv, igHdlv;
};

template <class R>
__device__ __forceinline__ PosScalars<R> pos_scalars(R v, R Sx, R Smx, R a2, R inv_s2, R halfDx, R miv) {
  PosScalars<R> P;
  P.e = exp(-v);
  P.Sx = Sx;
  P.Smx = Smx;
  P.gv = -v * inv_s2 - halfDx + R(0.5) * P.e * Sx;
  P.L = R(1) + a2 * (P.e * P.e * Smx + miv * P.gv * P.gv);  // normalizing_constant (metrics.py:193-197)
  P.sL = sqrt(P.L);
  P.isL = R(1) / P.sL;
  P.dlv = miv * P.gv * P.isL;
  // Hdl_v = (e (x . dl_x) - (1/sigma^2 + e Sx / 2) dl_v) / sqrt(L),  x . dl_x = -e Smx / sqrt(L)
  P.Hdlv = (-P.e * P.e * Smx * P.isL - (inv_s2 + R(0.5) * P.e * Sx) * P.dlv) * P.isL;
  P.igHdlv = miv * P.Hdlv;
  return P;
}

// One half_step_omega (geomjax/lmcmonge/integrators.py:158-194) at the
// position described by P, updating the velocity (vx, vv) and dJ.
template <class R, int LPC, int EPL>
__device__ __forceinline__ void half_step(const PosScalars<R>& P, const R (&x)[EPL], const R (&mi)[EPL], R (&vx)[EPL],
                                          R& vv, double& dJ, R a2, R inv_s2, R h, bool exact_sm) {
  R dl[EPL], dphig[EPL];
  R s0 = R(0), s1 = R(0), s2 = R(0), s3 = R(0), s4 = R(0), s5 = R(0), s6 = R(0), s7 = R(0);
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    const R gn = -P.e * x[k] * P.isL;
    dl[k] = mi[k] * gn;
    const R Hdl = P.e * (-dl[k] + x[k] * P.dlv) * P.isL;
    const R igH = mi[k] * Hdl;
    const R dphi = -gn + a2 * Hdl * P.isL;
    dphig[k] = -dl[k] + a2 * igH * P.isL;
    const R hv = P.e * (-vx[k] + x[k] * vv) * P.isL;
    s0 += hv * dl[k];
    s1 += dphi * dl[k];
    s2 += x[k] * vx[k];
    s3 += vx[k] * hv;
    s4 += dphig[k] * hv;
    s5 += Hdl * vx[k];
    s6 += Hdl * dphig[k];
    s7 += Hdl * dl[k];
  }
  s0 = group_sum<LPC>(s0);
  s1 = group_sum<LPC>(s1);
  s2 = group_sum<LPC>(s2);
  s3 = group_sum<LPC>(s3);
  s4 = group_sum<LPC>(s4);
  s5 = group_sum<LPC>(s5);
  s6 = group_sum<LPC>(s6);
  s7 = group_sum<LPC>(s7);
  // log-scale coordinate contributions
  const R hv_v = (P.e * s2 - (inv_s2 + R(0.5) * P.e * P.Sx) * vv) * P.isL;
  const R gn_v = P.gv * P.isL;
  const R dphi_v = -gn_v + a2 * P.Hdlv * P.isL;
  const R dphig_v = -P.dlv + a2 * P.igHdlv * P.isL;
  const R P1 = s0 + hv_v * P.dlv;     // hv . dl
  const R P2 = s1 + dphi_v * P.dlv;   // dphi . dl
  const R P3 = s3 + vv * hv_v;        // v . hv
  const R P4 = s4 + dphig_v * hv_v;   // dphi_ig . hv
  const R P5 = s5 + P.Hdlv * vv;      // Hdl . v
  const R P6 = s6 + P.Hdlv * dphig_v; // Hdl . dphi_ig
  const R P7 = s7 + P.Hdlv * P.dlv;   // Hdl . dl
  // 1st volume adjustment (:178-179)
  const R d1 = R(1) + h * a2 * P1;
  dJ -= log(fabs((double)val(d1)));
  // v_t = v - (eps/2) sqrt(L) (dphi_ig - alpha2 (dphi.dl) dl)   (:182-185)
  const R vtH = P3 - h * P.sL * (P4 - a2 * P2 * P1);
  R coef = h * vtH / d1;  // (:186-189); the reference has no alpha2 here
  if (exact_sm) coef *= a2;
  const R HdlVn = P5 - h * P.sL * (P6 - a2 * P2 * P7) - coef * P7;
  // 2nd volume adjustment (:192)
  dJ += log(fabs((double)val(R(1) - h * a2 * HdlVn)));
  const R kq = h * P.sL, kd = a2 * P2;
#pragma unroll
  for (int k = 0; k < EPL; ++k) vx[k] = vx[k] - kq * (dphig[k] - kd * dl[k]) - coef * dl[k];
  vv = vv - kq * (dphig_v - kd * P.dlv) - coef * P.dlv;
}

// Scalars of the position (x, v) and the inclusive prefix of mi x^2 over the
// chain's coordinates in lane order (the velocity draw's running sums).
template <class R, int LPC, int EPL>
__device__ __forceinline__ PosScalars<R> position_scalars(const R (&x)[EPL], R v, const R (&mi)[EPL], R (&loc_inc)[EPL],
                                                          int lg, R a2, R inv_s2, R halfDx, R miv) {
  R sx = R(0), smx = R(0);
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    sx += x[k] * x[k];
    smx += mi[k] * x[k] * x[k];
    loc_inc[k] = smx;
  }
  const R off = group_exclusive_scan<LPC>(smx, lg);
#pragma unroll
  for (int k = 0; k < EPL; ++k) loc_inc[k] += off;
  sx = group_sum<LPC>(sx);
  smx = group_sum<LPC>(smx);
  return pos_scalars<R>(v, sx, smx, a2, inv_s2, halfDx, miv);
}

// The transition's standard normals for the velocity draw and its accept
// uniform: host-fed ([T, C, D] in storage type T, [T, C]) or the device Philox
// streams at counter + t.
template <class T, int EPL>
__device__ __forceinline__ void transition_noise(const FunnelLmcArgs& a, size_t c, int t, int j0, int Dx, bool live,
                                                 T (&z)[EPL], T& zv, double& uacc) {
  const int D = Dx + 1;
  const uint64_t tt = a.counter + (uint64_t)t;
  const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
  if (a.z != nullptr) {
    const size_t base = ((size_t)t * a.C + c) * D;
#pragma unroll
    for (int k = 0; k < EPL; ++k) z[k] = (live && j0 + k < Dx) ? ld<T>(a.z, base + j0 + k) : T(0);
    zv = live ? ld<T>(a.z, base + Dx) : T(0);
    uacc = (live && a.u != nullptr) ? a.u[(size_t)t * a.C + c] : 1.0;
    return;
  }
#pragma unroll
  for (int k = 0; k < EPL; k += 2) {
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(j0 + k), (uint32_t)(tt >> 32)},
                            k0, k1);
    T n0, n1;
    normal2(r, n0, n1);
    z[k] = j0 + k < Dx ? n0 : T(0);
    if (k + 1 < EPL) z[k + 1] = j0 + k + 1 < Dx ? n1 : T(0);
  }
  {
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentumV, (uint32_t)(tt >> 32)}, k0, k1);
    T n0, n1;
    normal2(r, n0, n1);
    zv = n0;
  }
  u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamAccept, (uint32_t)(tt >> 32)}, k0, k1);
  uacc = u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
}

// Velocity v0 = chol(M^-1 - alpha2 dl dl^T) z (geomjax/lmcmonge/metrics.py:155-166)
// with the exact lower factor of a diagonal-minus-rank-one matrix.
template <class R, int LPC, int EPL>
__device__ __forceinline__ void draw_velocity(const PosScalars<R>& P, const R (&x)[EPL], const R (&mi)[EPL],
                                              const R (&loc_inc)[EPL], const R (&z)[EPL], R zv, R a2, R alpha,
                                              R miv, int lg, int j0, int Dx, R (&vx)[EPL], R& vv) {
  // a_j^2 = alpha2 dl_j^2 / mi_j = (alpha2 e^2 / L) mi_j x_j^2
  const R w = a2 * P.e * P.e / P.L;
  R cj[EPL];
  R acc2 = R(0);
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    const R tj = R(1) - w * loc_inc[k];
    const R tjm = R(1) - w * (loc_inc[k] - mi[k] * x[k] * x[k]);
    const R aj = -alpha * P.e * sqrt(mi[k]) * x[k] * P.isL;
    cj[k] = aj * z[k] / sqrt(tj * tjm);
    vx[k] = sqrt(tj / tjm) * z[k];
    acc2 += cj[k];
  }
  const R off2 = group_exclusive_scan<LPC>(acc2, lg);
  const R tot2 = group_sum<LPC>(acc2);
  R run = off2;
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    const R aj = -alpha * P.e * sqrt(mi[k]) * x[k] * P.isL;
    vx[k] = j0 + k < Dx ? sqrt(mi[k]) * (vx[k] - aj * run) : R(0);
    run += cj[k];
  }
  // t after all x coordinates is 1 - alpha2 e^2 Smx / L; then the log-scale coordinate
  const R txe = R(1) - w * P.Smx;
  const R av = alpha * sqrt(miv) * P.gv * P.isL;
  const R tv = txe - av * av;
  vv = sqrt(miv) * (sqrt(tv / txe) * zv - av * tot2);
}

// One lan_integrator step (geomjax/lmcmonge/integrators.py:52-156): half step of
// the velocity, full step of the position, half step of the velocity.
template <class R, int LPC, int EPL>
__device__ __forceinline__ void lmc_step(PosScalars<R>& P, R (&xc)[EPL], R& vc, const R (&mi)[EPL], R (&vx)[EPL],
                                         R& vv, double& dJ, R eps, R h, R a2, R inv_s2, R halfDx, R miv,
                                         bool exact_sm) {
  half_step<R, LPC, EPL>(P, xc, mi, vx, vv, dJ, a2, inv_s2, h, exact_sm);
  R sx = R(0), smx = R(0);
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    xc[k] = xc[k] + eps * vx[k];
    sx += xc[k] * xc[k];
    smx += mi[k] * xc[k] * xc[k];
  }
  vc = vc + eps * vv;
  sx = group_sum<LPC>(sx);
  smx = group_sum<LPC>(smx);
  P = pos_scalars<R>(vc, sx, smx, a2, inv_s2, halfDx, miv);
  half_step<R, LPC, EPL>(P, xc, mi, vx, vv, dJ, a2, inv_s2, h, exact_sm);
}


// Kinetic energy of the lmcmonge metric (geomjax/lmcmonge/metrics.py:168-186) in
// float64 sums: -(log L - sum log mi)/2 + |v|_M^2 / 2 + alpha2 (v . g)^2 / 2,
// g = (-e x, gv) the full gradient.
template <class R, int LPC, int EPL>
__device__ __forceinline__ double lmc_kinetic(const PosScalars<R>& Q, const R (&xx)[EPL], const R (&vel)[EPL], R velv,
                                              const R (&mi)[EPL], R miv, int j0, int Dx, double sum_log_mi,
                                              double alpha2) {
  double svm = 0.0, xv = 0.0;
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    if (j0 + k < Dx) svm += (double)val(vel[k]) * (double)val(vel[k]) / (double)val(mi[k]);
    xv += (double)val(xx[k]) * (double)val(vel[k]);
  }
  svm = group_sum<LPC>(svm) + (double)val(velv) * (double)val(velv) / (double)val(miv);
  xv = group_sum<LPC>(xv);
  const double vg = -(double)val(Q.e) * xv + (double)val(velv) * (double)val(Q.gv);  // v . g
  return -0.5 * (log((double)val(Q.L)) - sum_log_mi) + 0.5 * svm + 0.5 * alpha2 * vg * vg;
}

}  // namespace lmc_ops
}  // namespace gjx
