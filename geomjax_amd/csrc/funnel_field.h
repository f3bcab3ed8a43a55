// Closed-form vector field of RMHMC with the Monge metric on Neal's funnel
// (shared by the sampling and the tangent kernels; R may be float, double or
// a dual number).  See funnel_rmhmc.hip for the derivation.
#pragma once

#include "common.h"

namespace gjx {

// Per-chain scalar coefficients of the vector field at one iterate.
template <class R>
struct FieldCoef {
  R ce;   // u_x = p_x + ce * x
  R A, B; // dH/dq_x = A x + B p_x
  R uv;   // u_v
  R dHv;  // dH/dq_v
};

template <class R>
__device__ __forceinline__ FieldCoef<R> field_coef(R v, R pv, R sx, R xpx, R a2, R inv_s2, R halfDx) {
  const R e = exp(-v);
  const R gv = -v * inv_s2 - halfDx + R(0.5) * e * sx;
  const R gg = e * e * sx + gv * gv;
  const R gp = -e * xpx + gv * pv;
  const R L = R(1) + a2 * gg;
  const R k = a2 / L;
  const R c = gp * k;
  const R ce = c * e;
  const R uv = pv - c * gv;
  const R wv = gv - gp * uv;
  // x . w_x = -e Sx - (g.p) (x . u_x) = -e Sx - (g.p)(Xp + ce Sx)
  const R xw = -e * sx - gp * (xpx + ce * sx);
  const R hwv = e * xw - (inv_s2 + R(0.5) * e * sx) * wv;
  FieldCoef<R> f;
  f.ce = ce;
  f.A = e * (R(1) + k * (e + wv + gp * ce));
  f.B = e * k * gp;
  f.uv = uv;
  f.dHv = k * hwv - gv;
  return f;
}

// Quadratic form p^T G^-1 p of the Monge metric G = I + a2 g g^T, written as
//   |p|^2 / L + (a2 |g|^2 / L) |p_perp|^2,   L = 1 + a2 |g|^2,
//   p_perp = p - (g.p / |g|^2) g   (formed component by component).
// Both terms are sums of squares, so the value is never negative and keeps its
// relative accuracy.  The textbook |p|^2 - a2 (g.p)^2 / L loses every digit to
// cancellation once a2 |g|^2 is large (a fixed-point iterate that ran away)
// and can come out hugely negative, which would make a run-away state look
// like a large energy gain.  The dense Cholesky form -log N(p; 0, G) never
// goes negative either.  With g = (-e x, gv) for the funnel,
// p_perp = (p_x + c e x, p_v - c gv) where c = monge_perp_coef(g.p, |g|^2).
template <class R>
__device__ __forceinline__ R monge_perp_coef(R gp, R gg) {
  return gg > R(0) ? gp / gg : R(0);
}
template <class R>
__device__ __forceinline__ R monge_quad(R pp, R perp2, R gg, R a2) {
  const R L = R(1) + a2 * gg;
  return pp / L + (a2 * gg / L) * perp2;
}

}  // namespace gjx
