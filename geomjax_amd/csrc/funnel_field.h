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

}  // namespace gjx
