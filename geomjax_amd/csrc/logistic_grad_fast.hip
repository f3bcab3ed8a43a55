// Tangent (forward-mode) derivative of one RMHMC transition with the Fisher
// metric on Bayesian logistic regression with respect to log step_size, fast
// path: the tuned kernel's float32 state and split-bf16 MFMA passes carried on
// (value, d/d log eps) pairs, with the energies, log-determinant sums, the
// acceptance probability and the objective in float64.  D <= 64, one wave per
// chain.
//
// Reference: the derivative the reference's adaptation takes with jax.jacfwd
// through the integrator (geomjax/adaptation/chees_adaptation_lmcmonge.py:
// 497-519) of geomjax/rmhmc/integrators.py:53-156 and
// geomjax/rmhmc/metrics.py:42-129, here for the RMHMC kernel with the Fisher
// metric (geomjax/rmhmc/rmhmc.py:131-174).
//
// The closed forms are those of the float64 tangent kernel (logistic_grad.hip):
// with eta = X q, s = sigmoid(eta), lambda = s(1-s), w = lambda (1-2s),
//   eta' = X q',  s' = lambda eta',  lambda' = w eta',  w' = lambda (1 - 6s + 6s^2) eta',
//   G' = X^T diag(w eta') X,  (G^-1)' = -G^-1 G' G^-1,  u' = G^-1 (p' - G' u),
//   a_n' = x_n^T (G^-1)' x_n,  h' = w' (a - b^2)/2 + w (a' - 2 b b')/2 + s',
//   (dH/dq)' = X^T h' + q'/prior_var,  (log det G)' = sum_k d_k'/d_k.
// Per evaluation: G' by the signed-weight Gram pass (weights w eta' of either
// sign) copied to a second LDS matrix, G by the Gram pass, one sweep carrying
// both (-G^-1 and -(G^-1)' with the pivots and their tangents), u and u' by
// refined solves, and one pass over the rows with two MFMA products (X G^-1 and
// X (G^-1)', split operands staged in LDS) for a, a', the fields and the next
// iterate's eta and eta' rows.  The momentum p0 = chol(G(q0)) z has no
// tangent.  Stop decisions use values only (the derivative of the executed
// computation).  Per-chain scratch rows (float32): eta and eta' at the step
// start and at the iterate, and the Gram weights.
#include "logistic_fast_ops.h"

namespace gjx {
namespace lfast {

// The tuned one-wave block plus the tangent matrix and vectors.
struct TSmem : Smem<1> {
  float GT[MAXD * LDG];  // G' (row stride MAXD), then (G^-1)' (row stride LDG); split staging in the field pass
  float bcT[2][64];      // sweep: tangent column broadcast
  float tdk[MAXD];       // pivot tangents
  float zq[MAXD], zp[MAXD], tzq[MAXD], tzp[MAXD];  // step start and its tangent
  float tq[MAXD], tp[MAXD], tu[MAXD], tdH[MAXD], r[MAXD];
};

// Sweep of (G, G') in (S.G, S.GT) (row stride MAXD, left intact): A = -G^-1,
// T = -(G^-1)', pivots in S.dk and their tangents in S.tdk.  With want_chol,
// p0j = (L z)_jl (values only).  The value recursion is that of sweep(); the
// tangent one differentiates each update.
template <int NB>
__device__ __forceinline__ void sweep_dual(TSmem& S, f32x32 (&A)[NB], f32x32 (&T)[NB], int D, bool want_chol,
                                           float& p0j) {
  constexpr int DP = 32 * NB;
  const int jl = threadIdx.x % DP;
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      A[b][i] = S.G[(32 * b + i) * MAXD + jl];
      T[b][i] = S.GT[(32 * b + i) * MAXD + jl];
    }
  p0j = 0.0f;
  // one pivot loop per 32-row tile, so the register-relative reads and writes
  // of row k index a fixed tile of A and T and stay in place (a runtime choice
  // of tile makes the compiler copy the tiles out and back for every pivot)
  auto pivot = [&](auto tile, int k) {
    constexpr int TL = decltype(tile)::value;
    const float ak = A[TL][k - 32 * TL], tk = T[TL][k - 32 * TL];
    float* bc = S.bc[k & 1];
    float* bt = S.bcT[k & 1];
    bc[jl] = ak;
    bt[jl] = tk;
    __syncthreads();
    const float d = bc[k], td = bt[k];
    const float rd = __builtin_amdgcn_rcpf(d);
    if (want_chol && jl >= k) p0j = fmaf(ak / sqrtf(d), S.z[k], p0j);
    const float f = (jl == k) ? 1.0f - rd : ak * rd;
    const float tf = (jl == k) ? td * rd * rd : (tk - ak * td * rd) * rd;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const float ci = bc[32 * b + i], ti = bt[32 * b + i];
        T[b][i] = fmaf(-ti, f, fmaf(-ci, tf, T[b][i]));
        A[b][i] = fmaf(-ci, f, A[b][i]);
      }
    A[TL][k - 32 * TL] = (jl == k) ? -rd : ak * rd;
    T[TL][k - 32 * TL] = (jl == k) ? td * rd * rd : (tk - ak * td * rd) * rd;
    if (threadIdx.x == 0) {
      S.dk[k] = d;
      S.tdk[k] = td;
    }
  };
  const int k0 = D < 32 ? D : 32;
  for (int k = 0; k < k0; ++k) pivot(std::integral_constant<int, 0>{}, k);
  if constexpr (NB == 2)
    for (int k = 32; k < D; ++k) pivot(std::integral_constant<int, 1>{}, k);
  for (int k = D; k < 32; ++k) {
    A[0][k] = jl == k ? -1.0f : 0.0f;
    T[0][k] = 0.0f;
  }
  if constexpr (NB == 2)
    for (int k = D > 32 ? D : 32; k < DP; ++k) {
      A[1][k - 32] = jl == k ? -1.0f : 0.0f;
      T[1][k - 32] = 0.0f;
    }
  if (threadIdx.x < DP && threadIdx.x >= D) {
    S.dk[threadIdx.x] = 1.0f;
    S.tdk[threadIdx.x] = 0.0f;
  }
  __syncthreads();
}

// One evaluation of the vector field and its tangent at (S.q, S.p; S.tq, S.tp)
// with G^-1 in S.G and (G^-1)' in S.GT (row stride LDG), S.u, S.tu: S.dH and
// S.tdH.  Advances the rows: E1, TE1 <- the next iterate's eta, eta' (base + h
// X u, tangent base' + h X u + h X u'; base E0 for a fixed-point iterate, E1
// for the explicit half step, which also sets E0, TE0).
template <int NB>
__device__ __forceinline__ void tangent_field_pass(TSmem& S, const float* __restrict__ Xr,
                                                   const float* __restrict__ Xs, const float* __restrict__ yf,
                                                   float* E0, float* E1, float* TE0, float* TE1, int npad, int D,
                                                   float hs, float inv_prior, bool explicit_step) {
  constexpr int KC = 2 * NB, NTK = NB * KC;
  const int l = threadIdx.x % 64, h = l / 32;
  bf16x8* const AH = reinterpret_cast<bf16x8*>(S.G);
  bf16x8* const AL = AH + NTK * 64;
  bf16x8* const TH = reinterpret_cast<bf16x8*>(S.GT);
  bf16x8* const TL = TH + NTK * 64;
  static_assert(2 * NTK * 64 * sizeof(bf16x8) <= sizeof(S.GT), "the split operands must fit");
  // stage the split A operands (rows permuted by phi, as the field pass) over
  // the matrices they come from: every read of a matrix before its writes
  auto stage = [&](float* M, bf16x8* H, bf16x8* Lo) {
    float v[NTK][8];
#pragma unroll
    for (int tk = 0; tk < NTK; ++tk) load8(M + phi((tk / KC) * 32 + (l % 32)) * LDG + (tk % KC) * 16 + h * 8, v[tk]);
    __syncthreads();
#pragma unroll
    for (int tk = 0; tk < NTK; ++tk) {
      bf16x8 hi, lo;
      split8(v[tk], hi, lo);
      H[tk * 64 + l] = hi;
      Lo[tk * 64 + l] = lo;
    }
  };
  stage(S.G, AH, AL);
  stage(S.GT, TH, TL);
  __syncthreads();
  float dacc[KC][8], tacc[KC][8];
#pragma unroll
  for (int kc = 0; kc < KC; ++kc)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      dacc[kc][i] = 0.0f;
      tacc[kc][i] = 0.0f;
    }
  const uint4* const sbase = reinterpret_cast<const uint4*>(Xs);
  const int nj = npad / 32;
  for (int rb = 0; rb < nj; ++rb) {
    float x[KC][8];
    bf16x8 xh[KC], xl[KC];
    {
      const float* const blk = Xr + (size_t)rb * (KC * 512);
      const uint4* const sblk = sbase + (size_t)rb * (KC * 128);
      // the MFMA operands first (the first product waits only for its own),
      // then the rows of X, first read after the products are issued
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
        xh[kc] = __builtin_bit_cast(bf16x8, sblk[kc * 128 + l * 2]);
        xl[kc] = __builtin_bit_cast(bf16x8, sblk[kc * 128 + l * 2 + 1]);
      }
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) load8(blk + kc * 512 + l * 8, x[kc]);
    }
    f32x16 acc[NB], tac[NB];
#pragma unroll
    for (int tk = 0; tk < NTK; ++tk) {
      const int tt = tk / KC, kc = tk % KC;
      if (kc == 0) {
        acc[tt] = f32x16{0};
        tac[tt] = f32x16{0};
      }
      acc[tt] = mfma3(AH[tk * 64 + l], AL[tk * 64 + l], xh[kc], xl[kc], acc[tt]);
      tac[tt] = mfma3(TH[tk * 64 + l], TL[tk * 64 + l], xh[kc], xl[kc], tac[tt]);
    }
    float bp = 0.0f, tbp = 0.0f;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      float u8[8], tu8[8];
      load8(S.u + kc * 16 + h * 8, u8);
      load8(S.tu + kc * 16 + h * 8, tu8);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        bp = fmaf(x[kc][i], u8[i], bp);
        tbp = fmaf(x[kc][i], tu8[i], tbp);
      }
    }
    const float b = bp + __shfl_xor(bp, 32, 64), tb = tbp + __shfl_xor(tbp, 32, 64);
    float ap = 0.0f, tap = 0.0f;
#pragma unroll
    for (int tt = 0; tt < NB; ++tt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = 16 * tt + r;
        ap = fmaf(acc[tt][r], x[j / 8][j % 8], ap);
        tap = fmaf(tac[tt][r], x[j / 8][j % 8], tap);
      }
    const float an = ap + __shfl_xor(ap, 32, 64), tan_ = tap + __shfl_xor(tap, 32, 64);
    const int row = rb * 32 + (l % 32);
    const float eta = E1[row], teta = TE1[row], yv = yf[row];
    const float s = sigmoidf(eta);
    const float lam = s * (1.0f - s);
    const float wv = lam * (1.0f - 2.0f * s);
    const float ts = lam * teta;
    const float tw = lam * (1.0f - 6.0f * s + 6.0f * s * s) * teta;
    const float q2 = an - b * b;
    const float hv = 0.5f * wv * q2 - (yv - s);
    const float thv = 0.5f * tw * q2 + 0.5f * wv * (tan_ - 2.0f * b * tb) + ts;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        dacc[kc][i] = fmaf(hv, x[kc][i], dacc[kc][i]);
        tacc[kc][i] = fmaf(thv, x[kc][i], tacc[kc][i]);
      }
    const float base = explicit_step ? eta : E0[row];
    const float tbase = explicit_step ? teta : TE0[row];
    const float en = fmaf(hs, b, base);
    const float ten = fmaf(hs, b + tb, tbase);  // d(hs)/d log eps = hs
    if (h == 0) {
      E1[row] = en;
      TE1[row] = ten;
      if (explicit_step) {
        E0[row] = en;
        TE0[row] = ten;
      }
    }
  }
  const float tot = reduce_features<NB, 1>(S, dacc);
  S.dH[l] = l < D ? tot + S.q[l] * inv_prior : 0.0f;
  const float ttot = reduce_features<NB, 1>(S, tacc);
  S.tdH[l] = l < D ? ttot + S.tq[l] * inv_prior : 0.0f;
  __syncthreads();
}

template <int NB>
__device__ __forceinline__ void run_tangent_chain(const LogisticArgs& a, const TangentOut& o, TSmem& S, int c,
                                                  int post) {
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
  float* const TE0 = W + 2 * ns;
  float* const TE1 = W + 3 * ns;
  float* const WT = W + 4 * ns;  // Gram weights
  const int Dp = a.Dmax;
  const bool own = l < D;
  const float h = (float)(0.5 * a.eps);  // and d h / d log eps
  const float inv_prior = (float)(1.0 / a.prior_var);
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;

  S.q[l] = own ? static_cast<const float*>(a.q_in)[(size_t)c * Dp + l] : 0.0f;
  S.q0[l] = S.q[l];
  S.tq[l] = 0.0f;
  S.tp[l] = 0.0f;
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
  const double logp = a.logp_in[c];
  __syncthreads();
  // eta at q0 in float64 from X, no tangent yet
  (void)fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, E1, WT, false, inv_prior);
  for (int row = l; row < npad; row += 64) {
    E0[row] = E1[row];
    TE0[row] = 0.0f;
    TE1[row] = 0.0f;
  }
  __syncthreads();

  // metric and its tangent at the iterate: G', G, the dual sweep, u, u', and
  // G^-1, (G^-1)' in LDS; returns (log det G, its tangent)
  auto metric = [&](bool draw, double& tlogdet) -> double {
    for (int row = l; row < npad; row += 64) {  // weights w eta' (either sign)
      const float s = sigmoidf(E1[row]);
      const float d = s * (1.0f - s) * (1.0f - 2.0f * s) * TE1[row];
      WT[row] = copysignf(sqrt_fast(fabsf(d)), d);
    }
    __syncthreads();
    gram_pass<NB, 1, true>(S, Xg, WT, npad, D, 0.0f);
    for (int i = 0; i < MAXD; ++i) S.GT[i * MAXD + l] = (i == l && i >= D) ? 0.0f : S.G[i * MAXD + l];
    for (int row = l; row < npad; row += 64) {  // sqrt(lambda)
      const float s = sigmoidf(E1[row]);
      WT[row] = sqrt_fast(s * (1.0f - s));
    }
    __syncthreads();
    gram_pass<NB, 1>(S, Xg, WT, npad, D, inv_prior);
    f32x32 A[NB], T[NB];
    float p0j;
    sweep_dual<NB>(S, A, T, D, draw, p0j);
    if (draw) {
      S.p[l] = own ? p0j : 0.0f;
      __syncthreads();
    }
    const double ld = logdet_from_pivots(S, D);
    tlogdet = group_sum<64>(own ? (double)S.tdk[l] / (double)S.dk[l] : 0.0);
    ginv_apply<NB, 1>(S, A);  // S.u = G^-1 p (refined against G in S.G)
    {
      float gu = 0.0f;  // (G' u)_l, G' symmetric (rows past the tile width hold no Gram)
      for (int i = 0; i < 32 * NB; ++i) gu = fmaf(S.GT[i * MAXD + l], S.u[i], gu);
      S.r[l] = own ? S.tp[l] - gu : 0.0f;
    }
    __syncthreads();
    solve_refined<NB, 1>(S, A, S.r, S.tu, S.dH);  // u' = G^-1 (p' - G' u)
    store_ginv<NB, 1>(S, A);
    if (l < 32 * NB) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int i = 0; i < 32; ++i) S.GT[(32 * b + i) * LDG + l] = -T[b][i];
    }
    __syncthreads();
    return ld;
  };

  double tlogdet = 0.0;
  double logdet = metric(true, tlogdet);
  const double H0 = -logp + 0.5 * group_sum<64>(own ? (double)S.p[l] * (double)S.u[l] : 0.0) + 0.5 * logdet + ke_const;
  bool have_ginv = true;
  int64_t n_fev = 0;
  // iterate (q, p) <- z0 + h F(x): tangents z0' + h F + h F'
  auto update = [&]() -> double {
    float dq = 0.0f;
    if (own) {
      const float qn = fmaf(h, S.u[l], S.zq[l]);
      const float pn = fmaf(-h, S.dH[l], S.zp[l]);
      const float tqn = fmaf(h, S.u[l] + S.tu[l], S.tzq[l]);
      const float tpn = fmaf(-h, S.dH[l] + S.tdH[l], S.tzp[l]);
      dq = nan_max(fabsf(qn - S.q[l]), fabsf(pn - S.p[l]));
      S.q[l] = qn;
      S.p[l] = pn;
      S.tq[l] = tqn;
      S.tp[l] = tpn;
    }
    const double norm = (double)group_reduce<64, NanMax>(dq);
    __syncthreads();
    return norm;
  };

  for (int s = 0; s < a.L; ++s) {
    S.zq[l] = S.q[l];
    S.zp[l] = S.p[l];
    S.tzq[l] = S.tq[l];
    S.tzp[l] = S.tp[l];
    __syncthreads();
    int mode = 0, n = 0;  // 0 first evaluation, 1 fixed-point iterations, 2 explicit half step
    for (;;) {
      if (!have_ginv) {
        logdet = metric(false, tlogdet);
        have_ginv = true;
      }
      tangent_field_pass<NB>(S, Xr, Xs, yf, E0, E1, TE0, TE1, npad, D, h, inv_prior, mode == 2);
      have_ginv = false;
      ++n_fev;
      if (mode == 2) {  // explicit half step from the last iterate (integrators.py:146-148)
        if (own) {
          S.tq[l] = fmaf(h, S.u[l] + S.tu[l], S.tq[l]);
          S.tp[l] = fmaf(-h, S.dH[l] + S.tdH[l], S.tp[l]);
          S.q[l] = fmaf(h, S.u[l], S.q[l]);
          S.p[l] = fmaf(-h, S.dH[l], S.p[l]);
        }
        __syncthreads();
        break;
      }
      const double norm = update();
      if (mode == 0) {
        mode = 1;
        n = 0;
      } else {
        ++n;
      }
      if (!(n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol)) mode = 2;
    }
  }

  // ---- end: energy at (q_L, p_L) and its tangent (eta and the log likelihood from X in float64) ----
  const double loglik = fresh_pass<NB, 1>(S, Xr, yf, N, npad, D, E1, WT, true, inv_prior);
  if (!have_ginv) logdet = metric(false, tlogdet);
  const double qq = group_sum<64>(own ? (double)S.q[l] * (double)S.q[l] : 0.0);
  const double pu = group_sum<64>(own ? (double)S.p[l] * (double)S.u[l] : 0.0);
  const double tpu = group_sum<64>(own ? (double)S.tp[l] * S.u[l] + (double)S.p[l] * S.tu[l] : 0.0);
  const double tlogp1 = group_sum<64>(own ? (double)S.g[l] * (double)S.tq[l] : 0.0);
  const double dqv = own ? (double)S.q[l] - (double)S.q0[l] : 0.0;
  const double jmp = group_sum<64>(dqv * dqv);
  const double tjmp = group_sum<64>(own ? 2.0 * dqv * (double)S.tq[l] : 0.0);
  const double logp1 = loglik - 0.5 * qq / prior + lp_const;
  const double H1 = -logp1 + 0.5 * pu + 0.5 * logdet + ke_const;
  const double tH1 = -tlogp1 + 0.5 * tpu + 0.5 * tlogdet;
  const double dHe = H0 - H1, tdHe = -tH1;
  double pacc, tpacc;
  if (isnan(dHe)) {
    pacc = 0.0;
    tpacc = 0.0;
  } else if (dHe < 0.0) {
    pacc = exp(dHe);
    tpacc = pacc * tdHe;
  } else {
    pacc = 1.0;
    tpacc = 0.0;
  }
  if (l == 0) {
    o.objective[c] = pacc * jmp;
    o.acceptance[c] = pacc;
    o.sq_jump[c] = jmp;
    o.d_objective[2 * c] = tpacc * jmp + pacc * tjmp;
    o.d_objective[2 * c + 1] = 0.0;
    o.d_acceptance[2 * c] = tpacc;
    o.d_acceptance[2 * c + 1] = 0.0;
    o.d_sq_jump[2 * c] = tjmp;
    o.d_sq_jump[2 * c + 1] = 0.0;
    if (o.num_fevals) o.num_fevals[c] = n_fev;
  }
}

template <int NB>
__global__ void __launch_bounds__(64) logistic_rmhmc_tangent_f32_kernel(const LogisticArgs a, const TangentOut o) {
  __shared__ TSmem S;
  const int per_xcd = gridDim.x / 8;
  const int slot = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (slot >= a.C) return;
  const int c = slot;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  if (a.packed_hdr[post].nb != NB) return;
  run_tangent_chain<NB>(a, o, S, c, post);
}

}  // namespace lfast

hipError_t launch_logistic_rmhmc_tangent_f32(const LogisticArgs& a, const TangentOut& o, int max_features,
                                             hipStream_t stream) {
  const int blocks = (a.C + 7) / 8 * 8;
  const bool nb1 = max_features <= 32 || a.P > 1;
  const bool nb2 = max_features > 32;
  if (nb2) {
    hipLaunchKernelGGL((lfast::logistic_rmhmc_tangent_f32_kernel<2>), dim3(blocks), dim3(64), 0, stream, a, o);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (nb1) hipLaunchKernelGGL((lfast::logistic_rmhmc_tangent_f32_kernel<1>), dim3(blocks), dim3(64), 0, stream, a, o);
  return hipGetLastError();
}

}  // namespace gjx
