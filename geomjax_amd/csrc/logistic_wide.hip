// RMHMC with the Fisher metric on Bayesian logistic regression, tuned form
// for posteriors of 65..128 features: the arithmetic of the one-wave kernel
// (float32 state, bf16 hi/lo MFMA contractions with float32 accumulation,
// float32 sweep inverse with one refinement step per solve, float64 log-det
// sums, energies, accept test and fixed-point norm) spread over four waves,
// because a 128 x 128 metric does not fit one wave's registers.
//
// Reference: geomjax/rmhmc/rmhmc.py:131-174,376-462 (kernel, proposal, accept),
// geomjax/rmhmc/integrators.py:53-156 (implicit midpoint, fixed point),
// geomjax/rmhmc/metrics.py:42-129 (kinetic energy -log N(p; 0, G)).
// Closed forms (w = s(1-s)(1-2s), r = y - s, a_n = x_n^T G^-1 x_n, b = X u):
//   dH/dp = u = G^-1 p,   dH/dq = X^T [ w (a - b^2) / 2 - r ] + q / prior_var.
//
// One workgroup of four waves per chain; the packed design matrix uses the
// same Gram / row / split-row fragment layouts as the one-wave kernel with
// four 32-feature blocks (eight 16-feature chunks).
//   Gram pass: the 10 lower 32 x 32 tiles are split over the waves (3, 3, 2,
//     2), each wave streaming only the feature blocks its tiles touch.  All
//     four split products (lo*lo included) are accumulated: at up to 128
//     features and a few hundred rows the metric's condition number reaches
//     ~20, and the dropped term alone would put u = G^-1 p ~1e-5 off.
//   Sweep: each lane holds an 8 x 8 block of G (16 x 16 lanes); one barrier
//     per pivot.  Its Schur-complement columns give the momentum draw.
//   Field pass: wave w computes the 32 output features phi(32 w + .) of
//     V^T = G^-1 X^T for each 32-row block, whose accumulator slots pair with
//     the lane's row-layout chunks 2w and 2w+1; the partial a_n and b_n of the
//     four waves meet in LDS (one barrier per block), and wave w accumulates
//     X^T h for its own 32 features.  Only wave 0 reads and writes the
//     per-row scratch (eta, sqrt(lambda)), so no row is read by one wave
//     while another rewrites it.
//   Fresh pass (float64 eta, log likelihood, gradient): row blocks dealt to
//     the waves round robin.
#include "common.h"
#include "logistic.h"
#include "logistic_f32.h"

namespace gjx {
namespace lwide {

using namespace lf32;

constexpr int NB = 4;            // 32-feature blocks of the packed layouts
constexpr int KC = 2 * NB;       // 16-feature chunks of the row layout
constexpr int MAXD = 32 * NB;    // 128
constexpr int NW = 4;            // waves per chain
constexpr int NT = 64 * NW;
constexpr int LDG = MAXD + 4;    // LDS row stride of G^-1 (rows stay 16-byte aligned)

struct Smem {
  float G[MAXD * MAXD];    // Gram G (row stride 128), kept for the refinement step
  float Gi[MAXD * LDG];    // G^-1; scratch of the feature reductions
  float bc[2][MAXD];       // sweep: column broadcast [pivot parity]
  float dk[MAXD];          // pivots
  float ab[2][NW][2][32];  // field pass: partial a_n, b_n [block parity][wave][a|b][row]
  float ey[2][3][32];      // field pass: eta, y, base of the block's rows [block parity]
  float half[2][MAXD];     // the two row halves' partial sums of the solves
  double red[NW];
  float q0[MAXD], p0[MAXD], q[MAXD], p[MAXD], u[MAXD], dH[MAXD], z[MAXD], g[MAXD];
};

__device__ __forceinline__ double block_sum(Smem& S, double v) {
  v = group_sum<64>(v);
  __syncthreads();
  if (threadIdx.x % 64 == 0) S.red[threadIdx.x / 64] = v;
  __syncthreads();
  return (S.red[0] + S.red[1]) + (S.red[2] + S.red[3]);
}

__device__ __forceinline__ double block_max(Smem& S, double v) {
  v = group_reduce<64, NanMax>(v);
  __syncthreads();
  if (threadIdx.x % 64 == 0) S.red[threadIdx.x / 64] = v;
  __syncthreads();
  return nan_max(nan_max(S.red[0], S.red[1]), nan_max(S.red[2], S.red[3]));
}

// accumulator slot r of lane l of a 32 x 32 MFMA result holds row crow(r), column l % 32
__device__ __forceinline__ int crow(int r, int l) { return (r / 4) * 8 + (l / 32) * 4 + (r % 4); }

// ---------------------------------------------------------------------------
// Gram pass.  Tile patterns: waves 0 and 1 stream feature blocks
// {2w, 2w+1} and own tiles (2w, 2w), (2w+1, 2w), (2w+1, 2w+1); waves 2 and 3
// stream blocks {w-2, 2, 3} and own (2, w-2), (3, w-2).  Tile t multiplies
// slot ta(t) (A operand) by slot tb(t) (B operand).
// ---------------------------------------------------------------------------
template <int PAT>
struct GramPat;
template <>
struct GramPat<0> {
  static constexpr int NS = 2, NTL = 3;
  __device__ static constexpr int ta(int t) { return t == 0 ? 0 : 1; }
  __device__ static constexpr int tb(int t) { return t == 2 ? 1 : 0; }
  __device__ static int block(int s, int w) { return 2 * w + s; }
};
template <>
struct GramPat<1> {
  static constexpr int NS = 3, NTL = 2;
  __device__ static constexpr int ta(int t) { return t + 1; }
  __device__ static constexpr int tb(int) { return 0; }
  __device__ static int block(int s, int w) { return s == 0 ? w - 2 : s + 1; }
};

template <int PAT>
__device__ __forceinline__ void gram_tiles(Smem& S, const float* __restrict__ Xg, const float* __restrict__ sl,
                                           int npad, int D, float inv_prior) {
  using P = GramPat<PAT>;
  constexpr int NS = P::NS, NTL = P::NTL;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  int blk[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) blk[s] = P::block(s, w);
  f32x16 acc[NTL];
#pragma unroll
  for (int t = 0; t < NTL; ++t) acc[t] = f32x16{0};
  const float* slp = sl + (l / 32) * 8;
  const float* xp = Xg + (size_t)l * 8;
  const int nkb = npad / 16;  // even: npad is a multiple of 64
  struct Frag {
    float s[8];
    float x[NS][8];
  };
  auto fetch = [&](Frag& f, int kb) {
    kb = kb < nkb ? kb : nkb - 1;  // the pipeline's tail re-reads the last step (in bounds, unused)
    load8(slp + kb * 16, f.s);
#pragma unroll
    for (int s = 0; s < NS; ++s) load8(xp + (size_t)(kb * NB + blk[s]) * 512, f.x[s]);
  };
  auto split_frag = [&](const Frag& f, bf16x8 (&hi)[NS], bf16x8 (&lo)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      float y8[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) y8[i] = f.x[s][i] * f.s[i];
      split8(y8, hi[s], lo[s]);
    }
  };
  auto tiles = [&](const bf16x8 (&hi)[NS], const bf16x8 (&lo)[NS]) {
#pragma unroll
    for (int t = 0; t < NTL; ++t)
      acc[t] = mfma4(hi[P::ta(t)], lo[P::ta(t)], hi[P::tb(t)], lo[P::tb(t)], acc[t]);
  };
  Frag fa, fb;
  bf16x8 ha[NS], la[NS], hb[NS], lb[NS];
  fetch(fa, 0);
  fetch(fb, 1);
  split_frag(fa, ha, la);
  for (int kb = 0; kb < nkb; kb += 2) {
    fetch(fa, kb + 2);
    tiles(ha, la);
    split_frag(fb, hb, lb);
    interleave_mfma_valu<4 * NTL, 6 * NS / (NTL) + 1>();
    fetch(fb, kb + 3);
    tiles(hb, lb);
    split_frag(fa, ha, la);
    interleave_mfma_valu<4 * NTL, 6 * NS / (NTL) + 1>();
  }
#pragma unroll
  for (int t = 0; t < NTL; ++t) {
    const int bi = blk[P::ta(t)], bj = blk[P::tb(t)];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = bi * 32 + crow(r, l);
      const int j = bj * 32 + (l % 32);
      // a diagonal tile holds (i, j) and (j, i) rounded differently: keep the lower one
      if (bi == bj && i < j) continue;
      float v = acc[t][r];
      if (i == j) v = i < D ? v + inv_prior : 1.0f;
      S.G[i * MAXD + j] = v;
      S.G[j * MAXD + i] = v;
    }
  }
}

// S.G = X^T diag(lambda) X + I / prior_var for the first D features,
// identity on the padding, from sqrt(lambda) in `sl`.
__device__ __forceinline__ void gram_pass(Smem& S, const float* __restrict__ Xg, const float* __restrict__ sl, int npad,
                                          int D, float inv_prior) {
  __syncthreads();  // earlier readers of S.G are done
  if (threadIdx.x / 64 < 2)
    gram_tiles<0>(S, Xg, sl, npad, D, inv_prior);
  else
    gram_tiles<1>(S, Xg, sl, npad, D, inv_prior);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Sweep of S.G (left intact): lane t holds the 8 x 8 block of rows
// [8 (t/16), +8) and columns [8 (t%16), +8) in registers.  Pivots run in
// groups of 8 with the pivot inside a group unrolled, so row k = 8 kk + a0 is
// register row a0 of the 16 lanes t / 16 == kk, which broadcast it through
// LDS (one barrier per pivot); by symmetry it is also column k.  On return
// S.Gi = G^-1, S.dk the pivots; with want_chol S.p0 = S.p = chol(G) S.z, from
// the Schur-complement column (row k before its elimination) of each pivot.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sweep(Smem& S, int D, bool want_chol) {
  const int t = threadIdx.x;
  const int i0 = 8 * (t / 16), j0 = 8 * (t % 16);
  float A[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    float v[8];
    load8(S.G + (i0 + a) * MAXD + j0, v);
#pragma unroll
    for (int b = 0; b < 8; ++b) A[a][b] = v[b];
  }
  float pacc = 0.0f;
  for (int kk = 0; kk * 8 < D; ++kk) {
#pragma unroll
    for (int a0 = 0; a0 < 8; ++a0) {
      const int k = kk * 8 + a0;
      if (k >= D) break;
      float* buf = S.bc[k & 1];
      if (t / 16 == kk) {
        *reinterpret_cast<float4*>(buf + j0) = float4{A[a0][0], A[a0][1], A[a0][2], A[a0][3]};
        *reinterpret_cast<float4*>(buf + j0 + 4) = float4{A[a0][4], A[a0][5], A[a0][6], A[a0][7]};
      }
      __syncthreads();
      const float d = buf[k];
      const float rd = __builtin_amdgcn_rcpf(d);  // 1 ulp, as the float32 sweep itself
      if (want_chol && t >= k && t < D) pacc = fmaf(buf[t] / sqrtf(d), S.z[k], pacc);
      float col[8], bjr[8];
      load8(buf + i0, col);  // A[i][k] = A[k][i]
      load8(buf + j0, bjr);
#pragma unroll
      for (int b = 0; b < 8; ++b) bjr[b] *= rd;
      // a_ij - a_ik a_kj / d everywhere, then the pivot row and column (register
      // row / column a0 of the lanes that hold them) become a_kj / d and
      // a_ik / d, and the pivot -1/d
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) A[a][b] = fmaf(-col[a], bjr[b], A[a][b]);
      const bool prow = t / 16 == kk, pcol = t % 16 == kk;
#pragma unroll
      for (int b = 0; b < 8; ++b) A[a0][b] = prow ? bjr[b] : A[a0][b];
#pragma unroll
      for (int a = 0; a < 8; ++a) A[a][a0] = pcol ? col[a] * rd : A[a][a0];
      if (prow && pcol) A[a0][a0] = -rd;
      if (t == 0) S.dk[k] = d;
    }
  }
  // G^-1 = -A; the padding block of G is I, so its inverse is I
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    float v[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) v[b] = (i0 + a == j0 + b && i0 + a >= D) ? 1.0f : -A[a][b];
    *reinterpret_cast<float4*>(S.Gi + (i0 + a) * LDG + j0) = float4{v[0], v[1], v[2], v[3]};
    *reinterpret_cast<float4*>(S.Gi + (i0 + a) * LDG + j0 + 4) = float4{v[4], v[5], v[6], v[7]};
  }
  if (t < MAXD && t >= D) S.dk[t] = 1.0f;
  if (want_chol && t < MAXD) {
    const float pv = t < D ? pacc : 0.0f;
    S.p0[t] = pv;
    S.p[t] = pv;
  }
  __syncthreads();
}

// log det G = sum log d_k in float64 (padding pivots are 1)
__device__ __forceinline__ double logdet_from_pivots(Smem& S, int D) {
  const int l = threadIdx.x;
  return block_sum(S, l < D ? log((double)S.dk[l]) : 0.0);
}

// (M v)_j for symmetric M (row stride ld) on lanes j < 128: each half of the
// lanes sums 64 rows of column j, the halves meet in LDS.
__device__ __forceinline__ float sym_times(Smem& S, const float* M, int ld, const float* v) {
  const int l = threadIdx.x, jl = l % MAXD, hh = l / MAXD;
  float s = 0.0f;
#pragma unroll 16
  for (int i = 0; i < 64; ++i) s = fmaf(M[(64 * hh + i) * ld + jl], v[64 * hh + i], s);
  S.half[hh][jl] = s;
  __syncthreads();
  const float r = l < MAXD ? S.half[0][l] + S.half[1][l] : 0.0f;
  __syncthreads();
  return r;
}

// G (in S.G) -> pivots in S.dk, S.Gi = G^-1, S.u = G^-1 S.p with one
// refinement step against G.  With want_chol the momentum p0 = chol(G) S.z is
// drawn first into S.p0 and S.p.
__device__ __forceinline__ void invert(Smem& S, int D, bool want_chol) {
  const int l = threadIdx.x;
  sweep(S, D, want_chol);
  const float u0 = sym_times(S, S.Gi, LDG, S.p);
  if (l < MAXD) S.u[l] = u0;
  __syncthreads();
  const float gu = sym_times(S, S.G, MAXD, S.u);
  if (l < MAXD) S.dH[l] = S.p[l] - gu;  // residual (dH is free here)
  __syncthreads();
  const float du = sym_times(S, S.Gi, LDG, S.dH);
  if (l < MAXD) S.u[l] = u0 + du;
  __syncthreads();
}

// ---------------------------------------------------------------------------
// One evaluation of the vector field at (S.q, S.p) with S.u = G^-1 p and
// G^-1 in S.Gi: S.dH = dH/dq.  Advances the scratch as the one-wave kernel:
// E1 <- eta of the next iterate (base + h X u; base = E0 for a fixed-point
// iterate, E1 for the explicit half step, which also sets E0), SL <-
// sqrt(lambda) there.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void field_pass(Smem& S, const float* __restrict__ Xr, const float* __restrict__ Xs,
                                           const float* __restrict__ yf, float* E0, float* E1, float* SL, int npad,
                                           int D, float hs, float inv_prior, bool explicit_step) {
  const int w = threadIdx.x / 64, l = threadIdx.x % 64, h = l / 32, rl = l % 32;
  // A operand: rows phi(32 w + m) of G^-1, all 128 columns
  bf16x8 Ah[KC], Al[KC];
  {
    const int row = phi(w * 32 + rl);
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      float v[8];
      load8(S.Gi + row * LDG + kc * 16 + h * 8, v);
      split8(v, Ah[kc], Al[kc]);
    }
  }
  float dacc[2][8];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int i = 0; i < 8; ++i) dacc[c][i] = 0.0f;
  const float* xp = Xr + (size_t)l * 8;
  const uint4* sp = reinterpret_cast<const uint4*>(Xs) + (size_t)l * 2;
  const int nrb = npad / 32;
  float* const e0_dst = explicit_step ? E0 : E1;
  bf16x8 xh[KC], xl[KC];
  float xa[2][8];
  float eta = 0.0f, base = 0.0f, yv = 0.0f;
  // The next block's operands load before the exchange barrier, so that the
  // barrier and the finish of this block cover their latency.  A value still
  // live when its next load lands is copied at the end of the loop behind a
  // wait that covers every load issued before it: the rows of X (still read
  // by X^T h, through a copy) therefore load first, the row scalars are loaded
  // by every wave (only wave 0 uses them, base only for a fixed-point iterate:
  // a load under a branch is such a copy too) and pass through LDS before the
  // fetch, and the MFMA operands load last.
  auto fetch = [&](int rb) {
#pragma unroll
    for (int c = 0; c < 2; ++c) load8(xp + (size_t)(rb * KC + 2 * w + c) * 512, xa[c]);
    const int row = rb * 32 + rl;
    eta = E1[row];
    base = E0[row];
    yv = yf[row];
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      const uint4* q = sp + (size_t)(rb * KC + kc) * 128;
      xh[kc] = __builtin_bit_cast(bf16x8, q[0]);
      xl[kc] = __builtin_bit_cast(bf16x8, q[1]);
    }
  };
  fetch(0);
  for (int rb = 0; rb < nrb; ++rb) {
    asm volatile("" ::: "memory");  // u is re-read from LDS per block (registers are scarce here)
    const int par = rb & 1;
    f32x16 acc = f32x16{0};
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) acc = mfma3(Ah[kc], Al[kc], xh[kc], xl[kc], acc);
    float bp = 0.0f;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      float u8[8];
      load8(S.u + (2 * w + c) * 16 + h * 8, u8);
#pragma unroll
      for (int i = 0; i < 8; ++i) bp = fmaf(xa[c][i], u8[i], bp);
    }
    interleave_mfma_valu<3 * KC, 2>();
    float ap = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) ap = fmaf(acc[r], xa[r / 8][r % 8], ap);
    ap += __shfl_xor(ap, 32, 64);
    bp += __shfl_xor(bp, 32, 64);
    if (h == 0) {
      S.ab[par][w][0][rl] = ap;
      S.ab[par][w][1][rl] = bp;
      if (w == 0) {
        S.ey[par][0][rl] = eta;
        S.ey[par][1][rl] = yv;
        S.ey[par][2][rl] = base;
      }
    }
    float xr[2][8];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 8; ++i) xr[c][i] = xa[c][i];
    // the next block's operands load across the exchange barrier and the
    // finish of this block (past the end the last block again: in bounds, unused)
    fetch(rb + 1 < nrb ? rb + 1 : rb);
    __syncthreads();
    const float an = (S.ab[par][0][0][rl] + S.ab[par][1][0][rl]) + (S.ab[par][2][0][rl] + S.ab[par][3][0][rl]);
    const float bn = (S.ab[par][0][1][rl] + S.ab[par][1][1][rl]) + (S.ab[par][2][1][rl] + S.ab[par][3][1][rl]);
    const float et = S.ey[par][0][rl];
    const float s = sigmoidf(et);
    const float wv = s * (1.0f - s) * (1.0f - 2.0f * s);
    const float hv = 0.5f * wv * (an - bn * bn) - (S.ey[par][1][rl] - s);
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 8; ++i) dacc[c][i] = fmaf(hv, xr[c][i], dacc[c][i]);
    if (w == 0 && h == 0) {
      const int row = rb * 32 + rl;
      const float en = fmaf(hs, bn, explicit_step ? et : S.ey[par][2][rl]);
      const float sn = sigmoidf(en);
      E1[row] = en;
      e0_dst[row] = en;  // the explicit step moves the step start; iterates rewrite E1 only
      SL[row] = sqrt_fast(sn * (1.0f - sn));
    }
  }
  // X^T h: wave w holds features (2w + c) 16 + 8h + i summed over its lane's rows
  __syncthreads();
  float* R = S.Gi;  // [wave][lane][17]
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int i = 0; i < 8; ++i) R[(w * 64 + l) * 17 + c * 8 + i] = dacc[c][i];
  __syncthreads();
  if (threadIdx.x < MAXD) {
    const int f = threadIdx.x, kc = f / 16, hf = (f / 8) % 2, i = f % 8;
    const int ww = kc / 2, c = kc % 2;
    float tot = 0.0f;
    for (int j = 0; j < 32; ++j) tot += R[(ww * 64 + hf * 32 + j) * 17 + c * 8 + i];
    S.dH[f] = f < D ? tot + S.q[f] * inv_prior : 0.0f;
  }
  __syncthreads();
}

// eta = X q (float64 accumulation) into Eout, sqrt(lambda) into SL; with
// want_grad also the log likelihood (float64, returned) and
// S.g = X^T (y - s) - q / prior_var.  Row blocks go to the waves round robin.
__device__ __forceinline__ double fresh_pass(Smem& S, const float* __restrict__ Xr, const float* __restrict__ yf, int N,
                                             int npad, int D, float* Eout, float* SL, bool want_grad,
                                             float inv_prior) {
  const int w = threadIdx.x / 64, l = threadIdx.x % 64, h = l / 32;
  float gacc[KC][8];
#pragma unroll
  for (int kc = 0; kc < KC; ++kc)
#pragma unroll
    for (int i = 0; i < 8; ++i) gacc[kc][i] = 0.0f;
  double ll = 0.0;
  const float* xp = Xr + (size_t)l * 8;
  const int nrb = npad / 32;
  for (int rb = w; rb < nrb; rb += NW) {
    asm volatile("" ::: "memory");  // q is re-read from LDS per block (hoisting it costs 64 registers)
    float x[KC][8];
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) load8(xp + (size_t)(rb * KC + kc) * 512, x[kc]);
    double ep = 0.0;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      float q8[8];
      load8(S.q + kc * 16 + h * 8, q8);
#pragma unroll
      for (int i = 0; i < 8; ++i) ep = fma((double)x[kc][i], (double)q8[i], ep);
    }
    const double eta = ep + __shfl_xor(ep, 32, 64);
    const int row = rb * 32 + (l % 32);
    const double s = 1.0 / (1.0 + exp(-eta));
    if (want_grad) {
      const bool ok = row < N;
      const double yv = (double)yf[row];
      if (ok && h == 0) ll += yv * eta - softplus(eta);
      const float rr = ok ? (float)(yv - s) : 0.0f;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int i = 0; i < 8; ++i) gacc[kc][i] = fmaf(rr, x[kc][i], gacc[kc][i]);
    }
    if (h == 0) {
      Eout[row] = (float)eta;
      SL[row] = (float)sqrt(s * (1.0 - s));
    }
  }
  if (want_grad) {
    constexpr int RS = KC * 8 + 1;
    float* R = S.Gi;  // [wave][lane][RS]
    __syncthreads();
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int i = 0; i < 8; ++i) R[(w * 64 + l) * RS + kc * 8 + i] = gacc[kc][i];
    __syncthreads();
    if (threadIdx.x < MAXD) {
      const int f = threadIdx.x, kc = f / 16, hf = (f / 8) % 2, i = f % 8;
      float tot = 0.0f;
      for (int ww = 0; ww < NW; ++ww)
        for (int j = 0; j < 32; ++j) tot += R[(ww * 64 + hf * 32 + j) * RS + kc * 8 + i];
      S.g[f] = f < D ? tot - S.q[f] * inv_prior : 0.0f;
    }
  }
  __syncthreads();
  return block_sum(S, ll);
}

// E0 = E1 = ES, SL = sqrt(lambda(ES))
__device__ __forceinline__ void prep_pass(const float* ES, float* E0, float* E1, float* SL, int npad) {
  for (int row = threadIdx.x; row < npad; row += NT) {
    const float e = ES[row];
    const float s = sigmoidf(e);
    E0[row] = e;
    E1[row] = e;
    SL[row] = sqrt_fast(s * (1.0f - s));
  }
  __syncthreads();
}

enum Stage { kInit, kNoise, kTraj, kEnd };

// All transitions of one chain: the one-wave kernel's flat stage loop with
// block-wide reductions (every wave takes the same branches).
template <bool STEP>
__device__ __forceinline__ void run_chain(const LogisticArgs& a, Smem& S, int c, int post) {
  const int l = threadIdx.x;
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
  float* const ESa = W + 3 * ns;
  float* const ESb = W + 4 * ns;
  int cur = 0;  // eta of the chain state lives in cur ? ESb : ESa
  const int Dp = a.Dmax;
  const float hs = (float)(0.5 * a.eps);
  const float inv_prior = (float)(1.0 / a.prior_var);
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;
  const int T_steps = STEP ? 1 : a.T;

  if (l < MAXD) {
    const float qv = l < D ? static_cast<const float*>(a.q_in)[(size_t)c * Dp + l] : 0.0f;
    S.q0[l] = qv;
    S.q[l] = qv;
  }
  double logp = a.logp_in[c];
  int n_acc = 0;
  int64_t n_fev = 0;
  __syncthreads();

  int stage = kInit, tr = 0, s = 0, mode = 0, n = 0, Lcur = 0;
  bool have_ginv = false;
  float q0v = 0.0f, p0v = 0.0f;
  double uacc = 0.0, H0 = 0.0, loglik = 0.0, norm = 0.0;
  for (;;) {
    // ---- eta, log likelihood, gradient from X (initial state / end of trajectory) ----
    if (stage == kInit || stage == kEnd) {
      const bool end = stage == kEnd;
      float* Eout = (end ? 1 - cur : cur) ? ESb : ESa;
      loglik = fresh_pass(S, Xr, yf, N, npad, D, Eout, SL, end || !STEP, inv_prior);
      if (!end) {
        // the state's gradient goes straight to final_g (rewritten on every acceptance)
        if (!STEP && a.final_g != nullptr && l < D) static_cast<float*>(a.final_g)[(size_t)c * Dp + l] = S.g[l];
        stage = kNoise;
      }
    }
    // ---- noise, chain state, scratch ----
    if (stage == kNoise) {
      const uint64_t tt = a.counter + (uint64_t)tr;
      if (a.z != nullptr) {
        if (l < MAXD) S.z[l] = l < D ? static_cast<const float*>(a.z)[((size_t)tr * a.C + c) * Dp + l] : 0.0f;
        uacc = a.u[(size_t)tr * a.C + c];
      } else {
        if (l < MAXD / 2) {
          u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * l),
                                        (uint32_t)(tt >> 32)}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
          double n0, n1;
          normal2(r, n0, n1);
          S.z[2 * l] = 2 * l < D ? (float)n0 : 0.0f;
          S.z[2 * l + 1] = 2 * l + 1 < D ? (float)n1 : 0.0f;
        }
        u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamAccept, (uint32_t)(tt >> 32)},
                                (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        uacc = u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
      }
      if (l < MAXD) S.q[l] = S.q0[l];
      __syncthreads();
      prep_pass(cur ? ESb : ESa, E0, E1, SL, npad);
    }
    // ---- metric at S.q: G, its sweep, u = G^-1 p (momentum draw first at kNoise) ----
    if (stage != kTraj || !have_ginv) {
      gram_pass(S, Xg, SL, npad, D, inv_prior);
      invert(S, D, stage == kNoise);
    }
    if (stage == kNoise) {
      if (STEP && a.momentum != nullptr && l < D) static_cast<float*>(a.momentum)[(size_t)c * Dp + l] = S.p0[l];
      const double logdet = logdet_from_pivots(S, D);
      const double pu = block_sum(S, l < D ? (double)S.p[l] * (double)S.u[l] : 0.0);
      H0 = -logp + 0.5 * pu + 0.5 * logdet + ke_const;
      have_ginv = true;  // S.Gi = G(S.q)^-1 and S.u = G^-1 S.p
      s = 0;
      mode = 0;
      Lcur = chain_steps(a.Lc, a.L, a.C, STEP ? 0 : tr, c);
      stage = Lcur > 0 ? kTraj : kEnd;
      if (stage == kEnd) continue;
    }
    // ---- energy, accept, outputs ----
    if (stage == kEnd) {
      const double logdet = logdet_from_pivots(S, D);
      const double qq = block_sum(S, l < D ? (double)S.q[l] * (double)S.q[l] : 0.0);
      const double pu = block_sum(S, l < D ? (double)S.p[l] * (double)S.u[l] : 0.0);
      const double logp1 = loglik - 0.5 * qq / prior + lp_const;
      const double H1 = -logp1 + 0.5 * pu + 0.5 * logdet + ke_const;
      double dHe = H0 - H1;
      if (isnan(dHe)) dHe = -INFINITY;
      const double p_acc = fmin(exp(dHe), 1.0);
      const bool accept = uacc < p_acc;
      if (STEP) {
        if (l == 0) {
          if (a.acc_rate) a.acc_rate[c] = p_acc;
          if (a.accepted) a.accepted[c] = (uint8_t)accept;
          if (a.divergent) a.divergent[c] = (uint8_t)(-dHe > a.div_threshold);
          if (a.energy) a.energy[c] = H1;
          if (a.prop_logp) a.prop_logp[c] = logp1;
          if (a.prop_weight) a.prop_weight[c] = dHe;
          if (a.prop_slpa) a.prop_slpa[c] = fmin(dHe, 0.0);
          a.logp_out[c] = accept ? logp1 : logp;
        }
        if (l < D) {
          const size_t i = (size_t)c * Dp + l;
          if (a.prop_q) static_cast<float*>(a.prop_q)[i] = S.q[l];
          if (a.prop_p) static_cast<float*>(a.prop_p)[i] = -S.p[l];
          if (a.prop_v) static_cast<float*>(a.prop_v)[i] = -S.u[l];
          if (a.prop_g) static_cast<float*>(a.prop_g)[i] = S.g[l];
          static_cast<float*>(a.q_out)[i] = accept ? S.q[l] : S.q0[l];
          if (a.g_out)
            static_cast<float*>(a.g_out)[i] = accept ? S.g[l] : static_cast<const float*>(a.g_in)[i];
        }
      }
      if (accept) {
        if (l < MAXD) S.q0[l] = S.q[l];
        if (!STEP && a.final_g != nullptr && l < D) static_cast<float*>(a.final_g)[(size_t)c * Dp + l] = S.g[l];
        logp = logp1;
        cur = 1 - cur;
        ++n_acc;
      }
      __syncthreads();
      if (!STEP) {
        if (a.samples != nullptr && l < D)
          static_cast<float*>(a.samples)[((size_t)tr * a.C + c) * Dp + l] = S.q0[l];
        if (a.sample_acc != nullptr && l == 0) a.sample_acc[(size_t)tr * a.C + c] = (float)p_acc;
      }
      if (++tr == T_steps) break;
      stage = kNoise;
      continue;
    }
    // ---- stage kTraj: one evaluation of the vector field ----
    if (mode == 0 && l < MAXD) {  // z0 of step s (integrators.py:84 evaluates once outside the loop)
      q0v = S.q[l];
      p0v = S.p[l];
    }
    field_pass(S, Xr, Xs, yf, E0, E1, SL, npad, D, hs, inv_prior, mode == 2);
    have_ginv = false;
    ++n_fev;
    if (mode == 2) {  // explicit half step from the last iterate (integrators.py:146-148)
      if (l < D) {
        S.q[l] = fmaf(hs, S.u[l], S.q[l]);
        S.p[l] = fmaf(-hs, S.dH[l], S.p[l]);
      }
      __syncthreads();
      mode = 0;
      if (++s == Lcur) stage = kEnd;
      continue;
    }
    {
      float dq = 0.0f;
      if (l < D) {
        const float qn = fmaf(hs, S.u[l], q0v);
        const float pn = fmaf(-hs, S.dH[l], p0v);
        dq = nan_max(fabsf(qn - S.q[l]), fabsf(pn - S.p[l]));
        S.q[l] = qn;
        S.p[l] = pn;
      }
      norm = block_max(S, (double)dq);
    }
    if (mode == 0) {
      n = 0;
      mode = 1;
    } else {
      ++n;
    }
    if (!(n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol)) {
      if (STEP && l == 0) {
        const size_t idx = (size_t)c * a.L + s;
        if (a.fp_iters) a.fp_iters[idx] = n;
        if (a.fp_norm) a.fp_norm[idx] = norm;
        if (a.fp_success) a.fp_success[idx] = (uint8_t)(isfinite(norm) && norm <= a.conv_tol);
      }
      mode = 2;
    }
  }

  if (!STEP) {
    if (l < D) static_cast<float*>(a.final_q)[(size_t)c * Dp + l] = S.q0[l];
    if (l == 0) {
      a.final_logp[c] = logp;
      if (a.num_accepted) a.num_accepted[c] = n_acc;
      if (a.num_fevals) a.num_fevals[c] = n_fev;
    }
  }
  zero_row_tails<float, STEP>(a, c, D, l, NT);
}

// One workgroup of four waves per chain of a posterior packed with four
// feature blocks (65..128 features); chains of narrower posteriors leave at
// once.  Blocks are dealt to the XCDs as in the one-wave kernel.
template <bool STEP>
__global__ void __launch_bounds__(NT, 1) logistic_rmhmc_f32_wide_kernel(const LogisticArgs a) {
  __shared__ Smem S;
  const int per_xcd = gridDim.x / 8;
  const int slot = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (slot >= a.C) return;
  const int c = a.dispatch_order ? a.dispatch_order[slot] : slot;
  if (c < 0 || c >= a.C) return;  // not a permutation entry: nothing runs in this slot
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;  // invalid posterior index: chain left untouched
  if (a.packed_hdr[post].nb != NB) return;
  run_chain<STEP>(a, S, c, post);
}

}  // namespace lwide

hipError_t launch_logistic_rmhmc_f32_wide(const LogisticArgs& a, hipStream_t stream) {
  const int blocks = (a.C + 7) / 8 * 8;
  if (a.q_out != nullptr)
    hipLaunchKernelGGL((lwide::logistic_rmhmc_f32_wide_kernel<true>), dim3(blocks), dim3(lwide::NT), 0, stream, a);
  else
    hipLaunchKernelGGL((lwide::logistic_rmhmc_f32_wide_kernel<false>), dim3(blocks), dim3(lwide::NT), 0, stream, a);
  return hipGetLastError();
}

}  // namespace gjx
