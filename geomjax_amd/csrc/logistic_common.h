// Device routines of the float64 logistic-regression kernels (Fisher metric):
// LDS layout, chunked passes over X with MFMA Gram / quadratic forms, the
// symmetric sweep inverse and the Cholesky used for momentum draws.
//
// One workgroup of NT = 256 lanes (4 waves) per chain.  Everything is
// parameterised by the feature capacity MD (64 or 128) through Cfg<MD>:
//   RCH  rows of X per LDS chunk (32, or 16 so that a 128 x 128 float64
//        inverse and the chunk fit the 160 KiB of LDS together),
//   LDX  padded row stride, TB 16-wide tiles per dimension, NLT lower tiles of
//        the Gram, TPW tiles per wave, BS the BS x BS block of the inverse each
//        lane holds during the sweep (16 x 16 lanes cover the matrix),
//   TPR  lanes per row in the per-row passes (CPT = 8 columns each), RG row
//        groups of the column-per-lane accumulations (X^T r, X^T h, G v).
#pragma once

#include "common.h"

namespace gjx {
namespace logistic {

constexpr int NT = 256;

template <int MD>
struct Cfg {
  static_assert(MD == 64 || MD == 128, "feature capacity 64 or 128");
  static constexpr int MAXD = MD;
  static constexpr int RCH = MD == 64 ? 32 : 16;
  static constexpr int LDX = MD + 1;
  static constexpr int TB = MD / 16;
  static constexpr int NLT = TB * (TB + 1) / 2;
  static constexpr int TPW = (NLT + 3) / 4;
  static constexpr int BS = MD / 16;
  static constexpr int TPR = NT / RCH;
  static constexpr int CPT = MD / TPR;
  static constexpr int RG = NT / MD;
  static constexpr int CB2 = TB / 2;          // column-block pairs of the quadratic-form pass
  static constexpr int RB = RCH / 16;         // 16-row blocks per chunk
  static_assert(RB * CB2 == 4, "the quadratic-form pass gives each wave one (row block, column pair)");
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

template <class C>
struct Smem {
  // the chunk of X; outside the chunk loops its space holds the sweep's row
  // broadcast and the partial sums of the column-per-lane accumulations
  union {
    double Xs[C::RCH * C::LDX];
    struct {
      double rowbuf[2][C::MAXD];
      double part[4][C::MAXD];
    };
  };
  double Gi[C::MAXD * C::LDX];  // G, then the Cholesky factor or G^-1
  double yv[C::RCH];
  double lam[C::RCH];  // lambda (pass 1) / w (pass 2)
  double rr[C::RCH];   // y - s
  double bb[C::RCH];   // x_n . u
  double aa[4][C::RCH];
  double hh[C::RCH];
  double q0[C::MAXD], p0[C::MAXD], q[C::MAXD], p[C::MAXD], u[C::MAXD], dH[C::MAXD], g[C::MAXD], z[C::MAXD],
      qs[C::MAXD];
  double gv[C::MAXD], vt[C::MAXD];
  double red[8];
};

// the lower-triangular 16 x 16 tiles (bi >= bj) in the order (0,0),(1,0),(1,1),(2,0),...
__device__ __forceinline__ void tile_of(int t, int& bi, int& bj) {
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  bi = i;
  bj = t - i * (i + 1) / 2;
}

template <class C>
__device__ __forceinline__ double block_sum(Smem<C>& S, double v) {
  v = group_sum<64>(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) S.red[w] = v;
  __syncthreads();
  return S.red[0] + S.red[1] + S.red[2] + S.red[3];
}

template <class C>
__device__ __forceinline__ double block_max(Smem<C>& S, double v) {
  v = group_reduce<64, NanMax>(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) S.red[w] = v;
  __syncthreads();
  return nan_max(nan_max(S.red[0], S.red[1]), nan_max(S.red[2], S.red[3]));
}

__device__ __forceinline__ double softplus(double x) { return fmax(x, 0.0) + log1p(exp(-fabs(x))); }
__device__ __forceinline__ double sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }

// Stage rows [n0, n0 + RCH) of X (zero-padded past N and D) and y into LDS.
template <class C>
__device__ __forceinline__ void load_chunk(Smem<C>& S, const double* X, const double* y, int N, int D, int n0) {
  const int t = threadIdx.x;
  const int r = t / (C::MAXD / 8), c0 = (t % (C::MAXD / 8)) * 8;
  const int n = n0 + r;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int col = c0 + e;
    S.Xs[r * C::LDX + col] = (n < N && col < D) ? X[(size_t)n * D + col] : 0.0;
  }
  if (t < C::RCH) S.yv[t] = (n0 + t < N) ? y[n0 + t] : 0.0;
}

// Sum of the RG row-group partials of a column-per-lane accumulation: lane
// t < MAXD receives column t's total.
template <class C>
__device__ __forceinline__ double column_total(Smem<C>& S, int t) {
  double s = 0.0;
#pragma unroll
  for (int g = 0; g < C::RG; ++g) s += S.part[g][t];
  return s;
}

// Gram pass at S.q: S.Gi = X^T diag(lambda + coef w (X vec)) X + I / prior_var
// (vec == nullptr: the Fisher metric G itself).  With want_lg it also returns
// the log-likelihood sum and leaves grad l = X^T (y - s) - q/prior_var in S.g;
// with want_gv it leaves G vec = X^T (lambda (X vec)) + vec/prior_var in S.gv.
template <class C>
__device__ void pass_gram(Smem<C>& S, const double* X, const double* y, int N, int D, double prior_var,
                          const double* vec, double coef, bool want_lg, bool want_gv, double& loglik) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int nb = (D + 15) / 16;
  f64x4 acc[C::TPW];
#pragma unroll
  for (int m = 0; m < C::TPW; ++m) acc[m] = f64x4{0, 0, 0, 0};
  double ll = 0.0, gacc = 0.0, gvacc = 0.0;
  for (int n0 = 0; n0 < N; n0 += C::RCH) {
    load_chunk(S, X, y, N, D, n0);
    __syncthreads();
    {
      const int r = t / C::TPR, c0 = (t % C::TPR) * C::CPT;
      double eta = 0.0, xv = 0.0;
#pragma unroll
      for (int e = 0; e < C::CPT; ++e) {
        eta += S.Xs[r * C::LDX + c0 + e] * S.q[c0 + e];
        if (vec) xv += S.Xs[r * C::LDX + c0 + e] * vec[c0 + e];
      }
      eta = group_sum<C::TPR>(eta);
      if (vec) xv = group_sum<C::TPR>(xv);
      if (t % C::TPR == 0) {
        const bool ok = n0 + r < N;
        const double s = sigmoid(eta);
        const double lam = s * (1.0 - s);
        S.lam[r] = ok ? (vec ? lam + coef * lam * (1.0 - 2.0 * s) * xv : lam) : 0.0;
        S.rr[r] = ok ? S.yv[r] - s : 0.0;
        S.bb[r] = ok ? lam * xv : 0.0;
        if (want_lg && ok) ll += S.yv[r] * eta - softplus(eta);
      }
    }
    __syncthreads();
    // Gram tiles on MFMA: K steps of 4 rows
#pragma unroll 2
    for (int ks = 0; ks < C::RCH / 4; ++ks) {
      const int nr = ks * 4 + l / 16;
      double xb[C::TB];
#pragma unroll
      for (int b = 0; b < C::TB; ++b) xb[b] = S.Xs[nr * C::LDX + b * 16 + l % 16];
      const double lm = S.lam[nr];
#pragma unroll
      for (int m = 0; m < C::TPW; ++m) {
        const int tt = w + 4 * m;
        if (tt >= C::NLT) continue;
        int bi, bj;
        tile_of(tt, bi, bj);
        if (bi >= nb) continue;
        acc[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(lm * xb[bi], xb[bj], acc[m], 0, 0, 0);
      }
    }
    if (want_lg || want_gv) {
      const int k = t % C::MAXD;
#pragma unroll
      for (int m = 0; m < C::RCH / C::RG; ++m) {
        const int r = t / C::MAXD + C::RG * m;
        const double xk = S.Xs[r * C::LDX + k];
        gacc += xk * S.rr[r];
        gvacc += xk * S.bb[r];
      }
    }
    __syncthreads();
  }
  // tiles -> S.Gi (both triangles), C[4r + l/16][l%16] layout
#pragma unroll
  for (int m = 0; m < C::TPW; ++m) {
    const int tt = w + 4 * m;
    if (tt >= C::NLT) continue;
    int bi, bj;
    tile_of(tt, bi, bj);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = bi * 16 + 4 * r + l / 16, j = bj * 16 + l % 16;
      // a diagonal tile holds (i, j) and (j, i) rounded differently: keep the
      // lower one for both so that G is exactly symmetric
      if (i < j) continue;
      double v = acc[m][r];
      if (i == j && i < D) v += 1.0 / prior_var;
      S.Gi[i * C::LDX + j] = v;
      S.Gi[j * C::LDX + i] = v;
    }
  }
  if (want_lg) {
    S.part[t / C::MAXD][t % C::MAXD] = gacc;
    loglik = block_sum(S, ll);  // contains the barrier that publishes part[]
    if (t < D) S.g[t] = column_total(S, t) - S.q[t] / prior_var;
  }
  if (want_gv) {
    __syncthreads();
    S.part[t / C::MAXD][t % C::MAXD] = gvacc;
    __syncthreads();
    if (t < D) S.gv[t] = column_total(S, t) + vec[t] / prior_var;
  }
  __syncthreads();
}

template <class C>
__device__ __forceinline__ void pass1(Smem<C>& S, const double* X, const double* y, int N, int D, double prior_var,
                                      bool want_lg, double& loglik) {
  pass_gram(S, X, y, N, D, prior_var, nullptr, 0.0, want_lg, false, loglik);
}

// Load this lane's BS x BS block of S.Gi into registers (entries past D are 0).
template <class C>
__device__ __forceinline__ void load_block(const Smem<C>& S, double (&A)[C::BS][C::BS]) {
  const int t = threadIdx.x;
  const int i0 = C::BS * (t / 16), j0 = C::BS * (t % 16);
#pragma unroll
  for (int a = 0; a < C::BS; ++a)
#pragma unroll
    for (int b = 0; b < C::BS; ++b) A[a][b] = S.Gi[(i0 + a) * C::LDX + j0 + b];
}

// Symmetric sweep of the first D pivots: A <- -A^-1.  Returns log |det A|.
// Pivots run in groups of BS with the pivot inside a group unrolled, so the
// broadcast row k = BS kk + a is register row a of its owners' blocks.
// With zc != nullptr (A positive definite) it also draws pout = chol(A) zc:
// at pivot k the unswept part of row k is the Schur complement's column, so
// L[j][k] = row_k[j] / sqrt(d_k) for j >= k.
template <class C>
__device__ __forceinline__ double sweep(Smem<C>& S, double (&A)[C::BS][C::BS], int D, const double* zc = nullptr,
                        double* pout = nullptr) {
  const int t = threadIdx.x;
  const int i0 = C::BS * (t / 16), j0 = C::BS * (t % 16);
  double logdet = 0.0;
  double pacc = 0.0;
  for (int kk = 0; kk * C::BS < D; ++kk) {
#pragma unroll
    for (int a0 = 0; a0 < C::BS; ++a0) {
      const int k = kk * C::BS + a0;
      if (k >= D) break;
      double* buf = S.rowbuf[k & 1];
      if (t / 16 == kk) {
#pragma unroll
        for (int b = 0; b < C::BS; ++b) buf[j0 + b] = A[a0][b];
      }
      __syncthreads();
      const double d = buf[k];
      const double rd = 1.0 / d;
      logdet += log(fabs(d));  // log |det|: G +- (eps/2) Omega~ may be indefinite
      if (zc != nullptr && t >= k && t < D) pacc = fma(buf[t] * zc[k], 1.0 / sqrt(d), pacc);
      double bjr[C::BS];
#pragma unroll
      for (int b = 0; b < C::BS; ++b) bjr[b] = buf[j0 + b] * rd;
      // branch-free: a_ij - a_ik a_kj / d off the pivot row and column,
      // a_ik / d on them, -1/d on the pivot
#pragma unroll
      for (int a = 0; a < C::BS; ++a) {
        const int i = i0 + a;
        const double aik = buf[i];
#pragma unroll
        for (int b = 0; b < C::BS; ++b) {
          const int j = j0 + b;
          const double v = A[a][b];
          const double upd = fma(-aik, bjr[b], v);
          const double scaled = v * rd;
          A[a][b] = (i != k && j != k) ? upd : ((i == k && j == k) ? -rd : scaled);
        }
      }
    }
  }
  if (pout != nullptr && t < C::MAXD) pout[t] = t < D ? pacc : 0.0;
  return logdet;
}

// Write G^-1 = -A into S.Gi.
template <class C>
__device__ __forceinline__ void store_neg_block(Smem<C>& S, const double (&A)[C::BS][C::BS]) {
  const int t = threadIdx.x;
  const int i0 = C::BS * (t / 16), j0 = C::BS * (t % 16);
#pragma unroll
  for (int a = 0; a < C::BS; ++a)
#pragma unroll
    for (int b = 0; b < C::BS; ++b) S.Gi[(i0 + a) * C::LDX + j0 + b] = -A[a][b];
}

// Right-looking Cholesky of the leading D x D block of S.Gi in place (lower).
// Column k is scaled after the elimination, so each step reads unscaled
// columns: one barrier per pivot.  Returns log det (= 2 sum log L_kk).
template <class C>
__device__ double cholesky(Smem<C>& S, int D) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  for (int k = 0; k < D; ++k) {
    const double rd = 1.0 / S.Gi[k * C::LDX + k];
    for (int i = k + 1 + w; i < D; i += 4) {
      const double lik = S.Gi[i * C::LDX + k] * rd;
      for (int j = k + 1 + l; j <= i; j += 64) S.Gi[i * C::LDX + j] -= lik * S.Gi[j * C::LDX + k];
    }
    __syncthreads();
  }
  // pivots d_k sit on the diagonal; scale the columns
  if (t < D) S.qs[t] = sqrt(S.Gi[t * C::LDX + t]);
  __syncthreads();
  double logdet = 0.0;
  for (int k = 0; k < D; ++k) logdet += 2.0 * log(S.qs[k]);
  for (int idx = t; idx < D * D; idx += NT) {
    const int i = idx / D, k = idx % D;
    if (i > k) S.Gi[i * C::LDX + k] /= S.qs[k];
    else if (i == k) S.Gi[i * C::LDX + k] = S.qs[k];
  }
  __syncthreads();
  return logdet;
}

// u = G^-1 p with G^-1 in S.Gi (vector p in LDS).
// Entries past D are written as zeros: the passes over X multiply whole padded
// rows by this vector, and 0 times a stale LDS value left by an earlier kernel
// (possibly Inf or NaN) would not be 0.
template <class C>
__device__ __forceinline__ void ginv_times(Smem<C>& S, const double* p, double* out, int D) {
  const int t = threadIdx.x;
  if (t < D) {
    double s = 0.0;
    for (int j = 0; j < D; ++j) s += S.Gi[t * C::LDX + j] * p[j];
    out[t] = s;
  } else if (t < C::MAXD) {
    out[t] = 0.0;
  }
  __syncthreads();
}

// Quadratic-form pass with G^-1 in S.Gi, a_n = x_n^T G^-1 x_n:
//   rmhmc = true : S.dH = X^T [ w (a - b^2)/2 - r ] + q / prior_var, b = X S.u
//   rmhmc = false: S.dH = X^T (w a) = grad log det G.
template <class C>
__device__ void pass_quad(Smem<C>& S, const double* X, const double* y, int N, int D, double prior_var, bool rmhmc) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int ksteps = (D + 3) / 4;
  const int rb = w / C::CB2;          // row block (16 rows) of this wave
  const int cb0 = 2 * (w % C::CB2);   // first of its two column blocks
  double dacc = 0.0;
  for (int n0 = 0; n0 < N; n0 += C::RCH) {
    load_chunk(S, X, y, N, D, n0);
    __syncthreads();
    {
      const int r = t / C::TPR, c0 = (t % C::TPR) * C::CPT;
      double eta = 0.0, bq = 0.0;
#pragma unroll
      for (int e = 0; e < C::CPT; ++e) {
        const double xv = S.Xs[r * C::LDX + c0 + e];
        eta += xv * S.q[c0 + e];
        bq += xv * S.u[c0 + e];
      }
      eta = group_sum<C::TPR>(eta);
      bq = group_sum<C::TPR>(bq);
      if (t % C::TPR == 0) {
        const bool ok = n0 + r < N;
        const double s = sigmoid(eta);
        S.lam[r] = ok ? s * (1.0 - s) * (1.0 - 2.0 * s) : 0.0;
        S.rr[r] = ok ? S.yv[r] - s : 0.0;
        S.bb[r] = bq;
      }
    }
    // V = X_chunk G^-1 on MFMA: this wave's 16 rows x 32 columns
    f64x4 v[2] = {f64x4{0, 0, 0, 0}, f64x4{0, 0, 0, 0}};
    for (int ks = 0; ks < ksteps; ++ks) {
      const double af = S.Xs[(rb * 16 + l % 16) * C::LDX + ks * 4 + l / 16];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double bf = S.Gi[(ks * 4 + l / 16) * C::LDX + (cb0 + c) * 16 + l % 16];
        v[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, v[c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rb * 16 + 4 * r + l / 16;
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 2; ++c) s += v[c][r] * S.Xs[row * C::LDX + (cb0 + c) * 16 + l % 16];
      s = group_sum<16>(s);
      if (l % 16 == 0) S.aa[w % C::CB2][row] = s;
    }
    __syncthreads();
    if (t < C::RCH) {
      double a = 0.0;
#pragma unroll
      for (int s = 0; s < C::CB2; ++s) a += S.aa[s][t];
      const double b = S.bb[t];
      S.hh[t] = rmhmc ? 0.5 * S.lam[t] * (a - b * b) - S.rr[t] : S.lam[t] * a;
    }
    __syncthreads();
    {
      const int k = t % C::MAXD;
#pragma unroll
      for (int m = 0; m < C::RCH / C::RG; ++m) {
        const int r = t / C::MAXD + C::RG * m;
        dacc += S.Xs[r * C::LDX + k] * S.hh[r];
      }
    }
    __syncthreads();
  }
  S.part[t / C::MAXD][t % C::MAXD] = dacc;
  __syncthreads();
  if (t < D) S.dH[t] = column_total(S, t) + (rmhmc ? S.q[t] / prior_var : 0.0);
  __syncthreads();
}

template <class C>
__device__ __forceinline__ void pass2(Smem<C>& S, const double* X, const double* y, int N, int D, double prior_var) {
  pass_quad(S, X, y, N, D, prior_var, true);
}

}  // namespace logistic
}  // namespace gjx
