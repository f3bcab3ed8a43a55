// Device routines of the logistic-regression kernels (Fisher metric): LDS
// layout, chunked passes over X with MFMA Gram / quadratic forms, the
// symmetric sweep inverse and the Cholesky used for momentum draws.
#pragma once

#include "common.h"

namespace gjx {
namespace logistic {

constexpr int NT = 256;
constexpr int MAXD = 64;
constexpr int RCH = 32;        // rows of X per chunk
constexpr int LDX = MAXD + 1;  // padded LDS row stride (doubles)

typedef double f64x4 __attribute__((ext_vector_type(4)));

struct Smem {
  double Xs[RCH * LDX];
  double Gi[MAXD * LDX];  // G, then the Cholesky factor or G^-1
  double yv[RCH];
  double lam[RCH];  // lambda (pass 1) / w (pass 2)
  double rr[RCH];   // y - s
  double bb[RCH];   // x_n . u
  double aa[2][RCH];
  double hh[RCH];
  double q0[MAXD], p0[MAXD], q[MAXD], p[MAXD], u[MAXD], dH[MAXD], g[MAXD], z[MAXD], qs[MAXD];
  double gv[MAXD], vt[MAXD];
  double rowbuf[2][MAXD];
  double part[4][MAXD];
  double red[8];
};

// the 10 lower-triangular 16x16 tiles (bi >= bj) of a 64x64 matrix
__device__ __forceinline__ void tile_of(int t, int& bi, int& bj) {
  // enumerate (0,0),(1,0),(1,1),(2,0),(2,1),(2,2),(3,0),(3,1),(3,2),(3,3)
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= t) ++i;
  bi = i;
  bj = t - i * (i + 1) / 2;
}

__device__ __forceinline__ double block_sum(Smem& S, double v) {
  v = group_sum<64>(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) S.red[w] = v;
  __syncthreads();
  return S.red[0] + S.red[1] + S.red[2] + S.red[3];
}

__device__ __forceinline__ double block_max(Smem& S, double v) {
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
__device__ __forceinline__ void load_chunk(Smem& S, const double* X, const double* y, int N, int D, int n0) {
  const int t = threadIdx.x;
  const int r = t / 8, c0 = (t % 8) * 8;
  const int n = n0 + r;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int col = c0 + e;
    S.Xs[r * LDX + col] = (n < N && col < D) ? X[(size_t)n * D + col] : 0.0;
  }
  if (t < RCH) S.yv[t] = (n0 + t < N) ? y[n0 + t] : 0.0;
}

// Gram pass at S.q: S.Gi = X^T diag(lambda + coef w (X vec)) X + I/prior_var
// (vec == nullptr: the Fisher metric G itself).  With want_lg it also returns
// the log-likelihood sum and leaves grad l = X^T (y - s) - q/prior_var in S.g;
// with want_gv it leaves G vec = X^T (lambda (X vec)) + vec/prior_var in S.gv.
__device__ void pass_gram(Smem& S, const double* X, const double* y, int N, int D, double prior_var,
                          const double* vec, double coef, bool want_lg, bool want_gv, double& loglik) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int nb = (D + 15) / 16;
  f64x4 acc[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) acc[m] = f64x4{0, 0, 0, 0};
  double ll = 0.0, gacc = 0.0, gvacc = 0.0;
  for (int n0 = 0; n0 < N; n0 += RCH) {
    load_chunk(S, X, y, N, D, n0);
    __syncthreads();
    {
      const int r = t / 8, c0 = (t % 8) * 8;
      double eta = 0.0, xv = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        eta += S.Xs[r * LDX + c0 + e] * S.q[c0 + e];
        if (vec) xv += S.Xs[r * LDX + c0 + e] * vec[c0 + e];
      }
      eta = group_sum<8>(eta);
      if (vec) xv = group_sum<8>(xv);
      if (t % 8 == 0) {
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
    for (int ks = 0; ks < RCH / 4; ++ks) {
      const int nr = ks * 4 + l / 16;
      double xb[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) xb[b] = S.Xs[nr * LDX + b * 16 + l % 16];
      const double lm = S.lam[nr];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const int tt = w + 4 * m;
        if (tt >= 10) continue;
        int bi, bj;
        tile_of(tt, bi, bj);
        if (bi >= nb) continue;
        acc[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(lm * xb[bi], xb[bj], acc[m], 0, 0, 0);
      }
    }
    if (want_lg || want_gv) {
      const int k = t % 64;
#pragma unroll
      for (int m = 0; m < RCH / 4; ++m) {
        const int r = t / 64 + 4 * m;
        const double xk = S.Xs[r * LDX + k];
        gacc += xk * S.rr[r];
        gvacc += xk * S.bb[r];
      }
    }
    __syncthreads();
  }
  // tiles -> S.Gi (both triangles), C[4r + l/16][l%16] layout
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int tt = w + 4 * m;
    if (tt >= 10) continue;
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
      S.Gi[i * LDX + j] = v;
      S.Gi[j * LDX + i] = v;
    }
  }
  if (want_lg) {
    S.part[w][t % 64] = gacc;
    loglik = block_sum(S, ll);  // contains the barrier that publishes part[]
    if (t < D) S.g[t] = S.part[0][t] + S.part[1][t] + S.part[2][t] + S.part[3][t] - S.q[t] / prior_var;
  }
  if (want_gv) {
    __syncthreads();
    S.part[w][t % 64] = gvacc;
    __syncthreads();
    if (t < D) S.gv[t] = S.part[0][t] + S.part[1][t] + S.part[2][t] + S.part[3][t] + vec[t] / prior_var;
  }
  __syncthreads();
}

__device__ __forceinline__ void pass1(Smem& S, const double* X, const double* y, int N, int D, double prior_var,
                                      bool want_lg, double& loglik) {
  pass_gram(S, X, y, N, D, prior_var, nullptr, 0.0, want_lg, false, loglik);
}

// Load this lane's 4x4 block of S.Gi into registers (entries past D are 0).
__device__ __forceinline__ void load_block(const Smem& S, double (&A)[4][4]) {
  const int t = threadIdx.x;
  const int i0 = 4 * (t / 16), j0 = 4 * (t % 16);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) A[a][b] = S.Gi[(i0 + a) * LDX + j0 + b];
}

// Symmetric sweep of the first D pivots: A <- -A^-1.  Returns log |det A|.
__device__ double sweep(Smem& S, double (&A)[4][4], int D) {
  const int t = threadIdx.x;
  const int i0 = 4 * (t / 16), j0 = 4 * (t % 16);
  double logdet = 0.0;
  for (int k = 0; k < D; ++k) {
    double* buf = S.rowbuf[k & 1];
    if (t / 16 == k / 4) {
#pragma unroll
      for (int b = 0; b < 4; ++b) buf[j0 + b] = A[k % 4][b];
    }
    __syncthreads();
    const double d = buf[k];
    const double rd = 1.0 / d;
    logdet += log(fabs(d));  // log |det|: G +- (eps/2) Omega~ may be indefinite
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = i0 + a;
      const double aik = buf[i];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = j0 + b;
        if (i != k && j != k)
          A[a][b] -= aik * buf[j] * rd;
        else if (i == k && j == k)
          A[a][b] = -rd;
        else
          A[a][b] *= rd;
      }
    }
  }
  return logdet;
}

// Write G^-1 = -A into S.Gi.
__device__ __forceinline__ void store_neg_block(Smem& S, const double (&A)[4][4]) {
  const int t = threadIdx.x;
  const int i0 = 4 * (t / 16), j0 = 4 * (t % 16);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) S.Gi[(i0 + a) * LDX + j0 + b] = -A[a][b];
}

// Right-looking Cholesky of the leading D x D block of S.Gi in place (lower).
// Column k is scaled after the elimination, so each step reads unscaled
// columns: one barrier per pivot.  Returns log det (= 2 sum log L_kk).
__device__ double cholesky(Smem& S, int D) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  for (int k = 0; k < D; ++k) {
    const double rd = 1.0 / S.Gi[k * LDX + k];
    for (int i = k + 1 + w; i < D; i += 4) {
      const double lik = S.Gi[i * LDX + k] * rd;
      const int j = k + 1 + l;
      if (j <= i) S.Gi[i * LDX + j] -= lik * S.Gi[j * LDX + k];
    }
    __syncthreads();
  }
  // pivots d_k sit on the diagonal; scale the columns
  if (t < D) S.qs[t] = sqrt(S.Gi[t * LDX + t]);
  __syncthreads();
  double logdet = 0.0;
  for (int k = 0; k < D; ++k) logdet += 2.0 * log(S.qs[k]);
  for (int idx = t; idx < D * D; idx += NT) {
    const int i = idx / D, k = idx % D;
    if (i > k) S.Gi[i * LDX + k] /= S.qs[k];
    else if (i == k) S.Gi[i * LDX + k] = S.qs[k];
  }
  __syncthreads();
  return logdet;
}

// u = G^-1 p with G^-1 in S.Gi (vector p in LDS).
__device__ __forceinline__ void ginv_times(Smem& S, const double* p, double* out, int D) {
  const int t = threadIdx.x;
  if (t < D) {
    double s = 0.0;
    for (int j = 0; j < D; ++j) s += S.Gi[t * LDX + j] * p[j];
    out[t] = s;
  }
  __syncthreads();
}

// Quadratic-form pass with G^-1 in S.Gi, a_n = x_n^T G^-1 x_n:
//   rmhmc = true : S.dH = X^T [ w (a - b^2)/2 - r ] + q / prior_var, b = X S.u
//   rmhmc = false: S.dH = X^T (w a) = grad log det G.
__device__ void pass_quad(Smem& S, const double* X, const double* y, int N, int D, double prior_var, bool rmhmc) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int ksteps = (D + 3) / 4;
  const int rb = w / 2;            // row block (16 rows) of this wave
  const int cb0 = 2 * (w % 2);     // first of its two column blocks
  double dacc = 0.0;
  for (int n0 = 0; n0 < N; n0 += RCH) {
    load_chunk(S, X, y, N, D, n0);
    __syncthreads();
    {
      const int r = t / 8, c0 = (t % 8) * 8;
      double eta = 0.0, bq = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const double xv = S.Xs[r * LDX + c0 + e];
        eta += xv * S.q[c0 + e];
        bq += xv * S.u[c0 + e];
      }
      eta = group_sum<8>(eta);
      bq = group_sum<8>(bq);
      if (t % 8 == 0) {
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
      const double af = S.Xs[(rb * 16 + l % 16) * LDX + ks * 4 + l / 16];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double bf = S.Gi[(ks * 4 + l / 16) * LDX + (cb0 + c) * 16 + l % 16];
        v[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, v[c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rb * 16 + 4 * r + l / 16;
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 2; ++c) s += v[c][r] * S.Xs[row * LDX + (cb0 + c) * 16 + l % 16];
      s = group_sum<16>(s);
      if (l % 16 == 0) S.aa[w % 2][row] = s;
    }
    __syncthreads();
    if (t < RCH) {
      const double a = S.aa[0][t] + S.aa[1][t];
      const double b = S.bb[t];
      S.hh[t] = rmhmc ? 0.5 * S.lam[t] * (a - b * b) - S.rr[t] : S.lam[t] * a;
    }
    __syncthreads();
    {
      const int k = t % 64;
#pragma unroll
      for (int m = 0; m < RCH / 4; ++m) {
        const int r = t / 64 + 4 * m;
        dacc += S.Xs[r * LDX + k] * S.hh[r];
      }
    }
    __syncthreads();
  }
  S.part[w][t % 64] = dacc;
  __syncthreads();
  if (t < D) S.dH[t] = S.part[0][t] + S.part[1][t] + S.part[2][t] + S.part[3][t] + (rmhmc ? S.q[t] / prior_var : 0.0);
  __syncthreads();
}

__device__ __forceinline__ void pass2(Smem& S, const double* X, const double* y, int N, int D, double prior_var) {
  pass_quad(S, X, y, N, D, prior_var, true);
}

}  // namespace logistic
}  // namespace gjx
