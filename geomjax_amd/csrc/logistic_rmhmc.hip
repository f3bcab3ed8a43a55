// RMHMC (implicit midpoint, fixed-point solve) for Bayesian logistic
// regression with the Fisher metric G(q) = X^T diag(s(1-s)) X + I / prior_var,
// s = sigmoid(X q).  Float64 path: the two dense contractions per evaluation
// of the vector field run on v_mfma_f64_16x16x4f64.
//
// Reference: geomjax/rmhmc/rmhmc.py:131-174,376-462,
// geomjax/rmhmc/integrators.py:53-156, geomjax/rmhmc/metrics.py:42-129
// (kinetic energy -log N(p; 0, G) through a Cholesky, differentiated by
// autodiff), geomjax/mcmc/proposal.py:87-185.  The model and metric are the
// user-supplied logdensity_fn / metric_fn of the logistic configurations.
//
// Closed forms (w_n = s(1-s)(1-2s) = d lambda / d eta, r_n = y_n - s_n):
//   dT/dp = u = G^-1 p
//   dT/dq = 1/2 X^T [ w (a - b^2) ],  a_n = x_n^T G^-1 x_n,  b_n = x_n . u
//   grad l = X^T r - q / prior_var
// so dH/dq = X^T [ w (a - b^2) / 2 - r ] + q / prior_var.
//
// One workgroup (256 lanes, 4 waves) owns one chain for the whole launch, so
// every chain runs its own fixed-point iteration count.  X streams through LDS
// in chunks of 32 rows.  Per evaluation:
//   pass 1: eta = X q, lambda, and the Gram X^T diag(lambda) X on MFMA
//           (10 lower 16x16 tiles of the 64x64 matrix, tiles owned by waves);
//   sweep : G^-1 by the symmetric sweep operator with G held in registers
//           (4x4 entries per lane, one barrier per pivot), log det G = sum of
//           the pivots' logs;
//   pass 2: eta, s, w, r, b = X u on VALU, V = X_chunk G^-1 on MFMA,
//           a_n = x_n . V_n, h_n, and X^T h.
// The momentum draw p0 = chol(G) z uses a right-looking Cholesky in LDS
// (once per transition).
#include "common.h"
#include "logistic.h"

namespace gjx {

namespace {

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

// Pass 1: Gram G = X^T diag(lambda) X + I/prior_var into S.Gi.  With want_lg
// also returns the log-likelihood sum and leaves X^T (y - s) in S.g.
__device__ void pass1(Smem& S, const double* X, const double* y, int N, int D, double prior_var, bool want_lg,
                      double& loglik) {
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int nb = (D + 15) / 16;
  f64x4 acc[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) acc[m] = f64x4{0, 0, 0, 0};
  double ll = 0.0, gacc = 0.0;
  for (int n0 = 0; n0 < N; n0 += RCH) {
    load_chunk(S, X, y, N, D, n0);
    __syncthreads();
    {
      const int r = t / 8, c0 = (t % 8) * 8;
      double eta = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) eta += S.Xs[r * LDX + c0 + e] * S.q[c0 + e];
      eta = group_sum<8>(eta);
      if (t % 8 == 0) {
        const bool ok = n0 + r < N;
        const double s = sigmoid(eta);
        S.lam[r] = ok ? s * (1.0 - s) : 0.0;
        S.rr[r] = ok ? S.yv[r] - s : 0.0;
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
    if (want_lg) {
      const int k = t % 64;
#pragma unroll
      for (int m = 0; m < RCH / 4; ++m) {
        const int r = t / 64 + 4 * m;
        gacc += S.Xs[r * LDX + k] * S.rr[r];
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
  __syncthreads();
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

// Symmetric sweep of the first D pivots: A <- -A^-1.  Returns log det A.
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
    logdet += log(d);
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

// Pass 2: S.dH = X^T [ w (a - b^2)/2 - r ] + q / prior_var, with G^-1 in S.Gi
// and u = G^-1 p in S.u.
__device__ void pass2(Smem& S, const double* X, const double* y, int N, int D, double prior_var) {
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
      S.hh[t] = 0.5 * S.lam[t] * (a - b * b) - S.rr[t];
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
  if (t < D) S.dH[t] = S.part[0][t] + S.part[1][t] + S.part[2][t] + S.part[3][t] + S.q[t] / prior_var;
  __syncthreads();
}

}  // namespace

template <bool STEP>
__global__ void __launch_bounds__(NT, 2) logistic_rmhmc_f64_kernel(const LogisticArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& S = *reinterpret_cast<Smem*>(smem_raw);
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  if (c >= a.C) return;  // whole workgroups leave together
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;  // invalid posterior index: chain left untouched
  const int N = a.post_N[post], D = a.post_D[post];
  const double* X = a.X + a.post_xoff[post];
  const double* y = a.y + a.post_yoff[post];
  const int Dp = a.Dmax;  // row stride of the chain-major state arrays
  const double h = 0.5 * a.eps;
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;
  const int T_steps = STEP ? 1 : a.T;

  // chain state
  if (t < MAXD) {
    const double qv = t < D ? static_cast<const double*>(a.q_in)[(size_t)c * Dp + t] : 0.0;
    S.q0[t] = qv;
    S.q[t] = qv;
    S.qs[t] = 0.0;
  }
  double logp = a.logp_in[c];
  int n_acc = 0;
  int64_t n_fev = 0;
  __syncthreads();

  for (int tr = 0; tr < T_steps; ++tr) {
    // ---- noise ----
    const uint64_t tt = a.counter + (uint64_t)tr;
    double uacc;
    if (a.z != nullptr) {
      if (t < MAXD) S.z[t] = t < D ? static_cast<const double*>(a.z)[((size_t)tr * a.C + c) * Dp + t] : 0.0;
      uacc = a.u[(size_t)tr * a.C + c];
    } else {
      if (t < MAXD / 2) {
        u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * t),
                                      (uint32_t)(tt >> 32)}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        double n0, n1;
        normal2(r, n0, n1);
        S.z[2 * t] = 2 * t < D ? n0 : 0.0;
        S.z[2 * t + 1] = 2 * t + 1 < D ? n1 : 0.0;
      }
      u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamAccept, (uint32_t)(tt >> 32)},
                              (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
      uacc = u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
    }
    // the current position S.q0 is the chain state; S.q = q0
    if (t < MAXD) S.q[t] = S.q0[t];
    __syncthreads();

    // ---- momentum p0 = chol(G(q0)) z and H0 ----
    double dummy = 0.0;
    pass1(S, X, y, N, D, prior, false, dummy);
    double A[4][4];
    load_block(S, A);
    __syncthreads();
    cholesky(S, D);
    if (t < MAXD) {
      double s = 0.0;
      for (int k = 0; k <= t && k < D; ++k) s += S.Gi[t * LDX + k] * S.z[k];
      S.p0[t] = t < D ? s : 0.0;
      S.p[t] = S.p0[t];
    }
    __syncthreads();
    if (STEP && a.momentum != nullptr && t < D) static_cast<double*>(a.momentum)[(size_t)c * Dp + t] = S.p0[t];
    double logdet = sweep(S, A, D);
    __syncthreads();
    store_neg_block(S, A);
    __syncthreads();
    ginv_times(S, S.p, S.u, D);
    double H0;
    {
      const double pu = block_sum(S, t < D ? S.p[t] * S.u[t] : 0.0);
      H0 = -logp + 0.5 * pu + 0.5 * logdet + ke_const;
    }
    bool have_ginv = true;  // S.Gi = G(S.q)^-1 and S.u = G^-1 S.p

    // one evaluation of the vector field at (S.q, S.p): S.u = dH/dp, S.dH = dH/dq
    auto feval = [&]() {
      if (!have_ginv) {
        double dd = 0.0;
        pass1(S, X, y, N, D, prior, false, dd);
        load_block(S, A);
        __syncthreads();
        sweep(S, A, D);
        __syncthreads();
        store_neg_block(S, A);
        __syncthreads();
        ginv_times(S, S.p, S.u, D);
      }
      pass2(S, X, y, N, D, prior);
      have_ginv = false;
      ++n_fev;
    };

    // ---- trajectory ----
    for (int s = 0; s < a.L; ++s) {
      // z0 of this step is (S.q, S.p), kept in registers of lanes t < D
      double q0v = 0.0, p0v = 0.0;
      if (t < D) {
        q0v = S.q[t];
        p0v = S.p[t];
      }
      // first evaluation outside the loop (integrators.py:84)
      feval();
      double norm;
      {
        double dq = 0.0;
        if (t < D) {
          const double qn = q0v + h * S.u[t];
          const double pn = p0v - h * S.dH[t];
          dq = nan_max(fabs(qn - S.q[t]), fabs(pn - S.p[t]));
          S.q[t] = qn;
          S.p[t] = pn;
        }
        norm = block_max(S, dq);
      }
      int n = 0;
      while (n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol) {
        feval();
        double dq = 0.0;
        if (t < D) {
          const double qn = q0v + h * S.u[t];
          const double pn = p0v - h * S.dH[t];
          dq = nan_max(fabs(qn - S.q[t]), fabs(pn - S.p[t]));
          S.q[t] = qn;
          S.p[t] = pn;
        }
        norm = block_max(S, dq);
        ++n;
      }
      if (STEP && t == 0) {
        const size_t idx = (size_t)c * a.L + s;
        if (a.fp_iters) a.fp_iters[idx] = n;
        if (a.fp_norm) a.fp_norm[idx] = norm;
        if (a.fp_success) a.fp_success[idx] = (uint8_t)(isfinite(norm) && norm <= a.conv_tol);
      }
      // explicit half step from the last iterate (integrators.py:146-148)
      feval();
      if (t < D) {
        S.q[t] = S.q[t] + h * S.u[t];
        S.p[t] = S.p[t] - h * S.dH[t];
      }
      __syncthreads();
    }

    // ---- end: l, grad l, G^-1, velocity and energy at (q_L, p_L) ----
    double loglik = 0.0;
    pass1(S, X, y, N, D, prior, true, loglik);  // S.g = grad l(q_L)
    load_block(S, A);
    __syncthreads();
    logdet = sweep(S, A, D);
    __syncthreads();
    store_neg_block(S, A);
    __syncthreads();
    ginv_times(S, S.p, S.u, D);
    const double qq = block_sum(S, t < D ? S.q[t] * S.q[t] : 0.0);
    const double pu = block_sum(S, t < D ? S.p[t] * S.u[t] : 0.0);
    const double logp1 = loglik - 0.5 * qq / prior + lp_const;
    const double H1 = -logp1 + 0.5 * pu + 0.5 * logdet + ke_const;

    double dHe = H0 - H1;
    if (isnan(dHe)) dHe = -INFINITY;
    const double p_acc = fmin(exp(dHe), 1.0);
    const bool accept = uacc < p_acc;

    if (STEP) {
      if (t == 0) {
        if (a.acc_rate) a.acc_rate[c] = p_acc;
        if (a.accepted) a.accepted[c] = (uint8_t)accept;
        if (a.divergent) a.divergent[c] = (uint8_t)(-dHe > a.div_threshold);
        if (a.energy) a.energy[c] = H1;
        if (a.prop_logp) a.prop_logp[c] = logp1;
        if (a.prop_weight) a.prop_weight[c] = dHe;
        if (a.prop_slpa) a.prop_slpa[c] = fmin(dHe, 0.0);
        a.logp_out[c] = accept ? logp1 : logp;
      }
      if (t < D) {
        const size_t i = (size_t)c * Dp + t;
        double* qo = static_cast<double*>(a.q_out);
        if (a.prop_q) static_cast<double*>(a.prop_q)[i] = S.q[t];
        if (a.prop_p) static_cast<double*>(a.prop_p)[i] = -S.p[t];
        if (a.prop_v) static_cast<double*>(a.prop_v)[i] = -S.u[t];
        if (a.prop_g) static_cast<double*>(a.prop_g)[i] = S.g[t];
        qo[i] = accept ? S.q[t] : S.q0[t];
        if (a.g_out)
          static_cast<double*>(a.g_out)[i] = accept ? S.g[t] : static_cast<const double*>(a.g_in)[i];
      }
    }
    if (accept) {
      if (t < MAXD) S.q0[t] = S.q[t];
      logp = logp1;
      ++n_acc;
    }
    __syncthreads();
    if (!STEP) {
      if (a.samples != nullptr && t < D)
        static_cast<double*>(a.samples)[((size_t)tr * a.C + c) * Dp + t] = S.q0[t];
      if (a.sample_acc != nullptr && t == 0) a.sample_acc[(size_t)tr * a.C + c] = (float)p_acc;
    }
  }

  if (!STEP) {
    if (t < D) static_cast<double*>(a.final_q)[(size_t)c * Dp + t] = S.q0[t];
    if (t == 0) {
      a.final_logp[c] = logp;
      if (a.num_accepted) a.num_accepted[c] = n_acc;
      if (a.num_fevals) a.num_fevals[c] = n_fev;
    }
  }
}

// value_and_grad of the logistic log density (the reference's init,
// geomjax/rmhmc/rmhmc.py:96-98): one workgroup per chain, pass 1 with the
// likelihood and gradient accumulations.
__global__ void __launch_bounds__(NT) logistic_value_and_grad_kernel(const LogisticArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem& S = *reinterpret_cast<Smem*>(smem_raw);
  const int t = threadIdx.x, c = blockIdx.x;
  if (c >= a.C) return;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  const int N = a.post_N[post], D = a.post_D[post];
  const double* X = a.X + a.post_xoff[post];
  const double* y = a.y + a.post_yoff[post];
  if (t < MAXD) S.q[t] = t < D ? static_cast<const double*>(a.q_in)[(size_t)c * a.Dmax + t] : 0.0;
  __syncthreads();
  double ll = 0.0;
  pass1(S, X, y, N, D, a.prior_var, true, ll);
  const double qq = block_sum(S, t < D ? S.q[t] * S.q[t] : 0.0);
  if (t == 0 && a.logp_out != nullptr)
    a.logp_out[c] = ll - 0.5 * qq / a.prior_var - 0.5 * (double)D * log(2.0 * M_PI * a.prior_var);
  if (t < D) static_cast<double*>(a.g_out)[(size_t)c * a.Dmax + t] = S.g[t];
}

hipError_t launch_logistic_value_and_grad(const LogisticArgs& a, hipStream_t stream) {
  const size_t smem = sizeof(Smem);
  (void)hipFuncSetAttribute((const void*)logistic_value_and_grad_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  hipLaunchKernelGGL(logistic_value_and_grad_kernel, dim3(a.C), dim3(NT), smem, stream, a);
  return hipGetLastError();
}

hipError_t launch_logistic_rmhmc_f64(const LogisticArgs& a, hipStream_t stream) {
  const size_t smem = sizeof(Smem);
  if (a.q_out != nullptr) {
    (void)hipFuncSetAttribute((const void*)logistic_rmhmc_f64_kernel<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipLaunchKernelGGL(logistic_rmhmc_f64_kernel<true>, dim3(a.C), dim3(NT), smem, stream, a);
  } else {
    (void)hipFuncSetAttribute((const void*)logistic_rmhmc_f64_kernel<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipLaunchKernelGGL(logistic_rmhmc_f64_kernel<false>, dim3(a.C), dim3(NT), smem, stream, a);
    if (a.final_g != nullptr) {
      // gradient at the final positions: value_and_grad on them (log density already carried)
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      LogisticArgs b = a;
      b.q_in = a.final_q;
      b.logp_out = nullptr;
      b.g_out = a.final_g;
      return launch_logistic_value_and_grad(b, stream);
    }
  }
  return hipGetLastError();
}

}  // namespace gjx
