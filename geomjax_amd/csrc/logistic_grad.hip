// Tangent (forward-mode) derivative of one RMHMC transition with the Fisher
// metric on Bayesian logistic regression with respect to log step_size.
//
// The plain kernel's computation (logistic_rmhmc.hip) carried on pairs
// (value, d/d log eps), seeded with d eps / d log eps = eps.  The momentum
// p0 = chol(G(q0)) z depends on q0 and z only and has no tangent.  Every
// executed fixed-point iterate carries its tangent and the stop decisions use
// values only, so the result is the exact derivative of the unrolled
// computation at the executed iteration counts -- what jax.jacfwd returns
// through the reference integrator (geomjax/rmhmc/integrators.py:53-156, the
// forward derivative the reference's adaptation takes in
// geomjax/adaptation/chees_adaptation_lmcmonge.py:497-519).
//
// Tangent closed forms at an iterate (q, p) with tangents (q', p'), eta = X q,
// s = sigmoid(eta), lambda = s(1-s), w = lambda (1-2s):
//   eta' = X q',  s' = lambda eta',  lambda' = w eta',  w' = lambda (1 - 6s + 6s^2) eta'
//   G' = X^T diag(lambda') X,  (G^-1)' = -G^-1 G' G^-1  (tangent of the sweep)
//   u' = (G^-1)' p + G^-1 p',  b' = X u',  a_n' = x_n^T (G^-1)' x_n
//   h' = w' (a - b^2)/2 + w (a' - 2 b b')/2 + s',   (dH/dq)' = X^T h' + q'/prior_var
//   (log det G)' = sum_k d_k'/d_k over the sweep pivots,  l(q)' = grad l . q'.
// Per chain it returns p = min(1, e^dH), |q_L - q_0|^2, their product, and
// the derivatives (column 0: d/d log step_size; column 1, alpha2, is zero for
// the Fisher metric).
//
// Layout as the plain kernel's: one workgroup of 256 lanes per chain, X
// through LDS in chunks of rows, MFMA Gram and X G^-1 products (twice: values
// and tangents), BS x BS blocks of the metric and of its tangent per lane
// during the sweep.  Two instances:
//   <double, capacity 64>:  the check path (v_mfma_f64_16x16x4f64), G^-1 and
//     its tangent 2 x 33 KiB of LDS;
//   <float, capacity 128>:  the tangent for 65..128 features (float32 state,
//     metric and tangents on v_mfma_f32_16x16x4f32 with float32 accumulation;
//     log likelihood, log-determinant sums, energies and the objective in
//     float64), whose G^-1 and tangent take 2 x 66 KiB of LDS -- the float64
//     pair (2 x 132 KiB) does not fit a CU.
// The two MFMAs place their 16 x 16 results differently in the accumulator
// slots (float64: slot r of lane l holds row 4 r + l/16; float32: row
// 4 (l/16) + r; column l % 16 in both).
#include "common.h"
#include "logistic.h"
#include "logistic_common.h"

namespace gjx {
namespace lgrad {

using namespace logistic;

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <class T>
struct Mf;
template <>
struct Mf<double> {
  typedef f64x4 acc;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int l) { return 4 * r + l / 16; }
};
template <>
struct Mf<float> {
  typedef f32x4 acc;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int l) { return 4 * (l / 16) + r; }
};

template <class T, class C>
struct Smem {
  union {
    T Xs[C::RCH * C::LDX];
    struct {
      T rowbuf[2][C::MAXD];
      T drowbuf[2][C::MAXD];
      T part[4][C::MAXD];
      T dpart[4][C::MAXD];
    };
  };
  T Gi[C::MAXD * C::LDX];   // G, then G^-1
  T dGi[C::MAXD * C::LDX];  // tangents
  T yv[C::RCH], lam[C::RCH], dlam[C::RCH], rr[C::RCH], drr[C::RCH], bb[C::RCH], dbb[C::RCH], hh[C::RCH], dhh[C::RCH];
  T aa[C::CB2][C::RCH], daa[C::CB2][C::RCH];
  // values: chain start, step start (z0), iterate, G^-1 p, dH/dq, grad l, noise
  T q0[C::MAXD], zq[C::MAXD], zp[C::MAXD], q[C::MAXD], p[C::MAXD], u[C::MAXD], dH[C::MAXD], g[C::MAXD], z[C::MAXD];
  // tangents
  T tzq[C::MAXD], tzp[C::MAXD], tq[C::MAXD], tp[C::MAXD], tu[C::MAXD], tdH[C::MAXD];
  double red[8];
};

template <class SM>
__device__ __forceinline__ double bsum(SM& S, double v) {
  v = group_sum<64>(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) S.red[w] = v;
  __syncthreads();
  return S.red[0] + S.red[1] + S.red[2] + S.red[3];
}

template <class SM>
__device__ __forceinline__ double bmax(SM& S, double v) {
  v = group_reduce<64, NanMax>(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) S.red[w] = v;
  __syncthreads();
  return nan_max(nan_max(S.red[0], S.red[1]), nan_max(S.red[2], S.red[3]));
}

template <class T, class C>
__device__ __forceinline__ void load_chunk(Smem<T, C>& S, const double* X, const double* y, int N, int D, int n0) {
  constexpr int MAXD = C::MAXD, LDX = C::LDX, RCH = C::RCH;
  constexpr int PER = MAXD * RCH / NT;  // entries per lane
  const int t = threadIdx.x;
  const int r = t / (MAXD / PER), c0 = (t % (MAXD / PER)) * PER;
  const int n = n0 + r;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int col = c0 + e;
    S.Xs[r * LDX + col] = (n < N && col < D) ? (T)X[(size_t)n * D + col] : T(0);
  }
  if (t < RCH) S.yv[t] = (n0 + t < N) ? (T)y[n0 + t] : T(0);
}

// Gram pass at S.q with tangent S.tq: S.Gi = G, S.dGi = G'.  With want_lg
// also the log-likelihood sum (returned, float64) and S.g = grad l(q).
template <class T, class C>
__device__ void gram_pass(Smem<T, C>& S, const double* X, const double* y, int N, int D, double prior_var,
                          bool want_lg, double& loglik) {
  constexpr int MAXD = C::MAXD, LDX = C::LDX, RCH = C::RCH, TB = C::TB, NLT = C::NLT, TPW = C::TPW, TPR = C::TPR,
                CPT = C::CPT, RG = C::RG;
  using A4 = typename Mf<T>::acc;
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int nb = (D + 15) / 16;
  A4 acc[TPW], tacc[TPW];
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    acc[m] = A4{0, 0, 0, 0};
    tacc[m] = A4{0, 0, 0, 0};
  }
  double ll = 0.0;
  T gacc = T(0);
  for (int n0 = 0; n0 < N; n0 += RCH) {
    load_chunk(S, X, y, N, D, n0);
    __syncthreads();
    {
      const int r = t / TPR, c0 = (t % TPR) * CPT;
      T eta = T(0), teta = T(0);
#pragma unroll
      for (int e = 0; e < CPT; ++e) {
        const T xv = S.Xs[r * LDX + c0 + e];
        eta += xv * S.q[c0 + e];
        teta += xv * S.tq[c0 + e];
      }
      eta = group_sum<TPR>(eta);
      teta = group_sum<TPR>(teta);
      if (t % TPR == 0) {
        const bool ok = n0 + r < N;
        const double s = sigmoid((double)eta);
        const double lam = s * (1.0 - s);
        S.lam[r] = ok ? (T)lam : T(0);
        S.dlam[r] = ok ? (T)(lam * (1.0 - 2.0 * s) * (double)teta) : T(0);
        S.rr[r] = ok ? (T)((double)S.yv[r] - s) : T(0);
        if (want_lg && ok) ll += (double)S.yv[r] * (double)eta - softplus((double)eta);
      }
    }
    __syncthreads();
#pragma unroll 2
    for (int ks = 0; ks < RCH / 4; ++ks) {
      const int nr = ks * 4 + l / 16;
      T xb[TB];
#pragma unroll
      for (int b = 0; b < TB; ++b) xb[b] = S.Xs[nr * LDX + b * 16 + l % 16];
      const T lm = S.lam[nr], tlm = S.dlam[nr];
#pragma unroll
      for (int m = 0; m < TPW; ++m) {
        const int tt = w + 4 * m;
        if (tt >= NLT) continue;
        int bi, bj;
        tile_of(tt, bi, bj);
        if (bi >= nb) continue;
        acc[m] = Mf<T>::mma(lm * xb[bi], xb[bj], acc[m]);
        tacc[m] = Mf<T>::mma(tlm * xb[bi], xb[bj], tacc[m]);
      }
    }
    if (want_lg) {
      const int k = t % MAXD;
#pragma unroll
      for (int m = 0; m < RCH / RG; ++m) {
        const int r = t / MAXD + RG * m;
        gacc += S.Xs[r * LDX + k] * S.rr[r];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < TPW; ++m) {
    const int tt = w + 4 * m;
    if (tt >= NLT) continue;
    int bi, bj;
    tile_of(tt, bi, bj);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = bi * 16 + Mf<T>::row(r, l), j = bj * 16 + l % 16;
      if (i < j) continue;  // keep the lower entry of diagonal tiles for both
      T v = acc[m][r];
      if (i == j && i < D) v += (T)(1.0 / prior_var);
      S.Gi[i * LDX + j] = v;
      S.Gi[j * LDX + i] = v;
      S.dGi[i * LDX + j] = tacc[m][r];
      S.dGi[j * LDX + i] = tacc[m][r];
    }
  }
  if (want_lg) {
    S.part[t / MAXD][t % MAXD] = gacc;
    loglik = bsum(S, ll);  // its barrier publishes part[]
    if (t < D) {
      T s = T(0);
#pragma unroll
      for (int gq = 0; gq < RG; ++gq) s += S.part[gq][t];
      S.g[t] = s - S.q[t] / (T)prior_var;
    }
  }
  __syncthreads();
}

// Sweep of (G, G') into (-G^-1, -(G^-1)'); returns (log det G, its tangent) in
// float64.  With zc, pout = chol(G) zc (values only).
template <class T, class C>
__device__ __forceinline__ void sweep(Smem<T, C>& S, T (&A)[C::BS][C::BS], T (&Tg)[C::BS][C::BS], int D,
                                      double& logdet, double& tlogdet, const T* zc, T* pout) {
  constexpr int BS = C::BS, MAXD = C::MAXD;
  const int t = threadIdx.x;
  const int i0 = BS * (t / 16), j0 = BS * (t % 16);
  logdet = 0.0;
  tlogdet = 0.0;
  T pacc = T(0);
  for (int kk = 0; kk * BS < D; ++kk) {
#pragma unroll
    for (int a0 = 0; a0 < BS; ++a0) {
      const int k = kk * BS + a0;
      if (k >= D) break;
      T* buf = S.rowbuf[k & 1];
      T* tbuf = S.drowbuf[k & 1];
      if (t / 16 == kk) {
#pragma unroll
        for (int b = 0; b < BS; ++b) {
          buf[j0 + b] = A[a0][b];
          tbuf[j0 + b] = Tg[a0][b];
        }
      }
      __syncthreads();
      const T d = buf[k], td = tbuf[k];
      const T rd = T(1) / d;
      const T trd = -td * rd * rd;
      logdet += log(fabs((double)d));
      tlogdet += (double)(td * rd);
      if (zc != nullptr && t >= k && t < D) pacc = fma(buf[t] * zc[k], T(1) / sqrt(d), pacc);
      T bjr[BS], tbjr[BS];
#pragma unroll
      for (int b = 0; b < BS; ++b) {
        bjr[b] = buf[j0 + b] * rd;
        tbjr[b] = tbuf[j0 + b] * rd + buf[j0 + b] * trd;
      }
#pragma unroll
      for (int a = 0; a < BS; ++a) {
        const int i = i0 + a;
        const T aik = buf[i], taik = tbuf[i];
#pragma unroll
        for (int b = 0; b < BS; ++b) {
          const int j = j0 + b;
          const T v = A[a][b], tv = Tg[a][b];
          const T upd = fma(-aik, bjr[b], v);
          const T tupd = tv - taik * bjr[b] - aik * tbjr[b];
          const bool off = i != k && j != k, piv = i == k && j == k;
          A[a][b] = off ? upd : (piv ? -rd : v * rd);
          Tg[a][b] = off ? tupd : (piv ? -trd : tv * rd + v * trd);
        }
      }
    }
  }
  if (pout != nullptr && t < MAXD) pout[t] = t < D ? pacc : T(0);
}

// G (and G') in S.Gi / S.dGi -> G^-1 and (G^-1)' there; u = G^-1 p,
// u' = G^-1 p' + (G^-1)' p.  With zc, the sweep first draws pout = chol(G) zc
// (pout may be S.p: it is written before the solve reads S.p).
template <class T, class C>
__device__ __forceinline__ void invert_solve(Smem<T, C>& S, int D, double& logdet, double& tlogdet, const T* zc,
                                             T* pout) {
  constexpr int BS = C::BS, LDX = C::LDX, MAXD = C::MAXD;
  const int t = threadIdx.x;
  const int i0 = BS * (t / 16), j0 = BS * (t % 16);
  T A[BS][BS], Tg[BS][BS];
#pragma unroll
  for (int a = 0; a < BS; ++a)
#pragma unroll
    for (int b = 0; b < BS; ++b) {
      A[a][b] = S.Gi[(i0 + a) * LDX + j0 + b];
      Tg[a][b] = S.dGi[(i0 + a) * LDX + j0 + b];
    }
  __syncthreads();
  sweep(S, A, Tg, D, logdet, tlogdet, zc, pout);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < BS; ++a)
#pragma unroll
    for (int b = 0; b < BS; ++b) {
      S.Gi[(i0 + a) * LDX + j0 + b] = -A[a][b];
      S.dGi[(i0 + a) * LDX + j0 + b] = -Tg[a][b];
    }
  __syncthreads();
  if (t < D) {
    T s = T(0), ts = T(0);
    for (int j = 0; j < D; ++j) {
      const T gij = S.Gi[t * LDX + j];
      s += gij * S.p[j];
      ts += gij * S.tp[j] + S.dGi[t * LDX + j] * S.p[j];
    }
    S.u[t] = s;
    S.tu[t] = ts;
  } else if (t < MAXD) {
    // the quadratic-form pass multiplies whole zero-padded rows of X by u and
    // u': entries past D must be 0, not stale LDS (Inf or NaN times 0 is NaN)
    S.u[t] = T(0);
    S.tu[t] = T(0);
  }
  __syncthreads();
}

// Quadratic-form pass with G^-1, (G^-1)' in LDS: S.dH = dH/dq at (S.q, S.p)
// and S.tdH its tangent.
template <class T, class C>
__device__ void pass_quad(Smem<T, C>& S, const double* X, const double* y, int N, int D, double prior_var) {
  constexpr int MAXD = C::MAXD, LDX = C::LDX, RCH = C::RCH, TPR = C::TPR, CPT = C::CPT, RG = C::RG, CB2 = C::CB2;
  using A4 = typename Mf<T>::acc;
  const int t = threadIdx.x, w = t / 64, l = t % 64;
  const int ksteps = (D + 3) / 4;
  const int rb = w / CB2;
  const int cb0 = 2 * (w % CB2);
  T dacc = T(0), tdacc = T(0);
  for (int n0 = 0; n0 < N; n0 += RCH) {
    load_chunk(S, X, y, N, D, n0);
    __syncthreads();
    {
      const int r = t / TPR, c0 = (t % TPR) * CPT;
      T eta = T(0), teta = T(0), bq = T(0), tbq = T(0);
#pragma unroll
      for (int e = 0; e < CPT; ++e) {
        const T xv = S.Xs[r * LDX + c0 + e];
        eta += xv * S.q[c0 + e];
        teta += xv * S.tq[c0 + e];
        bq += xv * S.u[c0 + e];
        tbq += xv * S.tu[c0 + e];
      }
      eta = group_sum<TPR>(eta);
      teta = group_sum<TPR>(teta);
      bq = group_sum<TPR>(bq);
      tbq = group_sum<TPR>(tbq);
      if (t % TPR == 0) {
        const bool ok = n0 + r < N;
        const double s = sigmoid((double)eta);
        const double lam = s * (1.0 - s);
        S.lam[r] = ok ? (T)(lam * (1.0 - 2.0 * s)) : T(0);                              // w
        S.dlam[r] = ok ? (T)(lam * (1.0 - 6.0 * s + 6.0 * s * s) * (double)teta) : T(0);  // w'
        S.rr[r] = ok ? (T)((double)S.yv[r] - s) : T(0);
        S.drr[r] = ok ? (T)(-lam * (double)teta) : T(0);
        S.bb[r] = bq;
        S.dbb[r] = tbq;
      }
    }
    A4 v[2] = {A4{0, 0, 0, 0}, A4{0, 0, 0, 0}};
    A4 tv[2] = {A4{0, 0, 0, 0}, A4{0, 0, 0, 0}};
    for (int ks = 0; ks < ksteps; ++ks) {
      const T af = S.Xs[(rb * 16 + l % 16) * LDX + ks * 4 + l / 16];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int bi = (ks * 4 + l / 16) * LDX + (cb0 + c) * 16 + l % 16;
        v[c] = Mf<T>::mma(af, S.Gi[bi], v[c]);
        tv[c] = Mf<T>::mma(af, S.dGi[bi], tv[c]);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rb * 16 + Mf<T>::row(r, l);
      T s = T(0), ts = T(0);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const T xv = S.Xs[row * LDX + (cb0 + c) * 16 + l % 16];
        s += v[c][r] * xv;
        ts += tv[c][r] * xv;
      }
      s = group_sum<16>(s);
      ts = group_sum<16>(ts);
      if (l % 16 == 0) {
        S.aa[w % CB2][row] = s;
        S.daa[w % CB2][row] = ts;
      }
    }
    __syncthreads();
    if (t < RCH) {
      T a = T(0), ta = T(0);
#pragma unroll
      for (int s = 0; s < CB2; ++s) {
        a += S.aa[s][t];
        ta += S.daa[s][t];
      }
      const T b = S.bb[t], tb = S.dbb[t];
      const T wv = S.lam[t], twv = S.dlam[t];
      S.hh[t] = T(0.5) * wv * (a - b * b) - S.rr[t];
      S.dhh[t] = T(0.5) * twv * (a - b * b) + T(0.5) * wv * (ta - T(2) * b * tb) - S.drr[t];
    }
    __syncthreads();
    {
      const int k = t % MAXD;
#pragma unroll
      for (int m = 0; m < RCH / RG; ++m) {
        const int r = t / MAXD + RG * m;
        const T xv = S.Xs[r * LDX + k];
        dacc += xv * S.hh[r];
        tdacc += xv * S.dhh[r];
      }
    }
    __syncthreads();
  }
  S.part[t / MAXD][t % MAXD] = dacc;
  S.dpart[t / MAXD][t % MAXD] = tdacc;
  __syncthreads();
  if (t < D) {
    T s = T(0), ts = T(0);
#pragma unroll
    for (int gq = 0; gq < RG; ++gq) {
      s += S.part[gq][t];
      ts += S.dpart[gq][t];
    }
    S.dH[t] = s + S.q[t] / (T)prior_var;
    S.tdH[t] = ts + S.tq[t] / (T)prior_var;
  }
  __syncthreads();
}

template <class T, class C>
__global__ void __launch_bounds__(NT, 1) logistic_rmhmc_tangent_kernel(const LogisticArgs a, const TangentOut o) {
  constexpr int MAXD = C::MAXD;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem<T, C>& S = *reinterpret_cast<Smem<T, C>*>(smem_raw);
  const int t = threadIdx.x;
  const int c = blockIdx.x;
  if (c >= a.C) return;
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;
  const int N = a.post_N[post], D = a.post_D[post];
  if (D > MAXD) return;
  // the float32 instance serves the posteriors past the float64 one's capacity
  if (sizeof(T) == sizeof(float) && D <= 64) return;
  const double* X = a.X + a.post_xoff[post];
  const double* y = a.y + a.post_yoff[post];
  const int Dp = a.Dmax;
  const T h = (T)(0.5 * a.eps);
  const T th = h;  // d h / d log eps
  const double prior = a.prior_var;
  const double lp_const = -0.5 * (double)D * log(2.0 * M_PI * prior);
  const double ke_const = 0.5 * (double)D * kLog2Pi;

  if (t < MAXD) {
    const T qv = t < D ? static_cast<const T*>(a.q_in)[(size_t)c * Dp + t] : T(0);
    S.q0[t] = qv;
    S.q[t] = qv;
    S.tq[t] = T(0);
    S.tp[t] = T(0);
    if (a.z != nullptr) {
      S.z[t] = t < D ? static_cast<const T*>(a.z)[(size_t)c * Dp + t] : T(0);
    }
  }
  if (a.z == nullptr && t < MAXD / 2) {
    const uint64_t tt = a.counter;
    u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * t), (uint32_t)(tt >> 32)},
                            (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
    double n0, n1;
    normal2(r, n0, n1);
    S.z[2 * t] = 2 * t < D ? (T)n0 : T(0);
    S.z[2 * t + 1] = 2 * t + 1 < D ? (T)n1 : T(0);
  }
  const double logp = a.logp_in[c];
  int64_t n_fev = 0;
  __syncthreads();

  // ---- momentum p0 = chol(G(q0)) z (no tangent) and H0 ----
  double logdet, tlogdet, dummy = 0.0;
  gram_pass(S, X, y, N, D, prior, false, dummy);
  invert_solve(S, D, logdet, tlogdet, S.z, S.p);  // the sweep writes p0 before the solve reads it
  const double H0 =
      -logp + 0.5 * bsum(S, t < D ? (double)S.p[t] * (double)S.u[t] : 0.0) + 0.5 * logdet + ke_const;
  bool have_ginv = true;

  auto feval = [&]() __attribute__((always_inline)) {
    if (!have_ginv) {
      double d0 = 0.0, ld = 0.0, tld = 0.0;
      gram_pass(S, X, y, N, D, prior, false, d0);
      invert_solve(S, D, ld, tld, (const T*)nullptr, (T*)nullptr);
    }
    pass_quad(S, X, y, N, D, prior);
    have_ginv = false;
    ++n_fev;
  };
  // iterate (q, p) <- z0 + h F(x):  tangents  z0' + h' F + h F'
  auto update = [&]() __attribute__((always_inline)) -> double {
    double dq = 0.0;
    if (t < D) {
      const T qn = S.zq[t] + h * S.u[t];
      const T pn = S.zp[t] - h * S.dH[t];
      const T tqn = S.tzq[t] + th * S.u[t] + h * S.tu[t];
      const T tpn = S.tzp[t] - th * S.dH[t] - h * S.tdH[t];
      dq = nan_max((double)fabs(qn - S.q[t]), (double)fabs(pn - S.p[t]));
      S.q[t] = qn;
      S.p[t] = pn;
      S.tq[t] = tqn;
      S.tp[t] = tpn;
    }
    return bmax(S, dq);
  };

  for (int s = 0; s < a.L; ++s) {
    if (t < MAXD) {
      S.zq[t] = S.q[t];
      S.zp[t] = S.p[t];
      S.tzq[t] = S.tq[t];
      S.tzp[t] = S.tp[t];
    }
    __syncthreads();
    feval();  // first evaluation outside the loop (integrators.py:84)
    double norm = update();
    int n = 0;
    while (n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol) {
      feval();
      norm = update();
      ++n;
    }
    // explicit half step from the last iterate (integrators.py:146-148)
    feval();
    if (t < D) {
      const T qv = S.q[t], pv = S.p[t];
      S.tq[t] = S.tq[t] + th * S.u[t] + h * S.tu[t];
      S.tp[t] = S.tp[t] - th * S.dH[t] - h * S.tdH[t];
      S.q[t] = qv + h * S.u[t];
      S.p[t] = pv - h * S.dH[t];
    }
    __syncthreads();
  }

  // ---- end: energy at (q_L, p_L) and its tangent ----
  double loglik = 0.0;
  gram_pass(S, X, y, N, D, prior, true, loglik);
  invert_solve(S, D, logdet, tlogdet, (const T*)nullptr, (T*)nullptr);
  const double qq = bsum(S, t < D ? (double)S.q[t] * (double)S.q[t] : 0.0);
  const double pu = bsum(S, t < D ? (double)S.p[t] * (double)S.u[t] : 0.0);
  const double tpu = bsum(S, t < D ? (double)S.tp[t] * (double)S.u[t] + (double)S.p[t] * (double)S.tu[t] : 0.0);
  const double tlogp1 = bsum(S, t < D ? (double)S.g[t] * (double)S.tq[t] : 0.0);
  const double dqv = t < D ? (double)S.q[t] - (double)S.q0[t] : 0.0;
  const double jmp = bsum(S, dqv * dqv);
  const double tjmp = bsum(S, t < D ? 2.0 * dqv * (double)S.tq[t] : 0.0);
  const double logp1 = loglik - 0.5 * qq / prior + lp_const;
  const double H1 = -logp1 + 0.5 * pu + 0.5 * logdet + ke_const;
  const double tH1 = -tlogp1 + 0.5 * tpu + 0.5 * tlogdet;
  double dHe = H0 - H1, tdHe = -tH1;
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
  if (t == 0) {
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

template <class T, class C>
hipError_t launch_tangent(const LogisticArgs& a, const TangentOut& o, hipStream_t stream) {
  const size_t smem = sizeof(Smem<T, C>);
  const hipError_t e = hipFuncSetAttribute((const void*)logistic_rmhmc_tangent_kernel<T, C>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((logistic_rmhmc_tangent_kernel<T, C>), dim3(a.C), dim3(NT), smem, stream, a, o);
  return hipGetLastError();
}

}  // namespace lgrad

hipError_t launch_logistic_rmhmc_tangent(const LogisticArgs& a, const TangentOut& o, hipStream_t stream) {
  return lgrad::launch_tangent<double, logistic::Cfg<64>>(a, o, stream);
}

hipError_t launch_logistic_rmhmc_tangent_f32_wide(const LogisticArgs& a, const TangentOut& o, hipStream_t stream) {
  return lgrad::launch_tangent<float, logistic::Cfg<128>>(a, o, stream);
}

}  // namespace gjx
