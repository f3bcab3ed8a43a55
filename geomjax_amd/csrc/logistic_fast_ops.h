// Device building blocks of the tuned logistic-regression kernels (float32
// state, split-bf16 MFMA contractions, float32 sweep) with one or two waves
// per chain: the per-chain LDS layout, the Gram pass (weights sqrt(lambda), or
// signed weights for matrices X^T diag(d) X with d of either sign), the
// symmetric sweep (inverse, pivots, Cholesky columns), the solve with one
// refinement step, and the row passes over X (vector field with
// a_n = x_n^T G^-1 x_n, eta / log likelihood / gradient).  Shared by the RMHMC
// kernel (logistic_fast.hip, which also describes the packed layouts of X) and
// the LMC kernel (logistic_lmc_fast.hip).
#pragma once

#include <type_traits>

#include "common.h"
#include "logistic.h"
#include "logistic_f32.h"

namespace gjx {
namespace lfast {

using namespace lf32;

constexpr int MAXD = 64;
constexpr int LDG = MAXD + 4;  // LDS row stride of G^-1 (floats; rows stay 16-byte aligned)

// LDS of one chain.  With one wave per chain G^-1 overwrites G once the
// refinement step has read G (the field pass needs only G^-1 and the next Gram
// pass rewrites G), which keeps the workgroup at 20 KiB: eight one-wave
// workgroups per CU, two chains per SIMD.  With two waves per chain G^-1 has
// its own buffer (the field pass stages its split operand over G).
template <int NW>
struct Smem {
  float G[MAXD * LDG];  // Gram G (row stride 64), kept for the refinement step; scratch of the
                        // partial-sum exchanges of a row pass (no Gram is live then)
  float Gi_[NW == 1 ? 4 : MAXD * LDG];
  float bc[2][64];  // sweep: column broadcast [pivot parity][row]
  float dk[MAXD];   // pivots
  float q0[MAXD], p0[MAXD], q[MAXD], p[MAXD], u[MAXD], dH[MAXD], z[MAXD], g[MAXD];
  double red[2];    // per-wave partial sums of the two-wave form
  __device__ __forceinline__ float* gi() {  // G^-1 (row stride LDG)
    if constexpr (NW == 1)
      return G;
    else
      return Gi_;
  }
};

// With NW waves per chain the rows of every pass over X are dealt to the
// waves in groups of 64 (group g to wave g % NW), so that each row's scratch
// (eta, sqrt(lambda)) is only ever read and written by one wave; the metric
// inverse runs on wave 0 alone.  Synchronisation of work that wave 0 does on
// its own: the whole workgroup for one wave, else the wave itself (the LDS
// accesses of one wave complete in order; the fences and the wave barrier
// keep the compiler from moving them across this point).
template <int NW>
__device__ __forceinline__ void wave0_sync() {
  if constexpr (NW == 1) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Index of the calling wave within the workgroup (wave-uniform: index
// arithmetic on it stays scalar).
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64); }

// Sum of one value per wave, identical in every wave (LDS scratch S.red).
template <int NW>
__device__ __forceinline__ double waves_sum(Smem<NW>& S, double v) {
  if constexpr (NW == 1) {
    return v;
  } else {
    if (threadIdx.x % 64 == 0) S.red[threadIdx.x / 64] = v;
    __syncthreads();
    const double t = S.red[0] + S.red[1];
    __syncthreads();
    return t;
  }
}

// K step j (0, 1, 2, ...) of wave w's share of the rows: 16-row steps 4g..4g+3
// of the wave's 64-row groups g = w, w + NW, ...
template <int NW>
__device__ __forceinline__ int own_kstep(int j, int w) { return 4 * (w + NW * (j >> 2)) + (j & 3); }
// 32-row block j of wave w's share: blocks 2g, 2g + 1 of its groups
template <int NW>
__device__ __forceinline__ int own_block(int j, int w) { return 2 * (w + NW * (j >> 1)) + (j & 1); }
// number of 64-row groups of wave w
template <int NW>
__device__ __forceinline__ int own_groups(int npad, int w) { return (npad / 64 - w + NW - 1) / NW; }

typedef uint32_t u32v4 __attribute__((ext_vector_type(4)));

// v with the sign of element i flipped where bit 16 i + 15 of m is set (exact)
__device__ __forceinline__ bf16x8 flip_signs(const bf16x8& v, const uint32_t (&m)[4]) {
  u32v4 u = __builtin_bit_cast(u32v4, v);
  u[0] ^= m[0];
  u[1] ^= m[1];
  u[2] ^= m[2];
  u[3] ^= m[3];
  return __builtin_bit_cast(bf16x8, u);
}

// ---------------------------------------------------------------------------
// Gram pass: S.G = X^T diag(lambda) X + I / prior_var for the first D
// features, identity on the padding, from sqrt(lambda) in `sl`.  With SIGNED
// the row weights d may have either sign and `sl` holds sign(d) sqrt|d|: the A
// operand carries sqrt|d| x, the B operand the same bf16 values with the sign
// of d (a bf16 negation is exact), so the tiles sum d x x^T.
// ---------------------------------------------------------------------------
template <int NB, int NW, bool SIGNED = false>
__device__ __forceinline__ void gram_pass(Smem<NW>& S, const float* __restrict__ Xg, const float* __restrict__ sl, int npad, int D,
                          float inv_prior) {
  constexpr int NTL = NB * (NB + 1) / 2;
  const int l = threadIdx.x % 64, w = wave_id();
  f32x16 acc[NTL];
#pragma unroll
  for (int t = 0; t < NTL; ++t) acc[t] = f32x16{0};
  const float* slp = sl + (l / 32) * 8;
  const float* xp = Xg + (size_t)l * 8;
  const int nj = 4 * own_groups<NW>(npad, w);  // this wave's K steps (none for a wave without rows)
  // fragments of K step kb: sqrt(lambda) of the lane's 8 rows and X of its NB feature blocks
  struct Frag {
    float s[8];
    float x[NB][8];
  };
  auto fetch = [&](Frag& f, int j) {
    // the pipeline's tail re-reads the wave's last step (in bounds, unused)
    const int jj = j < nj ? j : nj - 1;
    const int kb = jj >= 0 ? own_kstep<NW>(jj, w) : 0;
    load8(slp + kb * 16, f.s);
#pragma unroll
    for (int b = 0; b < NB; ++b) load8(xp + (size_t)(kb * NB + b) * 512, f.x[b]);
  };
  auto split_frag = [&](const Frag& f, bf16x8 (&hi)[NB], bf16x8 (&lo)[NB], uint32_t (&m)[4]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      float y8[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) y8[i] = f.x[b][i] * (SIGNED ? fabsf(f.s[i]) : f.s[i]);
      split8(y8, hi[b], lo[b]);
    }
    if constexpr (SIGNED) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        m[k] = ((__float_as_uint(f.s[2 * k]) >> 16) & 0x8000u) | (__float_as_uint(f.s[2 * k + 1]) & 0x80000000u);
    }
  };
  auto tiles = [&](const bf16x8 (&hi)[NB], const bf16x8 (&lo)[NB], const uint32_t (&m)[4]) {
    bf16x8 bh[NB], bl[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      bh[b] = SIGNED ? flip_signs(hi[b], m) : hi[b];
      bl[b] = SIGNED ? flip_signs(lo[b], m) : lo[b];
    }
    int t = 0;
#pragma unroll
    for (int bi = 0; bi < NB; ++bi)
#pragma unroll
      for (int bj = 0; bj <= bi; ++bj) {
        acc[t] = mfma3(hi[bi], lo[bi], bh[bj], bl[bj], acc[t]);
        ++t;
      }
  };
  // Two-stage pipeline: the MFMAs of step k run while the operands of step
  // k+1 are scaled and split on VALU; loads run four steps ahead (a ring of
  // four fragments: the sqrt(lambda) rows come from the per-chain scratch,
  // whose latency two steps did not cover).  nj is a multiple of 4.
  Frag f0, f1, f2, f3;
  bf16x8 ha[NB], la[NB], hb[NB], lb[NB];
  uint32_t ma[4], mb[4];  // row signs (SIGNED only)
  fetch(f0, 0);
  fetch(f1, 1);
  fetch(f2, 2);
  fetch(f3, 3);
  split_frag(f0, ha, la, ma);
  for (int j = 0; j < nj; j += 4) {
    fetch(f0, j + 4);
    tiles(ha, la, ma);
    split_frag(f1, hb, lb, mb);
    interleave_mfma_valu<3 * NTL, 64 / (3 * NTL) + 1>();
    fetch(f1, j + 5);
    tiles(hb, lb, mb);
    split_frag(f2, ha, la, ma);
    interleave_mfma_valu<3 * NTL, 64 / (3 * NTL) + 1>();
    fetch(f2, j + 6);
    tiles(ha, la, ma);
    split_frag(f3, hb, lb, mb);
    interleave_mfma_valu<3 * NTL, 64 / (3 * NTL) + 1>();
    fetch(f3, j + 7);
    tiles(hb, lb, mb);
    split_frag(f0, ha, la, ma);
    interleave_mfma_valu<3 * NTL, 64 / (3 * NTL) + 1>();
  }
  __syncthreads();
  if constexpr (NW == 2) {
    // wave 1 hands its partial tiles to wave 0 through G^-1's buffer (a
    // two-wave workgroup has its own, not needed again before the next sweep)
    float* P = S.gi();
    if (w == 1) {
#pragma unroll
      for (int t = 0; t < NTL; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) P[(t * 16 + r) * 64 + l] = acc[t][r];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int t = 0; t < NTL; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] += P[(t * 16 + r) * 64 + l];
    }
  }
  if (w == 0) {
    int t = 0;
#pragma unroll
    for (int bi = 0; bi < NB; ++bi)
#pragma unroll
      for (int bj = 0; bj <= bi; ++bj) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = bi * 32 + (r / 4) * 8 + (l / 32) * 4 + (r % 4);
          const int j = bj * 32 + (l % 32);
          // a diagonal tile holds (i, j) and (j, i) rounded differently: keep the lower one
          if (bi == bj && i < j) continue;
          float v = acc[t][r];
          if (i == j) v = i < D ? v + inv_prior : 1.0f;
          S.G[i * MAXD + j] = v;
          S.G[j * MAXD + i] = v;
        }
        ++t;
      }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Sweep of the DP x DP matrix in S.G (left intact).  Lane jl = l % DP holds
// column jl in registers A[b][i] = A[32b + i][jl]; the pivot loop is rolled
// and reads/writes A[k] with a wave-uniform k (register-relative moves), the
// update of the other entries is unrolled.  On return A = -G^-1 and S.dk the
// pivots.  With want_chol, p0j = (L z)_jl with L = chol(G) and z in S.z.
// ---------------------------------------------------------------------------
template <int NB, int NW>
__device__ __forceinline__ void sweep(Smem<NW>& S, f32x32 (&A)[NB], int D, bool want_chol, float& p0j,
                                      bool store_l = false) {
  constexpr int DP = 32 * NB;
  const int jl = threadIdx.x % DP;
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 32; ++i) A[b][i] = S.G[(32 * b + i) * MAXD + jl];
  p0j = 0.0f;
  // Pivot k needs column k in every lane: entry i is lane i's register k
  // (A[k][i] = A[i][k]), broadcast through LDS.  Padding pivots (k >= D) are
  // 1 with zero off-diagonals: nothing to do.
  float(*bcast)[64] = S.bc;
  // One pivot; its row k lives in tile T (k in [32 T, 32 T + 32)), so the
  // register-relative read and write of A[k] index one fixed tile: with one
  // loop per tile the write stays in place (a runtime choice of tile made the
  // compiler copy the whole tile out and back around the write).
  auto pivot = [&](auto tile, int k) {
    constexpr int T = decltype(tile)::value;
    const float ak = A[T][k - 32 * T];  // A[k][jl] = A[jl][k]
    float* bc = bcast[k & 1];
    bc[jl] = ak;
    wave0_sync<NW>();
    float col[DP];
#pragma unroll
    for (int i = 0; i < DP; i += 4) {
      const float4 v = *reinterpret_cast<const float4*>(bc + i);
      col[i] = v.x; col[i + 1] = v.y; col[i + 2] = v.z; col[i + 3] = v.w;
    }
    const float d = bc[k];
    const float rd = __builtin_amdgcn_rcpf(d);  // 1 ulp, as the float32 sweep itself
    if (want_chol && jl >= k) p0j = fmaf(ak / sqrtf(d), S.z[k], p0j);
    // with store_l, L[jl][k] into row jl of S.G (stride LDG; S.G is not read
    // after the column loads above, and one wave's LDS accesses run in order)
    if (store_l && jl >= k) S.G[jl * LDG + k] = ak / sqrtf(d);
    // lane k's own column equals col, so f = 1 - 1/d turns its update into A[i][k] / d
    const float f = (jl == k) ? 1.0f - rd : ak * rd;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 32; ++i) A[b][i] = fmaf(-col[32 * b + i], f, A[b][i]);
    A[T][k - 32 * T] = (jl == k) ? -rd : ak * rd;
    if (threadIdx.x == 0) S.dk[k] = d;
  };
  const int k0 = D < 32 ? D : 32;
  for (int k = 0; k < k0; ++k) pivot(std::integral_constant<int, 0>{}, k);
  if constexpr (NB == 2)
    for (int k = 32; k < D; ++k) pivot(std::integral_constant<int, 1>{}, k);
  // -G^-1 on the padding block is -I
  for (int k = D; k < 32; ++k) A[0][k] = jl == k ? -1.0f : 0.0f;
  if constexpr (NB == 2)
    for (int k = D > 32 ? D : 32; k < DP; ++k) A[1][k - 32] = jl == k ? -1.0f : 0.0f;
  if (threadIdx.x < DP && threadIdx.x >= D) S.dk[threadIdx.x] = 1.0f;
  wave0_sync<NW>();
}

// log det G = sum log d_k in float64 (padding pivots are 1)
template <class SM>
__device__ __forceinline__ double logdet_from_pivots(const SM& S, int D) {
  const int l = threadIdx.x % 64;
  double v = l < D ? log((double)S.dk[l]) : 0.0;
  return group_sum<64>(v);
}

// log |det M| = sum log |d_k| in float64, for a symmetric M of either sign
template <class SM>
__device__ __forceinline__ double logabsdet_from_pivots(const SM& S, int D) {
  const int l = threadIdx.x % 64;
  double v = l < D ? log(fabs((double)S.dk[l])) : 0.0;
  return group_sum<64>(v);
}

// out = M^-1 rhs with one refinement step, for M in S.G (row stride MAXD) and
// A = -M^-1 from its sweep; tmp is overwritten.  rhs and out are MAXD-vectors
// in LDS whose entries from D on are zero (on return, out's are too).
template <int NB, int NW>
__device__ __forceinline__ void solve_refined(Smem<NW>& S, const f32x32 (&A)[NB], const float* rhs, float* out,
                                              float* tmp) {
  constexpr int DP = 32 * NB;
  const int l = threadIdx.x, jl = l % DP;
  float uj = 0.0f;
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 32; ++i) uj = fmaf(-A[b][i], rhs[32 * b + i], uj);
  if (l < DP) out[jl] = uj;
  wave0_sync<NW>();
  float mu = 0.0f;
#pragma unroll
  for (int i = 0; i < DP; ++i) mu = fmaf(S.G[i * MAXD + jl], out[i], mu);
  if (l < DP) tmp[jl] = rhs[jl] - mu;
  wave0_sync<NW>();
  float du = 0.0f;
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 32; ++i) du = fmaf(-A[b][i], tmp[32 * b + i], du);
  wave0_sync<NW>();
  if (l < DP) out[jl] = uj + du;
  else if (l < MAXD) out[l] = 0.0f;
  wave0_sync<NW>();
}

// S.u = G^-1 S.p with one refinement step against the Gram still in S.G.
template <int NB, int NW>
__device__ __forceinline__ void ginv_apply(Smem<NW>& S, const f32x32 (&A)[NB]) {
  constexpr int DP = 32 * NB;
  const int l = threadIdx.x, jl = l % DP;
  float uj = 0.0f;
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 32; ++i) uj = fmaf(-A[b][i], S.p[32 * b + i], uj);
  if (l < DP) S.u[jl] = uj;
  wave0_sync<NW>();
  float gu = 0.0f;
#pragma unroll
  for (int i = 0; i < DP; ++i) gu = fmaf(S.G[i * MAXD + jl], S.u[i], gu);
  if (l < DP) S.dH[jl] = S.p[jl] - gu;  // residual (dH is free here)
  wave0_sync<NW>();
  float du = 0.0f;
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 32; ++i) du = fmaf(-A[b][i], S.dH[32 * b + i], du);
  wave0_sync<NW>();
  if (l < DP) S.u[jl] = uj + du;
  wave0_sync<NW>();
}

// G^-1 into its buffer (over G with one wave per chain: call after the last read of G)
template <int NB, int NW>
__device__ __forceinline__ void store_ginv(Smem<NW>& S, const f32x32 (&A)[NB]) {
  const int l = threadIdx.x;
  if (l < 32 * NB) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 32; ++i) S.gi()[(32 * b + i) * LDG + l] = -A[b][i];
  }
  wave0_sync<NW>();
}

// G (in S.G) -> pivots in S.dk, S.u = G^-1 S.p (refined), and with
// keep_inverse G^-1 in its buffer.  With want_chol the momentum p0 = chol(G) S.z is
// drawn first into S.p0 and S.p.
template <int NB, int NW>
__device__ __forceinline__ void invert(Smem<NW>& S, int D, bool want_chol, bool keep_inverse) {
  f32x32 A[NB];
  float p0j;
  sweep<NB, NW>(S, A, D, want_chol, p0j);
  if (want_chol) {
    const int l = threadIdx.x;
    const float pv = l < D ? p0j : 0.0f;
    S.p0[l] = pv;
    S.p[l] = pv;
    wave0_sync<NW>();
  }
  ginv_apply<NB, NW>(S, A);
  if (keep_inverse) store_ginv<NB, NW>(S, A);
}

// Sum the per-lane feature partials of a row pass over the 32 lanes of each
// half wave: lane f < DP receives the total for feature f.
template <int NB, int NW>
__device__ __forceinline__ float reduce_features(Smem<NW>& S, const float (&acc)[2 * NB][8]) {
  constexpr int KC = 2 * NB, RS = KC * 8 + 1;
  static_assert(NW * 64 * RS <= MAXD * LDG, "the partials must fit S.G");
  const int t = threadIdx.x, l = t % 64;
  float* R = S.G;  // no Gram is live during a row pass, and G^-1 is no longer read
  __syncthreads();
#pragma unroll
  for (int kc = 0; kc < KC; ++kc)
#pragma unroll
    for (int i = 0; i < 8; ++i) R[t * RS + kc * 8 + i] = acc[kc][i];
  __syncthreads();
  float s = 0.0f;
  if (l < 32 * NB) {  // every wave forms the same totals in the same order
    const int kc = l / 16, h = (l / 8) % 2, i = l % 8;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww)
      for (int j = 0; j < 32; ++j) s += R[(ww * 64 + h * 32 + j) * RS + kc * 8 + i];
  }
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// One evaluation of the vector field at (S.q, S.p) with S.u = G^-1 p and
// G^-1 in S.gi(): S.dH = dH/dq.  Also advances the scratch: E1 <- eta of the
// next iterate (base + h X u, base = E0 for a fixed-point iterate, E1 for
// the explicit half step, which also sets E0), SL <- sqrt(lambda) there.
// ---------------------------------------------------------------------------
//
// The A operand (rows of G^-1 permuted by phi, split) is staged once in LDS
// over S.G, which is not read again before the next Gram pass rewrites it
// (with one wave per chain G^-1 itself lives there: every wave reads its share
// before any is written), and each wave runs its own blocks one at a time: the
// other wave resident on the SIMD covers the MFMA latency, and the registers of
// a second block in flight are not needed.
template <int NB, int NW>
__device__ __forceinline__ void field_pass(Smem<NW>& S, const float* __restrict__ Xr, const float* __restrict__ Xs,
                                           const float* __restrict__ yf, float* E0, float* E1, float* SL, int npad,
                                           int D, float hs, float inv_prior, bool explicit_step) {
  constexpr int KC = 2 * NB, NTK = NB * KC, PER = (NTK + NW - 1) / NW;
  const int t = threadIdx.x, l = t % 64, w = wave_id(), h = l / 32;
  bf16x8* const AH = reinterpret_cast<bf16x8*>(S.G);  // [NTK][64 lanes]
  bf16x8* const AL = AH + NTK * 64;
  static_assert(2 * NTK * 64 * sizeof(bf16x8) <= sizeof(S.G), "the split G^-1 operand must fit S.G");
  {
    float v[PER][8];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int tk = w + NW * j;
      if (tk < NTK) load8(S.gi() + phi((tk / KC) * 32 + (l % 32)) * LDG + (tk % KC) * 16 + h * 8, v[j]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int tk = w + NW * j;
      if (tk < NTK) {
        bf16x8 hi, lo;
        split8(v[j], hi, lo);
        AH[tk * 64 + l] = hi;
        AL[tk * 64 + l] = lo;
      }
    }
  }
  __syncthreads();
  float dacc[KC][8];
#pragma unroll
  for (int kc = 0; kc < KC; ++kc)
#pragma unroll
    for (int i = 0; i < 8; ++i) dacc[kc][i] = 0.0f;
  const uint32_t xoff = (uint32_t)l * 8u;
  const uint4* const sbase = reinterpret_cast<const uint4*>(Xs);
  const uint32_t soff = (uint32_t)l * 2u;
  const int nj = 2 * own_groups<NW>(npad, w);  // this wave's 32-row blocks
  // block of step j; past the end the wave's last block again (read only, unused)
  auto blk_of = [&](int j) {
    const int jj = j < nj ? j : nj - 1;
    return __builtin_amdgcn_readfirstlane(jj >= 0 ? own_block<NW>(jj, w) : 0);
  };
  struct XF {
    bf16x8 xh[KC], xl[KC];
  };
  struct Blk {
    float x[KC][8];
    float eta, base, yv, b;
    f32x16 acc[NB];
  };
  auto fetch_xf = [&](XF& f, int rb) {
    const uint4* const blk = sbase + (size_t)rb * (KC * 128);
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      const uint32_t i = (uint32_t)kc * 128u + soff;
      f.xh[kc] = __builtin_bit_cast(bf16x8, blk[i]);
      f.xl[kc] = __builtin_bit_cast(bf16x8, blk[i + 1u]);
    }
  };
  auto fetch_blk = [&](Blk& B, int rb) {
    const float* const blk = Xr + (size_t)rb * (KC * 512);
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) load8(blk + ((uint32_t)kc * 512u + xoff), B.x[kc]);
    const uint32_t lr = (uint32_t)(l % 32);
    B.eta = (E1 + (size_t)rb * 32)[lr];
    B.base = (E0 + (size_t)rb * 32)[lr];  // unused by the explicit step (no branch in the loop)
    B.yv = (yf + (size_t)rb * 32)[lr];
  };
  // The split G^-1 fragments of step tk + 1 are read from LDS while the three
  // MFMAs of step tk run (one step ahead: two fragments in flight).
  auto mfma_blk = [&](const XF& f, Blk& B) {
    bf16x8 ah = AH[l], al = AL[l];
#pragma unroll
    for (int tk = 0; tk < NTK; ++tk) {
      const int tt = tk / KC, kc = tk % KC;
      bf16x8 nh = ah, nl = al;
      if (tk + 1 < NTK) {
        nh = AH[(tk + 1) * 64 + l];
        nl = AL[(tk + 1) * 64 + l];
      }
      if (kc == 0) B.acc[tt] = f32x16{0};
      B.acc[tt] = mfma3(ah, al, f.xh[kc], f.xl[kc], B.acc[tt]);
      if (tk + 1 < NTK) {
        __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // the next step's two LDS reads
        __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);  // then this step's three MFMAs
      }
      ah = nh;
      al = nl;
    }
  };
  auto bdot = [&](Blk& B) {
    float bp = 0.0f;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      float u8[8];
      load8(S.u + kc * 16 + h * 8, u8);
#pragma unroll
      for (int i = 0; i < 8; ++i) bp = fmaf(B.x[kc][i], u8[i], bp);
    }
    B.b = bp + __shfl_xor(bp, 32, 64);
  };
  float* const e0_dst = explicit_step ? E0 : E1;
  auto post = [&](const Blk& B, int rb) {
    const int row = rb * 32 + (l % 32);
    float ap = 0.0f;
#pragma unroll
    for (int tt = 0; tt < NB; ++tt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = 16 * tt + r;
        ap = fmaf(B.acc[tt][r], B.x[j / 8][j % 8], ap);
      }
    const float an = ap + __shfl_xor(ap, 32, 64);
    const float s = sigmoidf(B.eta);
    const float wv = s * (1.0f - s) * (1.0f - 2.0f * s);
    const float hv = 0.5f * wv * (an - B.b * B.b) - (B.yv - s);
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int i = 0; i < 8; ++i) dacc[kc][i] = fmaf(hv, B.x[kc][i], dacc[kc][i]);
    const float en = fmaf(hs, B.b, explicit_step ? B.eta : B.base);
    const float sn = sigmoidf(en);
    E1[row] = en;
    e0_dst[row] = en;  // E1 again for a fixed-point iterate (no branch in the loop)
    SL[row] = sqrt_fast(sn * (1.0f - sn));
  };
  XF f;
  Blk B;
  fetch_xf(f, blk_of(0));
  fetch_blk(B, blk_of(0));
  for (int j = 0; j < nj; ++j) {
    // the A operand is read from LDS per block: holding it in registers
    // across the loop would cost 64 of them
    asm volatile("" ::: "memory");
    const int rb = blk_of(j);
    mfma_blk(f, B);
    fetch_xf(f, blk_of(j + 1));
    bdot(B);
    post(B, rb);
    // The next block's rows load into the registers this block's rows leave
    // (the loop body is one basic block, so this holds): hoisted above post's
    // last uses of B they would land in temporaries, and the copy into the
    // loop-carried registers would wait for every load in flight (3 % on
    // 2048 x 64 at 4096 chains).
    __builtin_amdgcn_sched_barrier(0);
    fetch_blk(B, blk_of(j + 1));
  }
  const float tot = reduce_features<NB, NW>(S, dacc);
  if (t < MAXD) S.dH[l] = l < D ? tot + S.q[l] * inv_prior : 0.0f;
  __syncthreads();
}

// eta = X q (float64 accumulation) into Eout, sqrt(lambda) into SL; with
// want_grad also the log likelihood (float64, returned) and
// S.g = X^T (y - s) - q / prior_var.
template <int NB, int NW>
__device__ __forceinline__ double fresh_pass(Smem<NW>& S, const float* __restrict__ Xr, const float* __restrict__ yf, int N, int npad,
                             int D, float* Eout, float* SL, bool want_grad, float inv_prior) {
  constexpr int KC = 2 * NB;
  const int t = threadIdx.x, l = t % 64, w = wave_id(), h = l / 32;
  double qv[KC][8];
  float gacc[KC][8];
#pragma unroll
  for (int kc = 0; kc < KC; ++kc)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      qv[kc][i] = (double)S.q[kc * 16 + h * 8 + i];
      gacc[kc][i] = 0.0f;
    }
  double ll = 0.0;
  const float* xp = Xr + (size_t)l * 8;
  const int nj = 2 * own_groups<NW>(npad, w);  // this wave's 32-row blocks
  for (int j = 0; j < nj; ++j) {
    const int rb = own_block<NW>(j, w);
    float x[KC][8];
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) load8(xp + (size_t)(rb * KC + kc) * 512, x[kc]);
    double ep = 0.0;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int i = 0; i < 8; ++i) ep = fma((double)x[kc][i], qv[kc][i], ep);
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
    const float tot = reduce_features<NB, NW>(S, gacc);
    if (t < MAXD) S.g[l] = l < D ? tot - S.q[l] * inv_prior : 0.0f;
  }
  __syncthreads();
  return waves_sum<NW>(S, group_sum<64>(ll));
}

// E0 = E1 = ES, SL = sqrt(lambda(ES))
template <int NW>
__device__ __forceinline__ void prep_pass(const float* ES, float* E0, float* E1, float* SL, int npad) {
  const int l = threadIdx.x % 64, w = wave_id();
  const int ng = own_groups<NW>(npad, w);
  for (int g = 0; g < ng; ++g) {
    const int row = 64 * (w + NW * g) + l;
    const float e = ES[row];
    const float s = sigmoidf(e);
    E0[row] = e;
    E1[row] = e;
    SL[row] = sqrt_fast(s * (1.0f - s));
  }
  __syncthreads();
}

// One pass over the rows of X with the velocity in S.p: b = X v, the signed
// square roots of the weights lambda + coef w b into SS; with want_gv also
// S.u = G v = X^T (lambda b) + v / prior_var; with step the position update
// q += eps v carried into the rows (ETA += eps b, SL = sqrt(lambda) there).
template <int NB>
__device__ __forceinline__ void lmc_row_pass(Smem<1>& S, const float* __restrict__ Xr, float* ETA, float* SL,
                                             float* SS, int npad, int D, float coef, float eps, bool want_gv,
                                             bool step, float inv_prior) {
  constexpr int KC = 2 * NB;
  const int l = threadIdx.x % 64, h = l / 32;
  float v[KC][8], gacc[KC][8];
#pragma unroll
  for (int kc = 0; kc < KC; ++kc) {
    load8(S.p + kc * 16 + h * 8, v[kc]);
#pragma unroll
    for (int i = 0; i < 8; ++i) gacc[kc][i] = 0.0f;
  }
  const float* xp = Xr + (size_t)l * 8;
  const int nj = npad / 32;  // even: npad is a multiple of 64
  auto fetch = [&](float (&x)[KC][8], float& eta, int rb) {
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) load8(xp + (size_t)(rb * KC + kc) * 512, x[kc]);
    eta = ETA[rb * 32 + (l % 32)];
  };
  auto block = [&](const float (&x)[KC][8], float eta, int rb) {
    float bp = 0.0f;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
      for (int i = 0; i < 8; ++i) bp = fmaf(x[kc][i], v[kc][i], bp);
    const float b = bp + __shfl_xor(bp, 32, 64);
    const int row = rb * 32 + (l % 32);
    const float s = sigmoidf(eta);
    const float lam = s * (1.0f - s);
    const float d = fmaf(coef * lam * (1.0f - 2.0f * s), b, lam);
    if (want_gv) {
      const float lb = lam * b;
#pragma unroll
      for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int i = 0; i < 8; ++i) gacc[kc][i] = fmaf(lb, x[kc][i], gacc[kc][i]);
    }
    if (h == 0) {
      SS[row] = copysignf(sqrt_fast(fabsf(d)), d);
      if (step) {
        const float en = fmaf(eps, b, eta);
        const float sn = sigmoidf(en);
        ETA[row] = en;
        SL[row] = sqrt_fast(sn * (1.0f - sn));
      }
    }
  };
  // two register sets in turn: each block's rows (and eta, whose row the
  // block itself rewrites only after reading it) load one block ahead
  float x0[KC][8], x1[KC][8], e0, e1;
  fetch(x0, e0, 0);
  for (int rb = 0; rb < nj; rb += 2) {
    fetch(x1, e1, rb + 1);
    block(x0, e0, rb);
    fetch(x0, e0, rb + 2 < nj ? rb + 2 : rb + 1);  // past the end the last block again (in bounds, unused)
    block(x1, e1, rb + 1);
  }
  if (want_gv) {
    const float tot = reduce_features<NB, 1>(S, gacc);
    S.u[l] = l < D ? tot + S.p[l] * inv_prior : 0.0f;
  }
  __syncthreads();
}

}  // namespace lfast
}  // namespace gjx
