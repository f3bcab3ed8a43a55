// RMHMC with the Fisher metric on Bayesian logistic regression, tuned form:
// float32 state, Gram and quadratic-form contractions on bf16 MFMA with float32
// accumulation, float32 sweep inverse with one refinement step per solve,
// float64 log-determinant sums, energies, accept test and fixed-point norm.
//
// Reference: geomjax/rmhmc/rmhmc.py:131-174,376-462 (kernel, proposal, accept),
// geomjax/rmhmc/integrators.py:53-156 (implicit midpoint, fixed point),
// geomjax/rmhmc/metrics.py:42-129 (kinetic energy -log N(p; 0, G));
// the model/metric are the logistic configurations' user-supplied functions.
// Closed forms as in the float64 kernel (w = s(1-s)(1-2s), r = y - s,
// a_n = x_n^T G^-1 x_n, b = X u):
//   dH/dp = u = G^-1 p,   dH/dq = X^T [ w (a - b^2) / 2 - r ] + q / prior_var.
//
// Layout.  One workgroup per chain, of one wave (64 lanes) or two (the rows of
// every pass over X split between them, the inverse on the first; chosen per
// problem by the caller's waves_per_chain): each chain runs its own
// fixed-point iteration count.  The design matrix is
// packed once into two float32 fragment layouts (logistic_pack_kernel):
//   Gram layout: for a 16-row K step, lane l holds X[k0 + (l/32)*8 + i][32b + l%32]
//     (i < 8) -- the A and B operands of v_mfma_f32_32x32x16_bf16 for the Gram
//     G = Y^T Y with Y = sqrt(lambda) X (both operands are the same registers);
//   row layout: for a 32-row block, lane l holds X[r0 + l%32][16kc + (l/32)*8 + i]
//     -- the B operand of V^T = G^-1 X^T and the rows for X u, X^T h.
// Each lane loads 32 contiguous bytes per fragment.
//
// Precision of the contractions: every float32 operand v is split into
// hi = bf16(v) and lo = bf16(v - hi); the products hi*hi + hi*lo + lo*hi
// carry ~16 significant bits per term and accumulate in float32.
//
// a_n without a second layout: the output feature index m of V^T is mapped
// through a fixed permutation phi (rows of the A operand G^-1 are permuted)
// so that the accumulator slot (t, r) of lane l holds V[row l%32][f] for the
// very feature f whose X value the lane holds in its row-layout fragment:
// a_n is 32 FMAs per lane plus one exchange with lane l^32.
//
// eta = X q is never recomputed inside a trajectory: the fixed-point update
// q' = q_0 + h u gives eta' = eta_0 + h (X u), and X u is formed anyway for
// the kinetic term; at the end of every trajectory eta is recomputed from X
// in float64 together with the log likelihood and gradient.  Per-chain
// scratch (global, L2-resident): eta at the step start, eta at the current
// iterate, sqrt(lambda) at the current iterate, eta of the chain state and of
// the proposal.
//
// Sweep.  G (64 x 64) is held one column per lane in registers; the symmetric
// sweep operator (Goodnight 1979) gives -G^-1 and the pivots d_k, whose logs
// sum (float64) to log det G.  At pivot k the not-yet-swept column k is the
// Schur complement, so the Cholesky factor column L[:, k] = A[:, k] / sqrt(d_k)
// is available for the momentum draw p0 = L z at no extra cost.  Every
// evaluation uses this sweep (a 32 x 32 blocked form with four MFMA products
// measured the same per evaluation on one wave and twice as slow on two).
#include <mutex>
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

// ---------------------------------------------------------------------------
// Gram pass: S.G = X^T diag(lambda) X + I / prior_var for the first D
// features, identity on the padding, from sqrt(lambda) in `sl`.
// ---------------------------------------------------------------------------
template <int NB, int NW>
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
  auto split_frag = [&](const Frag& f, bf16x8 (&hi)[NB], bf16x8 (&lo)[NB]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      float y8[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) y8[i] = f.x[b][i] * f.s[i];
      split8(y8, hi[b], lo[b]);
    }
  };
  auto tiles = [&](const bf16x8 (&hi)[NB], const bf16x8 (&lo)[NB]) {
    int t = 0;
#pragma unroll
    for (int bi = 0; bi < NB; ++bi)
#pragma unroll
      for (int bj = 0; bj <= bi; ++bj) {
        acc[t] = mfma3(hi[bi], lo[bi], hi[bj], lo[bj], acc[t]);
        ++t;
      }
  };
  // Two-stage pipeline: the MFMAs of step k run while the operands of step
  // k+1 are scaled and split on VALU; loads run four steps ahead (a ring of
  // four fragments: the sqrt(lambda) rows come from the per-chain scratch,
  // whose latency two steps did not cover).  nj is a multiple of 4.
  Frag f0, f1, f2, f3;
  bf16x8 ha[NB], la[NB], hb[NB], lb[NB];
  fetch(f0, 0);
  fetch(f1, 1);
  fetch(f2, 2);
  fetch(f3, 3);
  split_frag(f0, ha, la);
  for (int j = 0; j < nj; j += 4) {
    fetch(f0, j + 4);
    tiles(ha, la);
    split_frag(f1, hb, lb);
    interleave_mfma_valu<3 * NTL, 64 / (3 * NTL) + 1>();
    fetch(f1, j + 5);
    tiles(hb, lb);
    split_frag(f2, ha, la);
    interleave_mfma_valu<3 * NTL, 64 / (3 * NTL) + 1>();
    fetch(f2, j + 6);
    tiles(ha, la);
    split_frag(f3, hb, lb);
    interleave_mfma_valu<3 * NTL, 64 / (3 * NTL) + 1>();
    fetch(f3, j + 7);
    tiles(hb, lb);
    split_frag(f0, ha, la);
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
template <int NB>
__device__ __forceinline__ float col_get(const f32x32 (&A)[NB], int k) {
  if constexpr (NB == 1) {
    return A[0][k];
  } else {
    float v;
    if (k < 32)
      v = A[0][k];
    else
      v = A[1][k - 32];
    return v;
  }
}
template <int NB>
__device__ __forceinline__ void col_set(f32x32 (&A)[NB], int k, float v) {
  if constexpr (NB == 1) {
    A[0][k] = v;
  } else {
    if (k < 32)
      A[0][k] = v;
    else
      A[1][k - 32] = v;
  }
}

template <int NB, int NW>
__device__ __forceinline__ void sweep(Smem<NW>& S, f32x32 (&A)[NB], int D, bool want_chol, float& p0j) {
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
  for (int k = 0; k < D; ++k) {
    const float ak = col_get<NB>(A, k);  // A[k][jl] = A[jl][k]
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
    // lane k's own column equals col, so f = 1 - 1/d turns its update into A[i][k] / d
    const float f = (jl == k) ? 1.0f - rd : ak * rd;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int i = 0; i < 32; ++i) A[b][i] = fmaf(-col[32 * b + i], f, A[b][i]);
    col_set<NB>(A, k, (jl == k) ? -rd : ak * rd);
    if (threadIdx.x == 0) S.dk[k] = d;
  }
  // -G^-1 on the padding block is -I
  for (int k = D; k < DP; ++k) col_set<NB>(A, k, jl == k ? -1.0f : 0.0f);
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
    B.base = explicit_step ? 0.0f : (E0 + (size_t)rb * 32)[lr];
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
    e0_dst[row] = en;
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

enum Stage { kInit, kNoise, kTraj, kEnd };

// All transitions of one chain.  The trajectory is a flat loop of vector-field
// evaluations with a single site for each pass (Gram, sweep, field, fresh),
// which keeps one copy of each unrolled pass in the kernel.
template <int NB, int NW, bool STEP>
__device__ __forceinline__ void run_chain(const LogisticArgs& a, Smem<NW>& S, int c, int post) {
  const int l = threadIdx.x % 64, w = wave_id();  // lane, wave
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

  {
    const float qv = l < D ? static_cast<const float*>(a.q_in)[(size_t)c * Dp + l] : 0.0f;
    if (w == 0) {
      S.q0[l] = qv;
      S.q[l] = qv;
    }
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
      loglik = fresh_pass<NB, NW>(S, Xr, yf, N, npad, D, Eout, SL, end || !STEP, inv_prior);
      if (!end) {
        // the state's gradient goes straight to final_g (rewritten on every acceptance)
        if (!STEP && a.final_g != nullptr && w == 0 && l < D)
          static_cast<float*>(a.final_g)[(size_t)c * Dp + l] = S.g[l];
        stage = kNoise;
      }
    }
    // ---- noise, chain state, scratch ----
    if (stage == kNoise) {
      const uint64_t tt = a.counter + (uint64_t)tr;
      if (a.z != nullptr) {
        const float zv = l < D ? static_cast<const float*>(a.z)[((size_t)tr * a.C + c) * Dp + l] : 0.0f;
        if (w == 0) S.z[l] = zv;
        uacc = a.u[(size_t)tr * a.C + c];
      } else {
        if (l < MAXD / 2) {
          u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamMomentum + (uint32_t)(2 * l),
                                        (uint32_t)(tt >> 32)}, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
          double n0, n1;
          normal2(r, n0, n1);
          if (w == 0) {
            S.z[2 * l] = 2 * l < D ? (float)n0 : 0.0f;
            S.z[2 * l + 1] = 2 * l + 1 < D ? (float)n1 : 0.0f;
          }
        }
        u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)tt, kStreamAccept, (uint32_t)(tt >> 32)},
                                (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        uacc = u01_f64(r.x, r.y) - 0.5 * (1.0 / 9007199254740992.0);
      }
      if (w == 0) S.q[l] = S.q0[l];
      __syncthreads();
      prep_pass<NW>(cur ? ESb : ESa, E0, E1, SL, npad);
    }
    // ---- metric at S.q: G, its sweep, u = G^-1 p (momentum draw first at kNoise) ----
    if (stage != kTraj || !have_ginv) {
      gram_pass<NB, NW>(S, Xg, SL, npad, D, inv_prior);
      // the inverse runs on wave 0; at the momentum draw its sweep also gives chol(G)
      if (w == 0) invert<NB, NW>(S, D, stage == kNoise, true);
      if constexpr (NW > 1) __syncthreads();
    }
    if (stage == kNoise) {
      if (STEP && a.momentum != nullptr && w == 0 && l < D)
        static_cast<float*>(a.momentum)[(size_t)c * Dp + l] = S.p0[l];
      const double logdet = logdet_from_pivots(S, D);
      const double pu = group_sum<64>(l < D ? (double)S.p[l] * (double)S.u[l] : 0.0);
      H0 = -logp + 0.5 * pu + 0.5 * logdet + ke_const;
      have_ginv = true;  // S.gi() = G(S.q)^-1 and S.u = G^-1 S.p
      s = 0;
      mode = 0;
      Lcur = chain_steps(a.Lc, a.L, a.C, STEP ? 0 : tr, c);
      stage = Lcur > 0 ? kTraj : kEnd;
      if (stage == kEnd) continue;
    }
    // ---- energy, accept, outputs ----
    if (stage == kEnd) {
      const double logdet = logdet_from_pivots(S, D);
      const double qq = group_sum<64>(l < D ? (double)S.q[l] * (double)S.q[l] : 0.0);
      const double pu = group_sum<64>(l < D ? (double)S.p[l] * (double)S.u[l] : 0.0);
      const double logp1 = loglik - 0.5 * qq / prior + lp_const;
      const double H1 = -logp1 + 0.5 * pu + 0.5 * logdet + ke_const;
      double dHe = H0 - H1;
      if (isnan(dHe)) dHe = -INFINITY;
      const double p_acc = fmin(exp(dHe), 1.0);
      const bool accept = uacc < p_acc;
      if (STEP && w == 0) {
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
      if (accept) {  // (only wave 0 reads S.q0)
        if (w == 0) S.q0[l] = S.q[l];
        if (!STEP && a.final_g != nullptr && w == 0 && l < D)
          static_cast<float*>(a.final_g)[(size_t)c * Dp + l] = S.g[l];
        logp = logp1;
        cur = 1 - cur;
        ++n_acc;
      }
      __syncthreads();
      if (!STEP && w == 0) {
        if (a.samples != nullptr && l < D)
          static_cast<float*>(a.samples)[((size_t)tr * a.C + c) * Dp + l] = S.q0[l];
        if (a.sample_acc != nullptr && l == 0) a.sample_acc[(size_t)tr * a.C + c] = (float)p_acc;
      }
      if (++tr == T_steps) break;
      stage = kNoise;
      continue;
    }
    // ---- stage kTraj: one evaluation of the vector field ----
    if (mode == 0) {  // z0 of step s (integrators.py:84 evaluates once outside the loop)
      q0v = S.q[l];
      p0v = S.p[l];
    }
    field_pass<NB, NW>(S, Xr, Xs, yf, E0, E1, SL, npad, D, hs, inv_prior, mode == 2);
    have_ginv = false;
    ++n_fev;
    if (mode == 2) {  // explicit half step from the last iterate (integrators.py:146-148)
      // every wave reads before wave 0 updates in place
      const float qn = l < D ? fmaf(hs, S.u[l], S.q[l]) : 0.0f;
      const float pn = l < D ? fmaf(-hs, S.dH[l], S.p[l]) : 0.0f;
      if constexpr (NW > 1) __syncthreads();
      if (w == 0 && l < D) {
        S.q[l] = qn;
        S.p[l] = pn;
      }
      __syncthreads();
      mode = 0;
      if (++s == Lcur) stage = kEnd;
      continue;
    }
    {
      float dq = 0.0f, qn = 0.0f, pn = 0.0f;
      if (l < D) {
        qn = fmaf(hs, S.u[l], q0v);
        pn = fmaf(-hs, S.dH[l], p0v);
        dq = nan_max(fabsf(qn - S.q[l]), fabsf(pn - S.p[l]));
      }
      if constexpr (NW > 1) __syncthreads();  // every wave has read the old iterate
      if (w == 0 && l < D) {
        S.q[l] = qn;
        S.p[l] = pn;
      }
      norm = (double)group_reduce<64, NanMax>(dq);
      __syncthreads();
    }
    if (mode == 0) {
      n = 0;
      mode = 1;
    } else {
      ++n;
    }
    if (!(n < a.max_iters && isfinite(norm) && norm < a.div_tol && norm > a.conv_tol)) {
      if (STEP && w == 0 && l == 0) {
        const size_t idx = (size_t)c * a.L + s;
        if (a.fp_iters) a.fp_iters[idx] = n;
        if (a.fp_norm) a.fp_norm[idx] = norm;
        if (a.fp_success) a.fp_success[idx] = (uint8_t)(isfinite(norm) && norm <= a.conv_tol);
      }
      mode = 2;
    }
  }

  if (!STEP && w == 0) {
    if (l < D) {
      static_cast<float*>(a.final_q)[(size_t)c * Dp + l] = S.q0[l];
    }
    if (l == 0) {
      a.final_logp[c] = logp;
      if (a.num_accepted) a.num_accepted[c] = n_acc;
      if (a.num_fevals) a.num_fevals[c] = n_fev;
    }
  }
  if (w == 0) zero_row_tails<float, STEP>(a, c, D, l, 64);
}

// One wave per chain; one instantiation per tile width NB (chains of the
// other width leave at once).  Workgroups are dealt to the 8 XCDs round
// robin, so block b runs chain (b % 8) * (blocks / 8) + b / 8: each XCD takes
// a contiguous range of chains -- for a multi-posterior batch, a few
// posteriors, whose packed X then stays in that XCD's L2.
template <int NB, int NW, bool STEP>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2)))
logistic_rmhmc_f32_kernel(const LogisticArgs a) {
  __shared__ Smem<NW> S;
  const int per_xcd = gridDim.x / 8;
  const int slot = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (slot >= a.C) return;
  const int c = a.dispatch_order ? a.dispatch_order[slot] : slot;
  if (c < 0 || c >= a.C) return;  // not a permutation entry: nothing runs in this slot
  const int post = a.chain_post ? a.chain_post[c] : 0;
  if (post < 0 || post >= a.P) return;  // invalid posterior index: chain left untouched
  if (a.packed_hdr[post].nb != NB) return;
  run_chain<NB, NW, STEP>(a, S, c, post);
}

// Pack posterior p (blockIdx.y) into the two fragment layouts and float labels.
__global__ void logistic_pack_kernel(const double* __restrict__ X, const double* __restrict__ y,
                                     const int32_t* __restrict__ Nv, const int32_t* __restrict__ Dv,
                                     const int64_t* __restrict__ xoff, const int64_t* __restrict__ yoff,
                                     const PackedPost* __restrict__ hdr, float* __restrict__ packed) {
  const int p = blockIdx.y;
  const PackedPost pp = hdr[p];
  const int N = Nv[p], D = Dv[p];
  const int DP = 32 * pp.nb, KC = 2 * pp.nb;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)pp.npad * DP) return;
  const int n = (int)(e / DP), f = (int)(e % DP);
  const float v = (n < N && f < D) ? (float)X[xoff[p] + (int64_t)n * D + f] : 0.0f;
  {
    const int kb = n / 16, w = n % 16, b = f / 32;
    const int lane = (w / 8) * 32 + f % 32, i = w % 8;
    packed[pp.xg + (((int64_t)kb * pp.nb + b) * 64 + lane) * 8 + i] = v;
  }
  {
    const int rb = n / 32, kc = f / 16;
    const int lane = ((f / 8) % 2) * 32 + n % 32, i = f % 8;
    const int64_t frag = ((int64_t)rb * KC + kc) * 64 + lane;
    packed[pp.xr + frag * 8 + i] = v;
    // the same split as split8: hi = bf16(v), lo = bf16(v - hi)
    const __bf16 hi = (__bf16)v;
    const __bf16 lo = (__bf16)(v - (float)hi);
    __bf16* xs = reinterpret_cast<__bf16*>(packed + pp.xs);
    xs[frag * 16 + i] = hi;
    xs[frag * 16 + 8 + i] = lo;
  }
  if (f == 0) packed[pp.yf + n] = n < N ? (float)y[yoff[p] + n] : 0.0f;
}

}  // namespace lfast

uint64_t logistic_pack_layout(int P, const int32_t* N_host, const int32_t* D_host, PackedPost* hdr_out) {
  const uint64_t hdr_floats = (((uint64_t)P * sizeof(PackedPost) + 255) / 256) * 64;
  uint64_t off = hdr_floats;
  for (int p = 0; p < P; ++p) {
    PackedPost h;
    h.npad = (N_host[p] + 63) / 64 * 64;
    h.nb = D_host[p] <= 32 ? 1 : (D_host[p] <= 64 ? 2 : 4);
    const uint64_t sz = (uint64_t)h.npad * 32 * h.nb;  // floats of one layout
    h.xg = (int64_t)off;
    off += sz;
    h.xr = (int64_t)off;
    off += sz;
    h.xs = (int64_t)off;  // bf16 pairs: as many bytes as the float layout
    off += sz;
    h.yf = (int64_t)off;
    off += ((uint64_t)h.npad + 63) / 64 * 64;
    if (hdr_out) hdr_out[p] = h;
  }
  return off * sizeof(float);
}

hipError_t launch_logistic_pack(const double* X, const double* y, const int32_t* N, const int32_t* D,
                                const int64_t* xoff, const int64_t* yoff, const PackedPost* hdr, float* packed,
                                int P, int max_npad, hipStream_t stream) {
  const int64_t elems = (int64_t)max_npad * 128;  // npad x 32 nb elements per posterior, nb <= 4
  const dim3 grid((unsigned)((elems + 255) / 256), (unsigned)P);
  hipLaunchKernelGGL(lfast::logistic_pack_kernel, grid, dim3(256), 0, stream, X, y, N, D, xoff, yoff, hdr, packed);
  return hipGetLastError();
}

namespace {

// Two side streams per device (default priority, created on first use and kept
// for the process) on which the per-width launches of a batch run concurrently.
struct SideStreams {
  hipStream_t s[2];
  bool ok;
};

SideStreams* side_streams() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static SideStreams cache[kMaxDevices];
  static bool tried[kMaxDevices];
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!tried[dev]) {
    tried[dev] = true;
    cache[dev].ok = hipStreamCreateWithFlags(&cache[dev].s[0], hipStreamNonBlocking) == hipSuccess &&
                    hipStreamCreateWithFlags(&cache[dev].s[1], hipStreamNonBlocking) == hipSuccess;
  }
  return cache[dev].ok ? &cache[dev] : nullptr;
}

}  // namespace

hipError_t launch_logistic_rmhmc_f32(const LogisticArgs& a, int max_features, hipStream_t stream, bool concurrent) {
  const int blocks = (a.C + 7) / 8 * 8;
  // one launch per packed width present: a single posterior needs only its
  // own; a batch runs every width up to its widest posterior's (chains of
  // other widths leave at once)
  const bool nb1 = max_features <= 32 || a.P > 1;
  const bool nb2 = max_features > 32 && (max_features <= 64 || a.P > 1);
  const bool nb4 = max_features > 64;
  const bool step = a.q_out != nullptr;
  auto launch_nb = [&](auto nb, auto nw, hipStream_t st) {
    constexpr int NB = decltype(nb)::value, NW = decltype(nw)::value;
    if (step)
      hipLaunchKernelGGL((lfast::logistic_rmhmc_f32_kernel<NB, NW, true>), dim3(blocks), dim3(64 * NW), 0, st, a);
    else
      hipLaunchKernelGGL((lfast::logistic_rmhmc_f32_kernel<NB, NW, false>), dim3(blocks), dim3(64 * NW), 0, st, a);
  };
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  auto launch = [&](int which, hipStream_t st) -> hipError_t {
    if (which == 2) return launch_logistic_rmhmc_f32_wide(a, st);
    const bool two = a.waves[which] == 2;
    if (which == 0) {
      if (two) launch_nb(I1{}, I2{}, st);
      else launch_nb(I1{}, I1{}, st);
    } else {
      if (two) launch_nb(I2{}, I2{}, st);
      else launch_nb(I2{}, I1{}, st);
    }
    return hipGetLastError();
  };
  int widths[3], n = 0;
  if (nb4) widths[n++] = 2;  // the widest first: it holds most of a mixed batch's work
  if (nb2) widths[n++] = 1;
  if (nb1) widths[n++] = 0;
  // The launches touch disjoint chains.  In a batch with several widths they
  // run concurrently: side streams fork from the caller's stream at an event
  // and the caller's stream waits for both before anything after this call
  // (the pattern also holds under stream capture).  The tail of one width's
  // launch then overlaps the others.
  SideStreams* ss = n > 1 && concurrent ? side_streams() : nullptr;
  if (ss == nullptr) {
    for (int i = 0; i < n; ++i) {
      const hipError_t e = launch(widths[i], stream);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
  for (int i = 0; e == hipSuccess && i < n - 1; ++i) e = hipEventCreateWithFlags(&join[i], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(fork, stream);
  for (int i = 1; e == hipSuccess && i < n; ++i) e = hipStreamWaitEvent(ss->s[i - 1], fork, 0);
  if (e == hipSuccess) e = launch(widths[0], stream);
  for (int i = 1; e == hipSuccess && i < n; ++i) {
    e = launch(widths[i], ss->s[i - 1]);
    if (e == hipSuccess) e = hipEventRecord(join[i - 1], ss->s[i - 1]);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, join[i - 1], 0);
  }
  // events may be destroyed once recorded and waited on (their resources go
  // when they complete)
  if (fork) (void)hipEventDestroy(fork);
  for (int i = 0; i < 2; ++i)
    if (join[i]) (void)hipEventDestroy(join[i]);
  return e;
}

}  // namespace gjx
