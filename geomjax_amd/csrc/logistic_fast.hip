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

#include "logistic_fast_ops.h"

namespace gjx {
namespace lfast {

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
