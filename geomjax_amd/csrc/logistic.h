// Internal launch interface of the logistic-regression kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nuts.h"

namespace gjx {

// One posterior of the packed float32 design-matrix table (offsets in floats
// from the start of the packed buffer).
struct PackedPost {
  int64_t xg;    // Gram layout: [npad/16][nb][64 lanes][8]
  int64_t xr;    // row layout:  [npad/32][2 nb][64 lanes][8]
  int64_t xs;    // row layout split into bf16: [npad/32][2 nb][64 lanes][8 hi, 8 lo]
  int64_t yf;    // labels [npad]
  int32_t npad;  // rows rounded up to 64 (padding rows are zero)
  int32_t nb;    // 32-feature blocks: 1 (D <= 32), 2 (D <= 64) or 4 (D <= 128)
};

struct LogisticArgs {
  int C, L, T, max_iters, Dmax, P;
  int max_features;  // host copy of max_p D_p
  double eps, prior_var, conv_tol, div_tol, div_threshold;
  uint64_t seed, counter;
  // data: posteriors p = 0..P-1 with X_p [N_p, D_p] row-major at X + post_xoff[p]
  const int32_t* chain_post;  // [C] or null (posterior 0)
  const int32_t* post_N;
  const int32_t* post_D;
  const int64_t* post_xoff;
  const int64_t* post_yoff;
  const double* X;
  const double* y;
  // inputs (chain-major, row stride Dmax)
  const void* q_in;
  const double* logp_in;
  const void* g_in;
  const void* z;
  const double* u;
  const double* J_in;  // LMC volume adjustment
  const int32_t* Lc;   // [T, C] per-chain integration steps (dynamic kernel) or null: L for all
  // single-transition outputs
  double* J_out;
  double* prop_J;
  double* final_J;
  void* q_out;
  double* logp_out;
  void* g_out;
  void* momentum;
  double* acc_rate;
  uint8_t* accepted;
  uint8_t* divergent;
  double* energy;
  void* prop_q;
  void* prop_p;
  void* prop_v;
  double* prop_logp;
  void* prop_g;
  double* prop_weight;
  double* prop_slpa;
  int32_t* fp_iters;
  double* fp_norm;
  uint8_t* fp_success;
  // multi-transition outputs
  void* final_q;
  double* final_logp;
  void* final_g;
  void* samples;
  float* sample_acc;
  int32_t* num_accepted;
  int64_t* num_fevals;
  // float32 path: packed design matrices (PackedPost header, then float data)
  // and per-chain scratch of work_stride floats (5 arrays of work_rows)
  const PackedPost* packed_hdr;
  const float* packed;
  float* work;
  int64_t work_rows;
  const int32_t* dispatch_order;  // [C] chain of each dispatch slot (tuned kernels) or null
  // tuned one-workgroup-per-chain kernels of posteriors up to 32 / 64
  // features: waves per chain (1 or 2)
  int waves[2];
  // manifold MALA: metrics [C, Dmax, Dmax] and Gamma vectors [C, Dmax] of the
  // returned state and of the proposal; mmala_exact selects +log alpha as the
  // proposal weight (0: the reference's pairing, -log alpha)
  double* metric_out;
  double* gamma_out;
  double* prop_metric;
  double* prop_gamma;
  int mmala_exact;
};

// Coordinates [D, Dmax) of chain c's rows (a posterior narrower than the
// state arrays) are written as zeros in every per-chain row output the call
// fills, so outputs never depend on what the caller's buffers held.  Lanes
// lane, lane + nlanes, ... of the chain's workgroup share the work.
template <class T, bool STEP>
__device__ __forceinline__ void zero_row_tails(const LogisticArgs& a, int c, int D, int lane, int nlanes) {
  const int Dp = a.Dmax;
  if (D >= Dp) return;
  const size_t row = (size_t)c * Dp;
  auto zero = [&](void* base, size_t r) {
    if (base == nullptr) return;
    for (int j = D + lane; j < Dp; j += nlanes) static_cast<T*>(base)[r + j] = T(0);
  };
  if (STEP) {
    zero(a.q_out, row);
    zero(a.g_out, row);
    zero(a.momentum, row);
    zero(a.prop_q, row);
    zero(a.prop_p, row);
    zero(a.prop_v, row);
    zero(a.prop_g, row);
  } else {
    zero(a.final_q, row);
    zero(a.final_g, row);
    if (a.samples != nullptr)
      for (int tr = 0; tr < a.T; ++tr) zero(a.samples, ((size_t)tr * a.C + c) * Dp);
  }
}

constexpr int kLogisticMaxDLmc = 64;   // LMC kernel
constexpr int kLogisticMaxD = 128;     // RMHMC kernels (float64 and float32), value_and_grad, packing

// per-chain outputs of the tangent kernel ([C] values, [C, 2] derivatives)
struct TangentOut {
  double* objective;    // acceptance * squared jump
  double* acceptance;   // min(1, e^dH)
  double* sq_jump;      // |q_L - q_0|^2
  double* d_objective;
  double* d_acceptance;
  double* d_sq_jump;
  int64_t* num_fevals;
};

hipError_t launch_logistic_rmhmc_f64(const LogisticArgs& a, hipStream_t stream);
// manifold MALA (Fisher metric, float64, D <= 64): one transition (q_out set) or
// T transitions (final_q set); the init kernel fills logp_out, g_out,
// metric_out and gamma_out at q_in
hipError_t launch_logistic_mmala_f64(const LogisticArgs& a, hipStream_t stream);
hipError_t launch_logistic_mmala_init(const LogisticArgs& a, hipStream_t stream);
hipError_t launch_logistic_rmhmc_tangent(const LogisticArgs& a, const TangentOut& o, hipStream_t stream);
// tangent for 65..128 features: float32, fp32 MFMA, 256 lanes per chain (logistic_grad.hip)
hipError_t launch_logistic_rmhmc_tangent_f32_wide(const LogisticArgs& a, const TangentOut& o, hipStream_t stream);
// tangent fast path: float32 state, split-bf16 MFMA, D <= 64 (logistic_grad_fast.hip)
hipError_t launch_logistic_rmhmc_tangent_f32(const LogisticArgs& a, const TangentOut& o, int max_features,
                                             hipStream_t stream);
// NUTS with the RMHMC integrator, float64, D <= 64 (logistic_nuts.hip)
hipError_t launch_logistic_nuts_f64(const LogisticArgs& a, const NutsArgs& n, hipStream_t stream);
// NUTS with the RMHMC integrator, tuned form (float32 state), D <= 64 (logistic_nuts_fast.hip)
hipError_t launch_logistic_nuts_f32(const LogisticArgs& a, const NutsArgs& n, int max_features, hipStream_t stream);
// NUTS with the general-metric LMC integrator, tuned form, D <= 64 (logistic_nuts_fast.hip)
hipError_t launch_logistic_nuts_lmc_f32(const LogisticArgs& a, const NutsArgs& n, int max_features,
                                        hipStream_t stream);
// NUTS with the general-metric LMC integrator, float64, D <= 64 (logistic_nuts.hip)
hipError_t launch_logistic_nuts_lmc_f64(const LogisticArgs& a, const NutsArgs& n, hipStream_t stream);
hipError_t launch_logistic_value_and_grad(const LogisticArgs& a, hipStream_t stream);
hipError_t launch_logistic_lmc_f64(const LogisticArgs& a, hipStream_t stream);
// LMC, tuned form (float32 state, split-bf16 MFMA Grams), D <= 64 (logistic_lmc_fast.hip)
hipError_t launch_logistic_lmc_f32(const LogisticArgs& a, int max_features, hipStream_t stream);
// concurrent = false launches every packed width on `stream` in turn (no side
// streams or events are created)
hipError_t launch_logistic_rmhmc_f32(const LogisticArgs& a, int max_features, hipStream_t stream,
                                     bool concurrent = true);
hipError_t launch_logistic_rmhmc_f32_wide(const LogisticArgs& a, hipStream_t stream);
// waves per chain of the tuned kernels for posteriors of up to 32 / 64
// features when the caller leaves the choice to the library (the faster shape
// of each class on the 2048 x 64 and uneven-batch workloads)
constexpr int kDefaultWavesPerChain[2] = {2, 1};
hipError_t launch_logistic_pack(const double* X, const double* y, const int32_t* N, const int32_t* D,
                                const int64_t* xoff, const int64_t* yoff, const PackedPost* hdr, float* packed,
                                int P, int max_npad, hipStream_t stream);

// Host-side layout of the packed buffer: header of P PackedPost entries
// (rounded up to 256 bytes), then per posterior the two fragment layouts of
// X and the float labels.  Returns the total size in bytes.
uint64_t logistic_pack_layout(int P, const int32_t* N_host, const int32_t* D_host, PackedPost* hdr_out);

}  // namespace gjx
