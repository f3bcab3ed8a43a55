// Internal launch interface of the logistic-regression kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gjx {

// One posterior of the packed float32 design-matrix table (offsets in floats
// from the start of the packed buffer).
struct PackedPost {
  int64_t xg;    // Gram layout: [npad/16][nb][64 lanes][8]
  int64_t xr;    // row layout:  [npad/32][2 nb][64 lanes][8]
  int64_t xs;    // row layout split into bf16: [npad/32][2 nb][64 lanes][8 hi, 8 lo]
  int64_t yf;    // labels [npad]
  int32_t npad;  // rows rounded up to 64 (padding rows are zero)
  int32_t nb;    // 32-feature blocks: 1 (D <= 32) or 2 (D <= 64)
};

struct LogisticArgs {
  int C, L, T, max_iters, Dmax, P;
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
};

constexpr int kLogisticMaxD = 64;

hipError_t launch_logistic_rmhmc_f64(const LogisticArgs& a, hipStream_t stream);
hipError_t launch_logistic_value_and_grad(const LogisticArgs& a, hipStream_t stream);
hipError_t launch_logistic_lmc_f64(const LogisticArgs& a, hipStream_t stream);
hipError_t launch_logistic_rmhmc_f32(const LogisticArgs& a, int max_features, hipStream_t stream);
hipError_t launch_logistic_pack(const double* X, const double* y, const int32_t* N, const int32_t* D,
                                const int64_t* xoff, const int64_t* yoff, const PackedPost* hdr, float* packed,
                                int P, int max_npad, hipStream_t stream);

// Host-side layout of the packed buffer: header of P PackedPost entries
// (rounded up to 256 bytes), then per posterior the two fragment layouts of
// X and the float labels.  Returns the total size in bytes.
uint64_t logistic_pack_layout(int P, const int32_t* N_host, const int32_t* D_host, PackedPost* hdr_out);

}  // namespace gjx
