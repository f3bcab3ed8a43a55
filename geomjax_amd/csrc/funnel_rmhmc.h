// Internal launch interface of the funnel integrator kernels (not part of the
// public C ABI; the extern "C" wrappers in capi.hip validate and forward).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nuts.h"

namespace gjx {

// Everything one launch needs, passed by value as the kernel argument.
// Pointers that an entry point does not use are null.
struct FunnelArgs {
  int C, D, L, T, max_iters;
  double eps, alpha2, sigma, conv_tol, div_tol, div_threshold;
  uint64_t seed, counter;
  // inputs
  const void* q_in;
  const double* logp_in;
  const void* g_in;
  const void* z;    // [T, C, D] or null (device PRNG)
  const double* u;  // [T, C] or null (device PRNG)
  const int32_t* Lc;  // [T, C] per-chain integration steps (dynamic kernel) or null: L for all
  // single-transition outputs
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
};

// Outputs of the tangent kernel (per chain; d_* are [C, 2]: d/d log eps, d/d log alpha2).
struct FunnelGradOut {
  double* objective;    // acceptance * squared jump
  double* acceptance;   // min(1, e^dH)
  double* sq_jump;      // |q_L - q_0|^2
  double* d_objective;
  double* d_acceptance;
  double* d_sq_jump;
  int64_t* num_fevals;
};

hipError_t launch_funnel_rmhmc_tangent(const FunnelArgs& a, const FunnelGradOut& o, int lpc, int epl, int dtype,
                                       hipStream_t stream);

// Smallest compiled elements-per-lane that covers the D-1 "x" coordinates
// with `lpc` lanes per chain, or -1 if none does.
int funnel_pick_epl(int dim, int lpc);

hipError_t launch_funnel_rmhmc(const FunnelArgs& a, int dtype, int lpc, int epl, hipStream_t stream);

// NUTS (iterative, multinomial, numpyro-style U-turn with the "euc" criterion)
hipError_t launch_funnel_nuts(const FunnelArgs& a, const NutsArgs& n, int dtype, hipStream_t stream);

hipError_t launch_funnel_value_and_grad(int C, int D, double sigma, int dtype, const void* q, double* logp,
                                        void* grad, hipStream_t stream);

}  // namespace gjx
