// Internal launch interface of the LMC-Monge funnel kernel (not part of the
// public C ABI; capi.hip validates and forwards).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nuts.h"

namespace gjx {

struct FunnelLmcArgs {
  int C, D, L, T;
  int exact_sm;        // 1: Sherman-Morrison correction with its alpha2 factor (reference omits it)
  int exact_momentum;  // 1: momentum = M v + alpha2 (v.g) g (reference scales the rank-1 term by L)
  double eps, alpha2, sigma, div_threshold;
  uint64_t seed, counter;
  const void* inv_mass;  // [D] dtype, diagonal inverse mass matrix
  // inputs: LMCState
  const void* q_in;
  const double* logp_in;
  const void* g_in;
  const double* J_in;
  const void* z;    // [T, C, D] or null
  const double* u;  // [T, C] or null
  const int32_t* Lc;  // [T, C] per-chain integration steps (dynamic kernel) or null: L for all
  // single-transition outputs
  void* q_out;
  double* logp_out;
  void* g_out;
  double* J_out;
  void* velocity;  // info.velocity (v0)
  double* acc_rate;
  uint8_t* accepted;
  uint8_t* divergent;
  double* energy;
  void* prop_q;
  void* prop_p;
  void* prop_v;
  double* prop_logp;
  void* prop_gnorm;
  void* prop_dl_ig;
  void* prop_Hdl_ig;
  void* prop_ig_Hdl_ig;
  void* prop_hvn;
  double* prop_L;
  double* prop_J;
  double* prop_weight;
  double* prop_slpa;
  // multi-transition outputs
  void* final_q;
  double* final_logp;
  void* final_g;
  double* final_J;
  void* samples;
  float* sample_acc;
  int32_t* num_accepted;
  // tangent kernel: d(proposal position)/d alpha2, [C, D] float64
  double* d_prop_q;
};

hipError_t launch_funnel_lmc(const FunnelLmcArgs& a, int dtype, int lpc, int epl, hipStream_t stream);
// NUTS with the LMC-Monge integrator, float64, one wave per chain (funnel_nuts.hip)
hipError_t launch_funnel_nuts_lmc(const FunnelLmcArgs& a, const NutsArgs& n, int dtype, hipStream_t stream);
// forward-mode derivative of the proposal position in alpha2 (float64)
hipError_t launch_funnel_lmc_alpha2_tangent(const FunnelLmcArgs& a, int lpc, int epl, hipStream_t stream);

}  // namespace gjx
