// Shared pieces of the NUTS kernels (funnel_nuts.hip, logistic_nuts.hip):
// the launch arguments of the tree building and its small device helpers.
//
// Reference: geomjax/rmhmc/nuts.py:78-161, geomjax/mcmc/trajectory.py:149-317,
// 517-668, geomjax/mcmc/termination.py:30-119, geomjax/mcmc/proposal.py:198-254.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gjx {

struct NutsArgs {
  int max_doublings;           // M
  int criterion;               // U-turn criterion: 0 "euc", 1 "bet", 2 "riem"
  const double* u_dir;         // [C, M] or null (device PRNG)
  const double* u_merge;       // [C, M]
  const double* u_step;        // [C, 2^M - 1]
  // outputs (optional unless noted)
  double* energy;              // [C]
  double* acceptance_rate;     // [C]
  uint8_t* is_divergent;       // [C]
  uint8_t* is_turning;         // [C]
  int32_t* num_expansions;     // [C]
  int32_t* num_steps;          // [C]
  void* left_position;         // [C, D] (row stride of the state arrays)
  void* right_position;        // [C, D]
};
// checkpoint arrays of the iterative U-turn check: velocities and velocity sums,
// plus momenta and momentum sums for the criteria that read them
__host__ __device__ inline int nuts_checkpoint_arrays(int criterion) { return criterion == 0 ? 2 : 4; }

namespace nuts {

// log(e^a + e^b) that keeps -inf + -inf = -inf
__device__ __forceinline__ double logaddexp(double a, double b) {
  const double m = fmax(a, b);
  if (m == -INFINITY) return -INFINITY;
  return m + log(exp(a - m) + exp(b - m));
}

// numpyro checkpoint indices of leaf `step` of a subtree (termination.py:71-87):
// idx_max = number of set bits of step >> 1, idx_min = idx_max - (trailing ones
// of step) + 1; a check runs over [idx_min, idx_max] (empty at even steps).
__device__ __forceinline__ int ckpt_idx_max(int step) { return __builtin_popcount((unsigned)step >> 1); }
__device__ __forceinline__ int ckpt_idx_min(int step) {
  return ckpt_idx_max(step) - __builtin_ctz(~(unsigned)step) + 1;
}

}  // namespace nuts
}  // namespace gjx
