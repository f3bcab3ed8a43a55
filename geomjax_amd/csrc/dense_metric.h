// Internal launch interface of the dense general-metric LMC kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace gjx {

constexpr int kDenseMaxD = 64;

struct DenseMetricArgs {
  int C, D;
  double step_size;
  const double* step_sizes;  // [C] per-chain signed step sizes (half step), or null: step_size
  const double* metric;      // [C, D, D] G(q) as the metric function returns it
  const double* metric_jac;  // [C, D, D, D]: dG_ab/dq_i at [c][a][b][i]
  const double* z;           // [C, D] standard normals of the velocity draw, or null
  const double* velocity;    // [C, D] (position: v for v^T G v, or null; half step: the input velocity)
  const double* logdensity_grad;  // [C, D] (half step)
  const double* grad_logdet_in;   // [C, D] (half step)
  double* velocity_out;      // [C, D] (position: L^-T z; half step: the new velocity; may alias velocity)
  double* logdet;            // [C] log det of (G + G^T)/2 (position)
  double* grad_logdet;       // [C, D] grad log det G (position)
  double* vgv;               // [C] v^T G v of the given velocity (position)
  double* volume_adjustment; // [C] in/out (half step)
  // manifold MALA (drift kernel): Gamma_i = sum_j d(G^-1)_ij/dq_j (needs metric_jac),
  // drift G^-1 grad l + Gamma, and chol(G^-1) z
  double* gamma;             // [C, D] or null
  double* drift;             // [C, D] or null
  double* chol_inv_z;        // [C, D] or null (needs z)
  // RMHMC field (field kernel): momentum p (or chol(G) z when null and z is
  // given), u = G^-1 p, dT/dq and p^T u; the momentum used goes to momentum_out
  const double* momentum;    // [C, D] or null
  double* momentum_out;      // [C, D] or null
  double* dtdq;              // [C, D] or null (needs metric_jac)
  double* pu;                // [C] or null
};

size_t dense_metric_smem_bytes();
hipError_t launch_dense_metric_position(const DenseMetricArgs& a, hipStream_t stream);
hipError_t launch_dense_lmc_half_step(const DenseMetricArgs& a, hipStream_t stream);
hipError_t launch_dense_mmala_drift(const DenseMetricArgs& a, hipStream_t stream);
hipError_t launch_dense_rmhmc_field(const DenseMetricArgs& a, hipStream_t stream);

}  // namespace gjx
