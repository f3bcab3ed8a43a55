// extern "C" entry points: argument validation, dispatch, error reporting.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/geomjax_amd.h"
#include "funnel_lmc.h"
#include "funnel_rmhmc.h"
#include "dense_metric.h"
#include "logistic.h"

namespace gjx {
int diagnostics_max_samples();
uint64_t diagnostics_workspace_bytes(int N, int C, int D);
hipError_t launch_diagnostics(const void* x, int dtype, int N, int C, int D, int64_t st, int64_t sc, int64_t sd,
                              void* workspace, double* ess, double* rhat, hipStream_t stream);
}  // namespace gjx

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(GJX_EHIP, "%s: %s", where, hipGetErrorString(e));
}

#define GJX_REQUIRE(cond, ...)                     \
  do {                                             \
    if (!(cond)) return fail(GJX_EINVAL, __VA_ARGS__); \
  } while (0)

// Lanes per chain when the caller does not choose: enough lanes that each
// holds at most 8 coordinates, at least 4 and at most 64.
int default_lpc(int dim) {
  int lpc = 4;
  while (lpc < 64 && (dim - 1 + lpc - 1) / lpc > 8) lpc <<= 1;
  return lpc;
}

int check_funnel_params(const gjx_funnel_params* p, bool sample, int* lpc, int* epl) {
  GJX_REQUIRE(p != nullptr, "params is NULL");
  GJX_REQUIRE(p->num_chains > 0, "num_chains must be > 0 (got %d)", p->num_chains);
  GJX_REQUIRE(p->dim >= 2, "the funnel needs dim >= 2 (got %d)", p->dim);
  GJX_REQUIRE(p->dtype == GJX_F64 || p->dtype == GJX_F32, "dtype must be GJX_F64 or GJX_F32 (got %d)",
              p->dtype);
  GJX_REQUIRE(p->num_integration_steps >= 0, "num_integration_steps must be >= 0 (got %d)",
              p->num_integration_steps);
  GJX_REQUIRE(p->max_iters >= 0, "max_iters must be >= 0 (got %d)", p->max_iters);
  GJX_REQUIRE(std::isfinite(p->step_size), "step_size must be finite");
  GJX_REQUIRE(std::isfinite(p->alpha2) && p->alpha2 >= 0.0, "alpha2 must be finite and >= 0");
  GJX_REQUIRE(std::isfinite(p->sigma) && p->sigma > 0.0, "sigma must be finite and > 0");
  if (sample) GJX_REQUIRE(p->num_transitions >= 1, "num_transitions must be >= 1 (got %d)", p->num_transitions);
  int l = p->lanes_per_chain == 0 ? default_lpc(p->dim) : p->lanes_per_chain;
  GJX_REQUIRE(l == 4 || l == 8 || l == 16 || l == 32 || l == 64,
              "lanes_per_chain must be 0, 4, 8, 16, 32 or 64 (got %d)", p->lanes_per_chain);
  int e = gjx::funnel_pick_epl(p->dim, l);
  if (e < 0)
    return fail(GJX_EUNSUPPORTED, "dim %d needs more than 16 coordinates per lane at %d lanes per chain",
                p->dim, l);
  *lpc = l;
  *epl = e;
  return GJX_OK;
}

gjx::FunnelArgs base_args(const gjx_funnel_params* p) {
  gjx::FunnelArgs a;
  std::memset(&a, 0, sizeof(a));
  a.C = p->num_chains;
  a.D = p->dim;
  a.L = p->num_integration_steps;
  a.max_iters = p->max_iters;
  a.eps = p->step_size;
  a.alpha2 = p->alpha2;
  a.sigma = p->sigma;
  a.conv_tol = p->convergence_tol;
  a.div_tol = p->divergence_tol;
  a.div_threshold = p->divergence_threshold;
  a.seed = p->seed;
  a.counter = p->counter;
  return a;
}

}  // namespace

extern "C" {

const char* gjx_version(void) { return "geomjax_amd 0.1.0 (gfx950)"; }

const char* gjx_last_error(void) { return g_last_error.c_str(); }

int gjx_rmhmc_monge_funnel_step(const gjx_funnel_params* params, const gjx_rmhmc_step_io* io, void* stream) {
  int lpc = 0, epl = 0;
  int rc = check_funnel_params(params, false, &lpc, &epl);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  GJX_REQUIRE(!io->new_logdensity_grad || io->logdensity_grad,
              "new_logdensity_grad needs logdensity_grad (copied through on rejection)");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  gjx::FunnelArgs a = base_args(params);
  a.T = 1;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.g_in = io->logdensity_grad;
  a.z = io->z;
  a.u = io->u;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.momentum = io->momentum;
  a.acc_rate = io->acceptance_rate;
  a.accepted = io->is_accepted;
  a.divergent = io->is_divergent;
  a.energy = io->energy;
  a.prop_q = io->proposal_position;
  a.prop_p = io->proposal_momentum;
  a.prop_v = io->proposal_velocity;
  a.prop_logp = io->proposal_logdensity;
  a.prop_g = io->proposal_logdensity_grad;
  a.prop_weight = io->proposal_weight;
  a.prop_slpa = io->proposal_sum_log_p_accept;
  a.fp_iters = io->fp_iters;
  a.fp_norm = io->fp_norm;
  a.fp_success = io->fp_success;
  a.Lc = io->num_integration_steps;
  hipError_t e = gjx::launch_funnel_rmhmc(a, params->dtype, lpc, epl, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_monge_funnel_step launch");
  return GJX_OK;
}

int gjx_rmhmc_monge_funnel_sample(const gjx_funnel_params* params, const gjx_rmhmc_sample_io* io,
                                  void* stream) {
  int lpc = 0, epl = 0;
  int rc = check_funnel_params(params, true, &lpc, &epl);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->final_position && io->final_logdensity, "final_position and final_logdensity are required");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  gjx::FunnelArgs a = base_args(params);
  a.T = params->num_transitions;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  a.u = io->u;
  a.final_q = io->final_position;
  a.final_logp = io->final_logdensity;
  a.final_g = io->final_logdensity_grad;
  a.samples = io->samples;
  a.sample_acc = io->acceptance_rate;
  a.num_accepted = io->num_accepted;
  a.num_fevals = io->num_fevals;
  a.Lc = io->num_integration_steps;
  hipError_t e = gjx::launch_funnel_rmhmc(a, params->dtype, lpc, epl, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_monge_funnel_sample launch");
  return GJX_OK;
}

int gjx_rmhmc_monge_funnel_tangent(const gjx_funnel_params* params, const gjx_tangent_io* io, void* stream) {
  GJX_REQUIRE(params != nullptr && io != nullptr, "params and io are required");
  GJX_REQUIRE(params->dtype == GJX_F64 || params->dtype == GJX_F32,
              "dtype must be GJX_F64 (check path) or GJX_F32 (fast path)");
  GJX_REQUIRE(params->alpha2 > 0.0 && std::isfinite(params->alpha2), "alpha2 must be > 0 (log alpha2 is a parameter)");
  GJX_REQUIRE(params->step_size > 0.0 && std::isfinite(params->step_size),
              "step_size must be > 0 (log step_size is a parameter)");
  gjx_funnel_params p = *params;
  if (p.lanes_per_chain == 0) p.lanes_per_chain = 32;
  GJX_REQUIRE(p.lanes_per_chain == 16 || p.lanes_per_chain == 32 || p.lanes_per_chain == 64,
              "the tangent kernel takes lanes_per_chain 16, 32 or 64");
  int lpc = 0, epl = 0;
  int rc = check_funnel_params(&p, false, &lpc, &epl);
  if (rc != GJX_OK) return rc;
  if (epl > 8) return fail(GJX_EUNSUPPORTED, "dim %d too large for the tangent kernel at %d lanes", p.dim, lpc);
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->objective && io->acceptance && io->sq_jump && io->d_objective && io->d_acceptance &&
                  io->d_sq_jump,
              "every output except num_fevals is required");
  gjx::FunnelArgs a = base_args(&p);
  a.T = 1;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  gjx::FunnelGradOut o{io->objective, io->acceptance, io->sq_jump, io->d_objective, io->d_acceptance,
                       io->d_sq_jump, io->num_fevals};
  hipError_t e = gjx::launch_funnel_rmhmc_tangent(a, o, lpc, epl, p.dtype, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_monge_funnel_tangent launch");
  return GJX_OK;
}

namespace {

int check_funnel_nuts(const gjx_funnel_params* params, int32_t max_num_doublings, int32_t stopping_criterion) {
  GJX_REQUIRE(params != nullptr, "params is required");
  GJX_REQUIRE(params->dtype == GJX_F64 || params->dtype == GJX_F32,
              "dtype must be GJX_F64 or GJX_F32 (the tuned form)");
  GJX_REQUIRE(params->num_chains > 0, "num_chains must be > 0 (got %d)", params->num_chains);
  GJX_REQUIRE(params->dim >= 2, "the funnel needs dim >= 2 (got %d)", params->dim);
  if (params->dim - 1 > 256) return fail(GJX_EUNSUPPORTED, "dim %d > 257 for the NUTS kernel", params->dim);
  GJX_REQUIRE(max_num_doublings >= 1 && max_num_doublings <= 12, "max_num_doublings must be in [1, 12] (got %d)",
              max_num_doublings);
  GJX_REQUIRE(stopping_criterion == GJX_UTURN_EUC || stopping_criterion == GJX_UTURN_BET ||
                  stopping_criterion == GJX_UTURN_RIEM,
              "stopping_criterion must be GJX_UTURN_EUC, GJX_UTURN_BET or GJX_UTURN_RIEM (got %d)",
              stopping_criterion);
  GJX_REQUIRE(params->max_iters >= 0, "max_iters must be >= 0");
  GJX_REQUIRE(std::isfinite(params->step_size), "step_size must be finite");
  GJX_REQUIRE(std::isfinite(params->alpha2) && params->alpha2 >= 0.0, "alpha2 must be finite and >= 0");
  GJX_REQUIRE(std::isfinite(params->sigma) && params->sigma > 0.0, "sigma must be finite and > 0");
  return GJX_OK;
}

}  // namespace

int gjx_rmhmc_monge_funnel_nuts_sample(const gjx_funnel_params* params, int32_t max_num_doublings,
                                       int32_t stopping_criterion, const gjx_nuts_sample_io* io, void* stream) {
  int rc = check_funnel_nuts(params, max_num_doublings, stopping_criterion);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io != nullptr, "io is required");
  GJX_REQUIRE(params->num_transitions >= 1, "num_transitions must be >= 1 (got %d)", params->num_transitions);
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->final_position && io->final_logdensity, "final_position and final_logdensity are required");
  const bool host = io->z != nullptr;
  GJX_REQUIRE((io->u_direction != nullptr) == host && (io->u_merge != nullptr) == host &&
                  (io->u_step != nullptr) == host,
              "z, u_direction, u_merge and u_step must all be given or all be NULL");
  gjx::FunnelArgs a = base_args(params);
  a.T = params->num_transitions;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  a.final_q = io->final_position;
  a.final_logp = io->final_logdensity;
  a.final_g = io->final_logdensity_grad;
  a.samples = io->samples;
  a.sample_acc = io->acceptance_rate;
  a.num_fevals = io->num_fevals;
  gjx::NutsArgs n;
  std::memset(&n, 0, sizeof(n));
  n.max_doublings = max_num_doublings;
  n.criterion = stopping_criterion;
  n.u_dir = io->u_direction;
  n.u_merge = io->u_merge;
  n.u_step = io->u_step;
  n.num_steps = io->num_integration_steps;
  hipError_t e = gjx::launch_funnel_nuts(a, n, params->dtype, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_monge_funnel_nuts_sample launch");
  return GJX_OK;
}

int gjx_rmhmc_monge_funnel_nuts(const gjx_funnel_params* params, int32_t max_num_doublings,
                                int32_t stopping_criterion, const gjx_nuts_io* io, void* stream) {
  GJX_REQUIRE(params != nullptr && io != nullptr, "params and io are required");
  GJX_REQUIRE(params->dtype == GJX_F64 || params->dtype == GJX_F32,
              "dtype must be GJX_F64 or GJX_F32 (the tuned form)");
  GJX_REQUIRE(params->num_chains > 0, "num_chains must be > 0 (got %d)", params->num_chains);
  GJX_REQUIRE(params->dim >= 2, "the funnel needs dim >= 2 (got %d)", params->dim);
  if (params->dim - 1 > 256) return fail(GJX_EUNSUPPORTED, "dim %d > 257 for the NUTS kernel", params->dim);
  GJX_REQUIRE(max_num_doublings >= 1 && max_num_doublings <= 12, "max_num_doublings must be in [1, 12] (got %d)",
              max_num_doublings);
  GJX_REQUIRE(stopping_criterion == GJX_UTURN_EUC || stopping_criterion == GJX_UTURN_BET ||
                  stopping_criterion == GJX_UTURN_RIEM,
              "stopping_criterion must be GJX_UTURN_EUC, GJX_UTURN_BET or GJX_UTURN_RIEM (got %d)",
              stopping_criterion);
  GJX_REQUIRE(params->max_iters >= 0, "max_iters must be >= 0");
  GJX_REQUIRE(std::isfinite(params->step_size), "step_size must be finite");
  GJX_REQUIRE(std::isfinite(params->alpha2) && params->alpha2 >= 0.0, "alpha2 must be finite and >= 0");
  GJX_REQUIRE(std::isfinite(params->sigma) && params->sigma > 0.0, "sigma must be finite and > 0");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  const bool host = io->z != nullptr;
  GJX_REQUIRE((io->u_direction != nullptr) == host && (io->u_merge != nullptr) == host &&
                  (io->u_step != nullptr) == host,
              "z, u_direction, u_merge and u_step must all be given or all be NULL");
  gjx::FunnelArgs a = base_args(params);
  a.T = 1;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.momentum = io->momentum;
  gjx::NutsArgs n;
  std::memset(&n, 0, sizeof(n));
  n.max_doublings = max_num_doublings;
  n.criterion = stopping_criterion;
  n.u_dir = io->u_direction;
  n.u_merge = io->u_merge;
  n.u_step = io->u_step;
  n.energy = io->energy;
  n.acceptance_rate = io->acceptance_rate;
  n.is_divergent = io->is_divergent;
  n.is_turning = io->is_turning;
  n.num_expansions = io->num_trajectory_expansions;
  n.num_steps = io->num_integration_steps;
  n.left_position = io->leftmost_position;
  n.right_position = io->rightmost_position;
  hipError_t e = gjx::launch_funnel_nuts(a, n, params->dtype, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_monge_funnel_nuts launch");
  return GJX_OK;
}

int gjx_funnel_value_and_grad(int32_t num_chains, int32_t dim, double sigma, int32_t dtype, const void* position,
                              double* logdensity, void* logdensity_grad, void* stream) {
  GJX_REQUIRE(num_chains > 0, "num_chains must be > 0 (got %d)", num_chains);
  GJX_REQUIRE(dim >= 2, "the funnel needs dim >= 2 (got %d)", dim);
  GJX_REQUIRE(dtype == GJX_F64 || dtype == GJX_F32, "dtype must be GJX_F64 or GJX_F32 (got %d)", dtype);
  GJX_REQUIRE(std::isfinite(sigma) && sigma > 0.0, "sigma must be finite and > 0");
  GJX_REQUIRE(position && logdensity && logdensity_grad, "position, logdensity and logdensity_grad are required");
  hipError_t e = gjx::launch_funnel_value_and_grad(num_chains, dim, sigma, dtype, position, logdensity,
                                                   logdensity_grad, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_funnel_value_and_grad launch");
  return GJX_OK;
}

namespace {

gjx::FunnelLmcArgs lmc_args(const gjx_funnel_params* p) {
  gjx::FunnelLmcArgs a;
  std::memset(&a, 0, sizeof(a));
  a.C = p->num_chains;
  a.D = p->dim;
  a.L = p->num_integration_steps;
  a.exact_sm = (p->flags & GJX_LMC_EXACT_SHERMAN_MORRISON) ? 1 : 0;
  a.exact_momentum = (p->flags & GJX_LMC_EXACT_MOMENTUM) ? 1 : 0;
  a.eps = p->step_size;
  a.alpha2 = p->alpha2;
  a.sigma = p->sigma;
  a.div_threshold = p->divergence_threshold;
  a.seed = p->seed;
  a.counter = p->counter;
  return a;
}

}  // namespace

int gjx_lmcmonge_funnel_step(const gjx_funnel_params* params, const gjx_lmc_step_io* io, void* stream) {
  int lpc = 0, epl = 0;
  int rc = check_funnel_params(params, false, &lpc, &epl);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE((params->flags & ~3) == 0, "unknown flags 0x%x", params->flags);
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->inverse_mass_matrix, "inverse_mass_matrix is required");
  GJX_REQUIRE(io->position && io->logdensity && io->volume_adjustment,
              "position, logdensity and volume_adjustment are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  GJX_REQUIRE(!io->new_logdensity_grad || io->logdensity_grad,
              "new_logdensity_grad needs logdensity_grad (copied through on rejection)");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  gjx::FunnelLmcArgs a = lmc_args(params);
  a.T = 1;
  a.inv_mass = io->inverse_mass_matrix;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.g_in = io->logdensity_grad;
  a.J_in = io->volume_adjustment;
  a.Lc = io->num_integration_steps;
  a.z = io->z;
  a.u = io->u;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.J_out = io->new_volume_adjustment;
  a.velocity = io->velocity;
  a.acc_rate = io->acceptance_rate;
  a.accepted = io->is_accepted;
  a.divergent = io->is_divergent;
  a.energy = io->energy;
  a.prop_q = io->proposal_position;
  a.prop_p = io->proposal_momentum;
  a.prop_v = io->proposal_velocity;
  a.prop_logp = io->proposal_logdensity;
  a.prop_gnorm = io->proposal_logdensity_grad_norm;
  a.prop_dl_ig = io->proposal_dl_ig;
  a.prop_Hdl_ig = io->proposal_Hdl_ig;
  a.prop_ig_Hdl_ig = io->proposal_ig_Hdl_ig;
  a.prop_hvn = io->proposal_logdensity_hvp_velocity_norm;
  a.prop_L = io->proposal_determinant_metric;
  a.prop_J = io->proposal_volume_adjustment;
  a.prop_weight = io->proposal_weight;
  a.prop_slpa = io->proposal_sum_log_p_accept;
  hipError_t e = gjx::launch_funnel_lmc(a, params->dtype, lpc, epl, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_lmcmonge_funnel_step launch");
  return GJX_OK;
}

int gjx_lmcmonge_funnel_nuts(const gjx_funnel_params* params, int32_t max_num_doublings, int32_t stopping_criterion,
                             const gjx_lmc_nuts_io* io, void* stream) {
  int lpc = 0, epl = 0;
  int rc = check_funnel_params(params, false, &lpc, &epl);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE((params->flags & ~3) == 0, "unknown flags 0x%x", params->flags);
  if (params->dim - 1 > 256) return fail(GJX_EUNSUPPORTED, "dim %d > 257 for the NUTS kernel", params->dim);
  GJX_REQUIRE(max_num_doublings >= 1 && max_num_doublings <= 12, "max_num_doublings must be in [1, 12] (got %d)",
              max_num_doublings);
  GJX_REQUIRE(stopping_criterion == GJX_UTURN_EUC || stopping_criterion == GJX_UTURN_BET ||
                  stopping_criterion == GJX_UTURN_RIEM,
              "stopping_criterion must be GJX_UTURN_EUC, GJX_UTURN_BET or GJX_UTURN_RIEM (got %d)",
              stopping_criterion);
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->inverse_mass_matrix, "inverse_mass_matrix is required");
  GJX_REQUIRE(io->position && io->logdensity && io->volume_adjustment,
              "position, logdensity and volume_adjustment are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  const bool host = io->z != nullptr;
  GJX_REQUIRE((io->u_direction != nullptr) == host && (io->u_merge != nullptr) == host &&
                  (io->u_step != nullptr) == host,
              "z, u_direction, u_merge and u_step must all be given or all be NULL");
  gjx::FunnelLmcArgs a = lmc_args(params);
  a.T = 1;
  a.inv_mass = io->inverse_mass_matrix;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.J_in = io->volume_adjustment;
  a.z = io->z;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.J_out = io->new_volume_adjustment;
  a.velocity = io->velocity;
  gjx::NutsArgs n;
  std::memset(&n, 0, sizeof(n));
  n.max_doublings = max_num_doublings;
  n.criterion = stopping_criterion;
  n.u_dir = io->u_direction;
  n.u_merge = io->u_merge;
  n.u_step = io->u_step;
  n.energy = io->energy;
  n.acceptance_rate = io->acceptance_rate;
  n.is_divergent = io->is_divergent;
  n.is_turning = io->is_turning;
  n.num_expansions = io->num_trajectory_expansions;
  n.num_steps = io->num_integration_steps;
  n.left_position = io->leftmost_position;
  n.right_position = io->rightmost_position;
  hipError_t e = gjx::launch_funnel_nuts_lmc(a, n, params->dtype, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_lmcmonge_funnel_nuts launch");
  return GJX_OK;
}

int gjx_lmcmonge_funnel_alpha2_tangent(const gjx_funnel_params* params, const gjx_lmc_alpha2_tangent_io* io,
                                       void* stream) {
  int lpc = 0, epl = 0;
  int rc = check_funnel_params(params, false, &lpc, &epl);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(params->dtype == GJX_F64, "the alpha2 tangent runs in float64");
  GJX_REQUIRE((params->flags & ~3) == 0, "unknown flags 0x%x", params->flags);
  GJX_REQUIRE(params->alpha2 > 0.0, "alpha2 must be > 0 for its derivative (got %g)", params->alpha2);
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->inverse_mass_matrix && io->position, "inverse_mass_matrix and position are required");
  GJX_REQUIRE(io->d_proposal_position_d_alpha2, "d_proposal_position_d_alpha2 is required");
  gjx::FunnelLmcArgs a = lmc_args(params);
  a.T = 1;
  a.inv_mass = io->inverse_mass_matrix;
  a.q_in = io->position;
  a.z = io->z;
  a.Lc = io->num_integration_steps;
  a.prop_q = io->proposal_position;
  a.d_prop_q = io->d_proposal_position_d_alpha2;
  hipError_t e = gjx::launch_funnel_lmc_alpha2_tangent(a, lpc, epl, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_lmcmonge_funnel_alpha2_tangent launch");
  return GJX_OK;
}

int gjx_lmcmonge_funnel_sample(const gjx_funnel_params* params, const gjx_lmc_sample_io* io, void* stream) {
  int lpc = 0, epl = 0;
  int rc = check_funnel_params(params, true, &lpc, &epl);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE((params->flags & ~3) == 0, "unknown flags 0x%x", params->flags);
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->inverse_mass_matrix, "inverse_mass_matrix is required");
  GJX_REQUIRE(io->position && io->logdensity && io->volume_adjustment,
              "position, logdensity and volume_adjustment are required");
  GJX_REQUIRE(io->final_position && io->final_logdensity, "final_position and final_logdensity are required");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  gjx::FunnelLmcArgs a = lmc_args(params);
  a.T = params->num_transitions;
  a.inv_mass = io->inverse_mass_matrix;
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.J_in = io->volume_adjustment;
  a.Lc = io->num_integration_steps;
  a.z = io->z;
  a.u = io->u;
  a.final_q = io->final_position;
  a.final_logp = io->final_logdensity;
  a.final_g = io->final_logdensity_grad;
  a.final_J = io->final_volume_adjustment;
  a.samples = io->samples;
  a.sample_acc = io->acceptance_rate;
  a.num_accepted = io->num_accepted;
  hipError_t e = gjx::launch_funnel_lmc(a, params->dtype, lpc, epl, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_lmcmonge_funnel_sample launch");
  return GJX_OK;
}

namespace {

void workspace_bytes(int C, int max_rows, uint64_t* bytes) {
  *bytes = (uint64_t)C * 5 * (uint64_t)((max_rows + 63) / 64 * 64) * sizeof(float);
}

int check_logistic(const gjx_logistic_params* p, const gjx_logistic_data* d, bool sample, gjx::LogisticArgs* a,
                   int max_d) {
  GJX_REQUIRE(p != nullptr && d != nullptr, "params and data are required");
  GJX_REQUIRE(p->num_chains > 0, "num_chains must be > 0 (got %d)", p->num_chains);
  GJX_REQUIRE(p->dtype == GJX_F64 || p->dtype == GJX_F32, "dtype must be GJX_F64 or GJX_F32 (got %d)", p->dtype);
  GJX_REQUIRE(p->num_integration_steps >= 0, "num_integration_steps must be >= 0");
  GJX_REQUIRE(p->max_iters >= 0, "max_iters must be >= 0");
  GJX_REQUIRE(std::isfinite(p->step_size), "step_size must be finite");
  GJX_REQUIRE(std::isfinite(p->prior_var) && p->prior_var > 0.0, "prior_var must be finite and > 0");
  if (sample) GJX_REQUIRE(p->num_transitions >= 1, "num_transitions must be >= 1");
  GJX_REQUIRE(d->num_posteriors >= 1, "num_posteriors must be >= 1");
  GJX_REQUIRE(d->max_rows >= 1, "max_rows must be >= 1");
  GJX_REQUIRE(d->max_features >= 1 && d->max_features <= p->dim,
              "max_features must be in [1, dim] (got %d, dim %d)", d->max_features, p->dim);
  if (d->max_features > max_d)
    return fail(GJX_EUNSUPPORTED, "D = %d exceeds this kernel's maximum of %d features", d->max_features, max_d);
  GJX_REQUIRE(d->num_rows && d->num_features && d->x_offset && d->y_offset && d->x && d->y,
              "data arrays are required");
  std::memset(a, 0, sizeof(*a));
  a->C = p->num_chains;
  a->L = p->num_integration_steps;
  a->T = sample ? p->num_transitions : 1;
  a->max_iters = p->max_iters;
  a->Dmax = p->dim;
  a->P = d->num_posteriors;
  a->max_features = d->max_features;
  {
    const int v = p->waves_per_chain;
    GJX_REQUIRE(v >= 0 && v <= 0xFF, "waves_per_chain must be in [0, 0xFF] (got %d)", v);
    const int w1 = v & 0xF, w2 = (v >> 4) & 0xF;
    GJX_REQUIRE(w1 <= 2 && w2 <= 2, "waves_per_chain: each 4-bit field must be 0, 1 or 2 (got 0x%x)", v);
    a->waves[0] = w1 ? w1 : gjx::kDefaultWavesPerChain[0];
    a->waves[1] = w2 ? w2 : gjx::kDefaultWavesPerChain[1];
  }
  a->eps = p->step_size;
  a->prior_var = p->prior_var;
  a->conv_tol = p->convergence_tol;
  a->div_tol = p->divergence_tol;
  a->div_threshold = p->divergence_threshold;
  a->seed = p->seed;
  a->counter = p->counter;
  a->chain_post = d->chain_posterior;
  a->dispatch_order = d->dispatch_order;
  a->post_N = d->num_rows;
  a->post_D = d->num_features;
  a->post_xoff = d->x_offset;
  a->post_yoff = d->y_offset;
  a->X = static_cast<const double*>(d->x);
  a->y = static_cast<const double*>(d->y);
  if (p->dtype == GJX_F32) {
    GJX_REQUIRE(d->packed != nullptr, "the float32 path needs data->packed (gjx_logistic_pack)");
    uint64_t need = 0;
    workspace_bytes(p->num_chains, d->max_rows, &need);
    GJX_REQUIRE(d->workspace != nullptr && d->workspace_bytes >= need,
                "the float32 path needs a workspace of %llu bytes (got %llu)", (unsigned long long)need,
                (unsigned long long)d->workspace_bytes);
    a->packed_hdr = static_cast<const gjx::PackedPost*>(d->packed);
    a->packed = static_cast<const float*>(d->packed);
    a->work = static_cast<float*>(d->workspace);
    a->work_rows = (d->max_rows + 63) / 64 * 64;
  }
  return GJX_OK;
}

// both RMHMC kernels (float64 and float32) serve D <= 128
int rmhmc_max_d(const gjx_logistic_params*) { return gjx::kLogisticMaxD; }

// the RMHMC and LMC entry points have a float32 path; these do not
int require_f64(const gjx_logistic_params* p, const char* what) {
  if (p != nullptr && p->dtype == GJX_F32) return fail(GJX_EUNSUPPORTED, "%s has no float32 path (use GJX_F64)", what);
  return GJX_OK;
}

}  // namespace

namespace {

int rmhmc_logistic_step(const gjx_logistic_params* params, const gjx_logistic_data* data, const gjx_rmhmc_step_io* io,
                        void* stream, bool concurrent) {
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, false, &a, rmhmc_max_d(params));
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  GJX_REQUIRE(!io->new_logdensity_grad || io->logdensity_grad,
              "new_logdensity_grad needs logdensity_grad (copied through on rejection)");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.g_in = io->logdensity_grad;
  a.z = io->z;
  a.u = io->u;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.momentum = io->momentum;
  a.acc_rate = io->acceptance_rate;
  a.accepted = io->is_accepted;
  a.divergent = io->is_divergent;
  a.energy = io->energy;
  a.prop_q = io->proposal_position;
  a.prop_p = io->proposal_momentum;
  a.prop_v = io->proposal_velocity;
  a.prop_logp = io->proposal_logdensity;
  a.prop_g = io->proposal_logdensity_grad;
  a.prop_weight = io->proposal_weight;
  a.prop_slpa = io->proposal_sum_log_p_accept;
  a.fp_iters = io->fp_iters;
  a.fp_norm = io->fp_norm;
  a.fp_success = io->fp_success;
  a.Lc = io->num_integration_steps;
  hipError_t e = params->dtype == GJX_F32
                     ? gjx::launch_logistic_rmhmc_f32(a, data->max_features, (hipStream_t)stream, concurrent)
                     : gjx::launch_logistic_rmhmc_f64(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_fisher_logistic_step launch");
  return GJX_OK;
}

// Scratch of gjx_rmhmc_fisher_logistic_step_noalloc: a file-scope device array,
// one copy per GPU, alive until the process ends.  Host code never uses the
// array's name as an address (on the host it is only a stand-in for the
// symbol): its device address is looked up once per device and cached.
constexpr uint64_t kNoAllocScratchBytes = GJX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) float scratch[kNoAllocScratchBytes / sizeof(float)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices];

// Device address of `scratch` on the device that runs `stream` (which must be
// the current device: the symbol is resolved in the current device's context).
int scratch_address(void* stream, void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(GJX_EUNSUPPORTED, "device %d is beyond the %d supported", dev,
                                                kMaxDevices);
  if (stream != nullptr) {
    hipDevice_t sdev = -1;
    e = hipStreamGetDevice((hipStream_t)stream, &sdev);
    if (e != hipSuccess) return hip_fail(e, "hipStreamGetDevice");
    GJX_REQUIRE(sdev == dev, "the stream belongs to device %d; call with that device current (current: %d)", sdev,
                dev);
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (p == nullptr) return fail(GJX_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return GJX_OK;
}

}  // namespace

int gjx_rmhmc_fisher_logistic_step(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   const gjx_rmhmc_step_io* io, void* stream) {
  return rmhmc_logistic_step(params, data, io, stream, true);
}

namespace {

int check_diagnostics(const gjx_diagnostics_args* a) {
  GJX_REQUIRE(a != nullptr, "args is NULL");
  GJX_REQUIRE(a->num_samples >= 2, "num_samples must be >= 2 (got %d)", a->num_samples);
  GJX_REQUIRE(a->num_chains >= 2, "the multi-chain estimators need num_chains >= 2 (got %d)", a->num_chains);
  GJX_REQUIRE(a->num_coords >= 1, "num_coords must be >= 1 (got %d)", a->num_coords);
  GJX_REQUIRE(a->dtype == GJX_F64 || a->dtype == GJX_F32, "dtype must be GJX_F64 or GJX_F32 (got %d)", a->dtype);
  if (a->num_samples > gjx::diagnostics_max_samples())
    return fail(GJX_EUNSUPPORTED, "num_samples %d exceeds the %d a chain's series can hold in LDS", a->num_samples,
                gjx::diagnostics_max_samples());
  return GJX_OK;
}

}  // namespace

int gjx_diagnostics_workspace_size(const gjx_diagnostics_args* args, uint64_t* bytes) {
  int rc = check_diagnostics(args);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(bytes != nullptr, "bytes is NULL");
  *bytes = gjx::diagnostics_workspace_bytes(args->num_samples, args->num_chains, args->num_coords);
  return GJX_OK;
}

int gjx_effective_sample_size(const gjx_diagnostics_args* args, const void* samples, void* workspace,
                              uint64_t workspace_bytes, double* ess, double* rhat, void* stream) {
  int rc = check_diagnostics(args);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(samples != nullptr, "samples is NULL");
  GJX_REQUIRE(ess != nullptr || rhat != nullptr, "nothing to compute: ess and rhat are both NULL");
  const uint64_t need = gjx::diagnostics_workspace_bytes(args->num_samples, args->num_chains, args->num_coords);
  GJX_REQUIRE(workspace != nullptr && workspace_bytes >= need, "the workspace needs %llu bytes (got %llu)",
              (unsigned long long)need, (unsigned long long)workspace_bytes);
  hipError_t e = gjx::launch_diagnostics(samples, args->dtype, args->num_samples, args->num_chains, args->num_coords,
                                         args->stride_sample, args->stride_chain, args->stride_coord, workspace, ess,
                                         rhat, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_effective_sample_size launch");
  return GJX_OK;
}

int gjx_rmhmc_fisher_logistic_step_noalloc(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                           const gjx_rmhmc_step_io* io, void* stream) {
  GJX_REQUIRE(params != nullptr && data != nullptr, "params and data are required");
  if (params->dtype != GJX_F32)  // the float64 kernel uses no scratch
    return rmhmc_logistic_step(params, data, io, stream, false);
  GJX_REQUIRE(params->num_chains > 0 && data->max_rows >= 1, "num_chains and max_rows must be >= 1");
  uint64_t need = 0;
  workspace_bytes(params->num_chains, data->max_rows, &need);
  if (need > kNoAllocScratchBytes)
    return fail(GJX_EUNSUPPORTED,
                "%d chains x %d rows need %llu bytes of scratch; the allocation-free entry point holds %llu "
                "(use gjx_rmhmc_fisher_logistic_step with a workspace)",
                params->num_chains, data->max_rows, (unsigned long long)need,
                (unsigned long long)kNoAllocScratchBytes);
  void* ws = nullptr;
  int rc = scratch_address(stream, &ws);
  if (rc != GJX_OK) return rc;
  gjx_logistic_data d = *data;
  d.workspace = ws;
  d.workspace_bytes = kNoAllocScratchBytes;
  return rmhmc_logistic_step(params, &d, io, stream, false);
}

int gjx_rmhmc_fisher_logistic_nuts(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   int32_t max_num_doublings, int32_t stopping_criterion, const gjx_nuts_io* io,
                                   void* stream) {
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, false, &a, 64);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(max_num_doublings >= 1 && max_num_doublings <= 12, "max_num_doublings must be in [1, 12] (got %d)",
              max_num_doublings);
  GJX_REQUIRE(stopping_criterion == GJX_UTURN_EUC || stopping_criterion == GJX_UTURN_BET ||
                  stopping_criterion == GJX_UTURN_RIEM,
              "stopping_criterion must be GJX_UTURN_EUC, GJX_UTURN_BET or GJX_UTURN_RIEM (got %d)",
              stopping_criterion);
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  const bool host = io->z != nullptr;
  GJX_REQUIRE((io->u_direction != nullptr) == host && (io->u_merge != nullptr) == host &&
                  (io->u_step != nullptr) == host,
              "z, u_direction, u_merge and u_step must all be given or all be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.momentum = io->momentum;
  gjx::NutsArgs n;
  std::memset(&n, 0, sizeof(n));
  n.max_doublings = max_num_doublings;
  n.criterion = stopping_criterion;
  n.u_dir = io->u_direction;
  n.u_merge = io->u_merge;
  n.u_step = io->u_step;
  n.energy = io->energy;
  n.acceptance_rate = io->acceptance_rate;
  n.is_divergent = io->is_divergent;
  n.is_turning = io->is_turning;
  n.num_expansions = io->num_trajectory_expansions;
  n.num_steps = io->num_integration_steps;
  n.left_position = io->leftmost_position;
  n.right_position = io->rightmost_position;
  hipError_t e = params->dtype == GJX_F32
                     ? gjx::launch_logistic_nuts_f32(a, n, data->max_features, (hipStream_t)stream)
                     : gjx::launch_logistic_nuts_f64(a, n, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_fisher_logistic_nuts launch");
  return GJX_OK;
}

int gjx_lmc_fisher_logistic_nuts(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                 int32_t max_num_doublings, int32_t stopping_criterion, const gjx_lmc_nuts_io* io,
                                 void* stream) {
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, false, &a, 64);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(max_num_doublings >= 1 && max_num_doublings <= 12, "max_num_doublings must be in [1, 12] (got %d)",
              max_num_doublings);
  GJX_REQUIRE(stopping_criterion == GJX_UTURN_EUC || stopping_criterion == GJX_UTURN_BET ||
                  stopping_criterion == GJX_UTURN_RIEM,
              "stopping_criterion must be GJX_UTURN_EUC, GJX_UTURN_BET or GJX_UTURN_RIEM (got %d)",
              stopping_criterion);
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity && io->volume_adjustment,
              "position, logdensity and volume_adjustment are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  const bool host = io->z != nullptr;
  GJX_REQUIRE((io->u_direction != nullptr) == host && (io->u_merge != nullptr) == host &&
                  (io->u_step != nullptr) == host,
              "z, u_direction, u_merge and u_step must all be given or all be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.J_in = io->volume_adjustment;
  a.z = io->z;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.J_out = io->new_volume_adjustment;
  a.momentum = io->velocity;  // NUTSInfo.velocity: the v0 drawn
  gjx::NutsArgs n;
  std::memset(&n, 0, sizeof(n));
  n.max_doublings = max_num_doublings;
  n.criterion = stopping_criterion;
  n.u_dir = io->u_direction;
  n.u_merge = io->u_merge;
  n.u_step = io->u_step;
  n.energy = io->energy;
  n.acceptance_rate = io->acceptance_rate;
  n.is_divergent = io->is_divergent;
  n.is_turning = io->is_turning;
  n.num_expansions = io->num_trajectory_expansions;
  n.num_steps = io->num_integration_steps;
  n.left_position = io->leftmost_position;
  n.right_position = io->rightmost_position;
  hipError_t e = params->dtype == GJX_F32
                     ? gjx::launch_logistic_nuts_lmc_f32(a, n, data->max_features, (hipStream_t)stream)
                     : gjx::launch_logistic_nuts_lmc_f64(a, n, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_lmc_fisher_logistic_nuts launch");
  return GJX_OK;
}

int gjx_rmhmc_fisher_logistic_sample(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                     const gjx_rmhmc_sample_io* io, void* stream) {
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, true, &a, rmhmc_max_d(params));
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->final_position && io->final_logdensity, "final_position and final_logdensity are required");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  a.u = io->u;
  a.final_q = io->final_position;
  a.final_logp = io->final_logdensity;
  a.final_g = io->final_logdensity_grad;
  a.samples = io->samples;
  a.sample_acc = io->acceptance_rate;
  a.num_accepted = io->num_accepted;
  a.num_fevals = io->num_fevals;
  a.Lc = io->num_integration_steps;
  hipError_t e = params->dtype == GJX_F32 ? gjx::launch_logistic_rmhmc_f32(a, data->max_features, (hipStream_t)stream)
                                           : gjx::launch_logistic_rmhmc_f64(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_fisher_logistic_sample launch");
  return GJX_OK;
}

int gjx_lmc_fisher_logistic_step(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                 const gjx_lmcmc_step_io* io, void* stream) {
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, false, &a, gjx::kLogisticMaxDLmc);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity && io->volume_adjustment,
              "position, logdensity and volume_adjustment are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  GJX_REQUIRE(!io->new_logdensity_grad || io->logdensity_grad,
              "new_logdensity_grad needs logdensity_grad (copied through on rejection)");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.g_in = io->logdensity_grad;
  a.J_in = io->volume_adjustment;
  a.Lc = io->num_integration_steps;
  a.z = io->z;
  a.u = io->u;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.J_out = io->new_volume_adjustment;
  a.momentum = io->velocity;
  a.acc_rate = io->acceptance_rate;
  a.accepted = io->is_accepted;
  a.divergent = io->is_divergent;
  a.energy = io->energy;
  a.prop_q = io->proposal_position;
  a.prop_p = io->proposal_momentum;
  a.prop_v = io->proposal_velocity;
  a.prop_logp = io->proposal_logdensity;
  a.prop_g = io->proposal_logdensity_grad;
  a.prop_J = io->proposal_volume_adjustment;
  a.prop_weight = io->proposal_weight;
  a.prop_slpa = io->proposal_sum_log_p_accept;
  hipError_t e = params->dtype == GJX_F32 ? gjx::launch_logistic_lmc_f32(a, data->max_features, (hipStream_t)stream)
                                           : gjx::launch_logistic_lmc_f64(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_lmc_fisher_logistic_step launch");
  return GJX_OK;
}

int gjx_lmc_fisher_logistic_sample(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   const gjx_lmcmc_sample_io* io, void* stream) {
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, true, &a, gjx::kLogisticMaxDLmc);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity && io->volume_adjustment,
              "position, logdensity and volume_adjustment are required");
  GJX_REQUIRE(io->final_position && io->final_logdensity, "final_position and final_logdensity are required");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.J_in = io->volume_adjustment;
  a.Lc = io->num_integration_steps;
  a.z = io->z;
  a.u = io->u;
  a.final_q = io->final_position;
  a.final_logp = io->final_logdensity;
  a.final_g = io->final_logdensity_grad;
  a.final_J = io->final_volume_adjustment;
  a.samples = io->samples;
  a.sample_acc = io->acceptance_rate;
  a.num_accepted = io->num_accepted;
  hipError_t e = params->dtype == GJX_F32 ? gjx::launch_logistic_lmc_f32(a, data->max_features, (hipStream_t)stream)
                                           : gjx::launch_logistic_lmc_f64(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_lmc_fisher_logistic_sample launch");
  return GJX_OK;
}

static int dense_args(const gjx_dense_metric_io* io, gjx::DenseMetricArgs* a, const char* fn) {
  GJX_REQUIRE(io != nullptr, "%s: io is NULL", fn);
  GJX_REQUIRE(io->num_chains > 0, "%s: num_chains must be > 0 (got %d)", fn, io->num_chains);
  GJX_REQUIRE(io->dim >= 1, "%s: dim must be >= 1 (got %d)", fn, io->dim);
  if (io->dim > gjx::kDenseMaxD)
    return fail(GJX_EUNSUPPORTED, "%s: dim %d exceeds the dense-metric kernels' maximum of %d", fn, io->dim,
                gjx::kDenseMaxD);
  GJX_REQUIRE(io->metric, "%s: metric is required", fn);
  std::memset(a, 0, sizeof(*a));
  a->C = io->num_chains;
  a->D = io->dim;
  a->step_size = io->step_size;
  a->metric = io->metric;
  a->metric_jac = io->metric_jacobian;
  a->z = io->z;
  a->velocity = io->velocity;
  a->logdensity_grad = io->logdensity_grad;
  a->grad_logdet_in = io->grad_logdet_in;
  a->velocity_out = io->velocity_out;
  a->logdet = io->logdet;
  a->grad_logdet = io->grad_logdet;
  a->vgv = io->vgv;
  a->volume_adjustment = io->volume_adjustment;
  a->gamma = io->gamma;
  a->drift = io->drift;
  a->chol_inv_z = io->chol_inv_z;
  a->momentum = io->momentum;
  a->momentum_out = io->momentum_out;
  a->dtdq = io->dtdq;
  a->pu = io->pu;
  a->step_sizes = io->step_sizes;
  return GJX_OK;
}

int gjx_dense_metric_position(const gjx_dense_metric_io* io, void* stream) {
  gjx::DenseMetricArgs a;
  int rc = dense_args(io, &a, "gjx_dense_metric_position");
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io->metric_jacobian, "metric_jacobian is required");
  GJX_REQUIRE(!io->velocity_out || io->z, "velocity_out (L^-T z) needs z");
  GJX_REQUIRE(!io->vgv || io->velocity, "vgv needs velocity");
  hipError_t e = gjx::launch_dense_metric_position(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_dense_metric_position launch");
  return GJX_OK;
}

int gjx_dense_lmc_half_step(const gjx_dense_metric_io* io, void* stream) {
  gjx::DenseMetricArgs a;
  int rc = dense_args(io, &a, "gjx_dense_lmc_half_step");
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io->metric_jacobian, "metric_jacobian is required");
  GJX_REQUIRE(std::isfinite(io->step_size), "step_size must be finite");
  GJX_REQUIRE(io->velocity && io->logdensity_grad && io->grad_logdet_in && io->velocity_out &&
                  io->volume_adjustment,
              "velocity, logdensity_grad, grad_logdet_in, velocity_out and volume_adjustment are required");
  hipError_t e = gjx::launch_dense_lmc_half_step(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_dense_lmc_half_step launch");
  return GJX_OK;
}

int gjx_dense_mmala_drift(const gjx_dense_metric_io* io, void* stream) {
  gjx::DenseMetricArgs a;
  int rc = dense_args(io, &a, "gjx_dense_mmala_drift");
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(!io->gamma || io->metric_jacobian, "gamma needs metric_jacobian");
  GJX_REQUIRE(!io->drift || io->logdensity_grad, "drift needs logdensity_grad");
  GJX_REQUIRE(!io->chol_inv_z || io->z, "chol_inv_z needs z");
  hipError_t e = gjx::launch_dense_mmala_drift(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_dense_mmala_drift launch");
  return GJX_OK;
}

int gjx_dense_rmhmc_field(const gjx_dense_metric_io* io, void* stream) {
  gjx::DenseMetricArgs a;
  int rc = dense_args(io, &a, "gjx_dense_rmhmc_field");
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(io->momentum || io->z, "momentum or z is required");
  GJX_REQUIRE(!io->dtdq || io->metric_jacobian, "dtdq needs metric_jacobian");
  hipError_t e = gjx::launch_dense_rmhmc_field(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_dense_rmhmc_field launch");
  return GJX_OK;
}

int gjx_mmala_fisher_logistic_init(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   const void* position, double* logdensity, void* logdensity_grad, double* metric,
                                   double* gamma, void* stream) {
  if (int rc0 = require_f64(params, "gjx_mmala_fisher_logistic_init")) return rc0;
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, false, &a, gjx::kLogisticMaxDLmc);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(position && logdensity && logdensity_grad, "position, logdensity and logdensity_grad are required");
  a.q_in = position;
  a.logp_out = logdensity;
  a.g_out = logdensity_grad;
  a.metric_out = metric;
  a.gamma_out = gamma;
  hipError_t e = gjx::launch_logistic_mmala_init(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_mmala_fisher_logistic_init launch");
  return GJX_OK;
}

int gjx_mmala_fisher_logistic_step(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   int reference_faithful, const gjx_mmala_step_io* io, void* stream) {
  if (int rc0 = require_f64(params, "gjx_mmala_fisher_logistic_step")) return rc0;
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, false, &a, gjx::kLogisticMaxDLmc);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(params->step_size > 0.0, "step_size must be > 0 (the proposal covariance is 2 step_size G^-1)");
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->new_position && io->new_logdensity, "new_position and new_logdensity are required");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.g_in = io->logdensity_grad;
  a.z = io->z;
  a.u = io->u;
  a.q_out = io->new_position;
  a.logp_out = io->new_logdensity;
  a.g_out = io->new_logdensity_grad;
  a.metric_out = io->new_metric;
  a.gamma_out = io->new_gamma;
  a.acc_rate = io->acceptance_rate;
  a.accepted = io->is_accepted;
  a.divergent = io->is_divergent;
  a.prop_weight = io->weight;
  a.prop_q = io->proposal_position;
  a.prop_logp = io->proposal_logdensity;
  a.prop_g = io->proposal_logdensity_grad;
  a.prop_metric = io->proposal_metric;
  a.prop_gamma = io->proposal_gamma;
  a.mmala_exact = reference_faithful ? 0 : 1;
  hipError_t e = gjx::launch_logistic_mmala_f64(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_mmala_fisher_logistic_step launch");
  return GJX_OK;
}

int gjx_mmala_fisher_logistic_sample(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                     int reference_faithful, const gjx_mmala_sample_io* io, void* stream) {
  if (int rc0 = require_f64(params, "gjx_mmala_fisher_logistic_sample")) return rc0;
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, true, &a, gjx::kLogisticMaxDLmc);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(params->step_size > 0.0, "step_size must be > 0 (the proposal covariance is 2 step_size G^-1)");
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->final_position && io->final_logdensity, "final_position and final_logdensity are required");
  GJX_REQUIRE((io->z == nullptr) == (io->u == nullptr), "z and u must both be given or both be NULL");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  a.u = io->u;
  a.final_q = io->final_position;
  a.final_logp = io->final_logdensity;
  a.final_g = io->final_logdensity_grad;
  a.metric_out = io->final_metric;
  a.gamma_out = io->final_gamma;
  a.samples = io->samples;
  a.sample_acc = io->acceptance_rate;
  a.num_accepted = io->num_accepted;
  a.mmala_exact = reference_faithful ? 0 : 1;
  hipError_t e = gjx::launch_logistic_mmala_f64(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_mmala_fisher_logistic_sample launch");
  return GJX_OK;
}

int gjx_logistic_value_and_grad(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                const void* position, double* logdensity, void* logdensity_grad, void* stream) {
  if (int rc0 = require_f64(params, "gjx_logistic_value_and_grad")) return rc0;
  gjx::LogisticArgs a;
  int rc = check_logistic(params, data, false, &a, gjx::kLogisticMaxD);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(position && logdensity && logdensity_grad, "position, logdensity and logdensity_grad are required");
  a.q_in = position;
  a.logp_out = logdensity;
  a.g_out = logdensity_grad;
  hipError_t e = gjx::launch_logistic_value_and_grad(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "gjx_logistic_value_and_grad launch");
  return GJX_OK;
}

int gjx_rmhmc_fisher_logistic_tangent(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                      const gjx_tangent_io* io, void* stream) {
  gjx::LogisticArgs a;
  // float64: D <= 64; float32: D <= 64 on the tuned one-wave kernel, 65..128 on the wide one
  int rc = check_logistic(params, data, false, &a,
                          params != nullptr && params->dtype == GJX_F32 ? gjx::kLogisticMaxD : gjx::kLogisticMaxDLmc);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(params->step_size > 0.0, "step_size must be > 0 (log step_size is the parameter)");
  GJX_REQUIRE(io != nullptr, "io is NULL");
  GJX_REQUIRE(io->position && io->logdensity, "position and logdensity are required");
  GJX_REQUIRE(io->objective && io->acceptance && io->sq_jump && io->d_objective && io->d_acceptance &&
                  io->d_sq_jump,
              "every output except num_fevals is required");
  a.q_in = io->position;
  a.logp_in = io->logdensity;
  a.z = io->z;
  gjx::TangentOut o{io->objective, io->acceptance, io->sq_jump, io->d_objective, io->d_acceptance,
                    io->d_sq_jump, io->num_fevals};
  hipError_t e = hipSuccess;
  if (params->dtype == GJX_F32) {
    if (data->max_features <= 64 || data->num_posteriors > 1)
      e = gjx::launch_logistic_rmhmc_tangent_f32(a, o, data->max_features, (hipStream_t)stream);
    if (e == hipSuccess && data->max_features > 64)
      e = gjx::launch_logistic_rmhmc_tangent_f32_wide(a, o, (hipStream_t)stream);
  } else {
    e = gjx::launch_logistic_rmhmc_tangent(a, o, (hipStream_t)stream);
  }
  if (e != hipSuccess) return hip_fail(e, "gjx_rmhmc_fisher_logistic_tangent launch");
  return GJX_OK;
}

int gjx_logistic_pack_size(int32_t num_posteriors, const int32_t* num_rows_host, const int32_t* num_features_host,
                           uint64_t* bytes) {
  GJX_REQUIRE(num_posteriors >= 1, "num_posteriors must be >= 1");
  GJX_REQUIRE(num_rows_host && num_features_host && bytes, "num_rows_host, num_features_host and bytes are required");
  for (int p = 0; p < num_posteriors; ++p) {
    GJX_REQUIRE(num_rows_host[p] >= 1, "posterior %d has %d rows", p, num_rows_host[p]);
    GJX_REQUIRE(num_features_host[p] >= 1, "posterior %d has %d features", p, num_features_host[p]);
    if (num_features_host[p] > gjx::kLogisticMaxD)
      return fail(GJX_EUNSUPPORTED, "posterior %d has D = %d > %d features", p, num_features_host[p],
                  gjx::kLogisticMaxD);
  }
  *bytes = gjx::logistic_pack_layout(num_posteriors, num_rows_host, num_features_host, nullptr);
  return GJX_OK;
}

int gjx_logistic_pack(const gjx_logistic_data* data, const int32_t* num_rows_host, const int32_t* num_features_host,
                      void* packed, uint64_t packed_bytes, void* stream) {
  GJX_REQUIRE(data != nullptr && packed != nullptr, "data and packed are required");
  GJX_REQUIRE(data->num_rows && data->num_features && data->x_offset && data->y_offset && data->x && data->y,
              "data arrays are required");
  uint64_t need = 0;
  int rc = gjx_logistic_pack_size(data->num_posteriors, num_rows_host, num_features_host, &need);
  if (rc != GJX_OK) return rc;
  GJX_REQUIRE(packed_bytes >= need, "packed buffer of %llu bytes is smaller than the %llu needed",
              (unsigned long long)packed_bytes, (unsigned long long)need);
  const int P = data->num_posteriors;
  std::vector<gjx::PackedPost> hdr(P);
  gjx::logistic_pack_layout(P, num_rows_host, num_features_host, hdr.data());
  int max_npad = 0;
  for (int p = 0; p < P; ++p) max_npad = hdr[p].npad > max_npad ? hdr[p].npad : max_npad;  // grid covers npad x 64
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(packed, hdr.data(), sizeof(gjx::PackedPost) * P, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "gjx_logistic_pack header copy");
  e = gjx::launch_logistic_pack(static_cast<const double*>(data->x), static_cast<const double*>(data->y),
                                data->num_rows, data->num_features, data->x_offset, data->y_offset,
                                static_cast<const gjx::PackedPost*>(packed), static_cast<float*>(packed), P, max_npad, s);
  if (e != hipSuccess) return hip_fail(e, "gjx_logistic_pack launch");
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "gjx_logistic_pack");
  return GJX_OK;
}

int gjx_logistic_workspace_size(const gjx_logistic_params* params, const gjx_logistic_data* data, uint64_t* bytes) {
  GJX_REQUIRE(params && data && bytes, "params, data and bytes are required");
  GJX_REQUIRE(params->num_chains > 0 && data->max_rows >= 1, "num_chains and max_rows must be >= 1");
  workspace_bytes(params->num_chains, data->max_rows, bytes);
  return GJX_OK;
}

}  // extern "C"
