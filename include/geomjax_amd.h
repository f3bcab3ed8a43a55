/*
 * geomjax_amd C ABI: the Riemannian integrator hot path of geomjax as
 * hand-written gfx950 HIP kernels.
 *
 * Conventions
 *  - Every buffer is device memory owned by the caller.  The library never
 *    allocates on these entry points and never frees caller memory.
 *  - Work is enqueued on `stream` (a hipStream_t passed as void*; NULL is
 *    the null stream).  No entry point synchronises the device.
 *  - Chain-major layouts: a [C, D] array holds chain c's vector at c*D.
 *  - "state" arrays (positions, momenta, velocities, gradients, samples)
 *    have the element type given by `dtype` (GJX_F64 or GJX_F32); every
 *    per-chain scalar (log densities, energies, weights, acceptance
 *    probabilities, uniforms) is float64 in both precisions.
 *  - Return value: 0 on success, a negative GJX_E* code otherwise; the text
 *    of the last error on the calling thread is gjx_last_error().
 *  - Numerical failure never raises, as in the reference: a NaN energy
 *    difference becomes -inf and the proposal is rejected
 *    (geomjax/mcmc/proposal.py:107-108).
 *  - Re-entrant; calls on different streams may overlap.  hipGraph capture of
 *    these calls is legal (they only launch kernels).
 */
#ifndef GEOMJAX_AMD_H
#define GEOMJAX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gjx_dtype { GJX_F64 = 0, GJX_F32 = 1 };

/* Bits of gjx_funnel_params.flags for the LMC-Monge entry points.  0 reproduces
 * the reference exactly, including two algebra slips:
 *  - half_step_omega's Sherman-Morrison correction lacks alpha2
 *    (geomjax/lmcmonge/integrators.py:186-189 vs :178, :192);
 *  - the momentum's rank-one term carries an extra factor L
 *    (geomjax/lmcmonge/lmc.py:202-204, metrics.py:243-255). */
enum gjx_lmc_flags { GJX_LMC_EXACT_SHERMAN_MORRISON = 1, GJX_LMC_EXACT_MOMENTUM = 2 };

enum gjx_status {
  GJX_OK = 0,
  GJX_EINVAL = -1,      /* invalid argument (shape, pointer, parameter) */
  GJX_EUNSUPPORTED = -2,/* size outside the compiled tile configurations */
  GJX_EHIP = -3         /* a HIP runtime call failed */
};

/* Sizes and scalars of one RMHMC call on Neal's funnel with the Monge metric
 * G(q) = I + alpha2 * grad l(q) grad l(q)^T.
 * Mirrors the arguments of the reference kernel
 *   kernel(rng_key, state, logdensity_fn, step_size, metric_fn,
 *          num_integration_steps)            geomjax/rmhmc/rmhmc.py:131-138
 * with the fixed-point solver's keywords            (integrators.py:57-59)
 * and build_kernel's divergence_threshold            (rmhmc.py:110-113). */
typedef struct {
  int32_t num_chains;            /* C */
  int32_t dim;                   /* D >= 2; coordinate D-1 is the log-scale v */
  int32_t dtype;                 /* enum gjx_dtype of the state arrays */
  int32_t num_integration_steps; /* L >= 0 */
  int32_t max_iters;             /* fixed-point iteration cap (>= 0) */
  int32_t num_transitions;       /* T >= 1 (sample entry points) */
  int32_t lanes_per_chain;       /* 0 = library default; else 4,8,16,32 or 64 */
  int32_t flags;                 /* LMC-Monge only: GJX_LMC_EXACT_* bits; 0 = as the reference */
  double step_size;              /* epsilon */
  double alpha2;                 /* Monge warp alpha^2 >= 0 */
  double sigma;                  /* funnel scale of v (examples/funnel/main.py:29) */
  double convergence_tol;        /* 1e-6 in the reference */
  double divergence_tol;         /* 1e10 in the reference */
  double divergence_threshold;   /* 1000 in the reference */
  uint64_t seed;                 /* device PRNG key (when z == NULL) */
  uint64_t counter;              /* device PRNG transition offset */
} gjx_funnel_params;

/* Buffers of one RMHMC transition (gjx_rmhmc_monge_funnel_step).
 * Inputs are RMHMCState (geomjax/rmhmc/rmhmc.py:30-41); outputs are the new
 * RMHMCState and RMHMCInfo (rmhmc.py:58-93) with its Proposal
 * (geomjax/mcmc/proposal.py:22-40) over IntegratorState
 * (geomjax/rmhmc/integrators.py:25-36).  Optional outputs may be NULL. */
typedef struct {
  const void* position;          /* [C, D] dtype */
  const double* logdensity;      /* [C] */
  const void* logdensity_grad;   /* [C, D] dtype */
  const void* z;                 /* [C, D] dtype standard normals, or NULL: device PRNG */
  const double* u;               /* [C] uniforms in [0,1), or NULL: device PRNG */
  void* new_position;            /* [C, D] dtype */
  double* new_logdensity;        /* [C] */
  void* new_logdensity_grad;     /* [C, D] dtype */
  void* momentum;                /* [C, D] info.momentum (the p0 drawn), optional */
  double* acceptance_rate;       /* [C] optional */
  uint8_t* is_accepted;          /* [C] optional */
  uint8_t* is_divergent;         /* [C] optional */
  double* energy;                /* [C] optional: energy of the proposal */
  void* proposal_position;       /* [C, D] optional */
  void* proposal_momentum;       /* [C, D] optional (flipped) */
  void* proposal_velocity;       /* [C, D] optional (flipped) */
  double* proposal_logdensity;   /* [C] optional */
  void* proposal_logdensity_grad;/* [C, D] optional */
  double* proposal_weight;       /* [C] optional */
  double* proposal_sum_log_p_accept; /* [C] optional */
  int32_t* fp_iters;             /* [C, L] optional: FixedPointIterationInfo.iters */
  double* fp_norm;               /* [C, L] optional: .norm */
  uint8_t* fp_success;           /* [C, L] optional: .success */
  /* [C] optional: per-chain number of integration steps (the dynamic kernel,
   * geomjax/rmhmc/rmhmc.py:179-244); each value is clamped to [0, L] with
   * L = params->num_integration_steps, which also stays the row stride of fp_*. */
  const int32_t* num_integration_steps;
} gjx_rmhmc_step_io;

/* Buffers of T consecutive transitions in one persistent launch
 * (gjx_rmhmc_monge_funnel_sample): the reference's user loop
 * jax.lax.scan(jax.vmap(kernel)) (examples/funnel/main.py:16-25). */
typedef struct {
  const void* position;          /* [C, D] dtype initial positions */
  const double* logdensity;      /* [C] */
  const void* z;                 /* [T, C, D] dtype, or NULL: device PRNG */
  const double* u;               /* [T, C], or NULL: device PRNG */
  void* final_position;          /* [C, D] dtype */
  double* final_logdensity;      /* [C] */
  void* final_logdensity_grad;   /* [C, D] dtype, optional */
  void* samples;                 /* [T, C, D] dtype positions after each transition, optional */
  float* acceptance_rate;        /* [T, C] optional */
  int32_t* num_accepted;         /* [C] optional: accepted transitions */
  int64_t* num_fevals;           /* [C] optional: evaluations of the vector field executed */
  /* [T, C] optional: per-transition, per-chain number of integration steps
   * (dynamic kernel), clamped to [0, params->num_integration_steps]. */
  const int32_t* num_integration_steps;
} gjx_rmhmc_sample_io;

/* One RMHMC transition per chain.  Replaces the reference kernel
 * geomjax/rmhmc/rmhmc.py:131-174 (momentum draw :161, velocity :162,
 * rmhmc_proposal.generate :416-438) with implicit_midpoint
 * (geomjax/rmhmc/integrators.py:92-156) and metric_fn = Monge on the funnel
 * logdensity (examples/funnel/main.py:34-39). */
int gjx_rmhmc_monge_funnel_step(const gjx_funnel_params* params, const gjx_rmhmc_step_io* io,
                                void* stream);

/* T transitions per chain in one launch; state stays in registers between
 * transitions.  Same per-transition semantics as gjx_rmhmc_monge_funnel_step. */
int gjx_rmhmc_monge_funnel_sample(const gjx_funnel_params* params, const gjx_rmhmc_sample_io* io,
                                  void* stream);

/* Buffers of one LMC-Monge transition (gjx_lmcmonge_funnel_step).  Inputs are
 * LMCState (geomjax/lmcmonge/lmc.py:32-44); outputs the new LMCState and
 * LMCInfo (:63-98) whose Proposal state is RiemannianIntegratorState
 * (geomjax/lmcmonge/integrators.py:26-44; alpha2 is params->alpha2).
 * Optional outputs may be NULL. */
typedef struct {
  const void* inverse_mass_matrix;  /* [D] dtype: diagonal inverse mass (metrics.py:44-60, "fixed") */
  const void* position;             /* [C, D] dtype */
  const double* logdensity;         /* [C] */
  const void* logdensity_grad;      /* [C, D] dtype */
  const double* volume_adjustment;  /* [C] */
  const void* z;                    /* [C, D] dtype standard normals, or NULL: device PRNG */
  const double* u;                  /* [C] uniforms, or NULL: device PRNG */
  void* new_position;               /* [C, D] */
  double* new_logdensity;           /* [C] */
  void* new_logdensity_grad;        /* [C, D] optional */
  double* new_volume_adjustment;    /* [C] optional */
  void* velocity;                   /* [C, D] optional: info.velocity (the v0 drawn) */
  double* acceptance_rate;          /* [C] optional */
  uint8_t* is_accepted;             /* [C] optional */
  uint8_t* is_divergent;            /* [C] optional */
  double* energy;                   /* [C] optional */
  void* proposal_position;          /* [C, D] optional */
  void* proposal_momentum;          /* [C, D] optional (flipped) */
  void* proposal_velocity;          /* [C, D] optional (flipped) */
  double* proposal_logdensity;      /* [C] optional */
  void* proposal_logdensity_grad_norm; /* [C, D] optional */
  void* proposal_dl_ig;             /* [C, D] optional */
  void* proposal_Hdl_ig;            /* [C, D] optional */
  void* proposal_ig_Hdl_ig;         /* [C, D] optional */
  void* proposal_logdensity_hvp_velocity_norm; /* [C, D] optional */
  double* proposal_determinant_metric; /* [C] optional */
  double* proposal_volume_adjustment;  /* [C] optional */
  double* proposal_weight;          /* [C] optional */
  double* proposal_sum_log_p_accept;/* [C] optional */
  const int32_t* num_integration_steps; /* [C] optional: per-chain counts (dynamic kernel), each in [0, params->num_integration_steps] */
} gjx_lmc_step_io;

/* Buffers of T LMC-Monge transitions in one persistent launch. */
typedef struct {
  const void* inverse_mass_matrix;  /* [D] dtype */
  const void* position;             /* [C, D] */
  const double* logdensity;         /* [C] */
  const double* volume_adjustment;  /* [C] */
  const void* z;                    /* [T, C, D] or NULL */
  const double* u;                  /* [T, C] or NULL */
  void* final_position;             /* [C, D] */
  double* final_logdensity;         /* [C] */
  void* final_logdensity_grad;      /* [C, D] optional */
  double* final_volume_adjustment;  /* [C] optional */
  void* samples;                    /* [T, C, D] optional */
  float* acceptance_rate;           /* [T, C] optional */
  int32_t* num_accepted;            /* [C] optional */
  const int32_t* num_integration_steps; /* [T, C] optional: per-chain, per-transition counts (dynamic kernel) */
} gjx_lmc_sample_io;

/* One LMC-Monge transition per chain: the reference kernel
 * geomjax/lmcmonge/lmc.py:151-235 (velocity draw :201, lmc_proposal
 * :512-534) with lan_integrator/half_step_omega
 * (geomjax/lmcmonge/integrators.py:52-194) on the funnel
 * (examples/funnel/main.py:34-39). */
int gjx_lmcmonge_funnel_step(const gjx_funnel_params* params, const gjx_lmc_step_io* io, void* stream);

/* T LMC-Monge transitions per chain in one launch. */
int gjx_lmcmonge_funnel_sample(const gjx_funnel_params* params, const gjx_lmc_sample_io* io, void* stream);

/* NUTS with the LMC-Monge integrator (geomjax/lmcmonge/nuts.py:80-207,300-404):
 * the iterative expansion of gjx_rmhmc_monge_funnel_nuts over lan_integrator
 * steps, with the energy -l + K - J (the volume adjustment J carried along the
 * trajectory, geomjax/lmcmonge/metrics.py:267-284).  Funnel, one wave per
 * chain; params->dtype GJX_F64, or GJX_F32 for the tuned form (float32 state
 * buffers, inverse mass, z and trajectory; float64 log density, kinetic-energy
 * sums, volume change and weights); D - 1 <= 256, max_num_doublings in [1, 12], stopping criterion
 * GJX_UTURN_*; flags and alpha2 as for gjx_lmcmonge_funnel_step.  The draws
 * of one transition per chain: z (velocity), u_direction[e], u_merge[e] for
 * expansion e, u_step[j] for integration step j (all NULL: device PRNG). */
typedef struct {
  const void* inverse_mass_matrix;   /* [D] of the params' dtype */
  const void* position;              /* [C, D] */
  const double* logdensity;          /* [C] */
  const double* volume_adjustment;   /* [C] */
  const void* z;                     /* [C, D] or NULL */
  const double* u_direction;         /* [C, M] or NULL */
  const double* u_merge;             /* [C, M] or NULL */
  const double* u_step;              /* [C, 2^M - 1] or NULL */
  void* new_position;                /* [C, D] */
  double* new_logdensity;            /* [C] */
  void* new_logdensity_grad;         /* [C, D] optional */
  double* new_volume_adjustment;     /* [C] optional */
  void* velocity;                    /* [C, D] optional: NUTSInfo.velocity (the v0 drawn) */
  double* energy;                    /* [C] optional */
  double* acceptance_rate;           /* [C] optional */
  uint8_t* is_divergent;             /* [C] optional */
  uint8_t* is_turning;               /* [C] optional */
  int32_t* num_trajectory_expansions; /* [C] optional */
  int32_t* num_integration_steps;    /* [C] optional */
  void* leftmost_position;           /* [C, D] optional */
  void* rightmost_position;          /* [C, D] optional */
} gjx_lmc_nuts_io;

int gjx_lmcmonge_funnel_nuts(const gjx_funnel_params* params, int32_t max_num_doublings, int32_t stopping_criterion,
                             const gjx_lmc_nuts_io* io, void* stream);

/* Derivative in alpha2 of the proposal position of one LMC-Monge transition
 * per chain: the jax.jacfwd(trajectory_proposal_position, argnums=0) that the
 * reference's ChEES adaptation for lmcmonge takes
 * (geomjax/adaptation/chees_adaptation_lmcmonge.py:475-519), through the
 * velocity draw and every integrator step, in forward mode.  float64
 * (params->dtype = GJX_F64), params->alpha2 > 0; the noise (z, or the device
 * PRNG at params->seed / counter) and flags are those of the step entry point,
 * so the result belongs to the trajectory gjx_lmcmonge_funnel_step integrates
 * with the same arguments. */
typedef struct {
  const void* inverse_mass_matrix;       /* [D] float64 */
  const void* position;                  /* [C, D] float64 */
  const void* z;                         /* [C, D] float64 standard normals, or NULL: device PRNG */
  const int32_t* num_integration_steps;  /* [C] optional: per-chain counts in [0, params->num_integration_steps] */
  void* proposal_position;               /* [C, D] optional: the proposal position itself */
  double* d_proposal_position_d_alpha2;  /* [C, D] */
} gjx_lmc_alpha2_tangent_io;

int gjx_lmcmonge_funnel_alpha2_tangent(const gjx_funnel_params* params, const gjx_lmc_alpha2_tangent_io* io,
                                       void* stream);

/* ---------------------------------------------------------------------------
 * Bayesian logistic regression with the Fisher metric
 *   l(q) = sum_n [y_n eta_n - log(1 + e^eta_n)] - |q|^2 / (2 prior_var)
 *          - (D/2) log(2 pi prior_var),   eta = X q,
 *   G(q) = X^T diag(s (1 - s)) X + I / prior_var,  s = sigmoid(eta)
 * (the user-supplied logdensity_fn / metric_fn of the logistic
 * configurations).  Several posteriors of different (N, D) may share one
 * launch; each chain names its posterior.
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t num_chains;            /* C */
  int32_t dim;                   /* row stride of the [C, dim] state arrays, >= every D_p */
  int32_t dtype;                 /* GJX_F64, or GJX_F32 for the tuned RMHMC path (see below) */
  int32_t num_integration_steps; /* L >= 0 */
  int32_t max_iters;             /* fixed-point iteration cap */
  int32_t num_transitions;       /* T >= 1 (sample entry point) */
  int32_t flags;                 /* 0 */
  /* Tuned RMHMC path, posteriors of up to 64 features: workgroup size per
   * chain in waves, bits 0-3 for posteriors of up to 32 features, bits 4-7
   * for 33..64 features; each 1 or 2, or 0 for the library's default
   * (geomjax_amd.autotune picks it per problem).  Results do not depend on it
   * beyond float32 summation order.  Ignored elsewhere. */
  int32_t waves_per_chain;
  double step_size;
  double prior_var;              /* 100 in the logistic configurations */
  double convergence_tol;
  double divergence_tol;
  double divergence_threshold;
  uint64_t seed;
  uint64_t counter;
} gjx_logistic_params;

typedef struct {
  int32_t num_posteriors;        /* P >= 1 */
  int32_t max_rows;              /* host copy of max_p N_p */
  int32_t max_features;          /* host copy of max_p D_p (<= 128 for RMHMC, <= 64 for LMC) */
  int32_t reserved;
  const int32_t* num_rows;       /* [P] device: N_p */
  const int32_t* num_features;   /* [P] device: D_p */
  const int64_t* x_offset;       /* [P] device: element offset of X_p ([N_p, D_p] row-major) in x */
  const int64_t* y_offset;       /* [P] device: element offset of y_p in y */
  const int32_t* chain_posterior;/* [C] device, or NULL: every chain on posterior 0 */
  const void* x;                 /* float64 design matrices */
  const void* y;                 /* float64 labels in {0, 1} */
  /* GJX_F32 only (NULL / 0 otherwise): */
  const void* packed;            /* the table built by gjx_logistic_pack */
  void* workspace;               /* per-chain scratch, gjx_logistic_workspace_size bytes */
  uint64_t workspace_bytes;
  /* GJX_F32 only, optional: [C] device permutation of the chains giving the
   * order in which the tuned kernels start them (dispatch slot s runs chain
   * dispatch_order[s]; slots are dealt to the 8 XCDs in ranges of
   * ceil(C/8) slots).  NULL: chain s in slot s.  Results do not depend on it. */
  const int32_t* dispatch_order;
} gjx_logistic_data;

/* Tuned RMHMC path (params->dtype == GJX_F32).  State arrays (positions,
 * momenta, velocities, gradients, samples, z) are float32; per-chain scalars
 * stay float64.  Gram and quadratic-form contractions run on bf16 MFMA with
 * each float32 operand split into two bf16 terms (about 16 significant bits
 * per product) and float32 accumulation; the metric inverse is a float32
 * sweep with one refinement step per solve; log det G is summed, and the
 * energies, energy difference, accept test and fixed-point norm are kept, in
 * float64.  D <= 128 per posterior (one wave per chain up to 64 features,
 * four waves above).  The design matrices must first be packed (once per
 * data set) into a device table: */

/* Bytes of the packed table for P posteriors of the given (host) sizes. */
int gjx_logistic_pack_size(int32_t num_posteriors, const int32_t* num_rows_host, const int32_t* num_features_host,
                           uint64_t* bytes);

/* Fill `packed` (device, packed_bytes >= gjx_logistic_pack_size) from
 * data->x / data->y.  Set-up call: returns after the table is written
 * (synchronises `stream`). */
int gjx_logistic_pack(const gjx_logistic_data* data, const int32_t* num_rows_host, const int32_t* num_features_host,
                      void* packed, uint64_t packed_bytes, void* stream);

/* Bytes of the per-chain scratch the tuned path needs for params->num_chains
 * chains and data->max_rows rows (5 float arrays of max_rows rounded up to 64
 * per chain).  Calls that overlap in time need separate workspaces. */
int gjx_logistic_workspace_size(const gjx_logistic_params* params, const gjx_logistic_data* data, uint64_t* bytes);

/* One RMHMC transition per chain (same semantics and buffers as
 * gjx_rmhmc_monge_funnel_step, with metric_fn = Fisher metric). */
int gjx_rmhmc_fisher_logistic_step(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   const gjx_rmhmc_step_io* io, void* stream);

/* Bytes of the device scratch the allocation-free entry point below holds
 * per GPU (a static array, alive until the process ends). */
#define GJX_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  gjx_rmhmc_fisher_logistic_step with the tuned path's
 * per-chain scratch taken from a static per-GPU array of
 * GJX_NOALLOC_SCRATCH_BYTES instead of data->workspace (ignored here):
 * 5 * 4 * ceil(max_rows / 64) * 64 bytes per chain must fit, e.g. 1638 chains
 * at 2048 rows; larger calls return GJX_EUNSUPPORTED.  The call allocates
 * nothing: no hipMalloc / hipFree, no streams or events (a batch's packed
 * widths run one after another on `stream`).  Call it with the stream's
 * device current.  Every call on a GPU shares that GPU's scratch, so calls
 * that overlap in time on one GPU race; a call needs nothing in the scratch
 * to start at zero (every chain writes its rows before reading them).
 * float64 calls (no scratch) behave as gjx_rmhmc_fisher_logistic_step. */
int gjx_rmhmc_fisher_logistic_step_noalloc(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                           const gjx_rmhmc_step_io* io, void* stream);

/* T RMHMC transitions per chain in one launch (one workgroup per chain). */
int gjx_rmhmc_fisher_logistic_sample(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                     const gjx_rmhmc_sample_io* io, void* stream);

/* Buffers of one LMC transition with the general metric (lmcmc).  Inputs
 * LMCState (geomjax/lmcmc/lmc.py:30-42); outputs the new LMCState and LMCInfo
 * (:60-95) whose Proposal state is IntegratorState(position, momentum,
 * velocity, logdensity, logdensity_grad, volume_adjustment)
 * (geomjax/lmcmc/integrators.py:26-38).  Optional outputs may be NULL. */
typedef struct {
  const void* position;             /* [C, dim] */
  const double* logdensity;         /* [C] */
  const void* logdensity_grad;      /* [C, dim] */
  const double* volume_adjustment;  /* [C] */
  const void* z;                    /* [C, dim] standard normals, or NULL: device PRNG */
  const double* u;                  /* [C], or NULL */
  void* new_position;               /* [C, dim] */
  double* new_logdensity;           /* [C] */
  void* new_logdensity_grad;        /* [C, dim] optional */
  double* new_volume_adjustment;    /* [C] optional */
  void* velocity;                   /* [C, dim] optional: info.velocity (the v0 drawn) */
  double* acceptance_rate;          /* [C] optional */
  uint8_t* is_accepted;             /* [C] optional */
  uint8_t* is_divergent;            /* [C] optional */
  double* energy;                   /* [C] optional */
  void* proposal_position;          /* [C, dim] optional */
  void* proposal_momentum;          /* [C, dim] optional (flipped) */
  void* proposal_velocity;          /* [C, dim] optional (flipped) */
  double* proposal_logdensity;      /* [C] optional */
  void* proposal_logdensity_grad;   /* [C, dim] optional */
  double* proposal_volume_adjustment; /* [C] optional */
  double* proposal_weight;          /* [C] optional */
  double* proposal_sum_log_p_accept;/* [C] optional */
  const int32_t* num_integration_steps; /* [C] optional: per-chain counts (dynamic kernel), each in [0, params->num_integration_steps] */
} gjx_lmcmc_step_io;

typedef struct {
  const void* position;             /* [C, dim] */
  const double* logdensity;         /* [C] */
  const double* volume_adjustment;  /* [C] */
  const void* z;                    /* [T, C, dim] or NULL */
  const double* u;                  /* [T, C] or NULL */
  void* final_position;             /* [C, dim] */
  double* final_logdensity;         /* [C] */
  void* final_logdensity_grad;      /* [C, dim] optional */
  double* final_volume_adjustment;  /* [C] optional */
  void* samples;                    /* [T, C, dim] optional */
  float* acceptance_rate;           /* [T, C] optional */
  int32_t* num_accepted;            /* [C] optional */
  const int32_t* num_integration_steps; /* [T, C] optional: per-chain, per-transition counts (dynamic kernel) */
} gjx_lmcmc_sample_io;

/* One LMC transition per chain with metric_fn = Fisher metric: the reference
 * kernel geomjax/lmcmc/lmc.py:135-180 with lan_integrator
 * (geomjax/lmcmc/integrators.py:51-144) and gaussian_riemannian
 * (geomjax/lmcmc/metrics.py:42-206).  D <= 64.  params->dtype GJX_F64: the
 * float64 kernel; GJX_F32: the tuned form (float32 state buffers, Gram
 * matrices on split-bf16 MFMA, float64 log-determinants, volume change and
 * energies), which needs data->packed and data->workspace as the tuned RMHMC
 * entry point does. */
int gjx_lmc_fisher_logistic_step(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                 const gjx_lmcmc_step_io* io, void* stream);

/* T LMC transitions per chain in one launch (dtype as for the step). */
int gjx_lmc_fisher_logistic_sample(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   const gjx_lmcmc_sample_io* io, void* stream);

/* Buffers of one manifold-MALA transition.  Input DiffusionState(position,
 * logdensity, logdensity_grad, metric, Gamma) (geomjax/lmcmc/diffusions.py:27-32):
 * only position and logdensity are read (the metric, its inverse and Gamma
 * are functions of the position and are recomputed), logdensity_grad is copied
 * through on rejection.  Outputs the new DiffusionState and MMALAInfo
 * (geomjax/lmcmc/mmala.py:32-50: acceptance_rate, is_accepted, proposal,
 * is_divergent, weight).  Metrics are [C, dim, dim] blocks (zeros past the
 * posterior's D).  Optional outputs may be NULL. */
typedef struct {
  const void* position;             /* [C, dim] */
  const double* logdensity;         /* [C] */
  const void* logdensity_grad;      /* [C, dim] optional (else recomputed for the rejected case) */
  const void* z;                    /* [C, dim] standard normals, or NULL: device PRNG */
  const double* u;                  /* [C], or NULL */
  void* new_position;               /* [C, dim] */
  double* new_logdensity;           /* [C] */
  void* new_logdensity_grad;        /* [C, dim] optional */
  double* new_metric;               /* [C, dim, dim] optional */
  double* new_gamma;                /* [C, dim] optional */
  double* acceptance_rate;          /* [C] optional */
  uint8_t* is_accepted;             /* [C] optional */
  uint8_t* is_divergent;            /* [C] optional */
  double* weight;                   /* [C] optional */
  void* proposal_position;          /* [C, dim] optional */
  double* proposal_logdensity;      /* [C] optional */
  void* proposal_logdensity_grad;   /* [C, dim] optional */
  double* proposal_metric;          /* [C, dim, dim] optional */
  double* proposal_gamma;           /* [C, dim] optional */
} gjx_mmala_step_io;

typedef struct {
  const void* position;             /* [C, dim] */
  const double* logdensity;         /* [C] */
  const void* z;                    /* [T, C, dim] or NULL */
  const double* u;                  /* [T, C] or NULL */
  void* final_position;             /* [C, dim] */
  double* final_logdensity;         /* [C] */
  void* final_logdensity_grad;      /* [C, dim] optional */
  double* final_metric;             /* [C, dim, dim] optional */
  double* final_gamma;              /* [C, dim] optional */
  void* samples;                    /* [T, C, dim] optional */
  float* acceptance_rate;           /* [T, C] optional */
  int32_t* num_accepted;            /* [C] optional */
} gjx_mmala_sample_io;

/* init of manifold MALA (geomjax/lmcmc/mmala.py:53-60): log density, gradient,
 * Fisher metric and Gamma_i = sum_j d(G^-1)_ij/dx_j (geomjax/lmcmc/diffusions.py:83-94)
 * at C positions (float64, D <= 64).  metric and gamma may be NULL. */
int gjx_mmala_fisher_logistic_init(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   const void* position, double* logdensity, void* logdensity_grad, double* metric,
                                   double* gamma, void* stream);

/* One manifold-MALA transition per chain with metric_fn = Fisher metric
 * (geomjax/lmcmc/mmala.py:63-148, geomjax/lmcmc/diffusions.py:35-80; float64,
 * D <= 64).  reference_faithful != 0 keeps the reference's energy pairing
 * (proposal weight = -log of the Metropolis-Hastings ratio); 0 uses +log. */
int gjx_mmala_fisher_logistic_step(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   int reference_faithful, const gjx_mmala_step_io* io, void* stream);

/* T manifold-MALA transitions per chain in one launch. */
int gjx_mmala_fisher_logistic_sample(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                     int reference_faithful, const gjx_mmala_sample_io* io, void* stream);

/* ---------------------------------------------------------------------------
 * General dense metrics for LMC (lmcmc with a user metric_fn).  The caller
 * evaluates G(q) [C, dim, dim] and its Jacobian [C, dim, dim, dim]
 * (dG_ab/dq_i at [c][a][b][i], the layout of jax.jacfwd / torch.func.jacfwd);
 * these entry points do the per-chain linear algebra of the integrator.
 * float64, dim <= 64.  Buffers are device memory, [C, dim] row-major.
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t num_chains;
  int32_t dim;
  double step_size;                 /* half step: the signed step size eps */
  const double* metric;             /* [C, dim, dim] */
  const double* metric_jacobian;    /* [C, dim, dim, dim] */
  const double* z;                  /* position: [C, dim] normals of the velocity draw, or NULL */
  const double* velocity;           /* position: [C, dim] velocity for v^T G v, or NULL; half step: input */
  const double* logdensity_grad;    /* half step: [C, dim] */
  const double* grad_logdet_in;     /* half step: [C, dim] grad log det G at the position */
  double* velocity_out;             /* position: L^-T z (needs z); half step: new velocity (may alias velocity) */
  double* logdet;                   /* position: [C] log det (G + G^T)/2, optional */
  double* grad_logdet;              /* position: [C, dim] grad log det G, optional */
  double* vgv;                      /* position: [C] v^T G v (needs velocity), optional */
  double* volume_adjustment;        /* half step: [C] in/out, J - log|det A+| + log|det A-| */
  double* gamma;                    /* mmala: [C, dim] Gamma_i = sum_j d(G^-1)_ij/dq_j (needs metric_jacobian), optional */
  double* drift;                    /* mmala: [C, dim] G^-1 grad l + Gamma (needs logdensity_grad), optional */
  double* chol_inv_z;               /* mmala: [C, dim] chol(G^-1) z (needs z), optional */
  const double* momentum;           /* rmhmc: [C, dim] p, or NULL: p = chol(G) z (needs z) */
  double* momentum_out;             /* rmhmc: [C, dim] the p used, optional */
  double* dtdq;                     /* rmhmc: [C, dim] dT/dq (needs metric_jacobian), optional */
  double* pu;                       /* rmhmc: [C] p^T G^-1 p, optional */
  const double* step_sizes;         /* half step: [C] per-chain signed step sizes, or NULL: step_size */
} gjx_dense_metric_io;

/* Position quantities of a dense metric (geomjax/lmcmc/metrics.py:75-112,181-189):
 * log det G, grad log det G = tr(G^-1 dG/dq_k), the velocity draw L^-T z
 * (L = chol((G + G^T)/2)) and v^T G v. */
int gjx_dense_metric_position(const gjx_dense_metric_io* io, void* stream);

/* One half step of lan_integrator for a dense metric
 * (geomjax/lmcmc/integrators.py:61-91 with omega_tilde_fn, geomjax/lmcmc/metrics.py:158-179):
 * v' = (G + (eps/2) Omega~(v))^-1 (G v - (eps/2)(-grad l + grad log det G / 2)) by LU with
 * partial pivoting, J += -log|det(G + (eps/2) Omega~(v))| + log|det(G - (eps/2) Omega~(v'))|. */
int gjx_dense_lmc_half_step(const gjx_dense_metric_io* io, void* stream);

/* Manifold-MALA quantities of a dense metric (geomjax/lmcmc/mmala.py:91-113,
 * geomjax/lmcmc/diffusions.py:55-94): log det G, Gamma (from the Jacobian),
 * the drift G^-1 grad l + Gamma and the draw chol(G^-1) z.  metric_jacobian
 * may be NULL when gamma is not requested (the drift then omits Gamma). */
int gjx_dense_mmala_drift(const gjx_dense_metric_io* io, void* stream);

/* The implicit-midpoint vector field of RMHMC for a dense metric
 * (geomjax/rmhmc/metrics.py:42-129 with integrators.py:125-135): the momentum
 * p (given, or drawn as chol(G) z), u = G^-1 p (into velocity_out), dT/dq_k =
 * tr(G^-1 dG/dq_k)/2 - u^T (dG/dq_k) u / 2, log det G and p^T u. */
int gjx_dense_rmhmc_field(const gjx_dense_metric_io* io, void* stream);

/* value_and_grad of the logistic log density for C positions (init). */
int gjx_logistic_value_and_grad(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                const void* position, double* logdensity, void* logdensity_grad, void* stream);

/* ---------------------------------------------------------------------------
 * Derivatives for step-size / alpha2 tuning (no counterpart entry point in
 * the reference; its analogue is jax.jacfwd of the proposed position with
 * respect to alpha2, geomjax/adaptation/chees_adaptation_lmcmonge.py:475-519,
 * consumed by the alpha2 gradient at :222-257).
 * Per chain, one RMHMC-Monge transition from `position` with momentum
 * chol(G(q0; alpha2)) z: acceptance p = min(1, e^dH), squared jump
 * |q_L - q_0|^2, their product, and the exact derivatives of all three with
 * respect to (log step_size, log alpha2) through the executed fixed-point
 * iterates (forward-mode tangents).  alpha2 > 0.  The funnel entry point
 * takes dtype GJX_F64 (the check path) or GJX_F32 (the fast path: float32
 * state, trajectory and tangents, float64 energies and outputs); the Fisher
 * entry point GJX_F64.
 * ------------------------------------------------------------------------- */
typedef struct {
  const void* position;      /* [C, D] of the params' dtype */
  const double* logdensity;  /* [C] */
  const void* z;             /* [C, D] standard normals of the params' dtype, or NULL: device PRNG (seed, counter) */
  double* objective;         /* [C] */
  double* acceptance;        /* [C] */
  double* sq_jump;           /* [C] */
  double* d_objective;       /* [C, 2] */
  double* d_acceptance;      /* [C, 2] */
  double* d_sq_jump;         /* [C, 2] */
  int64_t* num_fevals;       /* [C] optional */
} gjx_tangent_io;

int gjx_rmhmc_monge_funnel_tangent(const gjx_funnel_params* params, const gjx_tangent_io* io, void* stream);

/* Derivatives for step-size tuning of RMHMC with the Fisher metric (the
 * quantities of gjx_rmhmc_monge_funnel_tangent; the analogue in the
 * reference is jax.jacfwd through its integrator,
 * geomjax/adaptation/chees_adaptation_lmcmonge.py:497-519).  One transition
 * per chain from io->position with momentum chol(G(q0)) z; column 0 of the
 * [C, 2] derivative arrays is d/d log step_size, column 1 is zero (the
 * Fisher metric has no alpha2).  io->z as for the step entry point (NULL:
 * device PRNG at params->counter).  params->dtype GJX_F64: the check path
 * (float64, D <= 64); GJX_F32 (float32 position buffer; data->packed and
 * data->workspace as for the tuned step; float64 energies and outputs): for
 * D <= 64 the tuned kernel's split-bf16 MFMA passes with their tangents, for
 * 65 <= D <= 128 a float32 form of the check path on fp32 MFMA. */
int gjx_rmhmc_fisher_logistic_tangent(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                      const gjx_tangent_io* io, void* stream);

/* ---------------------------------------------------------------------------
 * NUTS with the RMHMC implicit-midpoint integrator (geomjax/rmhmc/nuts.py:78-161,
 * 253-348): iterative trajectory doubling in a random direction, progressive
 * multinomial sampling (uniform inside a subtree, biased at the merge), the
 * numpyro-style iterative U-turn check with the stopping criterion
 * (geomjax/rmhmc/metrics.py:100-117: GJX_UTURN_EUC "euc", the default; BET
 * "bet"; RIEM "riem").  Monge metric
 * on Neal's funnel, float64, one wave per chain, D - 1 <= 256, M =
 * max_num_doublings in [1, 12] (the trajectory holds at most 2^M - 1 steps
 * past its start).  The random
 * draws of one transition per chain: z (momentum), u_direction[e] and u_merge[e]
 * for expansion e < M, u_step[j] for the j-th integration step (all NULL:
 * device PRNG at (params->seed, params->counter)).
 * ------------------------------------------------------------------------- */
enum { GJX_UTURN_EUC = 0, GJX_UTURN_BET = 1, GJX_UTURN_RIEM = 2 };

typedef struct {
  const void* position;          /* [C, D] */
  const double* logdensity;      /* [C] */
  const void* z;                 /* [C, D] or NULL */
  const double* u_direction;     /* [C, M] or NULL */
  const double* u_merge;         /* [C, M] or NULL */
  const double* u_step;          /* [C, 2^M - 1] or NULL */
  void* new_position;            /* [C, D] */
  double* new_logdensity;        /* [C] */
  void* new_logdensity_grad;     /* [C, D] optional */
  void* momentum;                /* [C, D] optional: NUTSInfo.momentum */
  double* energy;                /* [C] optional */
  double* acceptance_rate;       /* [C] optional */
  uint8_t* is_divergent;         /* [C] optional */
  uint8_t* is_turning;           /* [C] optional */
  int32_t* num_trajectory_expansions; /* [C] optional */
  int32_t* num_integration_steps;     /* [C] optional: states in the trajectory */
  void* leftmost_position;       /* [C, D] optional */
  void* rightmost_position;      /* [C, D] optional */
} gjx_nuts_io;

int gjx_rmhmc_monge_funnel_nuts(const gjx_funnel_params* params, int32_t max_num_doublings,
                                int32_t stopping_criterion, const gjx_nuts_io* io, void* stream);

/* T = params->num_transitions NUTS transitions of every chain in one launch:
 * the reference's sampling loop over the kernel above (a lax.scan over
 * jax.vmap(step), examples/funnel/main.py:16-25).  Each chain keeps its state
 * on chip between transitions, so the launch takes as long as the largest
 * sum of T trajectory lengths, not T times the longest single trajectory.
 * Transition t draws at (params->seed, params->counter + t), or from the
 * host arrays, each with a leading [T] axis. */
typedef struct {
  const void* position;          /* [C, D] */
  const double* logdensity;      /* [C] */
  const void* z;                 /* [T, C, D] or NULL */
  const double* u_direction;     /* [T, C, M] or NULL */
  const double* u_merge;         /* [T, C, M] or NULL */
  const double* u_step;          /* [T, C, 2^M - 1] or NULL */
  void* final_position;          /* [C, D] */
  double* final_logdensity;      /* [C] */
  void* final_logdensity_grad;   /* [C, D] optional */
  void* samples;                 /* [T, C, D] optional */
  float* acceptance_rate;        /* [T, C] optional */
  int32_t* num_integration_steps;/* [T, C] optional */
  int64_t* num_fevals;           /* [C] optional: integration steps of the launch */
} gjx_nuts_sample_io;

int gjx_rmhmc_monge_funnel_nuts_sample(const gjx_funnel_params* params, int32_t max_num_doublings,
                                       int32_t stopping_criterion, const gjx_nuts_sample_io* io, void* stream);

/* NUTS with the RMHMC integrator for Bayesian logistic regression with the
 * Fisher metric (geomjax/rmhmc/nuts.py:78-161 over the logistic integrator of
 * gjx_rmhmc_fisher_logistic_step): params->dtype GJX_F64 (one 256-lane
 * workgroup per chain) or GJX_F32, the tuned form (float32 state buffers and
 * trajectory, split-bf16 MFMA passes, float64 energies and tree weights, one
 * wave per chain; needs data->packed and data->workspace as the tuned step);
 * every posterior D <= 64; a batch of posteriors as for the step entry point.
 * io as for the funnel entry point, with rows of stride params->dim;
 * new_logdensity_grad is recomputed at the chosen state.
 * params->num_integration_steps and num_transitions are not used. */
int gjx_rmhmc_fisher_logistic_nuts(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                   int32_t max_num_doublings, int32_t stopping_criterion, const gjx_nuts_io* io,
                                   void* stream);

/* NUTS with the general-metric LMC integrator (geomjax/lmcmc/nuts.py:78-161)
 * for Bayesian logistic regression with the Fisher metric: the iterative
 * expansion over the lan_integrator steps of gjx_lmc_fisher_logistic_step, with
 * the energy -l - log det G / 2 + v^T G v / 2 - J (geomjax/lmcmc/metrics.py:
 * 209-221).  params->dtype GJX_F64 (one 256-lane workgroup per chain) or
 * GJX_F32, the tuned form (float32 state buffers and trajectory, split-bf16
 * MFMA Gram passes, float64 log densities, log-determinants, volume change,
 * energies and weights, one wave per chain; needs data->packed and
 * data->workspace as the tuned step); every posterior D <= 64; io as
 * gjx_lmc_nuts_io with rows of stride params->dim (inverse_mass_matrix unused;
 * velocity = NUTSInfo.velocity, the v0 drawn). */
int gjx_lmc_fisher_logistic_nuts(const gjx_logistic_params* params, const gjx_logistic_data* data,
                                 int32_t max_num_doublings, int32_t stopping_criterion, const gjx_lmc_nuts_io* io,
                                 void* stream);

/* Log density and its gradient of Neal's funnel for C positions: the
 * jax.value_and_grad(logdensity_fn)(position) of the reference's init
 * (geomjax/rmhmc/rmhmc.py:96-98, geomjax/lmcmonge/lmc.py:101-109) with the
 * funnel of examples/funnel/main.py:34-39. */
int gjx_funnel_value_and_grad(int32_t num_chains, int32_t dim, double sigma, int32_t dtype,
                              const void* position, double* logdensity, void* logdensity_grad,
                              void* stream);

/* Multi-chain convergence diagnostics of a sample array on the device
 * (geomjax/diagnostics.py:25-75 potential_scale_reduction and :78-209
 * effective_sample_size, float64 arithmetic).  Element (t, c, d) of the
 * samples is at samples[t * stride_sample + c * stride_chain +
 * d * stride_coord] (element strides: any axis order, views included). */
typedef struct {
  int32_t num_samples;   /* N >= 2 (at most 20352) */
  int32_t num_chains;    /* M >= 2 */
  int32_t num_coords;    /* D >= 1 */
  int32_t dtype;         /* GJX_F64 or GJX_F32 */
  int64_t stride_sample;
  int64_t stride_chain;
  int64_t stride_coord;
} gjx_diagnostics_args;

/* Bytes of device scratch gjx_effective_sample_size needs. */
int gjx_diagnostics_workspace_size(const gjx_diagnostics_args* args, uint64_t* bytes);

/* ess [D] and/or rhat [D] (either may be NULL) of every coordinate over all
 * chains.  Stream-ordered; workspace of gjx_diagnostics_workspace_size bytes. */
int gjx_effective_sample_size(const gjx_diagnostics_args* args, const void* samples, void* workspace,
                              uint64_t workspace_bytes, double* ess, double* rhat, void* stream);

/* Library identification and diagnostics. */
const char* gjx_version(void);
const char* gjx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GEOMJAX_AMD_H */
