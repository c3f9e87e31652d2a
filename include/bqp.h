/* bqp — MI355X-native batched orbit propagator: C ABI.
 *
 * Plain C types only (pointers, sizes, POD structs); no exceptions cross the
 * boundary; every function returns an int32 status (0 = ok, see BQP_E_*).
 * A context owns one HIP stream and the device copy of one gravity field; it
 * is thread-compatible (one thread at a time), distinct contexts are
 * independent.  Units are SI: m, m/s, s, m^2/kg.
 *
 * Reference interfaces replaced (tdude78/bqp, dust_transfer/):
 *   bqp_load_field            crates/lightyear_odeint_rs/src/config.rs:65-223
 *                             (load_constants / packed_constants_from_reader) and
 *                             crates/satpy_core/src/gravity.rs:1570 pack_gravity_coeffs
 *   bqp_gravity_accel_batch   crates/satpy_core/src/gravity.rs:2638
 *                             spherical_gravity_impl_frame_packed (and the unlaunched
 *                             CUDA mirror crates/satpy_core/kernels/cuda/gravity.cu:114)
 *   bqp_propagate_batch_final crates/lightyear_odeint_rs/src/session.rs:2325
 *                             LightyearSession::integrate_variable_final_into
 *                             (VariableFinalBatchRequest, session.rs:2028-2035) and
 *                             crates/lightyear_odeint_rs/src/batch.rs:247
 *                             integrate_batch_native_into (final-only shape)
 *   bqp_propagate_batch_sampled  crates/lightyear_odeint_rs/src/batch.rs:247
 *                             integrate_batch_native_into with t_eval (n, m, 6)
 *   bqp_config                crates/lightyear_odeint_rs/src/types.rs:332-520 ForceConfig
 *                             (hot-path subset) + odesolve/solver.rs:37-62 IntegratorConfig
 *   row status codes          odesolve/solver.rs:65-95 IntegrationStatus and
 *                             session.rs:1822-1850 VariableFinalRowFailure
 */
#ifndef BQP_H
#define BQP_H

#include <stdint.h>

#include "../bqp_amd/csrc/tableaus.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (function return values) */
#define BQP_OK 0
#define BQP_E_INVALID 1      /* bad argument / shape contract */
#define BQP_E_NO_FIELD 2     /* no gravity field loaded (reference: CONSTANTS_NOT_LOADED) */
#define BQP_E_DEVICE 3       /* HIP runtime error */
#define BQP_E_ROWS_FAILED 4  /* call completed, at least one row failed (see row_status) */
#define BQP_E_UNSUPPORTED 5

/* ---- per-row status (mirrors IntegrationStatus / VariableFinalRowFailure) */
#define BQP_ROW_OK 0
#define BQP_ROW_MAX_STEPS 1
#define BQP_ROW_STEP_UNDERFLOW 2
#define BQP_ROW_NON_FINITE 3
#define BQP_ROW_MAX_REJECTS 4
#define BQP_ROW_INVALID_INPUT 5
#define BQP_ROW_BELOW_MIN_RADIUS 6

/* ---- enums */
#define BQP_CONTROL_SCALED 0   /* WRMS, atol + rtol*max(|y|,|y'|) (solver.rs ErrorControl::Scaled) */
#define BQP_CONTROL_ABSOLUTE 1 /* max-norm vs eps (ErrorControl::Absolute) */
#define BQP_CONTROL_COMPAT 2   /* lightyear_compat.rs DOPRI5 controller preset */

#define BQP_DRAG_NONE 0
#define BQP_DRAG_VALLADO 1     /* Vallado 2013 Table 8-4 piecewise exponential */
#define BQP_DRAG_COMPAT 2      /* reference AtmModel::Exponential (rhs.rs:1778-1788) */

#define BQP_PRECISION_FP64 0   /* plain FP64 (the correctness oracle form) */
#define BQP_PRECISION_SPLIT 1  /* degree <= 4 in FP64, higher degrees FP32 + compensated sums */
/* Equations of motion.  Cowell: the inertial state is integrated.  Encke:
 * the deviation from a two-body baseline is integrated with Battin's
 * correction and the baseline is rebased every 5,400 s, each segment with a
 * fresh controller -- the reference's formulation (LY/src/rhs.rs:3885-4118,
 * LY/src/integrator.rs:3418-3563); inputs and outputs stay inertial states.
 * Encke needs precision FP64 and elliptic states. */
#define BQP_FORMULATION_COWELL 0
#define BQP_FORMULATION_ENCKE 1

#define BQP_PRECISION_SURROGATE 2 /* degree <= nlow in FP64, the rest from the learned network
                                     (bqp_surrogate_*), exact-field fallback */

typedef struct bqp_ctx bqp_ctx;

typedef struct bqp_config {
  int32_t method;      /* BQP_METHOD_DOPRI5 / _DOP853 / _RK87 (tableaus.h) */
  int32_t control;     /* BQP_CONTROL_* */
  int32_t precision;   /* BQP_PRECISION_* */
  int32_t drag;        /* BQP_DRAG_* */
  double rtol;         /* scaled control */
  double atol_pos;     /* m */
  double atol_vel;     /* m/s */
  double eps;          /* absolute / compat control */
  double h0;           /* <= 0: reference default (span/2 if span <= 60 s, else span/100) */
  double h_min;        /* <= 0: 1e-12 s */
  double h_max;        /* <= 0: unlimited (the row's span) */
  int64_t max_steps;   /* <= 0: 500000 */
  int32_t max_rejects; /* <= 0: 50 */
  int32_t formulation;  /* BQP_FORMULATION_COWELL (0) or BQP_FORMULATION_ENCKE (1) */
  double omega;        /* body-frame rotation rate about +z (rad/s) */
  double theta0;       /* body-frame angle at t = 0 (rad) */
  double r_atm;        /* altitude = |r| - r_atm (m) */
  double r_min;        /* rows whose |r| drops below r_min fail (0 = off) */
  double ground_flattening; /* 0: the r_min test is the sphere |r| >= r_min.  f > 0: it is
                               the reference's ground event (lightyear_odeint_rs/src/
                               events.rs:87-145): altitude above the ellipsoid of
                               equatorial radius r_atm and flattening f, along the
                               position's direction, at least r_min - r_atm */
} bqp_config;

typedef struct bqp_stats {
  int64_t steps;       /* accepted steps, summed over rows */
  int64_t rejects;
  int64_t evals;       /* RHS evaluations */
  int64_t rows_ok;
  int64_t rows_failed;
  int64_t first_failed_row; /* -1 if none */
  double kernel_ms;    /* device time of the propagation kernel (HIP events) */
  double flops;        /* algorithmic FLOPs of the call: evals x bqp_flops_per_rhs + steps x stage sums */
} bqp_stats;

/* Launch shape (tuning); zero fields mean "default".  Chosen by the
 * autotuner in bqp_amd.autotune when not given. */
typedef struct bqp_launch {
  int32_t waves_per_group;  /* G: waves sharing one 64-trajectory group (0 = auto, 1, 2, 4, 8, 16).
                               Auto (with teams_per_block 0) picks G from the batch size; FP64,
                               split and Encke calls on fields of degree >= 60 whose batch is one
                               round of G = 4 groups take the pair kernel instead (results equal
                               G = 8) */
  int32_t use_graph;        /* 1 = capture the call's launch sequence in a HIP graph once and
                               replay it while buffers, sizes and settings are unchanged */
  int32_t blocks;           /* persistent grid size; 0 = fill the device */
  int32_t sort_by_cost;     /* hand-out order of the rows (final-state calls):
                               0 = input order;
                               1 = descending estimated evaluations, the estimate being
                                   revolutions x evaluations per revolution learned from
                                   the context's earlier calls (bqp_get_cost_model);
                               2 = descending revolutions (nothing learned) */
  int32_t teams_per_block;  /* 0: the fixed-G kernel.  1, 2 or 4 (with waves_per_group = 0, FP64 or
                               Encke, field degree >= 30): the team kernel, one 16-wave workgroup
                               per CU carrying that many 64-trajectory groups, whose waves take the
                               gravity work of whichever groups are still running; its results
                               do not depend on grid, group width or batch composition (bit for
                               bit), but it measures slower on the 65,536-row headline batch.
                               2 with waves_per_group = 8 (FP64, split or Encke): the pair kernel, two
                               groups per 8-wave workgroup, a group's gravity in the 8 column
                               shares of G = 8 so that it runs on all 8 waves once its partner is
                               done; results equal waves_per_group = 8 (bit for bit) */
} bqp_launch;

/* Cells of the learned cost model: BQP_COST_ALT_BINS log-spaced bins of the
 * altitude of the semi-major axis above the field radius, 100 km to
 * 100,000 km, times BQP_COST_INC_BINS equal bins of cos(inclination) over
 * [-1, 1]; cell = alt_bin * BQP_COST_INC_BINS + inc_bin. */
#define BQP_COST_ALT_BINS 96
#define BQP_COST_INC_BINS 16

int32_t bqp_default_config(bqp_config* out);
int32_t bqp_create(const bqp_config* cfg, bqp_ctx** out);
int32_t bqp_destroy(bqp_ctx* ctx);
int32_t bqp_set_config(bqp_ctx* ctx, const bqp_config* cfg);
int32_t bqp_set_launch(bqp_ctx* ctx, const bqp_launch* launch);
const char* bqp_last_error(const bqp_ctx* ctx);

/* Normalised Cbar/Sbar, triangular row-major: idx(n, m) = n(n+1)/2 + m,
 * n = 0..nmax.  Cbar[0] should be 1 (central term); mmax <= nmax. */
int32_t bqp_load_field(bqp_ctx* ctx, int32_t nmax, int32_t mmax, const double* cbar,
                       const double* sbar, double mu, double radius);

/* Body-fixed gravity acceleration for n body-fixed positions (n x 3). */
int32_t bqp_gravity_accel_batch(bqp_ctx* ctx, const double* pos_bf, int64_t n, double* acc);
int32_t bqp_gravity_accel_batch_dev(bqp_ctx* ctx, const double* d_pos, int64_t n, double* d_acc);

/* Ragged final-state batch.  Row i: inertial state y0[6i..6i+5] at t0[i],
 * propagated to tf[i] >= t0[i] with ballistic coefficient bc[i] = Cd*A/m
 * (NULL = no drag coefficient).  Failed rows are NaN-filled and flagged in
 * row_status.  row_status / row_steps / row_evals may be NULL.  Returns
 * BQP_E_ROWS_FAILED (after writing every row) if any row failed. */
int32_t bqp_propagate_batch_final(bqp_ctx* ctx, const double* y0, const double* t0,
                                  const double* tf, const double* bc, int64_t n, double* yf,
                                  int32_t* row_status, int32_t* row_steps, int32_t* row_evals,
                                  bqp_stats* stats);

/* Same with device-resident buffers (no host staging; the caller's
 * buffers must stay valid until the call returns).  Optional ones may be
 * NULL. */
int32_t bqp_propagate_batch_final_dev(bqp_ctx* ctx, const double* d_y0, const double* d_t0,
                                      const double* d_tf, const double* d_bc, int64_t n,
                                      double* d_yf, int32_t* d_row_status, int32_t* d_row_steps,
                                      int32_t* d_row_evals, bqp_stats* stats);

/* Same as bqp_propagate_batch_final_dev, allocating nothing: no hipMalloc,
 * hipFree or allocating library call, for callers that cannot allocate
 * (code under HIP-graph capture, tight serving loops, harnesses that forbid
 * it).  No counterpart in the reference.  Its work queue, totals, hand-out
 * tables and per-row counters live in a static device array of the library
 * (BQP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends), so:
 *   - n must be at most BQP_NOALLOC_MAX_ROWS (BQP_E_INVALID otherwise);
 *   - calls that overlap on one GPU (two contexts, two threads, two streams)
 *     race on that array: serialise them;
 *   - every launch is direct (bqp_launch.use_graph is ignored), and a checked
 *     surrogate keeps all its checks in the kernel (no up-front partition);
 *     results equal bqp_propagate_batch_final_dev's.
 * The context must come from bqp_create (which looks the array's address
 * up once per device, with the device current); the call runs on the
 * context's device and stream. */
#define BQP_NOALLOC_MAX_ROWS 262144
#define BQP_NOALLOC_SCRATCH_BYTES (1024 + 4096 * 4 + BQP_NOALLOC_MAX_ROWS * 22 + 256)
int32_t bqp_propagate_batch_final_dev_noalloc(bqp_ctx* ctx, const double* d_y0, const double* d_t0,
                                              const double* d_tf, const double* d_bc, int64_t n, double* d_yf,
                                              int32_t* d_row_status, int32_t* d_row_steps,
                                              int32_t* d_row_evals, bqp_stats* stats);

/* Fixed-grid sampled batch: every row starts at t0 and is sampled at the
 * m increasing times t_eval[0..m-1] (>= t0), output (n, m, 6) inertial states.
 * Samples are hit exactly (the step is shortened onto each sample time). */
int32_t bqp_propagate_batch_sampled(bqp_ctx* ctx, const double* y0, double t0,
                                    const double* t_eval, int64_t m, const double* bc, int64_t n,
                                    double* out, int32_t* row_status, bqp_stats* stats);

/* Butcher tableau of a method (for host-side checks). */
int32_t bqp_get_tableau(int32_t method, bqp_tableau* out);

/* Algorithmic FLOPs of one RHS evaluation at the loaded field / config. */
double bqp_flops_per_rhs(const bqp_ctx* ctx);

/* The learned cost model behind sort_by_cost = 1 (no counterpart in the
 * reference, which runs rows in input order): per cell, the evaluations,
 * revolutions and row count of every row of at least one revolution that
 * finished ok in final-state calls with sort_by_cost = 1 since the field or
 * configuration was last set.  Each array holds
 * BQP_COST_ALT_BINS * BQP_COST_INC_BINS values; any may be NULL. */
int32_t bqp_get_cost_model(bqp_ctx* ctx, int64_t* evals, double* revolutions, int64_t* rows);

/* Launch shape of the context's last propagation call (no counterpart in the
 * reference, which has no launch shape): which persistent kernel ran and
 * how.  A checked-surrogate call reports its exact-row launch when it had
 * one, else its surrogate launch. */
#define BQP_KERNEL_NONE 0   /* no propagation call yet */
#define BQP_KERNEL_FIXED 1  /* fixed-G kernel: waves_per_group waves per 64-trajectory group */
#define BQP_KERNEL_TEAM 2   /* team kernel: teams groups per 16-wave workgroup */
#define BQP_KERNEL_PAIR 3   /* pair kernel: 2 groups per 8-wave workgroup, results equal G = 8 */
typedef struct bqp_launch_info {
  int32_t kernel;           /* BQP_KERNEL_* */
  int32_t waves_per_group;  /* G of the fixed-G kernel; 8 for the pair kernel; 0 for the team kernel */
  int32_t teams;            /* groups per workgroup (team and pair kernels), else 0 */
  int32_t blocks;           /* persistent grid */
  int32_t threads;          /* workgroup size */
} bqp_launch_info;
int32_t bqp_last_launch(const bqp_ctx* ctx, bqp_launch_info* out);

/* ------------------------------------------------------------------------
 * Learned gravity surrogate.  No counterpart in the reference (which always
 * evaluates every harmonic term, SC/src/gravity.rs:2638); these calls exist
 * for the learned form of the propagation path.
 *
 * Network: 16 features of the body-fixed position -> 4 x [256, SiLU] -> 3,
 * predicting (a_field - a_{degree <= nlow}) / out_scale.  FP32 master
 * parameters (flat: W_l row-major (out x in), then b_l, for l = 0..4) are
 * trained on the device with Adam; inference runs on MFMA in BF16 or with
 * the three 256 x 256 layers in FP8 E4M3 (per-layer scales), FP32 accumulate.
 * ---------------------------------------------------------------------- */
#define BQP_MLP_FP32 0  /* FP32 vector path (training forward), reference form */
#define BQP_MLP_BF16 1
#define BQP_MLP_FP8 2

typedef struct bqp_surrogate_stats {
  int64_t net_evals;       /* RHS evaluations that used the network */
  int64_t exact_evals;     /* evaluations that used the exact field */
  int64_t checks;          /* sampled network-vs-exact comparisons */
  int64_t check_failures;  /* comparisons above the threshold (row switched to exact) */
  int64_t band_fallbacks;  /* evaluations outside the radius band */
} bqp_surrogate_stats;

/* Number of FP32 parameters of the network (202,499). */
int64_t bqp_surrogate_param_count(void);

/* Set up the surrogate for the loaded field: exact part up to degree nlow
 * (1..8), radius band [r_lo, r_hi] (m) for features and fallback, seeded
 * He-uniform initial parameters, output scales = per-component RMS of the
 * residual over a seeded pilot sample.  Resets the optimiser. */
int32_t bqp_surrogate_init(bqp_ctx* ctx, int32_t nlow, double r_lo, double r_hi, uint64_t seed);

/* Copy the FP32 parameters (and the 3 output scales) out / in. */
int32_t bqp_surrogate_get_params(bqp_ctx* ctx, float* params, double* out_scale);
int32_t bqp_surrogate_set_params(bqp_ctx* ctx, const float* params, const double* out_scale);

/* Seeded training data on the device: n body-fixed positions (uniform
 * directions, |r| uniform in [r_lo, r_hi]), their 16 features, and the
 * scaled residual targets computed with the exact FP64 field.  d_pos
 * (n x 3 f64) may be NULL. */
int32_t bqp_surrogate_dataset_dev(bqp_ctx* ctx, int64_t n, uint64_t seed, double* d_pos, float* d_feat,
                                  float* d_targ);

/* `steps` Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8) on mean squared
 * error over minibatches of `batch` rows drawn (seeded, with replacement)
 * from the n device rows.  loss (host, may be NULL) receives each step's
 * minibatch loss. */
int32_t bqp_surrogate_train_dev(bqp_ctx* ctx, const float* d_feat, const float* d_targ, int64_t n,
                                int32_t steps, int32_t batch, double lr, uint64_t seed, float* loss);

/* Mean squared error over all n rows and its gradient (host, param_count
 * floats, may be NULL), FP32. */
int32_t bqp_surrogate_loss_grad_dev(bqp_ctx* ctx, const float* d_feat, const float* d_targ, int64_t n,
                                    double* loss, float* grad);

/* Pack the current parameters for MFMA inference in `format` (BF16 or
 * FP8); FP8 activation scales are calibrated on n_calib feature rows. */
int32_t bqp_surrogate_pack(bqp_ctx* ctx, int32_t format, const float* d_calib_feat, int64_t n_calib);

/* Network outputs (n x 3, scaled units) for n feature rows in `format`
 * (FP32 = vector path on the master parameters; BF16 / FP8 = MFMA path on
 * the packed weights).  kernel_ms (may be NULL) = device time. */
int32_t bqp_surrogate_infer_dev(bqp_ctx* ctx, int32_t format, const float* d_feat, int64_t n, float* d_out,
                                double* kernel_ms);

/* In-integrator use (precision BQP_PRECISION_SURROGATE): every
 * `check_every`-th evaluation of a row (0 = never) also evaluates the exact
 * field; a difference above `threshold` (m/s^2) switches that row to the
 * exact field for the rest of its arc.  Positions outside the radius band
 * always use the exact field. */
int32_t bqp_surrogate_set_check(bqp_ctx* ctx, int32_t check_every, double threshold);

/* Counters of the last propagation call. */
int32_t bqp_surrogate_last_stats(const bqp_ctx* ctx, bqp_surrogate_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* BQP_H */
