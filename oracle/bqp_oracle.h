/* CPU FP64 oracle for the batched orbit propagator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against, and the CPU baseline that bench.py times.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (bqp_amd/) never calls it and has no CPU fallback.
 *
 * It restates, in plain C, the reference algorithm of
 *   dust_transfer/crates/satpy_core/src/gravity.rs   (V/W fill + M-G summation)
 *   dust_transfer/crates/lightyear_odeint_rs/src/rhs.rs (density, drag, RHS)
 *   dust_transfer/crates/lightyear_odeint_rs/src/odesolve/solver.rs (rk_step,
 *       integrate_internal, PI controller)
 *   dust_transfer/crates/lightyear_odeint_rs/src/odesolve/lightyear_compat.rs
 *   dust_transfer/crates/lightyear_odeint_rs/src/session.rs (ragged batch rows)
 * with these deviations from the reference: fully normalised coefficients, SI
 * units, a Cowell (inertial) state by default with the reference's Encke
 * formulation as an option, a uniform z-rotation for the body frame, and the
 * Vallado piecewise exponential atmosphere besides the reference's single-scale
 * one.
 *
 * Pinning status of each function (what checks it):
 *   orc_gravity_bf (normalised)   Orekit d/o-5 vectors, mpmath brute force
 *                                  N = 8..120, point-mass and J2 closed forms,
 *                                  Laplace / rotation invariance.
 *   orc_gravity_bf (compat)       equality with the normalised form (1e-13).
 *   orc_density                   closed form at the table nodes, continuity.
 *   orc_tableau                   order conditions of every tableau.
 *   orc_propagate                 two-body Kepler closed form, Jacobi
 *                                  integral, tolerance convergence, controller
 *                                  unit behaviours (PI memory, h_min, rejects).
 *   orc_kep2eci                   round trip with an independent numpy
 *                                  inverse and the vis-viva / angular-momentum
 *                                  identities.
 */
#ifndef BQP_ORACLE_H
#define BQP_ORACLE_H

#include <stdint.h>
#include "../bqp_amd/csrc/tableaus.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NMAX 200

/* Integration status codes (shared meaning with include/bqp.h). */
enum {
  ORC_OK = 0,
  ORC_MAX_STEPS = 1,
  ORC_STEP_UNDERFLOW = 2,
  ORC_NON_FINITE = 3,
  ORC_MAX_REJECTS = 4,
  ORC_INVALID_INPUT = 5,
  ORC_BELOW_MIN_RADIUS = 6
};

enum { ORC_CONTROL_SCALED = 0, ORC_CONTROL_ABSOLUTE = 1, ORC_CONTROL_COMPAT = 2 };
enum { ORC_DRAG_NONE = 0, ORC_DRAG_VALLADO = 1, ORC_DRAG_COMPAT = 2 };
enum { ORC_GRAVITY_NORMALISED = 0, ORC_GRAVITY_COMPAT = 1 };

/* Prepared field: normalised Cbar/Sbar (triangular, idx = n(n+1)/2 + m) plus
 * the recursion and summation constants derived from them. */
typedef struct orc_field orc_field;

orc_field* orc_field_create(int nmax, int mmax, const double* cbar, const double* sbar,
                            double mu, double radius, int form);
void orc_field_destroy(orc_field* f);

/* Body-fixed acceleration (m/s^2) of a body-fixed position (m). */
int orc_gravity_bf(const orc_field* f, const double pos[3], double acc[3]);
int orc_gravity_bf_batch(const orc_field* f, const double* pos, int64_t n, double* acc,
                         int threads);

double orc_density(int model, double h_m);

typedef struct orc_model {
  const orc_field* field;
  double omega;     /* body rotation rate about +z, rad/s */
  double theta0;    /* body rotation angle at t = 0, rad */
  int drag;         /* ORC_DRAG_* */
  double r_atm;     /* radius subtracted from |r| to get altitude, m */
  double r_min;     /* minimum admissible radius, m (0 disables) */
  int formulation;  /* ORC_COWELL or ORC_ENCKE */
  double ground_flattening; /* 0: |r| >= r_min; f > 0: height above the
                               ellipsoid (equatorial radius r_atm, flattening
                               f) along r at least r_min - r_atm */
} orc_model;

/* Equations of motion: Cowell (inertial state) or Encke (deviation from a
 * two-body baseline, Battin's correction, rebased every ORC_ENCKE_SEGMENT_S). */
enum { ORC_COWELL = 0, ORC_ENCKE = 1 };
/* Test system for the step controller (not an orbit): y' = y in all six
 * components, exact solution y0 exp(t - t0) (the reference's ExpSystem,
 * LY/src/odesolve/basic_tests.rs). */
#define ORC_TEST_EXP 2
#define ORC_ENCKE_SEGMENT_S 5400.0

typedef struct orc_integ {
  int method;       /* bqp_method_id */
  int control;      /* ORC_CONTROL_* */
  double rtol;
  double atol_pos, atol_vel;
  double eps;       /* absolute / compat tolerance */
  double h0;        /* <= 0: reference default */
  double h_min;     /* <= 0: 1e-12 */
  double h_max;     /* <= 0: span */
  int64_t max_steps;
  int32_t max_rejects;
  int32_t no_hcarry;  /* 1: Encke segments always open with the default step
                         (test switch; the reference carries it, see
                         orc_propagate) */
} orc_integ;

typedef struct orc_stats {
  int64_t steps;    /* accepted steps (compat: attempted) */
  int64_t rejects;
  int64_t evals;
  int32_t status;
} orc_stats;

void orc_rhs(const orc_model* m, double bc, double t, const double y[6], double dy[6]);

int orc_propagate(const orc_model* m, const orc_integ* cfg, double bc, const double y0[6],
                  double t0, double tf, double yf[6], orc_stats* st);

/* Sampled propagation from t0 to the last of the m increasing sample times
 * t_eval (>= t0), steps shortened onto every sample (the controller runs
 * through them); out receives m x 6 inertial states, NaN from the first
 * unreached sample of a failed row on. */
int orc_propagate_sampled(const orc_model* m, const orc_integ* cfg, double bc, const double y0[6], double t0,
                          const double* t_eval, int64_t m_eval, double* out, orc_stats* st);
int orc_propagate_sampled_batch(const orc_model* m, const orc_integ* cfg, const double* y0, double t0,
                                const double* t_eval, int64_t m_eval, const double* bc, int64_t n, double* out,
                                int32_t* status, int64_t* steps, int64_t* evals, int threads);

/* Ragged batch: row i runs from t0[i] to tf[i] with ballistic coefficient
 * bc[i] (Cd*A/m, m^2/kg; NULL = 0).  Failed rows are NaN-filled. */
int orc_propagate_batch(const orc_model* m, const orc_integ* cfg, const double* y0,
                        const double* t0, const double* tf, const double* bc, int64_t n,
                        double* yf, int32_t* status, int64_t* steps, int64_t* evals,
                        int threads);

/* Trace of the calling thread's step attempts (test hook): after
 * orc_trace_begin(buf, cap), every attempt of orc_propagate on this thread
 * appends ORC_TRACE_FIELDS doubles {t, h, err_norm, threshold, accepted
 * (1/0), next h} until cap records; orc_trace_end returns the number of
 * attempts made (which may exceed cap) and stops tracing. */
#define ORC_TRACE_FIELDS 6
void orc_trace_begin(double* buf, int64_t cap_records);
int64_t orc_trace_end(void);

void orc_kep2eci(double mu, const double kep[6], double eci[6]);

/* Two-body (Keplerian) state dt seconds after the elliptic state s0 (m, m/s).
 * *x is the eccentric-anomaly change: on input a starting guess (or NaN),
 * on output the solution.  Returns ORC_OK or ORC_INVALID_INPUT (not an
 * ellipse, or no convergence). */
int orc_two_body(double mu, const double s0[6], double dt, double out[6], double* x);

/* Battin's form of mu (rb/|rb|^3 - r/|r|^3), r = rb + dr, added to acc. */
void orc_battin(double mu, const double rb[3], const double dr[3], double acc[3]);
int orc_tableau(int method, bqp_tableau* out);

#ifdef __cplusplus
}
#endif
#endif
