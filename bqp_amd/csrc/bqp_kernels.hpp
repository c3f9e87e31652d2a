// Persistent batched propagation kernel (gfx950).
//
// One trajectory per lane.  A workgroup is G waves that share one group of
// 64 trajectories.  Each wave computes the gravity columns it owns
// (col_owner) and the G partial accelerations are summed through LDS in a
// fixed order.  In the FP64, split and Encke forms wave 0 alone carries the
// integrator state of the 64 lanes (stage sums, error norm, controller,
// refill) and hands each evaluation's body-frame positions to the other
// waves (grav_worker); in the surrogate forms every wave carries the state,
// bit-identical, and takes identical (wave-uniform) control flow, because the
// network tiles are dealt to all G waves.
//
// Lanes pull rows from a device work queue (one atomic per wave per refill,
// ranks by mbcnt), so trajectories of very different lengths share one launch;
// a lane whose row finishes writes its result and takes the next row.
//
// The stepper restates the reference's explicit-RK engine:
//   rk_step           LY/src/odesolve/solver.rs:1965-2173 (fma-axpy stage sums
//                     with h*a_ij, skipped zero coefficients, Kahan-compensated
//                     error sums)
//   integrate_internal  solver.rs:1099-1925, PI controller macros :445-529
//   compat preset     LY/src/odesolve/lightyear_compat.rs:100-600
// Differences: after a rejected step k1 = f(t, y) is kept instead of being
// re-evaluated (the value is bit-identical; only the eval count differs); and
// the stage sums use the second-order rearrangement of the same method.  For
// y = (r, v), y' = (v, a(t, r, v)) the stage velocities are
// V_s = v + h sum_j a_sj A_j, so
//   R_s = r + h c^_s v + h^2 sum_j (A.A)_sj A_j        (c^_s = sum_j a_sj)
//   r'  = r + h (sum b) v + h^2 sum_j (b.A)_j A_j,  v' = v + h sum_j b_j A_j
// (and the same for the error weights).  Only the stage accelerations A_j
// (3 doubles per stage, not 6) stay live, which halves the register
// footprint of the 13-stage method.
#pragma once

#include "bqp_device.hpp"
#include "mlp_device.hpp"
#include "second_order.h"
#include "second_order_tables.h"

#ifndef BQP_WAVES_PER_EU
#define BQP_WAVES_PER_EU 4
#endif

namespace bqp {

static constexpr bqp_tableau c_tab[3] = {BQP_DOPRI5_TABLEAU, BQP_DOP853_TABLEAU, BQP_RK87_TABLEAU};

// Hand-out position (in groups) of block b's first group.  Every lane starts
// free, so the first round of groups is dealt by block index instead of
// through the queue: blocks b, b + C, b + 2C, ... (C = deal_cus) are the ones
// the dispatcher places on one CU when the grid fills the device, and a
// back-and-forth deal over the longest-first order (round j takes positions
// jC .. jC + C - 1, in reverse on odd rounds) gives each CU groups of ranks
// c, 2C-1-c, 2C+c, 4C-1-c, ..., so the CUs carry nearly equal work.  Rounds
// not complete keep block order.  A permutation of [0, nb) in every case.
__device__ __forceinline__ unsigned first_group(unsigned b, unsigned nb, unsigned C) {
  if (C == 0u || nb <= C) return b;
  const unsigned j = b / C, c = b - j * C;
  if (j < nb / C && (j & 1u)) return j * C + (C - 1u - c);
  return b;
}

// Second-order rearrangement of each tableau (second_order.h), generated on
// the host at build time (gen_second_order.cpp -> second_order_tables.h) and
// compiled in as constants: coefficients read at compile-time indices fold
// into the instructions, and zero weights drop out of the code.
static constexpr SecondOrderTableau c_so[3] = BQP_SECOND_ORDER_TABLES;

template <int METHOD>
struct MethodTraits;
template <>
struct MethodTraits<BQP_METHOD_DOPRI5> {
  static constexpr int S = 7;
  static constexpr bool FSAL = true;
  static constexpr bool E3 = false;
};
template <>
struct MethodTraits<BQP_METHOD_DOP853> {
  static constexpr int S = 12;
  static constexpr bool FSAL = false;
  static constexpr bool E3 = true;
};
template <>
struct MethodTraits<BQP_METHOD_RK87> {
  static constexpr int S = 13;
  static constexpr bool FSAL = false;
  static constexpr bool E3 = false;
};

// Stage rows of a method staged in LDS at kernel start: row s holds c^_s,
// c_s and the packed chunks of SecondOrderTableau::sc.  The scalar cache is
// streamed through by the gravity table, so these per-stage reads would
// otherwise go to L2 on every evaluation; from LDS they arrive together with
// the stage accelerations of the same chunk.
constexpr int STAGE_ROW = 2 + STAGE_CHUNKS * 2 * STAGE_CHUNK;

// Fills rows[s][*] for every stage; called by all threads of the workgroup
// (thread tid of nthreads) before a barrier.
template <int METHOD>
__device__ __forceinline__ void stage_rows_to_lds(double (*rows)[STAGE_ROW], int tid, int nthreads) {
  constexpr int S = MethodTraits<METHOD>::S;
  for (int k = tid; k < S * STAGE_ROW; k += nthreads) {
    const int st = k / STAGE_ROW, i = k - st * STAGE_ROW;
    double v;
    if (i == 0) {
      v = c_so[METHOD].csum[st];
    } else if (i == 1) {
      v = c_tab[METHOD].c[st];
    } else {
      const int q = i - 2;
      v = c_so[METHOD].sc[st][q / (2 * STAGE_CHUNK)][q % (2 * STAGE_CHUNK)];
    }
    rows[st][i] = v;
  }
}

// Stage state of stage s in the second-order form (the rearranged rk_step
// stage sums, LY/src/odesolve/solver.rs:1965-2173):
//   R_s = r + h c^_s v + h^2 sum_{j<s} (A.A)_sj A_j,   V_s = v + h sum_{j<s} a_sj A_j,
// each component summed in ascending j.  The terms of a chunk of four
// earlier stages are read together (eight coefficients from the LDS stage
// rows, twelve stage-acceleration reads) and added without branches: a zero coefficient adds an
// exact zero, so for finite stage accelerations the sums equal those that
// skip zero coefficients bit for bit, and slots j >= s (stale) are replaced
// by zero before use.  A non-finite stage acceleration makes the step's
// solution non-finite either way.  This keeps the stage sums off the chain
// of dependent scalar loads and branches that one term at a time costs.
template <int METHOD>
__device__ __forceinline__ void stage_sums(int s, double hs, const double* y, const double (*sA)[3][64], int lane,
                                           const double (*rows)[STAGE_ROW], double* ys) {
  constexpr int S = MethodTraits<METHOD>::S;
  constexpr int NCH = (S - 1 + STAGE_CHUNK - 1) / STAGE_CHUNK;
  const double hh = hs * hs;
  const double hc = hs * rows[s][0];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ys[i] = fma(y[3 + i], hc, y[i]);
    ys[3 + i] = y[3 + i];
  }
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    if (STAGE_CHUNK * k < s) {
      const double* cf = rows[s] + 2 + 2 * STAGE_CHUNK * k;
      double A[STAGE_CHUNK][3];
#pragma unroll
      for (int m = 0; m < STAGE_CHUNK; ++m) {
        const int j = STAGE_CHUNK * k + m;
        if (j < S - 1) {
          const bool use = j < s;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const double v = sA[j][i][lane];
            A[m][i] = use ? v : 0.0;
          }
        }
      }
#pragma unroll
      for (int m = 0; m < STAGE_CHUNK; ++m) {
        const int j = STAGE_CHUNK * k + m;
        if (j < S - 1) {
          const double sv = hs * cf[2 * m], sr = hh * cf[2 * m + 1];
#pragma unroll
          for (int i = 0; i < 3; ++i) ys[3 + i] = fma(A[m][i], sv, ys[3 + i]);
#pragma unroll
          for (int i = 0; i < 3; ++i) ys[i] = fma(A[m][i], sr, ys[i]);
        }
      }
    }
  }
}

enum { CONTROL_SCALED = 0, CONTROL_ABSOLUTE = 1, CONTROL_COMPAT = 2 };
enum {
  ROW_OK = 0,
  ROW_MAX_STEPS = 1,
  ROW_STEP_UNDERFLOW = 2,
  ROW_NON_FINITE = 3,
  ROW_MAX_REJECTS = 4,
  ROW_INVALID_INPUT = 5,
  ROW_BELOW_MIN_RADIUS = 6
};

struct PropParams {
  const double* y0;
  const double* t0;   // per row (final mode) or nullptr (sampled: t0c)
  const double* tf;   // per row (final mode)
  const double* bc;   // per row or nullptr
  const int32_t* order;  // queue position -> row, or nullptr
  const double* t_eval;  // sampled mode: m_eval sample times, else nullptr
  double* yf;            // final mode (n x 6)
  double* ysamp;         // sampled mode (n x m_eval x 6)
  int32_t* status;
  int32_t* steps;
  int32_t* evals;
  int32_t* rejects;
  unsigned int* queue;   // next queue position; starts after the first round
  int n;
  int deal_cus;          // CUs the first round of groups is dealt over (0: block order)
  int teams;             // team kernel: trajectory groups per workgroup (1, 2 or 4)
  int m_eval;
  double t0c;
  FieldView field;
  ModelView model;
  double mu;                      // GM of the field (m^3/s^2): Encke baseline
  int control;
  int max_steps;
  int max_rejects;
  double rtol, atol_pos, atol_vel, eps, h0, h_min, h_max;
  // learned surrogate (PREC 2 and 3 only)
  FieldView low;                  // exact part, degree <= nlow
  MlpPacked mlp;                  // packed network
  double sur_r_lo, sur_r_hi;      // radius band (m)
  double sur_thresh2;             // squared check threshold (m^2/s^4)
  int sur_check_every;            // 0: never check
  unsigned long long* sur_totals; // net, exact, checks, failures, band
};

// Per-lane surrogate bookkeeping (PREC 2 and 3).
struct SurLane {
  bool exact = false;       // row switched to the exact field after a failed check
  bool check_now = false;   // this evaluation is a sampled check
  unsigned net = 0, exact_n = 0, checks = 0, fails = 0, band = 0;
};

__device__ __forceinline__ double clampd(double v, double lo, double hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ unsigned long long wave_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Network residual of the 64 positions of a trajectory group (PREC 2: BF16,
// PREC 3: FP8 hidden layers).  Wave 0 writes the bf16 features of its 64
// lanes to LDS; the four 16-sample tiles are dealt to the group's waves
// (wave g runs tiles g, g+G, ...; with G = 8 waves 4..7 idle); the outputs
// come back through LDS to wave 0.  Called by all G waves (barriers).
template <int PREC, int G>
__device__ __forceinline__ void sur_residual(const PropParams& p, int g, int lane, uint4* sX, float* sO, double xb,
                                             double yb, double zb, double& rx, double& ry, double& rz) {
  if (g == 0) {
    float f[MLP_IN];
    mlp_features((float)xb, (float)yb, (float)zb, p.mlp.r_mid, p.mlp.r_inv_half, f);
    sX[lane * 4 + 0] = make_uint4(pack_bf16x2(f[0], f[1]), pack_bf16x2(f[2], f[3]), pack_bf16x2(f[4], f[5]),
                                  pack_bf16x2(f[6], f[7]));
    sX[lane * 4 + 1] = make_uint4(pack_bf16x2(f[8], f[9]), pack_bf16x2(f[10], f[11]), pack_bf16x2(f[12], f[13]),
                                  pack_bf16x2(f[14], f[15]));
    sX[lane * 4 + 2] = make_uint4(0u, 0u, 0u, 0u);
    sX[lane * 4 + 3] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();
  constexpr int FMT = (PREC == 3) ? MLP_FMT_FP8 : MLP_FMT_BF16;
#pragma unroll 1
  for (int t = g; t < 4; t += G) {
    const uint4 xin[1] = {sX[(16 * t + (lane & 15)) * 4 + (lane >> 4)]};
    f32x4 o[1];
    mlp_forward_wave<FMT, 1>(p.mlp, lane, xin, o);
    if (lane < 16) {
      sO[(16 * t + lane) * 3 + 0] = o[0][0];
      sO[(16 * t + lane) * 3 + 1] = o[0][1];
      sO[(16 * t + lane) * 3 + 2] = o[0][2];
    }
  }
  __syncthreads();
  if (g == 0) {
    rx = (double)(sO[lane * 3 + 0] * p.mlp.out_scale[0]);
    ry = (double)(sO[lane * 3 + 1] * p.mlp.out_scale[1]);
    rz = (double)(sO[lane * 3 + 2] * p.mlp.out_scale[2]);
  }
}

// Wave 0 of a surrogate group: exact low-degree part + residual, or the
// exact field (gx, gy, gz on entry: the full FP64 field where needed).  A
// sampled check whose difference exceeds the threshold moves the row to the
// exact field for the rest of its arc.
__device__ __forceinline__ void sur_combine(const PropParams& p, SurLane& sl, bool store, bool band, bool use_net,
                                            bool check, double xb, double yb, double zb, double rx, double ry,
                                            double rz, double& gx, double& gy, double& gz) {
  double lx, ly, lz;
  grav_fp64<1>(p.low, 0, xb, yb, zb, lx, ly, lz);
  if (!store) return;
  if (use_net) {
    const double nx = lx + rx, ny = ly + ry, nz = lz + rz;
    if (check) {
      ++sl.checks;
      const double dx = nx - gx, dy = ny - gy, dz = nz - gz;
      if (fma(dx, dx, fma(dy, dy, dz * dz)) > p.sur_thresh2) {
        ++sl.fails;
        sl.exact = true;
      }
      ++sl.exact_n;  // the exact value, already computed, is used
    } else {
      gx = nx;
      gy = ny;
      gz = nz;
      ++sl.net;
    }
  } else {
    if (!band) ++sl.band;
    ++sl.exact_n;
  }
}

// ------------------------------------------------------------- Encke
// PREC 4: Encke formulation in FP64 (the reference's native one,
// LY/src/rhs.rs:3885-4118, segment driver LY/src/integrator.rs:3418-3563).
// The integrated state is the deviation d from a two-body baseline started
// from the total state at the segment start; segments are at most
// ENCKE_SEGMENT_S long (MAX_RECTIFICATION_SEGMENT_S, integrator.rs:236) and
// each starts a fresh controller.
constexpr double ENCKE_SEGMENT_S = 5400.0;
// Step caps at or below this do not carry the step size across rebases
// (HCARRY_UNCLAMPED_DT_MAX_S, integrator.rs:370).
constexpr double HCARRY_MIN_CAP_S = 60.0;

struct EnckeLane {
  double base[6];  // total state at the segment start (inertial, m, m/s)
  double t0;       // segment start time
  double x;        // eccentric-anomaly change of the last baseline solve (warm start)
  double seg_end;  // segment end time
};

// Two-body state dt after s0 by Lagrange coefficients in the
// eccentric-anomaly change x (Battin 1999 §4.5):
//   n dt = x + s1 (1 - cos x) - c1 sin x,  s1 = r0.v0/sqrt(mu a),  c1 = 1 - r0/a,
// Halley iterations from the last solution (or n dt), |dx| <= 1e-15 max(1,|x|).
// Returns false if s0 is not an ellipse or the solve does not converge.
__device__ __forceinline__ bool two_body(double mu, const double* s0, double dt, double* out, double& xio) {
  const double r0 = sqrt(fma(s0[0], s0[0], fma(s0[1], s0[1], s0[2] * s0[2])));
  const double v2 = fma(s0[3], s0[3], fma(s0[4], s0[4], s0[5] * s0[5]));
  const double rv = fma(s0[0], s0[3], fma(s0[1], s0[4], s0[2] * s0[5]));
  const double ainv = 2.0 / r0 - v2 / mu;
  if (!(ainv > 0.0) || !isfinite(ainv)) {
    for (int i = 0; i < 6; ++i) out[i] = __builtin_nan("");
    return false;
  }
  const double a = 1.0 / ainv;
  const double sqa = sqrt(a);
  const double n = sqrt(mu * ainv * ainv * ainv);
  const double s1 = rv / (sqrt(mu) * sqa);
  const double c1 = 1.0 - r0 * ainv;
  const double M = n * dt;
  double x = isfinite(xio) ? xio : M;
  bool ok = false;
  for (int it = 0; it < 50; ++it) {
    double sx, cx;
    sincos(x, &sx, &cx);
    const double F = x + s1 * (1.0 - cx) - c1 * sx - M;
    const double F1 = 1.0 + s1 * sx - c1 * cx;
    const double F2 = s1 * cx + c1 * sx;
    const double d = -2.0 * F * F1 / (2.0 * F1 * F1 - F * F2);
    x += d;
    if (fabs(d) <= 1e-15 * fmax(1.0, fabs(x))) {
      ok = true;
      break;
    }
  }
  double sx, cx;
  sincos(x, &sx, &cx);
  const double r = a + (r0 - a) * cx + s1 * a * sx;
  const double f = 1.0 - (a / r0) * (1.0 - cx);
  const double gg = dt - (x - sx) / n;
  const double fd = -sqrt(mu * a) * sx / (r * r0);
  const double gd = 1.0 - (a / r) * (1.0 - cx);
  for (int i = 0; i < 3; ++i) {
    out[i] = fma(f, s0[i], gg * s0[3 + i]);
    out[3 + i] = fma(fd, s0[i], gd * s0[3 + i]);
  }
  xio = x;
  return ok && isfinite(x);
}

// Battin's form of mu (rb/|rb|^3 - r/|r|^3), r = rb + dr (LY/src/rhs.rs:1071-1104).
__device__ __forceinline__ void battin(double mu, const double* rb, const double* dr, double& ax, double& ay,
                                       double& az) {
  const double rb2 = fma(rb[0], rb[0], fma(rb[1], rb[1], rb[2] * rb[2]));
  const double rx = rb[0] + dr[0], ry = rb[1] + dr[1], rz = rb[2] + dr[2];
  const double r2 = fma(rx, rx, fma(ry, ry, rz * rz));
  if (!(rb2 > 0.0 && r2 > 0.0)) return;
  const double drb = fma(dr[0], rb[0], fma(dr[1], rb[1], dr[2] * rb[2]));
  const double dd = fma(dr[0], dr[0], fma(dr[1], dr[1], dr[2] * dr[2]));
  const double q = (dd + 2.0 * drb) / rb2;
  const double q1 = 1.0 + q;
  const double fq = q * (3.0 + q * (3.0 + q)) / (1.0 + q1 * sqrt(q1));
  const double k = mu / (r2 * sqrt(r2));
  ax += k * fma(fq, rb[0], -dr[0]);
  ay += k * fma(fq, rb[1], -dr[1]);
  az += k * fma(fq, rb[2], -dr[2]);
}

// Adds the partial sums of waves 1 .. G-1 (lds[q], written before a
// barrier) to (gx, gy, gz) in wave order.  The reads of up to four partials
// are issued before their adds, so the sum waits for one LDS round trip per
// four partials instead of one per read.
template <int G>
__device__ __forceinline__ void add_partials(const double (*lds)[3][64], int lane, double& gx, double& gy,
                                             double& gz) {
  constexpr int B = 4;
#pragma unroll
  for (int q0 = 1; q0 < G; q0 += B) {
    double P[B][3];
#pragma unroll
    for (int k = 0; k < B; ++k)
      if (q0 + k < G) {
        P[k][0] = lds[q0 + k][0][lane];
        P[k][1] = lds[q0 + k][1][lane];
        P[k][2] = lds[q0 + k][2][lane];
      }
#pragma unroll
    for (int k = 0; k < B; ++k)
      if (q0 + k < G) {
        gx += P[k][0];
        gy += P[k][1];
        gz += P[k][2];
      }
  }
}

// Wave 0 alone runs the integrator when the group has more than one wave and
// no network (PREC 0, 1, 4): the other waves only compute gravity columns.
template <int PREC, int G>
constexpr bool leader_mode() {
  return G > 1 && PREC != 2 && PREC != 3;
}

// A non-leading wave of a leader-mode group.  Each round wave 0 publishes the
// body-frame positions of its 64 lanes in sP and passes barrier A; this wave
// computes its gravity columns, leaves the partial sum in lds[g] and passes
// barrier B, after which wave 0 adds the G partials in wave order.  sCmd = 0
// at barrier A ends the loop.  sP is rewritten and lds[g] read by wave 0
// only between B and the next A, so two barriers per evaluation suffice.
template <int PREC, int G>
__device__ __forceinline__ void grav_worker(const PropParams& p, int g, int lane, double (*lds)[3][64],
                                            const double (*sP)[64], const int* sCmd) {
  while (true) {
    __syncthreads();  // A
    if (__builtin_amdgcn_readfirstlane(*sCmd) == 0) return;
    const double xb = sP[0][lane], yb = sP[1][lane], zb = sP[2][lane];
    double gx, gy, gz;
    if (PREC == 1)
      grav_split<G>(p.field, g, xb, yb, zb, gx, gy, gz);
    else
      grav_fp64<G>(p.field, g, xb, yb, zb, gx, gy, gz);
    lds[g][0][lane] = gx;
    lds[g][1][lane] = gy;
    lds[g][2][lane] = gz;
    __syncthreads();  // B
  }
}

// Inertial Cowell RHS: uniform z-rotation to the body frame, gravity, rotate
// back, exponential drag (LY/src/rhs.rs:1119-1163 compute_drag).  The
// acceleration is computed by wave 0 only and, for lanes with `store`,
// written into the group's LDS stage slot sA[slot].  In leader mode the
// other waves sit in grav_worker and see only sP and their lds[g].  Without
// it (surrogate forms) every wave carries the integrator state and calls
// this function; the trailing barrier then both frees the partial-sum buffer
// and publishes sA[slot] to the other waves, which only contribute gravity
// partial sums.
template <int PREC, int G, int S>
__device__ __forceinline__ void accel_eval(const PropParams& p, int g, int lane,
                                           double (*lds)[3][64], double (*sA)[3][64], int slot,
                                           bool store, double bc, double t, const double* y,
                                           double* acc, uint4* sX, float* sO, int* sFlag, SurLane& sl,
                                           EnckeLane& el, double (*sP)[64], const ValladoTable& vt) {
  constexpr bool SUR = (PREC == 2 || PREC == 3);
  constexpr bool ENCKE = (PREC == 4);
  constexpr bool LEAD = leader_mode<PREC, G>();
  // total (inertial) state: y itself, or baseline + deviation (Encke)
  double yt[6], bst[6];
  if constexpr (ENCKE) {
    two_body(p.mu, el.base, t - el.t0, bst, el.x);
#pragma unroll
    for (int i = 0; i < 6; ++i) yt[i] = bst[i] + y[i];
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) yt[i] = y[i];
  }
  const double th = fma(p.model.omega, t, p.model.theta0);
  double s, c;
  sincos(th, &s, &c);
  const double xb = fma(c, yt[0], s * yt[1]);
  const double yb = fma(c, yt[1], -(s * yt[0]));
  const double zb = yt[2];
  if constexpr (LEAD) {
    sP[0][lane] = xb;
    sP[1][lane] = yb;
    sP[2][lane] = zb;
    __syncthreads();  // A: the workers start their columns
  }
  double gx, gy, gz;
  // surrogate decisions (identical in every wave of the group)
  bool band = true, use_net = false, check = false;
  double rx = 0.0, ry = 0.0, rz = 0.0;
  if (PREC == 0 || ENCKE) {
    grav_fp64<G>(p.field, g, xb, yb, zb, gx, gy, gz);
  } else if (PREC == 1) {
    grav_split<G>(p.field, g, xb, yb, zb, gx, gy, gz);
  } else {
    const double r = sqrt(fma(xb, xb, fma(yb, yb, zb * zb)));
    band = (r >= p.sur_r_lo) && (r <= p.sur_r_hi);
    use_net = store && band && !sl.exact;
    check = use_net && sl.check_now;
    const bool need_exact = store && (!use_net || check);
    if (wave_ballot(use_net) != 0ull) sur_residual<PREC, G>(p, g, lane, sX, sO, xb, yb, zb, rx, ry, rz);
    gx = gy = gz = 0.0;
    if (wave_ballot(need_exact) != 0ull) grav_fp64<G>(p.field, g, xb, yb, zb, gx, gy, gz);
  }
  if constexpr (LEAD) {
    __syncthreads();  // B: the workers' partials are in lds[1..G-1]
  } else if (G > 1) {
    lds[g][0][lane] = gx;
    lds[g][1][lane] = gy;
    lds[g][2][lane] = gz;
    __syncthreads();
    gx = lds[0][0][lane];
    gy = lds[0][1][lane];
    gz = lds[0][2][lane];
  }
  if (G > 1) add_partials<G>(lds, lane, gx, gy, gz);
  // only wave 0 publishes the stage acceleration; the other waves' copies
  // would be identical, so they skip the inertial rotation and the drag
  if (!LEAD && G > 1 && g != 0) {
    __syncthreads();
    if (SUR) sl.exact = sFlag[lane] != 0;  // wave 0's verdict of this evaluation's check
    return;
  }
  if (SUR) {
    sur_combine(p, sl, store, band, use_net, check, xb, yb, zb, rx, ry, rz, gx, gy, gz);
    if (G > 1) sFlag[lane] = sl.exact ? 1 : 0;
  }
  double ax = fma(c, gx, -(s * gy));
  double ay = fma(s, gx, c * gy);
  double az = gz;
  if (p.model.drag != DRAG_NONE && bc != 0.0) {
    const double r = sqrt(fma(yt[0], yt[0], fma(yt[1], yt[1], yt[2] * yt[2])));
    const double rho = density(vt, p.model.drag, r - p.model.r_atm);
    if (rho > 0.0) {
      const double vrx = fma(p.model.omega, yt[1], yt[3]);
      const double vry = fma(-p.model.omega, yt[0], yt[4]);
      const double vrz = yt[5];
      const double vr = sqrt(fma(vrx, vrx, fma(vry, vry, vrz * vrz)));
      const double k = -0.5 * bc * rho * vr;
      ax = fma(k, vrx, ax);
      ay = fma(k, vry, ay);
      az = fma(k, vrz, az);
    }
  }
  if constexpr (ENCKE) {
    // the baseline carries the central term: add it back, then Battin's
    // correction of the baseline's two-body acceleration
    const double r2 = fma(yt[0], yt[0], fma(yt[1], yt[1], yt[2] * yt[2]));
    const double k = p.mu / (r2 * sqrt(r2));
    ax = fma(k, yt[0], ax);
    ay = fma(k, yt[1], ay);
    az = fma(k, yt[2], az);
    battin(p.mu, bst, y, ax, ay, az);
  }
  acc[0] = ax;
  acc[1] = ay;
  acc[2] = az;
  if (store) {
    sA[slot][0][lane] = ax;
    sA[slot][1][lane] = ay;
    sA[slot][2][lane] = az;
  }
  if (!LEAD && G > 1) __syncthreads();
}

template <int METHOD, int PREC, int G>
__global__ void __launch_bounds__(64 * G) __attribute__((amdgpu_waves_per_eu(BQP_WAVES_PER_EU)))
propagate_kernel(const PropParams p) {
  using MT = MethodTraits<METHOD>;
  constexpr int S = MT::S;
  constexpr bool SUR = (PREC == 2 || PREC == 3);
  constexpr bool ENCKE = (PREC == 4);
  const bqp_tableau& T = c_tab[METHOD];
  const int lane = threadIdx.x & 63;
  // wave index within the group; readfirstlane tells the compiler it is
  // wave-uniform, so column ownership stays a scalar branch and the
  // coefficient entries stay scalar loads
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  __shared__ double lds[G][3][64];
  __shared__ unsigned int lds_base;
  // surrogate: bf16 features of the wave's 64 lanes (32 k-slots each) and outputs
  __shared__ uint4 sX[SUR ? 64 * 4 : 1];
  __shared__ float sO[SUR ? 64 * 3 : 1];
  __shared__ int sFlag[SUR && G > 1 ? 64 : 1];
  constexpr bool LEAD = leader_mode<PREC, G>();
  // leader mode: body-frame positions wave 0 hands to the other waves, and
  // its go (1) / stop (0) word
  __shared__ double sP[LEAD ? 3 : 1][64];
  __shared__ int sCmd;
  __shared__ double sRows[S][STAGE_ROW];
  __shared__ ValladoTable sVal;
  stage_rows_to_lds<METHOD>(sRows, threadIdx.x, 64 * G);
  vallado_to_lds(&sVal, threadIdx.x, 64 * G);
  __syncthreads();
  if constexpr (LEAD) {
    if (g != 0) {
      grav_worker<PREC, G>(p, g, lane, lds, sP, &sCmd);
      return;
    }
    if (lane == 0) sCmd = 1;
  }
  SurLane sl;
  EnckeLane el{};

  const bool compat = (p.control == CONTROL_COMPAT);
  const bool sampled = (p.t_eval != nullptr);
  const double q = compat ? 5.0 : (double)(T.order_err + 1);
  const double inv_q = 1.0 / q;

  bool active = false, exhausted = false;
  int row = 0, steps = 0, evals = 0, rejects = 0, rej_run = 0, sidx = 0;
  double t = 0.0, tseg = 0.0, tend = 0.0, h = 0.0, tcomp = 0.0, eprev = 0.0, bc = 0.0;
  double hmin = 0.0, hmax = 0.0, eps = p.eps;
  bool have_prev = false, just_rej = false, have_k1 = false;
  double y[6] = {0, 0, 0, 0, 0, 0};
  const SecondOrderTableau& SO = c_so[METHOD];
  // stage accelerations of the group's 64 trajectories, shared by its G waves
  __shared__ double sA[S][3][64];

  // Writes the lane's row result and frees the lane.
  auto finish = [&](int st) {
    if (st != ROW_OK) {
      if (sampled) {
        for (int m = sidx; m < p.m_eval; ++m)
          for (int i = 0; i < 6; ++i) p.ysamp[((size_t)row * p.m_eval + m) * 6 + i] = __builtin_nan("");
      } else {
        for (int i = 0; i < 6; ++i) p.yf[(size_t)row * 6 + i] = __builtin_nan("");
      }
    } else if (!sampled) {
      for (int i = 0; i < 6; ++i) p.yf[(size_t)row * 6 + i] = y[i];
    }
    if (p.status) p.status[row] = st;
    if (p.steps) p.steps[row] = steps;
    if (p.evals) p.evals[row] = evals;
    if (p.rejects) p.rejects[row] = rejects;
    active = false;
  };

  // Step-size bounds and initial step of an integration over `span`
  // (solver.rs:385-391 default h0; lightyear_compat.rs:203-213 for compat).
  auto start_controller = [&](double span) {
    if (compat) {
      eps = fmax(p.eps, 1e-12);
      hmin = 1e-10;
      hmax = (p.h_max > 0.0 && isfinite(p.h_max)) ? fmax(p.h_max, hmin) : span;
      h = clampd(fabs(span / 100.0), 1e-3, fmax(hmax, 1e-3));
    } else {
      hmax = (p.h_max > 0.0 && isfinite(p.h_max)) ? p.h_max : span;
      hmin = (p.h_min > 0.0) ? p.h_min : 1e-12;
      if (hmax < hmin) hmax = hmin;
      const double h0 = (p.h0 > 0.0) ? p.h0 : ((span <= 60.0) ? span / 2.0 : span / 100.0);
      h = clampd(h0, hmin, hmax);
    }
  };

  // Encke: start a segment at the current time from total state `tot`
  // (deviation 0, fresh controller), integrator.rs:3418-3563.  `carry` > 0
  // is the step size the controller asked for when the previous segment of
  // the same row ended: the reference opens a rebased segment with it
  // instead of the span/100 guess (hcarry_take / hcarry_store,
  // integrator.rs:341-402), on its adaptive-solver path only (not the DOPRI5
  // compat loop) and only when the step cap is above 60 s.
  auto start_segment = [&](const double* tot, double carry) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      el.base[i] = tot[i];
      y[i] = 0.0;
    }
    el.t0 = t;
    el.x = __builtin_nan("");
    el.seg_end = (tend - t > ENCKE_SEGMENT_S) ? t + ENCKE_SEGMENT_S : tend;
    start_controller(el.seg_end - t);
    if (carry > 0.0 && !compat && !(p.h_max > 0.0 && p.h_max <= HCARRY_MIN_CAP_S)) h = clampd(carry, hmin, hmax);
    have_prev = just_rej = have_k1 = false;
    eprev = 0.0;
    rej_run = 0;
    tcomp = 0.0;
  };

  // Total (inertial) state of the lane at the current time.
  auto total_state = [&](double* tot) {
    if constexpr (ENCKE) {
      double b[6];
      two_body(p.mu, el.base, t - el.t0, b, el.x);
#pragma unroll
      for (int i = 0; i < 6; ++i) tot[i] = b[i] + y[i];
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) tot[i] = y[i];
    }
  };

  bool first_fill = true;
  while (true) {
    // ---------------------------------------------------------- refill
    const unsigned long long need = wave_ballot(!active);
    if (need != 0ull && !exhausted) {
      const int cnt = __popcll(need);
      unsigned int base;
      const bool dealt = first_fill;
      if (first_fill) {
        // every lane is free: the block's first group is dealt, not queued
        base = 64u * first_group(blockIdx.x, gridDim.x, (unsigned)p.deal_cus);
        first_fill = false;
      } else if (G == 1 || LEAD) {
        unsigned int b = 0;
        if (lane == __ffsll((long long)need) - 1) b = atomicAdd(p.queue, (unsigned)cnt);
        base = __shfl(b, __ffsll((long long)need) - 1);
      } else {
        if (g == 0 && lane == __ffsll((long long)need) - 1) lds_base = atomicAdd(p.queue, (unsigned)cnt);
        __syncthreads();
        base = lds_base;
        __syncthreads();
      }
      if (dealt)  // the queue starts after the dealt round
        exhausted = 64ull * gridDim.x >= (unsigned long long)p.n;
      else if (base + (unsigned)cnt >= (unsigned)p.n)
        exhausted = true;
      if (!active) {
        const unsigned pos = base + (unsigned)lane_rank(need);
        if (pos < (unsigned)p.n) {
          row = p.order ? p.order[pos] : (int)pos;
          active = true;
          steps = evals = rejects = rej_run = 0;
          sidx = 0;
          tcomp = 0.0;
          eprev = 0.0;
          have_prev = just_rej = have_k1 = false;
          sl.exact = false;
          bc = p.bc ? p.bc[row] : 0.0;
          bool ok = isfinite(bc);
          for (int i = 0; i < 6; ++i) {
            y[i] = p.y0[(size_t)row * 6 + i];
            ok = ok && isfinite(y[i]);
          }
          t = sampled ? p.t0c : p.t0[row];
          tend = sampled ? p.t_eval[p.m_eval - 1] : p.tf[row];
          ok = ok && isfinite(t) && isfinite(tend) && (tend >= t);
          if (!ok) {
            finish(ROW_INVALID_INPUT);
          } else if (p.model.r_min > 0.0 && !above_ground(p.model, y[0], y[1], y[2])) {
            finish(ROW_BELOW_MIN_RADIUS);
          } else {
            // samples that coincide with the start are written immediately
            if (sampled) {
              while (sidx < p.m_eval && p.t_eval[sidx] == t) {
                for (int i = 0; i < 6; ++i) p.ysamp[((size_t)row * p.m_eval + sidx) * 6 + i] = y[i];
                ++sidx;
              }
            }
            const double span = tend - t;
            if (span == 0.0 || (sampled && sidx >= p.m_eval)) {
              finish(ROW_OK);
            } else if constexpr (ENCKE) {
              double tot[6], probe[6], x0 = 0.0;
#pragma unroll
              for (int i = 0; i < 6; ++i) tot[i] = y[i];
              if (!two_body(p.mu, tot, 0.0, probe, x0)) {
                finish(ROW_INVALID_INPUT);  // the baseline must be an ellipse
              } else {
                start_segment(tot, 0.0);  // a row's first segment has no predecessor
                tseg = sampled ? fmin(p.t_eval[sidx], el.seg_end) : el.seg_end;
              }
            } else {
              tseg = sampled ? p.t_eval[sidx] : tend;
              start_controller(span);
            }
          }
        }
      }
    }
    if (wave_ballot(active) == 0ull) {
      if (exhausted) break;
      continue;
    }

    // ------------------------------------------------- one step attempt
    bool stepping = active;
    const double rem = tseg - t;
    const double hs = (h < rem) ? h : rem;
    const bool lands = (hs == rem);
    if (stepping) {
      const int budget = compat ? steps + rejects : steps;
      if (budget >= p.max_steps || steps + rejects >= 2 * p.max_steps) {
        finish(ROW_MAX_STEPS);
        stepping = false;
      } else if (!compat && hs < hmin && !lands) {
        finish(ROW_STEP_UNDERFLOW);
        stepping = false;
      }
    }
    if (wave_ballot(stepping) == 0ull) continue;
    // stage 1 (k1) is kept across rejects and reused after FSAL accepts, so
    // only lanes without it evaluate stage 0; the stage loop holds the one
    // inlined copy of the RHS (stage 0: c = 0 and no a terms, so ys = y, t)
    const bool needk = stepping && !have_k1;
    const double hh = hs * hs;
#pragma clang loop unroll(disable)
    for (int s = 0; s < S; ++s) {
      const bool st = (s == 0) ? needk : stepping;
      if (s == 0 && wave_ballot(needk) == 0ull) continue;
      double ys[6];
      stage_sums<METHOD>(s, hs, y, sA, lane, sRows, ys);
      double as[3];
      if (SUR) sl.check_now = p.sur_check_every > 0 && (evals % p.sur_check_every) == 0;
      accel_eval<PREC, G, S>(p, g, lane, lds, sA, s, st, bc, t + sRows[s][1] * hs, ys, as, sX, sO, sFlag, sl, el, sP, sVal);
      if (st) ++evals;
    }
    if (needk) have_k1 = true;
    if (!stepping) continue;

    // solution and error estimates (Kahan-compensated error sums)
    double yn[6], err[6], comp[6], err3[6], comp3[6];
    {
      const double hb = hs * SO.bsum, he = hs * SO.esum, he3 = hs * SO.e3sum;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        yn[i] = fma(y[3 + i], hb, y[i]);
        yn[3 + i] = y[3 + i];
        err[i] = he * y[3 + i];
        err[3 + i] = 0.0;
        err3[i] = MT::E3 ? he3 * y[3 + i] : 0.0;
        err3[3 + i] = 0.0;
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) comp[i] = comp3[i] = 0.0;
    }
    auto kahan = [](double& sum, double& c, double term) {
      const double yk = term - c;
      const double acc = sum + yk;
      c = (acc - sum) - yk;
      sum = acc;
    };
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const double bj = T.b[j], ej = T.e[j], baj = SO.ba[j], eaj = SO.ea[j];
      if (bj != 0.0) {
        const double hbj = hs * bj;
#pragma unroll
        for (int i = 0; i < 3; ++i) yn[3 + i] = fma(sA[j][i][lane], hbj, yn[3 + i]);
      }
      if (baj != 0.0) {
        const double hba = hh * baj;
#pragma unroll
        for (int i = 0; i < 3; ++i) yn[i] = fma(sA[j][i][lane], hba, yn[i]);
      }
      if (ej != 0.0) {
        const double hej = hs * ej;
#pragma unroll
        for (int i = 0; i < 3; ++i) kahan(err[3 + i], comp[3 + i], hej * sA[j][i][lane]);
      }
      if (eaj != 0.0) {
        const double hea = hh * eaj;
#pragma unroll
        for (int i = 0; i < 3; ++i) kahan(err[i], comp[i], hea * sA[j][i][lane]);
      }
      if (MT::E3) {
        const double e3j = T.e3[j], e3aj = SO.e3a[j];
        if (e3j != 0.0) {
          const double h3 = hs * e3j;
#pragma unroll
          for (int i = 0; i < 3; ++i) kahan(err3[3 + i], comp3[3 + i], h3 * sA[j][i][lane]);
        }
        if (e3aj != 0.0) {
          const double h3a = hh * e3aj;
#pragma unroll
          for (int i = 0; i < 3; ++i) kahan(err3[i], comp3[i], h3a * sA[j][i][lane]);
        }
      }
    }
    // FSAL: on accept, wave 0 overwrites sA[0] with the last stage for the
    // next step; every wave must have read sA[0] for the sums above first
    // (a wave that lags behind wave 0 would otherwise read the new value)
    if (MT::FSAL && G > 1 && !LEAD) __syncthreads();
    bool finite = true;
#pragma unroll
    for (int i = 0; i < 6; ++i) finite = finite && isfinite(yn[i]) && isfinite(err[i]);
    double en = 0.0, thresh = 1.0;
    if (finite) {
      const bool blend3 = MT::E3 && (p.control == CONTROL_SCALED ||
                                     (p.control == CONTROL_ABSOLUTE && eps >= 1e-6));
      if (p.control == CONTROL_SCALED) {
        double s5 = 0.0, s3 = 0.0;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          const double atol = (i < 3) ? p.atol_pos : p.atol_vel;
          const double sc = atol + p.rtol * fmax(fabs(y[i]), fabs(yn[i]));
          const double den = (sc > 0.0) ? sc : atol;
          const double e5 = err[i] / den;
          s5 = fma(e5, e5, s5);
          if (blend3) {
            const double e3 = err3[i] / den;
            s3 = fma(e3, e3, s3);
          }
        }
        const double rms5 = sqrt(s5 / 6.0);
        if (blend3) {
          const double rms3 = sqrt(s3 / 6.0);
          en = sqrt(fma(rms5, rms5, 0.01 * rms3 * rms3));
        } else {
          en = rms5;
        }
      } else {
        double m5 = 0.0, m3 = 0.0;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          m5 = fmax(m5, fabs(err[i]));
          if (blend3 && !compat) m3 = fmax(m3, fabs(err3[i]));
        }
        en = (blend3 && !compat) ? sqrt(fma(m5, m5, 0.01 * m3 * m3)) : m5;
        thresh = eps;
      }
      finite = isfinite(en);
    }
    if (!finite) {
      ++rejects;
      if (++rej_run > p.max_rejects) {
        finish(ROW_NON_FINITE);
      } else {
        h = clampd(hs * 0.1, hmin, hmax);
      }
    } else if (en <= thresh || hs <= hmin) {
      double tn;
      if (lands) {
        tn = tseg;
      } else {
        const double ky = hs - tcomp;
        tn = t + ky;
        tcomp = (tn - t) - ky;
      }
      ++steps;
      rej_run = 0;
      if (MT::FSAL) {
        if (g == 0)
          for (int i = 0; i < 3; ++i) sA[0][i][lane] = sA[S - 1][i][lane];
        have_k1 = true;
      } else {
        have_k1 = false;
      }
      t = tn;
#pragma unroll
      for (int i = 0; i < 6; ++i) y[i] = yn[i];
      const double eff = (en > 0.0) ? en : thresh * 0.1;
      double fac, maxg;
      if (compat) {
        fac = (have_prev && eprev > 0.0) ? 0.9 * pow(eps / eff, 0.7 / 5.0) * pow(eprev / eps, 0.4 / 5.0)
                                         : 0.9 * pow(eps / eff, 0.2);
        maxg = just_rej ? 2.0 : 5.0;
      } else {
        if (p.control == CONTROL_SCALED)
          fac = (have_prev && eprev > 0.0) ? 0.9 * pow(1.0 / eff, 0.7 / q) * pow(eprev, 0.4 / q)
                                           : 0.9 * pow(eff, -inv_q);
        else
          fac = (have_prev && eprev > 0.0)
                    ? 0.9 * pow(eps / eff, 0.7 / q) * pow(eprev / eps, 0.4 / q)
                    : 0.9 * pow(eps / eff, inv_q);
        const double base = fmin(1.0 + 4.0 * (5.0 / q), 5.0);
        maxg = just_rej ? fmin(base, 2.0) : base;
      }
      fac = clampd(fac, 0.2, maxg);
      h = clampd(hs * fac, hmin, hmax);
      eprev = en;
      have_prev = true;
      just_rej = false;
      double rmin_tot[6];
      if (p.model.r_min > 0.0) total_state(rmin_tot);
      if (p.model.r_min > 0.0 && !above_ground(p.model, rmin_tot[0], rmin_tot[1], rmin_tot[2])) {
        finish(ROW_BELOW_MIN_RADIUS);
      } else if (ENCKE && t == tseg) {
        double tot[6];
        total_state(tot);
        bool done = false;
        if (sampled) {
          if (sidx < p.m_eval && t == p.t_eval[sidx]) {
            for (int i = 0; i < 6; ++i) p.ysamp[((size_t)row * p.m_eval + sidx) * 6 + i] = tot[i];
            ++sidx;
          }
          done = (sidx >= p.m_eval);
        } else {
          done = (t == tend);
        }
        if (done) {
#pragma unroll
          for (int i = 0; i < 6; ++i) y[i] = tot[i];
          finish(ROW_OK);
        } else {
          if (t == el.seg_end) start_segment(tot, h);  // rebase, carrying the controller's step
          tseg = sampled ? fmin(p.t_eval[sidx], el.seg_end) : el.seg_end;
          tcomp = 0.0;
        }
      } else if (t == tseg) {
        if (sampled) {
          for (int i = 0; i < 6; ++i) p.ysamp[((size_t)row * p.m_eval + sidx) * 6 + i] = y[i];
          ++sidx;
          if (sidx >= p.m_eval) {
            finish(ROW_OK);
          } else {
            tseg = p.t_eval[sidx];
            tcomp = 0.0;
          }
        } else {
          finish(ROW_OK);
        }
      }
    } else {
      ++rejects;
      if (++rej_run > p.max_rejects) {
        finish(ROW_MAX_REJECTS);
      } else {
        double fac;
        if (p.control == CONTROL_SCALED)
          fac = 0.9 * pow(en, -inv_q);
        else
          fac = 0.9 * pow(eps / en, compat ? 0.2 : inv_q);
        fac = clampd(fac, 0.1, 0.5);
        h = clampd(hs * fac, hmin, hmax);
        if (compat) {
          eprev = en;
          have_prev = true;
        }
        just_rej = true;
      }
    }
    // publish the FSAL copy of sA[0] before the next step reads it
    if (MT::FSAL && G > 1 && !LEAD) __syncthreads();
  }
  if constexpr (LEAD) {
    if (lane == 0) sCmd = 0;
    __syncthreads();  // A: releases the workers
  }
  if (SUR && g == 0) {
    unsigned long long v[5] = {sl.net, sl.exact_n, sl.checks, sl.fails, sl.band};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off);
      if (lane == 0 && v[k] != 0ull) atomicAdd(p.sur_totals + k, v[k]);
    }
  }
}

}  // namespace bqp
