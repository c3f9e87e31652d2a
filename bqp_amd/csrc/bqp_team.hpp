// Shared-workgroup propagation kernels (gfx950): several 64-trajectory groups
// in one workgroup, whose waves move to the groups still running as the
// others finish.
//
// Both kernels run the per-row integrator of propagate_kernel
// (bqp_kernels.hpp) in a leader wave per group, as a per-evaluation state
// machine (GroupLeader below): rk_step and the controller restated from
// LY/src/odesolve/solver.rs:445-529, 1099-1925, 1965-2173 and
// LY/src/odesolve/lightyear_compat.rs:100-600, the Encke segment driver of
// LY/src/integrator.rs:341-402, 3418-3563.  Each loop iteration is one
// right-hand-side evaluation of every live group:
//   A  each leader starts a step if its group has none in flight (refilling
//      rows) and publishes the group's next stage in LDS;
//   -- barrier --
//   B  every wave, leaders included, computes gravity work items of the live
//      groups into LDS slots;
//   -- barrier --
//   C  each leader combines its group's slots in a fixed order, adds drag
//      (and the Encke terms), stores the stage and, after the last stage,
//      completes the step (solution, error norm, controller, row results).
// A group whose rows are done leaves the pool, so its waves work on the
// groups still running until the workgroup's last row ends.  The price is
// lockstep: each evaluation waits for the last item and for the slowest
// leader's serial work.
//
// Team kernel (propagate_team_kernel): one 16-wave workgroup per CU, up to
// four groups ("teams").  A work item is one of 16 fixed gravity units of a
// team (grav_unit), or an aligned pair of units when more than two teams are
// live, taken from an LDS counter; each unit is turned back to the inertial
// frame on its own and the units are combined along a fixed pairwise tree,
// so the results do not depend on the number of live teams, the grid, the
// hand-out order or the batch (bit for bit).
//
// Pair kernel (propagate_pair_kernel): 8-wave workgroups, two per CU, each
// carrying two groups.  A group's gravity is the eight column shares of the
// fixed-G kernel at G = 8 (col_owner<8>), summed in share order in the body
// frame, so its results equal waves_per_group = 8 bit for bit.  While both
// groups run each wave computes two shares of one group (the work of a wave
// of the G = 4 kernel); when one is done its four waves join the other,
// which then runs at the G = 8 latency.
#pragma once

#include "bqp_kernels.hpp"

namespace bqp {

constexpr int TEAM_WAVES = 16;
constexpr int TEAM_MAX = 4;
// LDS slots for the iteration's gravity work items: the 16 units of each of
// up to TEAM_PAIRS_ABOVE live teams, else 8 aligned unit pairs per team
constexpr int TEAM_PAIRS_ABOVE = 2;
constexpr int TEAM_SLOTS = 2 * GRAV_UNITS;

// Pair kernel: waves per workgroup = column shares per group, and groups.
constexpr int PAIR_WAVES = 8;
constexpr int PAIR_GROUPS = 2;
// Lowest field degree at which an automatic launch takes the pair kernel.
constexpr int PAIR_AUTO_MIN_N = 60;

// The integrator state of one group's 64 rows, held by the group's leader
// wave, and its per-evaluation steps.  sA is the group's LDS stage storage.
template <int METHOD, int PREC>
struct GroupLeader {
  using MT = MethodTraits<METHOD>;
  static constexpr int S = MT::S;
  static constexpr bool ENCKE = (PREC == 4);

  const PropParams& p;
  double (*sA)[3][64];
  const double (*rows)[STAGE_ROW];  // the method's stage rows in LDS (stage_rows_to_lds)
  int lane;
  bool compat, sampled;
  double q, inv_q;
  bool alive, in_step = false, exhausted = false, first_fill = true, active = false;
  int row = 0, steps = 0, evals = 0, rejects = 0, rej_run = 0, sidx = 0;
  double t = 0.0, tseg = 0.0, tend = 0.0, h = 0.0, tcomp = 0.0, eprev = 0.0, bc = 0.0;
  double hmin = 0.0, hmax = 0.0, eps;
  bool have_prev = false, just_rej = false, have_k1 = false;
  double y[6] = {0, 0, 0, 0, 0, 0};
  EnckeLane el{};
  // step attempt and stage in flight
  int s_cur = 0;
  bool stepping = false, needk = false, lands = false, st = false;
  double hs = 0.0;
  double bst[3] = {0, 0, 0}, ysr[3] = {0, 0, 0};  // Encke: stage baseline position and deviation

  __device__ __forceinline__ GroupLeader(const PropParams& p_, double (*sA_)[3][64], const double (*rows_)[STAGE_ROW],
                                         int lane_, bool alive_)
      : p(p_), sA(sA_), rows(rows_), lane(lane_), alive(alive_), eps(p_.eps) {
    compat = (p.control == CONTROL_COMPAT);
    sampled = (p.t_eval != nullptr);
    q = compat ? 5.0 : (double)(c_tab[METHOD].order_err + 1);
    inv_q = 1.0 / q;
  }

  // Writes the lane's row result and frees the lane.
  __device__ __forceinline__ void finish(int code) {
    if (code != ROW_OK) {
      if (sampled) {
        for (int m = sidx; m < p.m_eval; ++m)
          for (int i = 0; i < 6; ++i) p.ysamp[((size_t)row * p.m_eval + m) * 6 + i] = __builtin_nan("");
      } else {
        for (int i = 0; i < 6; ++i) p.yf[(size_t)row * 6 + i] = __builtin_nan("");
      }
    } else if (!sampled) {
      for (int i = 0; i < 6; ++i) p.yf[(size_t)row * 6 + i] = y[i];
    }
    if (p.status) p.status[row] = code;
    if (p.steps) p.steps[row] = steps;
    if (p.evals) p.evals[row] = evals;
    if (p.rejects) p.rejects[row] = rejects;
    active = false;
  }

  // Step-size bounds and initial step of an integration over `span`
  // (solver.rs:385-391 default h0; lightyear_compat.rs:203-213 for compat).
  __device__ __forceinline__ void start_controller(double span) {
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
  }

  // Encke: a segment from total state `tot` at the current time, opened with
  // the carried step size `carry` when there is one (integrator.rs:341-402).
  __device__ __forceinline__ void start_segment(const double* tot, double carry) {
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
  }

  // Total (inertial) state of the lane at the current time.
  __device__ __forceinline__ void total_state(double* tot) {
    if constexpr (ENCKE) {
      double b[6];
      two_body(p.mu, el.base, t - el.t0, b, el.x);
#pragma unroll
      for (int i = 0; i < 6; ++i) tot[i] = b[i] + y[i];
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) tot[i] = y[i];
    }
  }

  // Takes rows for free lanes (the dealt first group `first_group`, then the
  // queue; the dealt round covers `dealt_rows` hand-out positions) and sets
  // each new row up.
  __device__ __forceinline__ void refill(unsigned first_group, unsigned long long dealt_rows) {
    const unsigned long long need = wave_ballot(!active);
    if (need == 0ull || exhausted) return;
    const int cnt = __popcll(need);
    unsigned int base;
    const bool dealt = first_fill;
    if (first_fill) {
      base = 64u * first_group;
      first_fill = false;
    } else {
      unsigned int b = 0;
      const int first = __ffsll((long long)need) - 1;
      if (lane == first) b = atomicAdd(p.queue, (unsigned)cnt);
      base = __shfl(b, first);
    }
    if (dealt)
      exhausted = dealt_rows >= (unsigned long long)p.n;
    else if (base + (unsigned)cnt >= (unsigned)p.n)
      exhausted = true;
    if (active) return;
    const unsigned pos = base + (unsigned)lane_rank(need);
    if (pos >= (unsigned)p.n) return;
    row = p.order ? p.order[pos] : (int)pos;
    active = true;
    steps = evals = rejects = rej_run = 0;
    sidx = 0;
    tcomp = 0.0;
    eprev = 0.0;
    have_prev = just_rej = have_k1 = false;
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
          finish(ROW_INVALID_INPUT);
        } else {
          start_segment(tot, 0.0);
          tseg = sampled ? fmin(p.t_eval[sidx], el.seg_end) : el.seg_end;
        }
      } else {
        tseg = sampled ? p.t_eval[sidx] : tend;
        start_controller(span);
      }
    }
  }

  // Phase A, first half: a group with no step in flight refills its lanes
  // and starts a step attempt; a group with no rows left dies.
  __device__ __forceinline__ void begin(unsigned first_group, unsigned long long dealt_rows) {
    if (!alive || in_step) return;
    while (true) {
      refill(first_group, dealt_rows);
      if (wave_ballot(active) == 0ull) {
        if (exhausted) {
          alive = false;
          return;
        }
        continue;
      }
      stepping = active;
      const double rem = tseg - t;
      hs = (h < rem) ? h : rem;
      lands = (hs == rem);
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
      // stage 1 (k1) is kept across rejects and reused after FSAL accepts:
      // the attempt starts at stage 0 only if some lane lacks it
      needk = stepping && !have_k1;
      s_cur = wave_ballot(needk) != 0ull ? 0 : 1;
      in_step = true;
      return;
    }
  }

  // Phase A, second half: the total (inertial) state of stage s_cur (second-
  // order form, as propagate_kernel) into yt; returns the stage time.
  __device__ __forceinline__ double stage_state(double* yt) {
    const int s = s_cur;
    st = (s == 0) ? needk : stepping;
    double ys[6];
    stage_sums<METHOD>(s, hs, y, sA, lane, rows, ys);
    const double tst = t + rows[s][1] * hs;
    if constexpr (ENCKE) {
      double b6[6];
      two_body(p.mu, el.base, tst - el.t0, b6, el.x);
#pragma unroll
      for (int i = 0; i < 6; ++i) yt[i] = b6[i] + ys[i];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        bst[i] = b6[i];
        ysr[i] = ys[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 6; ++i) yt[i] = ys[i];
    }
    return tst;
  }

  // Phase C: the stage's acceleration (central term included) is stored;
  // after the last stage the attempt is completed.
  __device__ __forceinline__ void store_stage(double ax, double ay, double az) {
    if (st) {
      sA[s_cur][0][lane] = ax;
      sA[s_cur][1][lane] = ay;
      sA[s_cur][2][lane] = az;
      ++evals;
    }
    if (s_cur == 0 && needk) have_k1 = true;
    ++s_cur;
    if (s_cur == S) {
      complete_step();
      in_step = false;
    }
  }

  // rk_step's sums, the error norm and the controller, writing rows that end.
  __device__ __forceinline__ void complete_step() {
    const bqp_tableau& TB = c_tab[METHOD];
    const SecondOrderTableau& SO = c_so[METHOD];
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
    const double hh = hs * hs;
    auto kahan = [](double& sum, double& c, double term) {
      const double yk = term - c;
      const double acc = sum + yk;
      c = (acc - sum) - yk;
      sum = acc;
    };
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const double bj = TB.b[j], ej = TB.e[j], baj = SO.ba[j], eaj = SO.ea[j];
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
        const double e3j = TB.e3[j], e3aj = SO.e3a[j];
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
    if (!stepping) return;
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
          if (t == el.seg_end) start_segment(tot, h);
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
  }
};

// First group (in hand-out positions) of team t of block b: the teams of the
// blocks are dealt in rounds, round t over blocks 0..C-1 in order, odd rounds
// in reverse, so each block's teams take ranks b, 2C-1-b, 2C+b, 4C-1-b of the
// longest-first order and the CUs carry nearly equal work.
__device__ __forceinline__ unsigned team_first_group(unsigned b, unsigned t, unsigned C) {
  return t * C + ((t & 1u) ? (C - 1u - b) : b);
}

template <int METHOD, int PREC>
__global__ void __launch_bounds__(64 * TEAM_WAVES) propagate_team_kernel(const PropParams p) {
  using MT = MethodTraits<METHOD>;
  constexpr int S = MT::S;
  constexpr bool ENCKE = (PREC == 4);
  static_assert(PREC == 0 || PREC == 4, "team kernel: plain FP64 Cowell or Encke");
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int NT = p.teams;  // 1, 2 or 4 teams per block
  const bool leader = w < NT;
  const int team = w;

  __shared__ double sA[TEAM_MAX][S][3][64];   // stage accelerations of each team
  __shared__ double sYt[TEAM_MAX][6][64];     // inertial stage states (total state for Encke)
  __shared__ double sTh[TEAM_MAX][64];        // body-frame angle at the stage time
  __shared__ double sBc[TEAM_MAX][64];        // Cd*A/m of the lanes that store the stage (else 0)
  __shared__ double sDrag[TEAM_MAX][3][64];   // drag accelerations of the stage
  __shared__ double sSlot[TEAM_SLOTS][3][64]; // gravity subtree sums of the iteration's work items
  __shared__ int sAlive[TEAM_MAX];
  __shared__ int sNext;                       // next work item of the iteration

  __shared__ double sRows[S][STAGE_ROW];
  __shared__ ValladoTable sVal;
  stage_rows_to_lds<METHOD>(sRows, threadIdx.x, blockDim.x);
  vallado_to_lds(&sVal, threadIdx.x, blockDim.x);
  __syncthreads();
  GroupLeader<METHOD, PREC> L(p, sA[leader ? team : 0], sRows, lane, leader);
  const unsigned first = leader ? team_first_group(blockIdx.x, (unsigned)team, gridDim.x) : 0u;
  const unsigned long long dealt_rows = 64ull * (unsigned long long)NT * gridDim.x;
  const bool drag_on = p.model.drag != DRAG_NONE;

  while (true) {
    // ------------------------------------------------------------ A
    if (leader) {
      L.begin(first, dealt_rows);
      if (L.alive) {
        double yt[6];
        const double tst = L.stage_state(yt);
#pragma unroll
        for (int i = 0; i < 6; ++i) sYt[team][i][lane] = yt[i];
        sTh[team][lane] = fma(p.model.omega, tst, p.model.theta0);
        sBc[team][lane] = L.st ? L.bc : 0.0;
      }
      if (lane == 0) sAlive[team] = L.alive ? 1 : 0;
    }
    if (w == 0 && lane == 0) sNext = 0;
    __syncthreads();
    // ------------------------------------------------------------ B
    // the live teams in index order; work items are the units (two or fewer
    // live teams) or aligned unit pairs (three or four) of every live team
    int live = 0, pos_of_team = -1;
    int live_id[TEAM_MAX];
#pragma unroll
    for (int k = 0; k < TEAM_MAX; ++k) {
      const int a = (k < NT) ? sAlive[k] : 0;
      if (leader && k == team && a) pos_of_team = live;
      live_id[k] = -1;
      if (a) {
#pragma unroll
        for (int m = 0; m < TEAM_MAX; ++m)
          if (m == live) live_id[m] = k;
        ++live;
      }
    }
    if (live == 0) break;
    const bool pairs = live > TEAM_PAIRS_ABOVE;
    const int per_team = pairs ? GRAV_UNITS / 2 : GRAV_UNITS;
    const int items = live * per_team;
    {
      int cur = -1;
      UnitPoint q{};
      while (true) {
        int it = 0;
        if (lane == 0) it = atomicAdd(&sNext, 1);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
        if (it >= items) break;
        const int lp = it / per_team, sub = it - lp * per_team;
        int k = -1;
#pragma unroll
        for (int m = 0; m < TEAM_MAX; ++m)
          if (m == lp) k = live_id[m];
        k = __builtin_amdgcn_readfirstlane(k);
        if (k != cur) {  // this team's stage position, turned into the body frame
          double sth, cth;
          sincos(sTh[k][lane], &sth, &cth);
          q = unit_point(p.field, sYt[k][0][lane], sYt[k][1][lane], sYt[k][2][lane], cth, sth);
          cur = k;
        }
        double px, py, pz;
        if (pairs) {
          double bx, by, bz;
          grav_unit_inertial(p.field, q, 2 * sub, px, py, pz);
          grav_unit_inertial(p.field, q, 2 * sub + 1, bx, by, bz);
          px += bx;
          py += by;
          pz += bz;
        } else {
          grav_unit_inertial(p.field, q, sub, px, py, pz);
        }
        sSlot[it][0][lane] = px;
        sSlot[it][1][lane] = py;
        sSlot[it][2][lane] = pz;
        // the wave that takes a team's first item also computes its drag
        // (LY/src/rhs.rs:1119-1163)
        if (drag_on && sub == 0) {
          const double b = sBc[k][lane];
          double dx = 0.0, dy = 0.0, dz = 0.0;
          if (b != 0.0) {
            const double y0 = sYt[k][0][lane], y1 = sYt[k][1][lane], y2 = sYt[k][2][lane];
            const double r = sqrt(fma(y0, y0, fma(y1, y1, y2 * y2)));
            const double rho = density(sVal, p.model.drag, r - p.model.r_atm);
            if (rho > 0.0) {
              const double vrx = fma(p.model.omega, y1, sYt[k][3][lane]);
              const double vry = fma(-p.model.omega, y0, sYt[k][4][lane]);
              const double vrz = sYt[k][5][lane];
              const double vr = sqrt(fma(vrx, vrx, fma(vry, vry, vrz * vrz)));
              const double kk = -0.5 * b * rho * vr;
              dx = kk * vrx;
              dy = kk * vry;
              dz = kk * vrz;
            }
          }
          sDrag[k][0][lane] = dx;
          sDrag[k][1][lane] = dy;
          sDrag[k][2][lane] = dz;
        }
      }
    }
    __syncthreads();
    // ------------------------------------------------------------ C
    if (leader && L.alive) {
      // the team's slots combined along the unit tree (in place): units in
      // pairs first when the items were single units, then pairs up to the
      // root
      const int base = pos_of_team * per_team;
      for (int stride = 1; stride < per_team; stride *= 2)
        for (int r = 0; r + stride < per_team; r += 2 * stride)
#pragma unroll
          for (int i = 0; i < 3; ++i) sSlot[base + r][i][lane] += sSlot[base + r + stride][i][lane];
      double ax = sSlot[base][0][lane], ay = sSlot[base][1][lane], az = sSlot[base][2][lane];
      if (drag_on) {
        ax += sDrag[team][0][lane];
        ay += sDrag[team][1][lane];
        az += sDrag[team][2][lane];
      }
      if constexpr (ENCKE) {
        const double y0 = sYt[team][0][lane], y1 = sYt[team][1][lane], y2 = sYt[team][2][lane];
        const double r2 = fma(y0, y0, fma(y1, y1, y2 * y2));
        const double k = p.mu / (r2 * sqrt(r2));
        ax = fma(k, y0, ax);
        ay = fma(k, y1, ay);
        az = fma(k, y2, az);
        battin(p.mu, L.bst, L.ysr, ax, ay, az);
      }
      L.store_stage(ax, ay, az);
    }
  }
}

// First group (in hand-out positions) of slot k of pair block b over nb
// blocks: slot 0 takes rank first_group(b) of the first nb (the fixed-G
// kernel's back-and-forth deal over the CUs), slot 1 the mirror rank
// 2 nb - 1 - that, so each block pairs an expensive group with a cheap one
// and CU c holds ranks c, 2C-1-c, 2C+c, 4C-1-c when nb = 2C.
__device__ __forceinline__ unsigned pair_first_group(unsigned b, unsigned k, unsigned nb, unsigned C) {
  const unsigned r0 = first_group(b, nb, C);
  return k == 0u ? r0 : 2u * nb - 1u - r0;
}

// Phase B of the pair kernel for wave w: with two live groups, shares
// 2(w/2) and 2(w/2)+1 of group w % 2; with one, share w of it.  Each share's
// partial gravity goes to sPart.  Returns false when no group is live.
template <int PREC>
__device__ __forceinline__ bool pair_shares(const PropParams& p, int w, int lane, const double (*sP)[3][64],
                                            double (*sPart)[PAIR_WAVES][3][64], const int* sAlive) {
  const int a0 = __builtin_amdgcn_readfirstlane(sAlive[0]);
  const int a1 = __builtin_amdgcn_readfirstlane(sAlive[1]);
  if (a0 + a1 == 0) return false;
  const bool both = (a0 + a1) == 2;
  const int g = both ? (w & 1) : (a0 ? 0 : 1);
  const int sh0 = both ? 2 * (w >> 1) : w;
  const int nsh = both ? 2 : 1;
  const double xb = sP[g][0][lane], yb = sP[g][1][lane], zb = sP[g][2][lane];
  for (int k = 0; k < nsh; ++k) {
    const int share = sh0 + k;
    double gx, gy, gz;
    if constexpr (PREC == 1)
      grav_split<PAIR_WAVES>(p.field, share, xb, yb, zb, gx, gy, gz);
    else
      grav_fp64<PAIR_WAVES>(p.field, share, xb, yb, zb, gx, gy, gz);
    sPart[g][share][0][lane] = gx;
    sPart[g][share][1][lane] = gy;
    sPart[g][share][2][lane] = gz;
  }
  return true;
}

template <int METHOD, int PREC>
__global__ void __launch_bounds__(64 * PAIR_WAVES) __attribute__((amdgpu_waves_per_eu(BQP_WAVES_PER_EU)))
propagate_pair_kernel(const PropParams p) {
  using MT = MethodTraits<METHOD>;
  constexpr int S = MT::S;
  constexpr bool ENCKE = (PREC == 4);
  static_assert(PREC == 0 || PREC == 1 || PREC == 4, "pair kernel: FP64 or split Cowell, or Encke");
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool leader = w < PAIR_GROUPS;  // wave g leads group g
  const int grp = w;

  __shared__ double sA[PAIR_GROUPS][S][3][64];              // stage accelerations of each group
  __shared__ double sP[PAIR_GROUPS][3][64];                 // body-frame stage positions
  __shared__ double sPart[PAIR_GROUPS][PAIR_WAVES][3][64];  // gravity of each column share
  __shared__ int sAlive[PAIR_GROUPS];

  __shared__ double sRows[S][STAGE_ROW];
  __shared__ ValladoTable sVal;
  stage_rows_to_lds<METHOD>(sRows, threadIdx.x, blockDim.x);
  vallado_to_lds(&sVal, threadIdx.x, blockDim.x);
  __syncthreads();
  if (!leader) {
    // waves 2..7 only compute column shares: the leaders' two barriers per
    // round, with no integrator state live across the column sweeps (so
    // the sweeps keep their scalar registers; in the shared loop the
    // compiler restored 16 spilled SGPRs from VGPR lanes after every column)
    while (true) {
      __syncthreads();  // A: positions and live flags published
      if (!pair_shares<PREC>(p, w, lane, sP, sPart, sAlive)) return;
      __syncthreads();  // B: partial sums written
    }
  }
  GroupLeader<METHOD, PREC> L(p, sA[grp], sRows, lane, true);
  const unsigned first = pair_first_group(blockIdx.x, (unsigned)grp, gridDim.x, (unsigned)p.deal_cus);
  const unsigned long long dealt_rows = 64ull * PAIR_GROUPS * gridDim.x;
  // the leader's stage state and frame rotation, kept from phase A to C
  double yt[6] = {0, 0, 0, 0, 0, 0};
  double cth = 1.0, sth = 0.0;

  while (true) {
    // ------------------------------------------------------------ A
    L.begin(first, dealt_rows);
    if (L.alive) {
      const double tst = L.stage_state(yt);
      sincos(fma(p.model.omega, tst, p.model.theta0), &sth, &cth);
      sP[grp][0][lane] = fma(cth, yt[0], sth * yt[1]);
      sP[grp][1][lane] = fma(cth, yt[1], -(sth * yt[0]));
      sP[grp][2][lane] = yt[2];
    }
    if (lane == 0) sAlive[grp] = L.alive ? 1 : 0;
    __syncthreads();  // A
    // ------------------------------------------------------------ B
    if (!pair_shares<PREC>(p, w, lane, sP, sPart, sAlive)) break;
    __syncthreads();  // B
    // ------------------------------------------------------------ C
    // the shares in share order, then the rest of accel_eval's arithmetic,
    // so each acceleration equals the fixed-G kernel's at G = 8
    if (L.alive) {
      double gx = sPart[grp][0][0][lane], gy = sPart[grp][0][1][lane], gz = sPart[grp][0][2][lane];
#pragma unroll
      for (int s = 1; s < PAIR_WAVES; ++s) {
        gx += sPart[grp][s][0][lane];
        gy += sPart[grp][s][1][lane];
        gz += sPart[grp][s][2][lane];
      }
      double ax = fma(cth, gx, -(sth * gy));
      double ay = fma(sth, gx, cth * gy);
      double az = gz;
      if (p.model.drag != DRAG_NONE && L.bc != 0.0) {
        const double r = sqrt(fma(yt[0], yt[0], fma(yt[1], yt[1], yt[2] * yt[2])));
        const double rho = density(sVal, p.model.drag, r - p.model.r_atm);
        if (rho > 0.0) {
          const double vrx = fma(p.model.omega, yt[1], yt[3]);
          const double vry = fma(-p.model.omega, yt[0], yt[4]);
          const double vrz = yt[5];
          const double vr = sqrt(fma(vrx, vrx, fma(vry, vry, vrz * vrz)));
          const double k = -0.5 * L.bc * rho * vr;
          ax = fma(k, vrx, ax);
          ay = fma(k, vry, ay);
          az = fma(k, vrz, az);
        }
      }
      if constexpr (ENCKE) {
        const double r2 = fma(yt[0], yt[0], fma(yt[1], yt[1], yt[2] * yt[2]));
        const double k = p.mu / (r2 * sqrt(r2));
        ax = fma(k, yt[0], ax);
        ay = fma(k, yt[1], ay);
        az = fma(k, yt[2], az);
        battin(p.mu, L.bst, L.ysr, ax, ay, az);
      }
      L.store_stage(ax, ay, az);
    }
  }
}

}  // namespace bqp
