/* Second-order rearrangement of an explicit Runge-Kutta tableau for
 * y = (r, v), y' = (v, a(t, r, v)).  The stage velocities are
 * V_s = v + h sum_j a_sj A_j, so
 *   R_s = r + h c^_s v + h^2 sum_j (A.A)_sj A_j        (c^_s = sum_j a_sj)
 *   r'  = r + h (sum b) v + h^2 sum_j (b.A)_j A_j,  v' = v + h sum_j b_j A_j
 * and the same for the error weights: only the stage accelerations A_j are
 * stored.  make_second_order forms the products in long double on the host;
 * gen_second_order.cpp prints them as the compiled-in constant table
 * second_order_tables.h.
 */
#ifndef BQP_SECOND_ORDER_H
#define BQP_SECOND_ORDER_H

#include "tableaus.h"

namespace bqp {

// Stage sums read the earlier stages j < s in chunks of four; chunk k of a
// stage row holds {a_sj, (A.A)_sj} for j = 4k .. 4k+3, interleaved, so one
// chunk is one 64-byte scalar load.
constexpr int STAGE_CHUNK = 4;
constexpr int STAGE_CHUNKS = (BQP_MAX_STAGES - 1 + STAGE_CHUNK - 1) / STAGE_CHUNK;

struct SecondOrderTableau {
  double csum[BQP_MAX_STAGES];                  // sum_j a_sj
  double a2[BQP_MAX_STAGES][BQP_MAX_STAGES];    // (A.A)_sj
  double ba[BQP_MAX_STAGES];                    // (b.A)_j
  double ea[BQP_MAX_STAGES];                    // (e.A)_j
  double e3a[BQP_MAX_STAGES];                   // (e3.A)_j
  double bsum, esum, e3sum;
  double sc[BQP_MAX_STAGES][STAGE_CHUNKS][2 * STAGE_CHUNK];  // packed stage rows (zero for j >= s)
};

inline void make_second_order(const bqp_tableau& t, SecondOrderTableau& o) {
  const int S = t.stages;
  o = SecondOrderTableau{};
  long double bs = 0, es = 0, e3s = 0;
  for (int i = 0; i < S; ++i) {
    long double cs = 0, ba = 0, ea = 0, e3a = 0;
    for (int j = 0; j < S; ++j) {
      cs += (long double)t.a[i][j];
      long double a2 = 0;
      for (int k = 0; k < S; ++k) a2 += (long double)t.a[i][k] * (long double)t.a[k][j];
      o.a2[i][j] = (double)a2;
      ba += (long double)t.b[j] * (long double)t.a[j][i];
      ea += (long double)t.e[j] * (long double)t.a[j][i];
      e3a += (long double)t.e3[j] * (long double)t.a[j][i];
    }
    o.csum[i] = (double)cs;
    o.ba[i] = (double)ba;
    o.ea[i] = (double)ea;
    o.e3a[i] = (double)e3a;
    bs += t.b[i];
    es += t.e[i];
    e3s += t.e3[i];
  }
  o.bsum = (double)bs;
  o.esum = (double)es;
  o.e3sum = (double)e3s;
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < i && j < BQP_MAX_STAGES - 1; ++j) {
      o.sc[i][j / STAGE_CHUNK][2 * (j % STAGE_CHUNK)] = t.a[i][j];
      o.sc[i][j / STAGE_CHUNK][2 * (j % STAGE_CHUNK) + 1] = o.a2[i][j];
    }
}

}  // namespace bqp

#endif
