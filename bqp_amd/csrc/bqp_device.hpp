// Device-side physics for the batched propagator (gfx950, one trajectory per lane).
//
// Gravity: fully normalised Cunningham V/W recursion (Montenbruck & Gill,
// "Satellite Orbits" Sec. 3.2.4-3.2.5, normalised per Holmes & Featherstone
// 2002) evaluated as an m-outer column sweep.  Each lane keeps only the two
// previous entries of the current column, so the per-lane state is O(1)
// instead of the reference's dense (N+2)^2 V/W table
// (SC/src/gravity.rs:758-933 fill, :1719-1856 summation).
//
// The summation is regrouped by V/W entry instead of by (C,S) term: entry
// (n, j) contributes to the acceleration through the coefficients of the
// three terms (n-1, j-1), (n-1, j), (n-1, j+1) that read it in the M-G sums
// (eq. 3.33).  The host folds those coefficients, the normalisation factors
// and GM/R^2 into 6 numbers per entry, so the lane consumes every V/W value
// the moment it is produced:
//
//   entry (n, j) = { a, b, XV, XW, YV, YW, ZV, ZW }      (64 bytes)
//   V(n,j) = a*z~*V(n-1,j) - b*rho~*V(n-2,j)   (same for W)
//   ax += XV*V + XW*W;  ay += YV*V + YW*W;  az += ZV*V + ZW*W
//
// For the column's first (sectoral) entry, a holds the sectoral factor.
// The FP64 kernels use the rescaled form of this table (GravEntryL) and run
// the recursion of columns j >= 1 once, on the factor Q(n) = V(n,j)/V(j,j)
// = W(n,j)/W(j,j), multiplying each column's sums by its sectoral value at
// the column's end (BQP_QSTEP).
// All entries are wave-uniform, so the table is read through scalar loads
// (SGPR operands), not per-lane memory traffic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tableaus.h"

namespace bqp {

struct GravEntry {
  double a, b, xv, xw, yv, yw, zv, zw;
};

struct GravEntryF {
  float a, b, xv, xw, yv, yw, zv, zw;
};

// Rescaled form of the FP64 table.  Within column m, with s_n the product of
// the recursion coefficients a_km (k = m+1..n; s = 1 on the column's first
// entry), U(n) = V(n)/s_n obeys
//   U(n) = zt U(n-1) - g_n rho U(n-2),   g_n = b_n / (a_n a_{n-1}),
// and the summation coefficients absorb s_n (xv = XV_n s_n, ...).  One
// multiply per entry fewer than the (a, b) form, and 7 doubles instead of 8.
struct GravEntryL {
  double g, xv, xw, yv, yw, zv, zw;
};

// FP32 copy of the rescaled table (split precision, degree 30..90: the
// running products s_n stay below ~1e28 there, so the rescaled values U =
// V/s_n stay far inside the FP32 normal range).
struct GravEntryLF {
  float g, xv, xw, yv, yw, zv, zw;
};

// The table is read through the constant address space: loads from it are
// invariant by definition, so uniform reads compile to scalar (SMEM) loads
// even in kernels that also store to global memory.
// (The host pass of hipcc parses device code without address spaces.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) GravEntry* CEntry;
typedef const __attribute__((address_space(4))) GravEntryF* CEntryF;
typedef const __attribute__((address_space(4))) GravEntryL* CEntryL;
typedef const __attribute__((address_space(4))) GravEntryLF* CEntryLF;
#else
typedef const GravEntry* CEntry;
typedef const GravEntryF* CEntryF;
typedef const GravEntryL* CEntryL;
typedef const GravEntryLF* CEntryLF;
#endif

// Offset of column j in the entry table of a degree-N field.  Column 0 holds
// n = 1..N+1 (N+1 entries); column j >= 1 holds n = j..N+1 (N+2-j entries).
__host__ __device__ inline int col_offset(int N, int j) {
  return j == 0 ? 0 : (N + 1) + (j - 1) * (N + 2) - ((j - 1) * j) / 2;
}

// Degree split of the split-precision form: entries with n <= SPLIT_N (field
// degree n-1 <= 4: central term and J2..J4) stay FP64, the rest run in FP32.
constexpr int SPLIT_N = 5;

struct FieldView {
  const GravEntry* tab;    // FP64 entries, all columns ((a, b) form: split precision head)
  const GravEntryF* tabf;  // FP32 copy of the same entries (split precision)
  const GravEntryL* tabl;  // FP64 entries in the rescaled form (plain FP64 recursion)
  const GravEntryLF* tablf;  // FP32 copy of tabl (split precision, degree 30..90), or null
  int nmax;                // field degree N
  int jmax;                // last column = min(N+1, M+1)
  double radius;
};

// Owner wave of column j when G waves share a trajectory group: columns are
// dealt in a back-and-forth ("snake") order so long and short columns mix.
template <int G>
__device__ __forceinline__ int col_owner(int j) {
  if (G == 1) return 0;
  const int k = j / G, r = j - k * G;
  return (k & 1) ? (G - 1 - r) : r;
}

// Sectoral factors s_1 = sqrt(3), s_j = sqrt((2j+1)/(2j)) depend only on j;
// the split-precision form's sectoral chain reads them from constant memory
// (filled at context creation).  The FP64 form has them folded into its
// table (products S_j per column).
constexpr int SECT_MAX = 208;
static __constant__ double c_sect[SECT_MAX];
static __constant__ float c_sectf[SECT_MAX];

// One recursion step of column 0 (W = 0) in the rescaled form: U(n) from
// U(n-1), U(n-2), then the entry's three terms.  Writes the new value over
// U(n-2)'s register so a pair of steps needs no moves.
#define BQP_COL0_STEP(E, VOUT, VPREV)    \
  {                                      \
    const double C_ = (E).g * rho;       \
    VOUT = fma(zt, VPREV, -(C_ * VOUT)); \
    sx = fma((E).xv, VOUT, sx);          \
    sy = fma((E).yv, VOUT, sy);          \
    sz = fma((E).zv, VOUT, sz);          \
  }

// Factored column sweep.  Within column j >= 1, V and W obey the same
// recursion in n with the same coefficients and start from V(j-1,j) =
// W(j-1,j) = 0, so V(n,j) = Q(n) V(j,j) and W(n,j) = Q(n) W(j,j), where
// Q(j) = 1, Q(j-1) = 0 and Q(n) = zt Q(n-1) - g_n rho Q(n-2) (the rescaled
// recursion, run once per column instead of once each for V and W).  The
// column's six sums run over Q alone and are multiplied by the sectoral
// value (V(j,j), W(j,j)) when the column ends: 9 FP64 operations per entry
// instead of 11.
#define BQP_QSTEP(E, QOUT, QPREV)        \
  {                                      \
    const double C_ = (E).g * rho;       \
    QOUT = fma(zt, QPREV, -(C_ * QOUT)); \
    sx = fma((E).xv, QOUT, sx);          \
    tx = fma((E).xw, QOUT, tx);          \
    sy = fma((E).yv, QOUT, sy);          \
    ty = fma((E).yw, QOUT, ty);          \
    sz = fma((E).zv, QOUT, sz);          \
    tz = fma((E).zw, QOUT, tz);          \
  }
// Opens a column's sums with its sectoral entry (Q(j) = 1).
#define BQP_QOPEN(E) \
  {                  \
    sx = (E).xv;     \
    tx = (E).xw;     \
    sy = (E).yv;     \
    ty = (E).yw;     \
    sz = (E).zv;     \
    tz = (E).zw;     \
  }
// Opens a column's sums with its first recursion entry (Q(j+1) = zt).
#define BQP_QFIRST(E)     \
  {                       \
    sx = (E).xv * zt;     \
    tx = (E).xw * zt;     \
    sy = (E).yv * zt;     \
    ty = (E).yw * zt;     \
    sz = (E).zv * zt;     \
    tz = (E).zw * zt;     \
  }
// Adds the sectoral entry's terms (Q(j) = 1) to open sums.
#define BQP_QSECT(E) \
  {                  \
    sx += (E).xv;    \
    tx += (E).xw;    \
    sy += (E).yv;    \
    ty += (E).yw;    \
    sz += (E).zv;    \
    tz += (E).zw;    \
  }
// Adds a closed column, sums times its sectoral value (VS, WS), to the
// running acceleration (AX, AY, AZ).
#define BQP_QFOLD(VS, WS, AX, AY, AZ)  \
  {                                    \
    AX = fma(VS, sx, fma(WS, tx, AX)); \
    AY = fma(VS, sy, fma(WS, ty, AY)); \
    AZ = fma(VS, sz, fma(WS, tz, AZ)); \
  }

// (ar + i ai) *= (br + i bi)
__device__ __forceinline__ void cmul(double& ar, double& ai, double br, double bi) {
  const double re = fma(ar, br, -(ai * bi));
  ai = fma(ar, bi, ai * br);
  ar = re;
}

// Software-pipelined table reads.  Scalar loads may return out of order, so
// a wave can only wait for all of its outstanding ones (lgkmcnt(0)); the
// loops below therefore wait, issue the loads of the NEXT batch of entries,
// and only then compute the batch that has arrived, so each batch's L2
// latency overlaps the previous batch's arithmetic.  The explicit waits and
// scheduling barriers keep the compiler from hoisting the next batch's loads
// above the wait (the compiler still inserts any wait a use needs, so the
// explicit ones only add waiting, never remove it).
#define BQP_WAIT_SMEM() __builtin_amdgcn_s_waitcnt(0xc07f) /* lgkmcnt(0), vmcnt/expcnt untouched */
#define BQP_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)

// Below this degree the columns are too short for the pipeline's prologue
// and epilogue to pay (measured: degree 20 is 13% slower pipelined, degree
// 70 is 17% faster, degree 120 17% faster).
constexpr int PIPE_MIN_N = 30;

// grav_fp64 with pipelined table reads (same terms and operation order per
// column as grav_fp64's unpipelined loop; the sectoral entry's terms are
// added before the column's recursion terms instead of after its first
// three).  Column 0 runs in 4-entry batches; every other column loads its
// sectoral entry with its first two recursion entries (one exposed wait per
// column), then 2-entry batches with the next batch in flight; its last one
// to three entries are issued with its last full batch.  Reads at most two
// entries past a column's end (the table carries padding entries).
template <int G>
__device__ __forceinline__ void grav_fp64_pipelined(const FieldView& f, int g, double x, double y, double z,
                                                    double& ax, double& ay, double& az) {
  const double R = f.radius;
  const double r2 = fma(x, x, fma(y, y, z * z));
  const double rr = 1.0 / r2;
  const double xt = x * R * rr, yt = y * R * rr, zt = z * R * rr, rho = R * R * rr;
  const double vs0 = R / sqrt(r2);
  double sx = 0.0, sy = 0.0, sz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
  const int top = f.nmax + 1;
  CEntryL e = (CEntryL)f.tabl;
  if (g == 0) {  // column 0 (top >= PIPE_MIN_N + 1 entries)
    double v1 = vs0, v2 = 0.0;
    CEntryL r = e;
    int k = top;
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    GravEntryL a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
    BQP_SCHED_FENCE();
    while (k >= 8) {
      BQP_WAIT_SMEM();
      BQP_SCHED_FENCE();
      const GravEntryL b0 = r[4], b1 = r[5], b2 = r[6], b3 = r[7];
      BQP_SCHED_FENCE();
      BQP_COL0_STEP(a0, v2, v1);
      BQP_COL0_STEP(a1, v1, v2);
      BQP_COL0_STEP(a2, v2, v1);
      BQP_COL0_STEP(a3, v1, v2);
      BQP_WAIT_SMEM();
      BQP_SCHED_FENCE();
      a0 = r[8];  // up to four entries past column 0: column 1 follows it
      a1 = r[9];
      a2 = r[10];
      a3 = r[11];
      BQP_SCHED_FENCE();
      BQP_COL0_STEP(b0, v2, v1);
      BQP_COL0_STEP(b1, v1, v2);
      BQP_COL0_STEP(b2, v2, v1);
      BQP_COL0_STEP(b3, v1, v2);
      r += 8;
      k -= 8;
    }
    // k in 0..7; a holds r[0..3]
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    GravEntryL b0, b1, b2;
    if (k >= 5) b0 = r[4];
    if (k >= 6) b1 = r[5];
    if (k >= 7) b2 = r[6];
    BQP_SCHED_FENCE();
    if (k >= 1) BQP_COL0_STEP(a0, v2, v1);
    if (k >= 2) BQP_COL0_STEP(a1, v1, v2);
    if (k >= 3) BQP_COL0_STEP(a2, v2, v1);
    if (k >= 4) BQP_COL0_STEP(a3, v1, v2);
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    if (k >= 5) BQP_COL0_STEP(b0, v2, v1);
    if (k >= 6) BQP_COL0_STEP(b1, v1, v2);
    if (k >= 7) BQP_COL0_STEP(b2, v2, v1);
  }
  // columns j >= 1, reached by sectoral jumps as in grav_fp64
  const int ga = 2 * G - 1 - 2 * g, gb = 2 * g + 1;
  double zgr = 1.0, zgi = 0.0;  // z^g
  for (int k = 0; k < g; ++k) cmul(zgr, zgi, xt, yt);
  double zbr = zgr, zbi = zgi;  // z^(2g+1)
  cmul(zbr, zbi, zgr, zgi);
  cmul(zbr, zbi, xt, yt);
  double zar = xt, zai = yt;  // z^(2G-1-2g)
  for (int k = 1; k < ga; ++k) cmul(zar, zai, xt, yt);
  CEntryL e1 = e + top;
  int j = g;
  bool next_a = g != 0;
  double vs = vs0 * zgr, ws = vs0 * zgi;
  if (g == 0) {
    j = ga;
    vs = vs0 * zar;
    ws = vs0 * zai;
    next_a = false;
  }
  // column 0's terms (wave 0) open the running sums; the other columns are
  // folded in as they close
  double gx = sx, gy = sy, gz = sz;
  while (j <= f.jmax) {
    const CEntryL ec = e1 + (col_offset(f.nmax, j) - top);
    int kk = top - j;  // recursion entries of column j
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    const GravEntryL es = ec[0];
    GravEntryL a0 = ec[1], a1 = ec[2];  // past the end of a column shorter than 3
    BQP_SCHED_FENCE();
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    double q1 = 1.0, q2 = 0.0;  // Q(j), Q(j-1)
    BQP_QOPEN(es);
    CEntryL r = ec + 3;
    if (kk >= 2) {
      kk -= 2;  // entries after a0, a1
      while (kk >= 4) {
        BQP_WAIT_SMEM();
        BQP_SCHED_FENCE();
        const GravEntryL b0 = r[0], b1 = r[1];
        BQP_SCHED_FENCE();
        BQP_QSTEP(a0, q2, q1);
        BQP_QSTEP(a1, q1, q2);
        BQP_WAIT_SMEM();
        BQP_SCHED_FENCE();
        a0 = r[2];
        a1 = r[3];
        BQP_SCHED_FENCE();
        BQP_QSTEP(b0, q2, q1);
        BQP_QSTEP(b1, q1, q2);
        r += 4;
        kk -= 4;
      }
      // the last 0..3 entries: two are loaded unconditionally while a0, a1
      // are computed (as in the loop above, four entries live at once; more
      // would spill scalar registers to VGPR lanes), a third afterwards.
      // Past a column's end these read the next column's entries or the
      // table's padding (at most two entries past it).
      BQP_WAIT_SMEM();
      BQP_SCHED_FENCE();
      const GravEntryL t0 = r[0], t1 = r[1];
      BQP_SCHED_FENCE();
      BQP_QSTEP(a0, q2, q1);
      BQP_QSTEP(a1, q1, q2);
      if (kk >= 1) {
        BQP_WAIT_SMEM();
        BQP_SCHED_FENCE();
        const GravEntryL t2 = r[2];
        BQP_SCHED_FENCE();
        BQP_QSTEP(t0, q2, q1);
        if (kk >= 2) BQP_QSTEP(t1, q1, q2);
        if (kk >= 3) {
          BQP_WAIT_SMEM();
          BQP_SCHED_FENCE();
          BQP_QSTEP(t2, q2, q1);
        }
      }
    } else if (kk == 1) {
      BQP_QSTEP(a0, q2, q1);
    }
    BQP_QFOLD(vs, ws, gx, gy, gz);
    const double pr = next_a ? zar : zbr, pi = next_a ? zai : zbi;
    j += next_a ? ga : gb;
    next_a = !next_a;
    cmul(vs, ws, pr, pi);
  }
  ax = gx;
  ay = gy;
  az = gz;
}

// FP64 gravity, body-fixed position (m) -> acceleration (m/s^2), for the
// columns owned by wave g of G.  The caller sums the G partial results.
// Recursion entries are consumed four per scalar-load batch; fields of
// degree >= PIPE_MIN_N take the pipelined loops above.
template <int G>
__device__ __forceinline__ void grav_fp64(const FieldView& f, int g, double x, double y, double z,
                                          double& ax, double& ay, double& az) {
  if (f.nmax >= PIPE_MIN_N) {
    grav_fp64_pipelined<G>(f, g, x, y, z, ax, ay, az);
    return;
  }
  const double R = f.radius;
  const double r2 = fma(x, x, fma(y, y, z * z));
  const double rr = 1.0 / r2;
  const double xt = x * R * rr, yt = y * R * rr, zt = z * R * rr, rho = R * R * rr;
  const double vs0 = R / sqrt(r2);
  // column sums: V-coefficient terms in s, W-coefficient terms in t (six
  // independent accumulation chains per column)
  double sx = 0.0, sy = 0.0, sz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
  const int top = f.nmax + 1;
  CEntryL e = (CEntryL)f.tabl;
  if (g == 0) {  // column 0 belongs to wave 0
    double v1 = vs0, v2 = 0.0;
    CEntryL r = e;
    int k = top;
    for (; k >= 4; k -= 4, r += 4) {
      const GravEntryL e0 = r[0], e1 = r[1], e2 = r[2], e3 = r[3];
      BQP_COL0_STEP(e0, v2, v1);
      BQP_COL0_STEP(e1, v1, v2);
      BQP_COL0_STEP(e2, v2, v1);
      BQP_COL0_STEP(e3, v1, v2);
    }
    for (; k >= 2; k -= 2, r += 2) {
      const GravEntryL e0 = r[0], e1 = r[1];
      BQP_COL0_STEP(e0, v2, v1);
      BQP_COL0_STEP(e1, v1, v2);
    }
    if (k) {
      const GravEntryL e0 = r[0];
      BQP_COL0_STEP(e0, v2, v1);
    }
  }
  // Columns j >= 1.  The sectoral value of column j is
  //   V(j,j) + i W(j,j) = S_j (x~ + i y~)^j R/r,  S_j = s_1 s_2 ... s_j,
  // and the host folds S_j into the column's coefficients, so a wave goes
  // from one owned column to the next by one complex multiply with
  // z^gap (z = x~ + i y~) instead of walking every column in between.
  // Snake dealing gives wave g the columns g, 2G-1-g, 2G+g, 4G-1-g, ...:
  // after the first, the gaps alternate ga = 2G-1-2g and gb = 2g+1.
  const int ga = 2 * G - 1 - 2 * g, gb = 2 * g + 1;
  double zgr = 1.0, zgi = 0.0;  // z^g
  for (int k = 0; k < g; ++k) cmul(zgr, zgi, xt, yt);
  double zbr = zgr, zbi = zgi;  // z^(2g+1)
  cmul(zbr, zbi, zgr, zgi);
  cmul(zbr, zbi, xt, yt);
  double zar = xt, zai = yt;  // z^(2G-1-2g)
  for (int k = 1; k < ga; ++k) cmul(zar, zai, xt, yt);
  CEntryL e1 = e + top;  // column 1
  int j = g;
  bool next_a = g != 0;  // wave 0 starts from column 0 and takes gap ga first
  double vs = vs0 * zgr, ws = vs0 * zgi;
  if (g == 0) {
    j = ga;
    vs = vs0 * zar;
    ws = vs0 * zai;
    next_a = false;
  }
  // column 0's terms (wave 0) open the running sums; the other columns are
  // folded in as they close
  double gx = sx, gy = sy, gz = sz;
  while (j <= f.jmax) {
    const int len = top + 1 - j;
    const CEntryL ec = e1 + (col_offset(f.nmax, j) - top);
    // The sectoral entry shares its scalar-load batch with the column's first
    // three recursion entries (one wait for four entries; its terms are added
    // after the three steps, so the compiler cannot hoist its load and wait
    // above the branch).  The first recursion entry has Q(j+1) = zt, so its
    // terms are plain products that open the sums.
    CEntryL r = ec + 1;
    int kk = len - 1;
    double q1, q2;
    if (kk >= 3) {
      const GravEntryL es = ec[0], e0 = r[0], e1q = r[1], e2 = r[2];
      q1 = zt;   // Q(j+1)
      q2 = 1.0;  // Q(j)
      BQP_QFIRST(e0);
      BQP_QSTEP(e1q, q2, q1);
      BQP_QSTEP(e2, q1, q2);
      BQP_QSECT(es);
      kk -= 3;
      r += 3;
    } else {
      const GravEntryL es = ec[0];
      BQP_QOPEN(es);
      q1 = 1.0;
      q2 = 0.0;
    }
    for (; kk >= 4; kk -= 4, r += 4) {
      const GravEntryL e0 = r[0], e1q = r[1], e2 = r[2], e3 = r[3];
      BQP_QSTEP(e0, q2, q1);
      BQP_QSTEP(e1q, q1, q2);
      BQP_QSTEP(e2, q2, q1);
      BQP_QSTEP(e3, q1, q2);
    }
    for (; kk >= 2; kk -= 2, r += 2) {
      const GravEntryL e0 = r[0], e1q = r[1];
      BQP_QSTEP(e0, q2, q1);
      BQP_QSTEP(e1q, q1, q2);
    }
    if (kk) {
      const GravEntryL e0 = r[0];
      BQP_QSTEP(e0, q2, q1);
    }
    BQP_QFOLD(vs, ws, gx, gy, gz);
    // next owned column (wave-uniform gap; the select is on a scalar flag)
    const double pr = next_a ? zar : zbr, pi = next_a ? zai : zbi;
    j += next_a ? ga : gb;
    next_a = !next_a;
    cmul(vs, ws, pr, pi);
  }
  ax = gx;
  ay = gy;
  az = gz;
}

// FP32 recursion step of the split form ((a, b) entries); the new value
// overwrites the n-2 register as in BQP_QSTEP.
#define BQP_COL_STEP_F(E, VOUT, VPREV, WOUT, WPREV) \
  {                                                 \
    const float A_ = (E).a * ztf, B_ = (E).b * rhof; \
    VOUT = fmaf(A_, VPREV, -(B_ * VOUT));           \
    WOUT = fmaf(A_, WPREV, -(B_ * WOUT));           \
    px = fmaf((E).xv, VOUT, px);                    \
    px = fmaf((E).xw, WOUT, px);                    \
    py = fmaf((E).yv, VOUT, py);                    \
    py = fmaf((E).yw, WOUT, py);                    \
    pz = fmaf((E).zv, VOUT, pz);                    \
    pz = fmaf((E).zw, WOUT, pz);                    \
  }

// Split-precision gravity: entries with n <= SPLIT_N (degree <= 4) in FP64,
// the rest in FP32.  Each FP32 column sum is short (<= N+1 terms) and is
// folded into an FP32 Neumaier-compensated accumulator, whose value and
// compensation are added to the FP64 sum at the end.
struct Comp32 {
  float s, c;
};
__device__ __forceinline__ void comp_add(Comp32& acc, float v) {
  const float t = acc.s + v;
  const float big = fabsf(acc.s) >= fabsf(v) ? acc.s : v;
  const float small = fabsf(acc.s) >= fabsf(v) ? v : acc.s;
  acc.c += (big - t) + small;
  acc.s = t;
}

// Split-precision gravity over the (a, b) table: walks every column's
// sectoral chain (FP64 to SPLIT_N, FP32 after).  Used for fields of degree
// below PIPE_MIN_N or above SPLIT_RESCALED_MAX_N.
template <int G>
__device__ __forceinline__ void grav_split_ab(const FieldView& f, int g, double x, double y, double z,
                                              double& ax, double& ay, double& az) {
  const double R = f.radius;
  const double r2 = fma(x, x, fma(y, y, z * z));
  const double rr = 1.0 / r2;
  const double xt = x * R * rr, yt = y * R * rr, zt = z * R * rr, rho = R * R * rr;
  const float xtf = (float)xt, ytf = (float)yt, ztf = (float)zt, rhof = (float)rho;
  double vs = R / sqrt(r2), ws = 0.0;
  float vsf = 0.0f, wsf = 0.0f;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  Comp32 cx = {0.f, 0.f}, cy = {0.f, 0.f}, cz = {0.f, 0.f};
  const int top = f.nmax + 1;
  CEntry e = (CEntry)f.tab;
  CEntryF ef = (CEntryF)f.tabf;
  for (int j = 0; j <= f.jmax; ++j) {
    const int nlo = j == 0 ? 1 : j;
    const int len = top + 1 - nlo;
    if (j > 0) {
      if (j <= SPLIT_N) {
        const double s = c_sect[j];
        const double nv = s * fma(xt, vs, -(yt * ws));
        const double nw = s * fma(xt, ws, yt * vs);
        vs = nv;
        ws = nw;
        if (j == SPLIT_N) {
          vsf = (float)vs;
          wsf = (float)ws;
        }
      } else {
        const float s = c_sectf[j];
        const float nv = s * fmaf(xtf, vsf, -(ytf * wsf));
        const float nw = s * fmaf(xtf, wsf, ytf * vsf);
        vsf = nv;
        wsf = nw;
      }
    }
    if (col_owner<G>(j) != g) {
      e += len;
      ef += len;
      continue;
    }
    int n = nlo;
    float v1f, w1f, v2f, w2f;
    if (j <= SPLIT_N) {
      // FP64 head of the column: n = nlo .. SPLIT_N
      double v1, w1, v2 = 0.0, w2 = 0.0;
      if (j == 0) {
        v1 = vs;
        w1 = 0.0;
      } else {
        v1 = vs;
        w1 = ws;
        sx = fma(e->xv, vs, sx);
        sx = fma(e->xw, ws, sx);
        sy = fma(e->yv, vs, sy);
        sy = fma(e->yw, ws, sy);
        sz = fma(e->zv, vs, sz);
        sz = fma(e->zw, ws, sz);
        ++e;
        ++ef;
        ++n;
      }
      for (; n <= SPLIT_N && n <= top; ++n, ++e, ++ef) {
        const double A = e->a * zt, B = e->b * rho;
        const double v = fma(A, v1, -(B * v2));
        const double w = fma(A, w1, -(B * w2));
        sx = fma(e->xv, v, sx);
        sx = fma(e->xw, w, sx);
        sy = fma(e->yv, v, sy);
        sy = fma(e->yw, w, sy);
        sz = fma(e->zv, v, sz);
        sz = fma(e->zw, w, sz);
        v2 = v1;
        v1 = v;
        w2 = w1;
        w1 = w;
      }
      v1f = (float)v1;
      w1f = (float)w1;
      v2f = (float)v2;
      w2f = (float)w2;
    } else {
      // FP32 sectoral entry
      float px = ef->xv * vsf, py = ef->yv * vsf, pz = ef->zv * vsf;
      px = fmaf(ef->xw, wsf, px);
      py = fmaf(ef->yw, wsf, py);
      pz = fmaf(ef->zw, wsf, pz);
      comp_add(cx, px);
      comp_add(cy, py);
      comp_add(cz, pz);
      ++e;
      ++ef;
      ++n;
      v1f = vsf;
      w1f = wsf;
      v2f = 0.0f;
      w2f = 0.0f;
    }
    // FP32 tail of the column, four entries per scalar-load batch
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    {
      int kk = top + 1 - n;
      CEntryF r = ef;
      for (; kk >= 4; kk -= 4, r += 4) {
        const GravEntryF f0 = r[0], f1 = r[1], f2 = r[2], f3 = r[3];
        BQP_COL_STEP_F(f0, v2f, v1f, w2f, w1f);
        BQP_COL_STEP_F(f1, v1f, v2f, w1f, w2f);
        BQP_COL_STEP_F(f2, v2f, v1f, w2f, w1f);
        BQP_COL_STEP_F(f3, v1f, v2f, w1f, w2f);
      }
      for (; kk >= 2; kk -= 2, r += 2) {
        const GravEntryF f0 = r[0], f1 = r[1];
        BQP_COL_STEP_F(f0, v2f, v1f, w2f, w1f);
        BQP_COL_STEP_F(f1, v1f, v2f, w1f, w2f);
      }
      if (kk) {
        const GravEntryF f0 = r[0];
        BQP_COL_STEP_F(f0, v2f, v1f, w2f, w1f);
      }
      const int used = top + 1 - n;
      e += used;
      ef += used;
    }
    comp_add(cx, px);
    comp_add(cy, py);
    comp_add(cz, pz);
  }
  ax = sx + ((double)cx.s + (double)cx.c);
  ay = sy + ((double)cy.s + (double)cy.c);
  az = sz + ((double)cz.s + (double)cz.c);
}

// Largest degree for which the FP32 rescaled table is used (the running
// products s_n of a column grow like 2^n: at degree 90 below 1e28, so U =
// V/s_n stays above 1e-28, far from FP32's 1.2e-38 normal minimum).
constexpr int SPLIT_RESCALED_MAX_N = 90;

// FP32 column-0 step of the rescaled table (as BQP_COL0_STEP) into the FP32
// column partial sums q.
#define BQP_COL0_STEP_LF(E, VOUT, VPREV)   \
  {                                        \
    const float C_ = (E).g * rhof;         \
    VOUT = fmaf(ztf, VPREV, -(C_ * VOUT)); \
    qx = fmaf((E).xv, VOUT, qx);           \
    qy = fmaf((E).yv, VOUT, qy);           \
    qz = fmaf((E).zv, VOUT, qz);           \
  }
// FP32 step of the factored sweep (as BQP_QSTEP): one recursion Q, the
// column's V-coefficient sums in q, W-coefficient sums in u.
#define BQP_QSTEP_LF(E, QOUT, QPREV)           \
  {                                            \
    const float C_ = (E).g * rhof;             \
    QOUT = fmaf(ztf, QPREV, -(C_ * QOUT));     \
    qx = fmaf((E).xv, QOUT, qx);               \
    ux = fmaf((E).xw, QOUT, ux);               \
    qy = fmaf((E).yv, QOUT, qy);               \
    uy = fmaf((E).yw, QOUT, uy);               \
    qz = fmaf((E).zv, QOUT, qz);               \
    uz = fmaf((E).zw, QOUT, uz);               \
  }
// One step of either form: column 0 (no W half) carries V itself, the other
// columns carry the factored Q (W arguments unused).
#define BQP_STEP_LF(E, VOUT, VPREV, WOUT, WPREV) \
  {                                              \
    if constexpr (COL0) {                        \
      BQP_COL0_STEP_LF(E, VOUT, VPREV);          \
    } else {                                     \
      BQP_QSTEP_LF(E, VOUT, VPREV);              \
    }                                            \
  }

// FP32 recursion entries r[0..cnt) of one column, four per batch with the
// next batch in flight.  a0..a3 hold r[0..3] on entry (loaded by the caller
// with the column's head, or past the column's end when cnt < 4: the table
// is padded).  v1 is the latest value (V in column 0, Q in the others), v2
// the one before; w1, w2 are not used.
template <bool COL0>
__device__ __forceinline__ void col_tail_lf(CEntryLF r, int cnt, GravEntryLF a0, GravEntryLF a1, GravEntryLF a2,
                                            GravEntryLF a3, float ztf, float rhof, float& v1, float& v2, float& w1,
                                            float& w2, float& qx, float& ux, float& qy, float& uy, float& qz,
                                            float& uz) {
  while (cnt >= 8) {
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    const GravEntryLF b0 = r[4], b1 = r[5], b2 = r[6], b3 = r[7];
    BQP_SCHED_FENCE();
    BQP_STEP_LF(a0, v2, v1, w2, w1);
    BQP_STEP_LF(a1, v1, v2, w1, w2);
    BQP_STEP_LF(a2, v2, v1, w2, w1);
    BQP_STEP_LF(a3, v1, v2, w1, w2);
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    a0 = r[8];  // up to three entries past the column's end
    a1 = r[9];
    a2 = r[10];
    a3 = r[11];
    BQP_SCHED_FENCE();
    BQP_STEP_LF(b0, v2, v1, w2, w1);
    BQP_STEP_LF(b1, v1, v2, w1, w2);
    BQP_STEP_LF(b2, v2, v1, w2, w1);
    BQP_STEP_LF(b3, v1, v2, w1, w2);
    r += 8;
    cnt -= 8;
  }
  // cnt in 0..7; a0..a3 hold r[0..3]; the last 0..3 entries are issued
  // before a0..a3 are computed
  BQP_WAIT_SMEM();
  BQP_SCHED_FENCE();
  GravEntryLF b0, b1, b2;
  if (cnt >= 5) b0 = r[4];
  if (cnt >= 6) b1 = r[5];
  if (cnt >= 7) b2 = r[6];
  BQP_SCHED_FENCE();
  if (cnt >= 1) BQP_STEP_LF(a0, v2, v1, w2, w1);
  if (cnt >= 2) BQP_STEP_LF(a1, v1, v2, w1, w2);
  if (cnt >= 3) BQP_STEP_LF(a2, v2, v1, w2, w1);
  if (cnt >= 4) BQP_STEP_LF(a3, v1, v2, w1, w2);
  BQP_WAIT_SMEM();
  BQP_SCHED_FENCE();
  if (cnt >= 5) BQP_STEP_LF(b0, v2, v1, w2, w1);
  if (cnt >= 6) BQP_STEP_LF(b1, v1, v2, w1, w2);
  if (cnt >= 7) BQP_STEP_LF(b2, v2, v1, w2, w1);
}

// Split-precision gravity over the rescaled tables (degree PIPE_MIN_N ..
// SPLIT_RESCALED_MAX_N): the same column sweep, sectoral jumps and factored
// columns as grav_fp64_pipelined.  Entries with n <= SPLIT_N (the central
// term and J2..J4) run in FP64 from the FP64 table; the recursion continues
// in FP32 from the FP32 copy (Q values converted where a column crosses
// SPLIT_N).  Each column's FP32 terms accumulate in six FP32 sums over Q;
// the column's result, those sums times its sectoral value, is formed and
// accumulated across columns in FP64 (rounding 2^-53 per column, below the
// Neumaier-compensated FP32 accumulator this replaced, and no FP32 range
// limit on the sectoral values).
template <int G>
__device__ __forceinline__ void grav_split_pipelined(const FieldView& f, int g, double x, double y, double z,
                                                     double& ax, double& ay, double& az) {
  const double R = f.radius;
  const double r2 = fma(x, x, fma(y, y, z * z));
  const double rr = 1.0 / r2;
  const double xt = x * R * rr, yt = y * R * rr, zt = z * R * rr, rho = R * R * rr;
  const double vs0 = R / sqrt(r2);
  const float ztf = (float)zt, rhof = (float)rho;
  double sx = 0.0, sy = 0.0, sz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
  double gx = 0.0, gy = 0.0, gz = 0.0;  // FP64 accumulation of the column results
  const int top = f.nmax + 1;
  CEntryL e = (CEntryL)f.tabl;
  CEntryLF ef = (CEntryLF)f.tablf;
  if (g == 0) {  // column 0: n = k + 1 at entry k; FP64 for n <= SPLIT_N
    double v1 = vs0, v2 = 0.0;
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    const GravEntryL h0 = e[0], h1 = e[1], h2 = e[2], h3 = e[3], h4 = e[4];
    BQP_SCHED_FENCE();
    BQP_COL0_STEP(h0, v2, v1);
    BQP_COL0_STEP(h1, v1, v2);
    BQP_COL0_STEP(h2, v2, v1);
    BQP_COL0_STEP(h3, v1, v2);
    BQP_COL0_STEP(h4, v2, v1);
    float fv1 = (float)v2, fv2 = (float)v1, fw1 = 0.f, fw2 = 0.f;  // latest in v2
    float qx = 0.f, ux = 0.f, qy = 0.f, uy = 0.f, qz = 0.f, uz = 0.f;
    CEntryLF r = ef + SPLIT_N;
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    const GravEntryLF a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
    BQP_SCHED_FENCE();
    col_tail_lf<true>(r, top - SPLIT_N, a0, a1, a2, a3, ztf, rhof, fv1, fv2, fw1, fw2, qx, ux, qy, uy, qz, uz);
    gx = sx + (double)qx;
    gy = sy + (double)qy;
    gz = sz + (double)qz;
  }
  // columns j >= 1 by sectoral jumps (FP64), as in grav_fp64, each as the
  // factored sweep: Q(n) with Q(j) = 1, the FP64 head (n <= SPLIT_N) and the
  // FP32 rest summed per column, then multiplied by the sectoral value in FP64
  const int ga = 2 * G - 1 - 2 * g, gb = 2 * g + 1;
  double zgr = 1.0, zgi = 0.0;  // z^g
  for (int k = 0; k < g; ++k) cmul(zgr, zgi, xt, yt);
  double zbr = zgr, zbi = zgi;  // z^(2g+1)
  cmul(zbr, zbi, zgr, zgi);
  cmul(zbr, zbi, xt, yt);
  double zar = xt, zai = yt;  // z^(2G-1-2g)
  for (int k = 1; k < ga; ++k) cmul(zar, zai, xt, yt);
  int j = g;
  bool next_a = g != 0;
  double vs = vs0 * zgr, ws = vs0 * zgi;
  if (g == 0) {
    j = ga;
    vs = vs0 * zar;
    ws = vs0 * zai;
    next_a = false;
  }
  while (j <= f.jmax) {
    const int off = col_offset(f.nmax, j);  // the column's sectoral entry
    const int kk = top - j;                 // its recursion entries
    float fv1, fv2, fw1 = 0.f, fw2 = 0.f;
    float qx = 0.f, ux = 0.f, qy = 0.f, uy = 0.f, qz = 0.f, uz = 0.f;
    if (j <= SPLIT_N) {
      // FP64 sectoral term and recursion entries n = j+1 .. SPLIT_N
      CEntryL ec = e + off;
      const int h = SPLIT_N - j;
      BQP_WAIT_SMEM();
      BQP_SCHED_FENCE();
      const GravEntryL es = ec[0];
      BQP_SCHED_FENCE();
      BQP_QOPEN(es);
      double q1 = 1.0, q2 = 0.0;
      for (int i = 0; i < h; ++i) {
        const GravEntryL E = ec[1 + i];
        BQP_QSTEP(E, q2, q1);
        const double tq = q1;  // latest back into q1
        q1 = q2;
        q2 = tq;
      }
      fv1 = (float)q1;
      fv2 = (float)q2;
      CEntryLF r = ef + off + 1 + h;
      BQP_WAIT_SMEM();
      BQP_SCHED_FENCE();
      const GravEntryLF a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
      BQP_SCHED_FENCE();
      col_tail_lf<false>(r, kk - h, a0, a1, a2, a3, ztf, rhof, fv1, fv2, fw1, fw2, qx, ux, qy, uy, qz, uz);
      gx = fma(vs, sx + (double)qx, fma(ws, tx + (double)ux, gx));
      gy = fma(vs, sy + (double)qy, fma(ws, ty + (double)uy, gy));
      gz = fma(vs, sz + (double)qz, fma(ws, tz + (double)uz, gz));
    } else {
      // all FP32: the sectoral entry (Q = 1) is loaded with the first four recursion entries
      CEntryLF ec = ef + off;
      BQP_WAIT_SMEM();
      BQP_SCHED_FENCE();
      const GravEntryLF es = ec[0];
      const GravEntryLF a0 = ec[1], a1 = ec[2], a2 = ec[3], a3 = ec[4];
      BQP_SCHED_FENCE();
      BQP_WAIT_SMEM();
      BQP_SCHED_FENCE();
      qx = es.xv;
      ux = es.xw;
      qy = es.yv;
      uy = es.yw;
      qz = es.zv;
      uz = es.zw;
      fv1 = 1.f;
      fv2 = 0.f;
      col_tail_lf<false>(ec + 1, kk, a0, a1, a2, a3, ztf, rhof, fv1, fv2, fw1, fw2, qx, ux, qy, uy, qz, uz);
      gx = fma(vs, (double)qx, fma(ws, (double)ux, gx));
      gy = fma(vs, (double)qy, fma(ws, (double)uy, gy));
      gz = fma(vs, (double)qz, fma(ws, (double)uz, gz));
    }
    const double pr = next_a ? zar : zbr, pi = next_a ? zai : zbi;
    j += next_a ? ga : gb;
    next_a = !next_a;
    cmul(vs, ws, pr, pi);
  }
  ax = gx;
  ay = gy;
  az = gz;
}

// Split-precision gravity: the rescaled pipelined form where its FP32 table
// exists (degree PIPE_MIN_N .. SPLIT_RESCALED_MAX_N), else the (a, b) form.
template <int G>
__device__ __forceinline__ void grav_split(const FieldView& f, int g, double x, double y, double z, double& ax,
                                           double& ay, double& az) {
  if (f.tablf != nullptr) {
    grav_split_pipelined<G>(f, g, x, y, z, ax, ay, az);
    return;
  }
  grav_split_ab<G>(f, g, x, y, z, ax, ay, az);
}

// ------------------------------------------------------ gravity in units
// The team kernel splits one evaluation's column sweep over a varying number
// of waves.  So that its result does not depend on that number, the columns
// are cut into GRAV_UNITS fixed units and every unit is summed the same way
// whichever wave computes it:
//   * unit u holds the columns u + 32k (path A) and 31 - u + 32k (path B),
//     k = 0, 1, ... (the back-and-forth deal over 32 columns, so every unit
//     carries about the same number of entries); column 0, the zonal one,
//     opens path A of unit 0;
//   * a path starts from the sectoral value vs0 z^j of its first column,
//     with z^j formed by square-and-multiply over the bits of j from the
//     lowest (always the same products), and steps to the next column of the
//     path by one complex multiply with z^32;
//   * each column's terms go through the same pipelined loop as
//     grav_fp64_pipelined; a unit's six running sums fold into (px, py, pz);
//   * units are combined by the fixed pairwise tree over 0..15: a wave that
//     owns an aligned block of 2^m units returns that block's subtree sum,
//     and the caller combines the blocks further up the same tree.
constexpr int GRAV_UNITS = 16;

// (re, im) = z^k, 0 <= k < 64, by square-and-multiply from the lowest bit.
__device__ __forceinline__ void cpow_bits(double zr, double zi, int k, double& re, double& im) {
  re = 1.0;
  im = 0.0;
  double br = zr, bi = zi;
  for (int b = 0; b < 6 && (k >> b) != 0; ++b) {
    if ((k >> b) & 1) cmul(re, im, br, bi);
    if ((k >> (b + 1)) != 0) cmul(br, bi, br, bi);
  }
}

// Terms of column j >= 1 with sectoral value (vs, ws), folded into (AX, AY,
// AZ): the factored pipelined entry loop of grav_fp64_pipelined (reads at
// most two entries past the column).
#define BQP_COL_PIPE(EC, KK, VS, WS, AX, AY, AZ)            \
  {                                                         \
    int kk = (KK);                                          \
    BQP_WAIT_SMEM();                                        \
    BQP_SCHED_FENCE();                                      \
    const GravEntryL es = (EC)[0];                          \
    GravEntryL a0 = (EC)[1], a1 = (EC)[2];                  \
    BQP_SCHED_FENCE();                                      \
    BQP_WAIT_SMEM();                                        \
    BQP_SCHED_FENCE();                                      \
    double q1 = 1.0, q2 = 0.0;                              \
    BQP_QOPEN(es);                                          \
    CEntryL r = (EC) + 3;                                   \
    if (kk >= 2) {                                          \
      kk -= 2;                                              \
      while (kk >= 4) {                                     \
        BQP_WAIT_SMEM();                                    \
        BQP_SCHED_FENCE();                                  \
        const GravEntryL b0 = r[0], b1 = r[1];              \
        BQP_SCHED_FENCE();                                  \
        BQP_QSTEP(a0, q2, q1);                              \
        BQP_QSTEP(a1, q1, q2);                              \
        BQP_WAIT_SMEM();                                    \
        BQP_SCHED_FENCE();                                  \
        a0 = r[2];                                          \
        a1 = r[3];                                          \
        BQP_SCHED_FENCE();                                  \
        BQP_QSTEP(b0, q2, q1);                              \
        BQP_QSTEP(b1, q1, q2);                              \
        r += 4;                                             \
        kk -= 4;                                            \
      }                                                     \
      BQP_WAIT_SMEM();                                      \
      BQP_SCHED_FENCE();                                    \
      GravEntryL t0, t1, t2;                                \
      if (kk >= 1) t0 = r[0];                               \
      if (kk >= 2) t1 = r[1];                               \
      if (kk >= 3) t2 = r[2];                               \
      BQP_SCHED_FENCE();                                    \
      BQP_QSTEP(a0, q2, q1);                                \
      BQP_QSTEP(a1, q1, q2);                                \
      BQP_WAIT_SMEM();                                      \
      BQP_SCHED_FENCE();                                    \
      if (kk >= 1) BQP_QSTEP(t0, q2, q1);                   \
      if (kk >= 2) BQP_QSTEP(t1, q1, q2);                   \
      if (kk >= 3) BQP_QSTEP(t2, q2, q1);                   \
    } else if (kk == 1) {                                   \
      BQP_QSTEP(a0, q2, q1);                                \
    }                                                       \
    BQP_QFOLD((VS), (WS), AX, AY, AZ);                      \
  }

// Column 0 (zonal, W = 0) in 4-entry batches with the next batch in flight
// (grav_fp64_pipelined's column 0); needs top >= PIPE_MIN_N + 1 entries.
__device__ __forceinline__ void grav_col0_pipe(CEntryL e, int top, double vs0, double zt, double rho, double& sx,
                                               double& sy, double& sz) {
  double v1 = vs0, v2 = 0.0;
  CEntryL r = e;
  int k = top;
  BQP_WAIT_SMEM();
  BQP_SCHED_FENCE();
  GravEntryL a0 = r[0], a1 = r[1], a2 = r[2], a3 = r[3];
  BQP_SCHED_FENCE();
  while (k >= 8) {
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    const GravEntryL b0 = r[4], b1 = r[5], b2 = r[6], b3 = r[7];
    BQP_SCHED_FENCE();
    BQP_COL0_STEP(a0, v2, v1);
    BQP_COL0_STEP(a1, v1, v2);
    BQP_COL0_STEP(a2, v2, v1);
    BQP_COL0_STEP(a3, v1, v2);
    BQP_WAIT_SMEM();
    BQP_SCHED_FENCE();
    a0 = r[8];
    a1 = r[9];
    a2 = r[10];
    a3 = r[11];
    BQP_SCHED_FENCE();
    BQP_COL0_STEP(b0, v2, v1);
    BQP_COL0_STEP(b1, v1, v2);
    BQP_COL0_STEP(b2, v2, v1);
    BQP_COL0_STEP(b3, v1, v2);
    r += 8;
    k -= 8;
  }
  BQP_WAIT_SMEM();
  BQP_SCHED_FENCE();
  GravEntryL b0, b1, b2;
  if (k >= 5) b0 = r[4];
  if (k >= 6) b1 = r[5];
  if (k >= 7) b2 = r[6];
  BQP_SCHED_FENCE();
  if (k >= 1) BQP_COL0_STEP(a0, v2, v1);
  if (k >= 2) BQP_COL0_STEP(a1, v1, v2);
  if (k >= 3) BQP_COL0_STEP(a2, v2, v1);
  if (k >= 4) BQP_COL0_STEP(a3, v1, v2);
  BQP_WAIT_SMEM();
  BQP_SCHED_FENCE();
  if (k >= 5) BQP_COL0_STEP(b0, v2, v1);
  if (k >= 6) BQP_COL0_STEP(b1, v1, v2);
  if (k >= 7) BQP_COL0_STEP(b2, v2, v1);
}

// Partial acceleration (body frame) of unit u at a body-fixed position whose
// scaled coordinates are given (xt = x R / r^2, ..., rho = R^2 / r^2,
// vs0 = R / r; (z32r, z32i) = (xt + i yt)^32).  FP64 rescaled table, degree
// >= PIPE_MIN_N.
__device__ __forceinline__ void grav_unit(const FieldView& f, int u, double xt, double yt, double zt, double rho,
                                          double vs0, double z32r, double z32i, double& px, double& py,
                                          double& pz) {
  double sx = 0.0, sy = 0.0, sz = 0.0, tx, ty, tz;  // column sums
  double gx = 0.0, gy = 0.0, gz = 0.0;              // the unit's running sum
  const int top = f.nmax + 1;
  CEntryL e = (CEntryL)f.tabl;
  CEntryL e1 = e + top;  // column 1
  for (int path = 0; path < 2; ++path) {
    int j = path == 0 ? u : 2 * GRAV_UNITS - 1 - u;
    double vs, ws;
    if (j == 0) {  // path A of unit 0 opens with the zonal column
      grav_col0_pipe(e, top, vs0, zt, rho, gx, gy, gz);
      j = 2 * GRAV_UNITS;
      vs = vs0 * z32r;
      ws = vs0 * z32i;
    } else {
      double pr, pi;
      cpow_bits(xt, yt, j, pr, pi);
      vs = vs0 * pr;
      ws = vs0 * pi;
    }
    for (; j <= f.jmax; j += 2 * GRAV_UNITS) {
      const CEntryL ec = e1 + (col_offset(f.nmax, j) - top);
      BQP_COL_PIPE(ec, top - j, vs, ws, gx, gy, gz);
      cmul(vs, ws, z32r, z32i);
    }
  }
  px = gx;
  py = gy;
  pz = gz;
}

// A position prepared for unit sums: the inertial position turned into the
// body frame by the angle whose cosine and sine are (c, s), its scaled
// coordinates and z^32.
struct UnitPoint {
  double xt, yt, zt, rho, vs0, z32r, z32i, c, s;
};

__device__ __forceinline__ UnitPoint unit_point(const FieldView& f, double xi, double yi, double zi, double c,
                                                double s) {
  UnitPoint q;
  const double x = fma(c, xi, s * yi), y = fma(c, yi, -(s * xi)), z = zi;
  const double R = f.radius;
  const double r2 = fma(x, x, fma(y, y, z * z));
  const double rr = 1.0 / r2;
  q.xt = x * R * rr;
  q.yt = y * R * rr;
  q.zt = z * R * rr;
  q.rho = R * R * rr;
  q.vs0 = R / sqrt(r2);
  cpow_bits(q.xt, q.yt, 2 * GRAV_UNITS, q.z32r, q.z32i);
  q.c = c;
  q.s = s;
  return q;
}

// Unit u's partial turned back to the inertial frame: the same value
// whichever wave computes it.
__device__ __forceinline__ void grav_unit_inertial(const FieldView& f, const UnitPoint& q, int u, double& px,
                                                   double& py, double& pz) {
  double gx, gy, gz;
  grav_unit(f, u, q.xt, q.yt, q.zt, q.rho, q.vs0, q.z32r, q.z32i, gx, gy, gz);
  px = fma(q.c, gx, -(q.s * gy));
  py = fma(q.s, gx, q.c * gy);
  pz = gz;
}

// ------------------------------------------------------------- atmosphere

// Vallado (2013) Table 8-4: band base altitude (km), nominal density
// (kg/m^3), scale height (km).
constexpr int VALLADO_BANDS = 28;
constexpr int kValladoH0Km[VALLADO_BANDS] = {0,   25,  30,  40,  50,  60,  70,  80,  90,  100,
                                             110, 120, 130, 140, 150, 180, 200, 250, 300, 350,
                                             400, 450, 500, 600, 700, 800, 900, 1000};

enum { DRAG_NONE = 0, DRAG_VALLADO = 1, DRAG_COMPAT = 2 };

// The table as the kernels use it: per band base altitude (km), nominal
// density (kg/m^3) and scale height (km), and the band of every 5-km step.
// Every band base is a multiple of 5 km, so the band of an altitude in
// [0, 1000) km is a function of floor(h / 5 km): band5[k] is the last band
// whose base is at most 5k km.  The kernels copy it to LDS at start
// (vallado_to_lds): the band and its values are read per lane, and from
// constant memory a lane-varying index would be two dependent vector-memory
// round trips per evaluation.
struct ValladoTable {
  double h0[VALLADO_BANDS];
  double rho0[VALLADO_BANDS];
  double H[VALLADO_BANDS];
  unsigned char band5[200];
};
constexpr ValladoTable make_vallado_table() {
  ValladoTable t{{0,   25,  30,  40,  50,  60,  70,  80,  90,  100, 110, 120, 130, 140,
                  150, 180, 200, 250, 300, 350, 400, 450, 500, 600, 700, 800, 900, 1000},
                 {1.225,     3.899e-2,  1.774e-2,  3.972e-3,  1.057e-3,  3.206e-4,  8.770e-5,
                  1.905e-5,  3.396e-6,  5.297e-7,  9.661e-8,  2.438e-8,  8.484e-9,  3.845e-9,
                  2.070e-9,  5.464e-10, 2.789e-10, 7.248e-11, 2.418e-11, 9.518e-12, 3.725e-12,
                  1.585e-12, 6.967e-13, 1.454e-13, 3.614e-14, 1.170e-14, 5.245e-15, 3.019e-15},
                 {7.249,  6.349,  6.682,  7.554,  8.382,  7.714,  6.549,  5.799,  5.382,  5.877,
                  7.263,  9.473,  12.636, 16.149, 22.523, 29.740, 37.105, 45.546, 53.628, 53.298,
                  58.515, 60.828, 63.822, 71.835, 88.667, 124.64, 181.05, 268.00},
                 {}};
  int i = 0;
  for (int k = 0; k < 200; ++k) {
    while (i + 1 < VALLADO_BANDS && kValladoH0Km[i + 1] <= 5 * k) ++i;
    t.band5[k] = (unsigned char)i;
  }
  return t;
}
static constexpr ValladoTable k_vallado = make_vallado_table();

// Copies the table to LDS; called by all threads of the workgroup (thread
// tid of nthreads) before a barrier.
__device__ __forceinline__ void vallado_to_lds(ValladoTable* t, int tid, int nthreads) {
  for (int k = tid; k < 3 * VALLADO_BANDS; k += nthreads) {
    const int b = k % VALLADO_BANDS;
    if (k < VALLADO_BANDS)
      t->h0[b] = k_vallado.h0[b];
    else if (k < 2 * VALLADO_BANDS)
      t->rho0[b] = k_vallado.rho0[b];
    else
      t->H[b] = k_vallado.H[b];
  }
  for (int k = tid; k < 200; k += nthreads) t->band5[k] = k_vallado.band5[k];
}

// Band of altitude hk (km): the last band whose base is at most hk (band 0
// below 0 km and for NaN; the top band from 1000 km on).  floor(hk / 5) is
// taken from a truncated product and corrected by exact comparisons with 5k
// (multiples of 5 up to 1000 are exact doubles), so it is exact.
__device__ __forceinline__ int vallado_band(const ValladoTable& t, double hk) {
  if (hk >= 1000.0) return VALLADO_BANDS - 1;
  if (!(hk >= 0.0)) return 0;
  int k = (int)(hk * 0.2);
  if (k > 0 && 5.0 * k > hk) --k;
  if (k < 199 && 5.0 * (k + 1) <= hk) ++k;
  return t.band5[k];
}
__device__ __forceinline__ double density(const ValladoTable& t, int model, double h_m) {
  const double hk = h_m * 1e-3;
  if (model == DRAG_VALLADO) {
    if (hk < 0.0) return t.rho0[0];
    const int b = vallado_band(t, hk);
    return t.rho0[b] * exp(-(hk - t.h0[b]) / t.H[b]);
  }
  if (model == DRAG_COMPAT) {
    if (hk <= 0.0) return 1.225;
    if (hk >= 1000.0) return 0.0;
    return 1.225 * exp(-hk / 8.5);
  }
  return 0.0;
}

struct ModelView {
  double omega, theta0, r_atm, r_min;
  int drag;
  double flat;  // ground ellipsoid flattening (0: the sphere r_min)
};

// The row's minimum-radius test at inertial position (x, y, z); r_min > 0.
// flat == 0: |r| >= r_min.  flat > 0: the reference's ground event
// (lightyear_odeint_rs/src/events.rs:87-145 check_ground): the height above
// the ellipsoid of equatorial radius r_atm and polar radius r_atm (1 - flat),
// measured along the position's direction, must be at least r_min - r_atm.
__device__ __forceinline__ bool above_ground(const ModelView& m, double x, double y, double z) {
  const double d = sqrt(fma(x, x, fma(y, y, z * z)));
  if (m.flat == 0.0) return d >= m.r_min;
  const double a = m.r_atm, b = m.r_atm * (1.0 - m.flat);
  const double den = sqrt(fma(a * a, z * z, (b * b) * fma(x, x, y * y)));
  const double surface = (den > 0.0) ? a * b * d / den : b;  // at the centre: the polar radius
  return d - surface >= m.r_min - m.r_atm;
}

}  // namespace bqp
