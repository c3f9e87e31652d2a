// Auxiliary kernels of the runtime (bqp.hip only): standalone gravity,
// surrogate training data, queue reset, longest-first ordering, row
// statistics.
#pragma once

#include "bqp_kernels.hpp"

namespace bqp {

// Standalone batched gravity: one body-fixed position per lane.
template <int PREC>
__global__ void __launch_bounds__(256) gravity_kernel(FieldView f, const double* __restrict__ pos,
                                                      double* __restrict__ acc, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = i < n;
  const long long r = ok ? i : 0;
  const double x = pos[3 * r], y = pos[3 * r + 1], z = pos[3 * r + 2];
  double ax, ay, az;
  if (PREC == 0)
    grav_fp64<1>(f, 0, x, y, z, ax, ay, az);
  else
    grav_split<1>(f, 0, x, y, z, ax, ay, az);
  if (ok) {
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
  }
}

// Surrogate training data: seeded body-fixed positions (uniform direction:
// normalised Gaussian triple; |r| uniform in [r_lo, r_hi]), their features,
// and the residual a_full - a_low (exact FP64 recursion) times inv_scale.
__global__ void __launch_bounds__(256) surrogate_data_kernel(FieldView full, FieldView low, long long n,
                                                             unsigned long long seed, double r_lo, double r_hi,
                                                             float r_mid, float r_inv_half, double is0, double is1,
                                                             double is2, double* __restrict__ pos,
                                                             float* __restrict__ feat, float* __restrict__ targ) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = i < n;
  const unsigned long long idx = ok ? (unsigned long long)i : 0ull;
  double u[5];
#pragma unroll
  for (int k = 0; k < 5; ++k)  // uniforms in (0, 1]
    u[k] = ((double)(mix64(seed ^ mix64(5ull * idx + (unsigned long long)k)) >> 11) + 1.0) * 0x1.0p-53;
  const double two_pi = 6.283185307179586;
  const double ra = sqrt(-2.0 * log(u[0])), rb = sqrt(-2.0 * log(u[2]));
  double nx = ra * cos(two_pi * u[1]), ny = ra * sin(two_pi * u[1]), nz = rb * cos(two_pi * u[3]);
  double nn = sqrt(fma(nx, nx, fma(ny, ny, nz * nz)));
  if (!(nn > 1e-300)) {
    nx = 0.0;
    ny = 0.0;
    nz = 1.0;
    nn = 1.0;
  }
  const double r = fma(r_hi - r_lo, u[4], r_lo);
  const double x = r * (nx / nn), y = r * (ny / nn), z = r * (nz / nn);
  double fx, fy, fz, lx, ly, lz;
  grav_fp64<1>(full, 0, x, y, z, fx, fy, fz);
  grav_fp64<1>(low, 0, x, y, z, lx, ly, lz);
  if (!ok) return;
  if (pos) {
    pos[3 * i] = x;
    pos[3 * i + 1] = y;
    pos[3 * i + 2] = z;
  }
  float f[MLP_IN];
  mlp_features((float)x, (float)y, (float)z, r_mid, r_inv_half, f);
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) feat[i * MLP_IN + k] = f[k];
  targ[3 * i] = (float)((fx - lx) * is0);
  targ[3 * i + 1] = (float)((fy - ly) * is1);
  targ[3 * i + 2] = (float)((fz - lz) * is2);
}

// Surrogate pre-check: the first evaluation of every row is a sampled check
// (evaluation 0 at the initial state), so it can be made for all rows before
// the propagation.  fail[row] = 1 when the row's initial position is inside
// the radius band and |low + network - exact| exceeds the threshold; such a
// row would run the exact field for its whole arc.  One wave per block.
template <int PREC>
__global__ void __launch_bounds__(64) sur_precheck_kernel(const PropParams p, int n, int* __restrict__ fail) {
  __shared__ uint4 sX[64 * 4];
  __shared__ float sO[64 * 3];
  const int lane = threadIdx.x;
  const int row = blockIdx.x * 64 + lane;
  const bool ok = row < n;
  const int r = ok ? row : 0;
  const double t = p.t0 ? p.t0[r] : p.t0c;
  const double y0 = p.y0[(size_t)r * 6], y1 = p.y0[(size_t)r * 6 + 1], y2 = p.y0[(size_t)r * 6 + 2];
  const double th = fma(p.model.omega, t, p.model.theta0);
  double s, c;
  sincos(th, &s, &c);
  const double xb = fma(c, y0, s * y1);
  const double yb = fma(c, y1, -(s * y0));
  const double zb = y2;
  double rx = 0.0, ry = 0.0, rz = 0.0;
  sur_residual<PREC, 1>(p, 0, lane, sX, sO, xb, yb, zb, rx, ry, rz);
  double fx, fy, fz, lx, ly, lz;
  grav_fp64<1>(p.field, 0, xb, yb, zb, fx, fy, fz);
  grav_fp64<1>(p.low, 0, xb, yb, zb, lx, ly, lz);
  const double rad = sqrt(fma(xb, xb, fma(yb, yb, zb * zb)));
  const bool band = (rad >= p.sur_r_lo) && (rad <= p.sur_r_hi);
  const double dx = (lx + rx) - fx, dy = (ly + ry) - fy, dz = (lz + rz) - fz;
  const bool bad = band && (fma(dx, dx, fma(dy, dy, dz * dz)) > p.sur_thresh2);
  if (ok) fail[row] = bad ? 1 : 0;
}

// Stable split of the rows, taken in hand-out order (order[i], or i when
// order is null), by fail[]: failing rows to exact_rows, the others to
// sur_rows; counts[0..1] = sizes.  One block of 1024.
__global__ void __launch_bounds__(1024) partition_kernel(const int* __restrict__ fail,
                                                        const int32_t* __restrict__ order, int n,
                                                        int* __restrict__ exact_rows, int* __restrict__ sur_rows,
                                                        int* __restrict__ counts) {
  __shared__ int scan[1024];
  __shared__ int base_e, base_s;
  if (threadIdx.x == 0) {
    base_e = 0;
    base_s = 0;
  }
  __syncthreads();
  for (int off = 0; off < n; off += 1024) {
    const int i = off + (int)threadIdx.x;
    const int row = (i < n) ? (order ? order[i] : i) : 0;
    const int f = (i < n) ? fail[row] : 0;
    const int valid = (i < n) ? 1 : 0;
    scan[threadIdx.x] = f;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan of the fail flags
      const int v = (threadIdx.x >= (unsigned)d) ? scan[threadIdx.x - d] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    const int incl = scan[threadIdx.x];
    const int chunk_fail = scan[1023];
    if (valid) {
      if (f)
        exact_rows[base_e + incl - 1] = row;
      else
        sur_rows[base_s + (i - off) - incl] = row;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int m = min(1024, n - off);
      base_e += chunk_fail;
      base_s += m - chunk_fail;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[0] = base_e;
    counts[1] = base_s;
  }
}

// Queue reset without hipMemsetAsync (graph-capture safe): the queue starts
// after the round of groups dealt by block index.
__global__ void reset_counter_kernel(unsigned int* q, unsigned int start) { q[0] = start; }

// Longest-first hand-out order (a counting sort made of plain kernels, so it
// can be captured in a HIP graph).  The key estimates a row's evaluations as
// revolutions over the span (span * mean motion from vis-viva) times
// evaluations per revolution.  The latter come from a table that every call
// on the context fills from its own per-row counts (CostTable below); with
// no table, or an untrained one, the key is the revolution count.  Row
// results do not depend on which lane runs them, so any order, and any tie
// order, gives the same results.
constexpr int SORT_BUCKETS = 4096;

// Learned evaluations per revolution over (altitude of the semi-major axis,
// cos of the inclination): COST_NA log-spaced altitude bins over
// [COST_ALT0, COST_ALT0 * 10^COST_DECADES] and COST_NI equal bins of cos i
// over [-1, 1].  The cost per revolution rises steeply as altitude falls (the
// high-degree terms grow as (R/r)^n) and by several percent from equatorial
// to polar orbits; the table captures both without a model of either, and
// the key interpolates it bilinearly between cell centres.  Sums are
// integers (evals, revolutions in 1/65536 units), so the learned values do
// not depend on the order of the atomic additions.
constexpr int COST_NA = BQP_COST_ALT_BINS, COST_NI = BQP_COST_INC_BINS, COST_CELLS = COST_NA * COST_NI;
constexpr double COST_ALT0 = 1e5, COST_DECADES = 3.0, COST_REV_UNIT = 65536.0;
constexpr unsigned long long COST_MIN_ROWS = 4;

struct CostTable {
  unsigned long long evals[COST_CELLS];
  unsigned long long revs[COST_CELLS];  // revolutions * COST_REV_UNIT
  unsigned long long rows[COST_CELLS];
  float log2_epr[COST_CELLS];  // filled log2(evals per revolution), read by the key
};

// Revolutions over the span and the row's continuous bin positions (altitude
// and inclination, in units of bins); false for unbound or degenerate states.
__device__ __forceinline__ bool row_orbit(const double* __restrict__ s, double span, double mu, double radius,
                                          double* rev, double* ualt, double* uinc) {
  const double r = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  const double v2 = s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
  const double en = 0.5 * v2 - mu / r;
  if (!(en < 0.0) || !(r > 0.0) || !(span > 0.0)) return false;
  const double a = -mu / (2.0 * en);
  *rev = span * sqrt(mu / (a * a * a));
  if (!isfinite(*rev) || !(*rev > 0.0)) return false;
  const double alt = fmax(a - radius, COST_ALT0);
  *ualt = log10(alt / COST_ALT0) * (COST_NA / COST_DECADES);
  const double hx = s[1] * s[5] - s[2] * s[4], hy = s[2] * s[3] - s[0] * s[5], hz = s[0] * s[4] - s[1] * s[3];
  const double hn = sqrt(hx * hx + hy * hy + hz * hz);
  const double ci = hn > 0.0 ? hz / hn : 1.0;
  *uinc = 0.5 * (ci + 1.0) * COST_NI;
  return isfinite(*ualt) && isfinite(*uinc);
}

__device__ __forceinline__ int cost_bin(double u, int nb) {
  const int b = (int)u;
  return b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
}

__device__ __forceinline__ int cost_cell(double ualt, double uinc) {
  return cost_bin(ualt, COST_NA) * COST_NI + cost_bin(uinc, COST_NI);
}

// Lower cell-centre index and weight of the upper one along one axis.
__device__ __forceinline__ void cost_lerp(double u, int nb, int* i0, int* i1, double* f) {
  const double uc = u - 0.5;
  int k = (int)floor(uc);
  k = k < 0 ? 0 : (k > nb - 1 ? nb - 1 : k);
  *i0 = k;
  *i1 = k + 1 < nb ? k + 1 : nb - 1;
  const double w = uc - (double)k;
  *f = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
}

__device__ __forceinline__ int cost_bucket(const double* __restrict__ s, double span, double mu, double radius,
                                           const CostTable* __restrict__ tab) {
  double rev = 0.0, ualt = 0.0, uinc = 0.0;
  double est = 0.0;
  if (row_orbit(s, span, mu, radius, &rev, &ualt, &uinc)) {
    est = rev;
    if (tab) {
      int a0, a1, b0, b1;
      double fa, fb;
      cost_lerp(ualt, COST_NA, &a0, &a1, &fa);
      cost_lerp(uinc, COST_NI, &b0, &b1, &fb);
      const float* L = tab->log2_epr;
      const double lo = (1.0 - fa) * (double)L[a0 * COST_NI + b0] + fa * (double)L[a1 * COST_NI + b0];
      const double hi = (1.0 - fa) * (double)L[a0 * COST_NI + b1] + fa * (double)L[a1 * COST_NI + b1];
      est = rev * exp2((1.0 - fb) * lo + fb * hi);
    }
  }
  // 128 buckets per octave of the estimate, most expensive first
  int b = isfinite(est) && est > 0.0 ? (int)(log2(1.0 + est) * 128.0) : 0;
  b = b < 0 ? 0 : (b > SORT_BUCKETS - 1 ? SORT_BUCKETS - 1 : b);
  return SORT_BUCKETS - 1 - b;
}

__global__ void cost_clear_kernel(CostTable* __restrict__ tab) {
  for (int i = threadIdx.x; i < COST_CELLS; i += blockDim.x) {
    tab->evals[i] = 0ull;
    tab->revs[i] = 0ull;
    tab->rows[i] = 0ull;
    tab->log2_epr[i] = 0.0f;
  }
}

// Adds every finished row of at least one revolution to its cell (a shorter
// row's fixed start-up cost, the first step's selection, would inflate its
// evaluations per revolution).
__global__ void cost_learn_kernel(const double* __restrict__ y0, const double* __restrict__ t0,
                                  const double* __restrict__ tf, const int32_t* __restrict__ status,
                                  const int32_t* __restrict__ evals, double mu, double radius, int n,
                                  CostTable* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || status[i] != ROW_OK || evals[i] <= 0) return;
  double rev = 0.0, ualt = 0.0, uinc = 0.0;
  if (!row_orbit(y0 + 6 * (size_t)i, tf[i] - t0[i], mu, radius, &rev, &ualt, &uinc)) return;
  if (!(rev >= 1.0) || rev > 1e9) return;
  const double rq = rev * COST_REV_UNIT;
  const int cell = cost_cell(ualt, uinc);
  atomicAdd(&tab->evals[cell], (unsigned long long)evals[i]);
  atomicAdd(&tab->revs[cell], (unsigned long long)(rq + 0.5));
  atomicAdd(&tab->rows[cell], 1ull);
}

// Evaluations per revolution of every cell with at least COST_MIN_ROWS rows.
// An untrained cell takes the nearest trained cell in altitude within its
// inclination bin; in an inclination bin with no trained cell, cells take
// the value at the same altitude in the nearest bin that has one.  An
// untrained table is 1 evaluation per revolution.  One block.
__global__ void __launch_bounds__(1024) cost_fill_kernel(CostTable* __restrict__ tab) {
  __shared__ float val[COST_CELLS];
  __shared__ int have[COST_CELLS];
  __shared__ int col_have[COST_NI];
  for (int t = threadIdx.x; t < COST_CELLS; t += blockDim.x) {
    const bool ok = tab->rows[t] >= COST_MIN_ROWS && tab->revs[t] > 0ull;
    have[t] = ok ? 1 : 0;
    val[t] = ok ? (float)log2((double)tab->evals[t] * COST_REV_UNIT / (double)tab->revs[t]) : 0.0f;
  }
  for (int t = threadIdx.x; t < COST_NI; t += blockDim.x) col_have[t] = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < COST_CELLS; t += blockDim.x)
    if (have[t]) col_have[t % COST_NI] = 1;  // benign race: every writer stores 1
  __syncthreads();
  // along altitude within the inclination bin
  float v1[(COST_CELLS + 255) / 256];  // launched with 256 to 1024 threads
  int k = 0;
  for (int t = threadIdx.x; t < COST_CELLS; t += blockDim.x, ++k) {
    const int ia = t / COST_NI, ii = t % COST_NI;
    float v = val[t];
    if (!have[t] && col_have[ii]) {
      for (int d = 1; d < COST_NA; ++d) {
        if (ia - d >= 0 && have[(ia - d) * COST_NI + ii]) {
          v = val[(ia - d) * COST_NI + ii];
          break;
        }
        if (ia + d < COST_NA && have[(ia + d) * COST_NI + ii]) {
          v = val[(ia + d) * COST_NI + ii];
          break;
        }
      }
    }
    v1[k] = v;
  }
  __syncthreads();
  k = 0;
  for (int t = threadIdx.x; t < COST_CELLS; t += blockDim.x, ++k) val[t] = v1[k];
  __syncthreads();
  // inclination bins with no trained cell: nearest bin that has one
  for (int t = threadIdx.x; t < COST_CELLS; t += blockDim.x) {
    const int ia = t / COST_NI, ii = t % COST_NI;
    float v = val[t];
    if (!col_have[ii]) {
      v = 0.0f;
      for (int d = 1; d < COST_NI; ++d) {
        if (ii - d >= 0 && col_have[ii - d]) {
          v = val[ia * COST_NI + ii - d];
          break;
        }
        if (ii + d < COST_NI && col_have[ii + d]) {
          v = val[ia * COST_NI + ii + d];
          break;
        }
      }
    }
    tab->log2_epr[t] = v;
  }
}

__global__ void sort_clear_kernel(unsigned int* __restrict__ hist) {
  for (int i = threadIdx.x; i < SORT_BUCKETS; i += blockDim.x) hist[i] = 0u;
}

__global__ void sort_hist_kernel(const double* __restrict__ y0, const double* __restrict__ t0,
                                 const double* __restrict__ tf, double mu, double radius,
                                 const CostTable* __restrict__ tab, int n, unsigned int* __restrict__ hist,
                                 unsigned short* __restrict__ bucket) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = cost_bucket(y0 + 6 * (size_t)i, tf[i] - t0[i], mu, radius, tab);
  bucket[i] = (unsigned short)b;
  atomicAdd(&hist[b], 1u);
}

// exclusive scan of the histogram in place (one block of 1024 threads)
__global__ void __launch_bounds__(1024) sort_scan_kernel(unsigned int* __restrict__ hist) {
  __shared__ unsigned int part[1024];
  constexpr int PER = SORT_BUCKETS / 1024;
  const int t = threadIdx.x;
  unsigned int loc[PER], sum = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    loc[k] = hist[t * PER + k];
    sum += loc[k];
  }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  unsigned int run = part[t] - sum;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    hist[t * PER + k] = run;
    run += loc[k];
  }
}

__global__ void sort_scatter_kernel(const unsigned short* __restrict__ bucket, int n,
                                    unsigned int* __restrict__ offs, int32_t* __restrict__ order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned int pos = atomicAdd(&offs[bucket[i]], 1u);
  order[pos] = i;
}

// Sum of per-row counters into 64-bit totals (stats).
__global__ void row_stats_kernel(const int32_t* __restrict__ status, const int32_t* __restrict__ steps,
                                 const int32_t* __restrict__ evals, const int32_t* __restrict__ rejects,
                                 int n, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long acc[5];
  if (threadIdx.x < 5) acc[threadIdx.x] = 0ull;
  __syncthreads();
  unsigned long long s = 0, e = 0, r = 0, ok = 0, bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    s += (unsigned)steps[i];
    e += (unsigned)evals[i];
    r += (unsigned)rejects[i];
    if (status[i] == 0) ++ok; else ++bad;
  }
  atomicAdd(&acc[0], s);
  atomicAdd(&acc[1], e);
  atomicAdd(&acc[2], r);
  atomicAdd(&acc[3], ok);
  atomicAdd(&acc[4], bad);
  __syncthreads();
  if (threadIdx.x < 5) atomicAdd(&out[threadIdx.x], acc[threadIdx.x]);
}

__global__ void first_failed_kernel(const int32_t* __restrict__ status, int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && status[i] != 0) atomicMin(out, i);
}

__global__ void fill_int_kernel(int* p, int v) { p[0] = v; }
__global__ void zero_u64_kernel(unsigned long long* p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace bqp
