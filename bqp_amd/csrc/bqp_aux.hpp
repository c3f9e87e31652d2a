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

// Queue reset without hipMemsetAsync (graph-capture safe).
__global__ void reset_counter_kernel(unsigned int* q) { q[0] = 0u; }

// Longest-first hand-out order (a counting sort made of plain kernels, so it
// can be captured in a HIP graph).  The cost key is the number of orbit
// revolutions over the span (span * mean motion from vis-viva); at a fixed
// tolerance the step count per revolution varies slowly, so rows with similar
// keys take similar step counts and share a wave.  Row results do not depend
// on which lane runs them, so any tie order is exact.
constexpr int SORT_BUCKETS = 4096;

__device__ __forceinline__ int cost_bucket(const double* __restrict__ s, double span, double mu) {
  const double r = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  const double v2 = s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
  const double en = 0.5 * v2 - mu / r;
  double rev = 0.0;
  if (en < 0.0 && r > 0.0 && span > 0.0) {
    const double a = -mu / (2.0 * en);
    rev = span * sqrt(mu / (a * a * a));
  }
  // 128 buckets per octave of revolutions, most expensive first
  int b = isfinite(rev) && rev > 0.0 ? (int)(log2(1.0 + rev) * 128.0) : 0;
  b = b < 0 ? 0 : (b > SORT_BUCKETS - 1 ? SORT_BUCKETS - 1 : b);
  return SORT_BUCKETS - 1 - b;
}

__global__ void sort_clear_kernel(unsigned int* __restrict__ hist) {
  for (int i = threadIdx.x; i < SORT_BUCKETS; i += blockDim.x) hist[i] = 0u;
}

__global__ void sort_hist_kernel(const double* __restrict__ y0, const double* __restrict__ t0,
                                 const double* __restrict__ tf, double mu, int n,
                                 unsigned int* __restrict__ hist, unsigned short* __restrict__ bucket) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = cost_bucket(y0 + 6 * (size_t)i, tf[i] - t0[i], mu);
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
