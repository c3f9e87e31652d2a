// Learned gravity surrogate: device training (FP32, Adam), packing for
// MFMA inference (BF16 / FP8 E4M3), batched inference, and the C ABI of
// include/bqp.h's bqp_surrogate_* section.  No reference counterpart.
//
// Training runs entirely on the GPU with our own kernels:
//   forward   Z_l = H_{l-1} W_l^T + b_l, H_l = SiLU(Z_l)    (tiled FP32 GEMM)
//   loss      L = mean over rows and 3 components of (Y - T)^2
//   backward  dW_l = dZ_l^T H_{l-1} (split-K, fixed-order reduction),
//             db_l = column sums of dZ_l, dZ_{l-1} = (dZ_l W_l) * SiLU'(Z_{l-1})
//   update    Adam with bias correction.
// Every reduction has a fixed order, so a training run is reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bqp_internal.hpp"

using namespace bqp;

namespace {

// ------------------------------------------------------------------ GEMM
enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_DSILU = 3 };

constexpr int GT = 64;  // output tile (GT x GT), 256 threads, 4 x 4 per thread
constexpr int GK = 16;

// C(M x N) = op(A)(M x K) op(B)(K x N), row-major storage.  op(A)(m,k) is
// A[m*lda+k], or A[k*lda+m] when TA; op(B)(k,n) is B[k*ldb+n], or B[n*ldb+k]
// when TB.  blockIdx.z selects a K chunk [z*kchunk, (z+1)*kchunk); with
// EPI_STORE its partial product goes to C + z*split_stride.
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) sgemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                    const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                    int ldc, const float* __restrict__ bias, float* __restrict__ aux,
                                                    int kchunk, long long split_stride) {
  __shared__ float As[GK][GT + 4];
  __shared__ float Bs[GK][GT + 4];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  for (int k0 = kb; k0 < ke; k0 += GK) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int mm = TA ? (idx & 63) : (idx >> 4);
      const int kk = TA ? (idx >> 6) : (idx & 15);
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < M && gk < ke) v = TA ? A[(long long)gk * lda + gm] : A[(long long)gm * lda + gk];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int nn = TB ? (idx >> 4) : (idx & 63);
      const int kk = TB ? (idx & 15) : (idx >> 6);
      const int gn = n0 + nn, gk = k0 + kk;
      float v = 0.f;
      if (gn < N && gk < ke) v = TB ? B[(long long)gn * ldb + gk] : B[(long long)gk * ldb + gn];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  float* Cz = C + (EPI == EPI_STORE ? (long long)blockIdx.z * split_stride : 0ll);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gm = m0 + ty * 4 + i;
    if (gm >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + tx * 4 + j;
      if (gn >= N) continue;
      const long long o = (long long)gm * ldc + gn;
      float v = acc[i][j];
      if (EPI == EPI_BIAS_SILU) {
        v += bias[gn];
        Cz[o] = v;
        aux[o] = mlp_silu(v);
      } else if (EPI == EPI_BIAS) {
        Cz[o] = v + bias[gn];
      } else if (EPI == EPI_DSILU) {
        const float z = aux[o];
        const float sg = 1.0f / (1.0f + __expf(-z));
        Cz[o] = v * (sg * fmaf(z, 1.0f - sg, 1.0f));
      } else {
        Cz[o] = v;
      }
    }
  }
}

// ------------------------------------------------------- small kernels
__global__ void __launch_bounds__(256) gather_kernel(const float* __restrict__ feat, const float* __restrict__ targ,
                                                     long long n, int batch, unsigned long long seed,
                                                     float* __restrict__ X, float* __restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch) return;
  const long long src = (long long)(mix64(seed ^ mix64((unsigned long long)i)) % (unsigned long long)n);
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) X[(long long)i * MLP_IN + k] = feat[src * MLP_IN + k];
#pragma unroll
  for (int k = 0; k < MLP_OUT; ++k) T[(long long)i * MLP_OUT + k] = targ[src * MLP_OUT + k];
}

// dY = 2 (Y - T) / (3 B) and per-block partial sums of (Y - T)^2
__global__ void __launch_bounds__(256) loss_kernel(const float* __restrict__ Y, const float* __restrict__ T, int B,
                                                   float* __restrict__ dY, double* __restrict__ partial) {
  __shared__ double red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < B) {
    const float sc = 2.0f / (3.0f * (float)B);
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) {
      const float d = Y[(long long)i * MLP_OUT + k] - T[(long long)i * MLP_OUT + k];
      dY[(long long)i * MLP_OUT + k] = sc * d;
      s += (double)d * (double)d;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) loss_final_kernel(const double* __restrict__ partial, int nparts, int B,
                                                         float* __restrict__ loss_out, double* __restrict__ loss_d) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double l = red[0] / (3.0 * (double)B);
    if (loss_out) *loss_out = (float)l;
    if (loss_d) *loss_d = l;
  }
}

// Column sums of dZ (B x N) -> db (N) in two fixed-order passes: block
// (cb, rs) sums rows rs, rs + RS, ... of 64 columns into part[rs][.], then
// colsum_final_kernel adds the RS partials.
constexpr int COLSUM_SPLITS = 64;
__global__ void __launch_bounds__(256) colsum_kernel(const float* __restrict__ dZ, int B, int N,
                                                     float* __restrict__ part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int rg = threadIdx.x >> 6;
  const int rs = blockIdx.y;
  float s = 0.f;
  if (c < N)
    for (int r = rs * 4 + rg; r < B; r += 4 * COLSUM_SPLITS) s += dZ[(long long)r * N + c];
  red[rg][threadIdx.x & 63] = s;
  __syncthreads();
  if (rg == 0 && c < N)
    part[(long long)rs * N + c] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

__global__ void __launch_bounds__(256) colsum_final_kernel(const float* __restrict__ part, int N,
                                                           float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int rs = 0; rs < COLSUM_SPLITS; ++rs) s += part[(long long)rs * N + c];
  db[c] = s;
}

// dW = sum over splits of the partial products (fixed order)
__global__ void __launch_bounds__(256) split_reduce_kernel(const float* __restrict__ part, int splits, long long count,
                                                           float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[(long long)z * count + i];
  out[i] = s;
}

__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, long long n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = fmaf(b1, m[i], (1.0f - b1) * gi);
  const float vi = fmaf(b2, v[i], (1.0f - b2) * gi * gi);
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// max |x| over n values -> per-block partials (non-negative floats)
__global__ void __launch_bounds__(256) absmax_kernel(const float* __restrict__ x, long long n,
                                                     float* __restrict__ partial) {
  __shared__ float red[256];
  float s = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s = fmaxf(s, fabsf(x[i]));
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------- MFMA inference
// Each wave runs NT tiles of 16 samples at a time (grid-stride over tiles).
template <int FMT, int NT>
__global__ void __launch_bounds__(256) mlp_infer_kernel(const MlpPacked P, const float* __restrict__ feat,
                                                        long long n, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4;
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long base = wave * 16 * NT; base < n; base += nwaves * 16 * NT) {
    uint4 xin[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const long long s = base + 16 * t + (lane & 15);
      xin[t] = make_uint4(0u, 0u, 0u, 0u);
      if (q < 2 && s < n) {
        const float4 f0 = *reinterpret_cast<const float4*>(feat + s * MLP_IN + 8 * q);
        const float4 f1 = *reinterpret_cast<const float4*>(feat + s * MLP_IN + 8 * q + 4);
        xin[t] = make_uint4(pack_bf16x2(f0.x, f0.y), pack_bf16x2(f0.z, f0.w), pack_bf16x2(f1.x, f1.y),
                            pack_bf16x2(f1.z, f1.w));
      }
    }
    f32x4 o[NT];
    mlp_forward_wave<FMT, NT>(P, lane, xin, o);
    if (q == 0) {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const long long s = base + 16 * t + lane;
        if (s < n) {
          out[s * MLP_OUT + 0] = o[t][0];
          out[s * MLP_OUT + 1] = o[t][1];
          out[s * MLP_OUT + 2] = o[t][2];
        }
      }
    }
  }
}

// ------------------------------------------------------------ host side
int fail(bqp_ctx* c, int code, const std::string& m) { return bqp_fail(c, code, m); }

#define TRY(expr)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(c, BQP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// training scratch for minibatches of up to `cap` rows
struct Work {
  float *X, *T, *Y, *dY, *dZa, *dZb, *part, *cpart;
  float* Z[4];
  float* H[4];
  double* lpart;
  float* lsteps;
  double* lossd;
  float* amax;
};
constexpr int SPLITS = 16;
constexpr int MAX_LOSS_STEPS = 1 << 16;

// Lays the scratch out from `base` (which may be null, to measure it);
// every piece is padded to 256 bytes.  *total = bytes used.
Work carve(char* base, int cap, size_t* total) {
  Work w{};
  size_t off = 0;
  auto take = [&](size_t nf) {
    float* r = (float*)(base + off);
    off += ((nf + 63) & ~(size_t)63) * sizeof(float);
    return r;
  };
  w.X = take((size_t)cap * MLP_IN);
  w.T = take((size_t)cap * MLP_OUT);
  w.Y = take((size_t)cap * MLP_OUT);
  w.dY = take((size_t)cap * MLP_OUT);
  w.dZa = take((size_t)cap * MLP_WIDTH);
  w.dZb = take((size_t)cap * MLP_WIDTH);
  for (int l = 0; l < 4; ++l) {
    w.Z[l] = take((size_t)cap * MLP_WIDTH);
    w.H[l] = take((size_t)cap * MLP_WIDTH);
  }
  w.part = take((size_t)SPLITS * MLP_WIDTH * MLP_WIDTH);
  w.cpart = take((size_t)COLSUM_SPLITS * MLP_WIDTH);
  w.lsteps = take(MAX_LOSS_STEPS);
  w.amax = take(1024);
  w.lpart = (double*)take(2 * ((size_t)cap / 256 + 2));
  w.lossd = (double*)take(16);
  *total = off;
  return w;
}

int ensure_work(bqp_ctx* c, int cap, Work& w) {
  bqp_mlp_state* s = c->mlp;
  size_t need = 0;
  (void)carve(nullptr, cap, &need);
  int rc = bqp_ensure(c, &s->d_work, &s->work_bytes, need);
  if (rc) return rc;
  size_t used = 0;
  w = carve((char*)s->d_work, cap, &used);
  return BQP_OK;
}

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
          const float* bias, float* aux, int splits = 1) {
  const int kchunk = splits > 1 ? ((K + splits - 1) / splits + GK - 1) / GK * GK : K;
  const int z = splits > 1 ? (K + kchunk - 1) / kchunk : 1;
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, z);
  hipLaunchKernelGGL((sgemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias,
                     aux, kchunk, (long long)M * N);
}

const float* Wp(const float* P, int l) { return P + mlp_w_off(l); }
const float* Bp(const float* P, int l) { return P + mlp_b_off(l); }

// forward of B rows in X; leaves Z_l, H_l (l < 4) and Y in the work buffers
void forward(hipStream_t st, const float* P, const float* X, int B, const Work& w) {
  const float* hin = X;
  for (int l = 0; l < 4; ++l) {
    const int in = mlp_in_dim(l);
    gemm<false, true, EPI_BIAS_SILU>(st, B, MLP_WIDTH, in, hin, in, Wp(P, l), in, w.Z[l], MLP_WIDTH, Bp(P, l),
                                     w.H[l]);
    hin = w.H[l];
  }
  gemm<false, true, EPI_BIAS>(st, B, MLP_OUT, MLP_WIDTH, hin, MLP_WIDTH, Wp(P, 4), MLP_WIDTH, w.Y, MLP_OUT, Bp(P, 4),
                              nullptr);
}

// loss and full gradient (into G) of the B rows in X / T
void loss_backward(hipStream_t st, const float* P, float* G, const float* X, const float* T, int B, const Work& w,
                   float* loss_step) {
  const int nb = (B + 255) / 256;
  hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(256), 0, st, w.Y, T, B, w.dY, w.lpart);
  hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(256), 0, st, w.lpart, nb, B, loss_step, w.lossd);
  const float* dZ = w.dY;
  int outd = MLP_OUT;
  for (int l = 4; l >= 0; --l) {
    const int in = mlp_in_dim(l);
    const float* hprev = l > 0 ? w.H[l - 1] : X;
    // dW_l = dZ^T (out x B) . H_{l-1} (B x in), split over B
    gemm<true, false, EPI_STORE>(st, outd, in, B, dZ, outd, hprev, in, w.part, in, nullptr, nullptr, SPLITS);
    const int kchunk = ((B + SPLITS - 1) / SPLITS + GK - 1) / GK * GK;
    const int z = (B + kchunk - 1) / kchunk;
    const long long cnt = (long long)outd * in;
    hipLaunchKernelGGL(split_reduce_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, w.part, z, cnt,
                       G + mlp_w_off(l));
    hipLaunchKernelGGL(colsum_kernel, dim3((outd + 63) / 64, COLSUM_SPLITS), dim3(256), 0, st, dZ, B, outd,
                       w.cpart);
    hipLaunchKernelGGL(colsum_final_kernel, dim3((outd + 255) / 256), dim3(256), 0, st, w.cpart, outd,
                       G + mlp_b_off(l));
    if (l > 0) {
      float* dzn = (dZ == w.dZa) ? w.dZb : w.dZa;
      // dZ_{l-1} = (dZ_l (B x out) . W_l (out x in)) * SiLU'(Z_{l-1})
      gemm<false, false, EPI_DSILU>(st, B, in, outd, dZ, outd, Wp(P, l), in, dzn, in, nullptr, w.Z[l - 1]);
      dZ = dzn;
      outd = in;
    }
  }
}

// ------------------------------------------------------------ packing
uint16_t bf16_rne(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// float -> OCP E4M3 (bias 7, max 448), round to nearest even, saturating
uint8_t e4m3_rne(float x) {
  const uint8_t sign = std::signbit(x) ? 0x80 : 0x00;
  double a = std::fabs((double)x);
  if (!(a == a)) return 0x7f;
  if (a >= 448.0) return sign | 0x7e;
  if (a < std::ldexp(1.0, -6)) {  // subnormal range: quantum 2^-9
    const double m = std::nearbyint(a * 512.0);
    return sign | (uint8_t)m;  // m == 8 is the smallest normal (code 0x08)
  }
  int e;
  const double fr = std::frexp(a, &e);  // a = fr * 2^e, fr in [0.5, 1)
  int ex = e - 1;                        // a = (2 fr) 2^ex
  double m = std::nearbyint((2.0 * fr - 1.0) * 8.0);
  if (m >= 8.0) {
    m = 0.0;
    ++ex;
  }
  const int be = ex + 7;
  if (be > 15 || (be == 15 && m > 6.0)) return sign | 0x7e;
  return sign | (uint8_t)((be << 3) | (int)m);
}

}  // namespace

// ----------------------------------------------------------- C ABI
void bqp_mlp_release(bqp_ctx* c) {
  if (!c || !c->mlp) return;
  bqp_mlp_state* s = c->mlp;
  (void)hipStreamSynchronize(c->stream);
  bqp_drop_graph(c);
  (void)hipFree(s->d_params);
  (void)hipFree(s->d_grad);
  (void)hipFree(s->d_m);
  (void)hipFree(s->d_v);
  (void)hipFree(s->d_tab_low);
  (void)hipFree(s->d_packed);
  (void)hipFree(s->d_bias);
  (void)hipFree(s->d_work);
  (void)hipFree(s->d_sur_totals);
  delete s;
  c->mlp = nullptr;
}

// Allocates parameters / optimiser state and writes seeded He-uniform
// initial parameters (zero biases).  Called by bqp_surrogate_init.
int bqp_mlp_alloc_init(bqp_ctx* c, uint64_t seed) {
  bqp_mlp_state* s = c->mlp;
  const size_t bytes = (size_t)MLP_NPARAMS * sizeof(float);
  if (!s->d_params) {
    TRY(hipMalloc(&s->d_params, bytes));
    TRY(hipMalloc(&s->d_grad, bytes));
    TRY(hipMalloc(&s->d_m, bytes));
    TRY(hipMalloc(&s->d_v, bytes));
    TRY(hipMalloc(&s->d_sur_totals, 8 * sizeof(unsigned long long)));
  }
  std::vector<float> h((size_t)MLP_NPARAMS, 0.0f);
  uint64_t k = 0;
  for (int l = 0; l < MLP_LAYERS; ++l) {
    const int in = mlp_in_dim(l), out = mlp_out_dim(l);
    const double a = std::sqrt(6.0 / in) * (l == MLP_LAYERS - 1 ? 0.1 : 1.0);
    float* W = h.data() + mlp_w_off(l);
    for (long i = 0; i < (long)in * out; ++i) {
      const double u = (double)(mix64(seed ^ mix64(++k)) >> 11) * 0x1.0p-53;
      W[i] = (float)((2.0 * u - 1.0) * a);
    }
  }
  TRY(hipMemcpy(s->d_params, h.data(), bytes, hipMemcpyHostToDevice));
  TRY(hipMemset(s->d_m, 0, bytes));
  TRY(hipMemset(s->d_v, 0, bytes));
  s->adam_t = 0;
  s->packed_fmt = -1;
  return BQP_OK;
}

extern "C" {

int64_t bqp_surrogate_param_count(void) { return MLP_NPARAMS; }

int32_t bqp_surrogate_get_params(bqp_ctx* c, float* params, double* out_scale) {
  if (!c) return BQP_E_INVALID;
  if (!c->mlp || !c->mlp->d_params) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (params) {
    TRY(hipStreamSynchronize(c->stream));
    TRY(hipMemcpy(params, c->mlp->d_params, (size_t)MLP_NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
  }
  if (out_scale)
    for (int i = 0; i < 3; ++i) out_scale[i] = c->mlp->out_scale[i];
  return BQP_OK;
}

int32_t bqp_surrogate_set_params(bqp_ctx* c, const float* params, const double* out_scale) {
  if (!c) return BQP_E_INVALID;
  if (!c->mlp || !c->mlp->d_params) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (params) {
    for (long i = 0; i < MLP_NPARAMS; ++i)
      if (!std::isfinite(params[i])) return fail(c, BQP_E_INVALID, "non-finite parameter");
    TRY(hipStreamSynchronize(c->stream));
    TRY(hipMemcpy(c->mlp->d_params, params, (size_t)MLP_NPARAMS * sizeof(float), hipMemcpyHostToDevice));
    c->mlp->packed_fmt = -1;
  }
  if (out_scale) {
    for (int i = 0; i < 3; ++i)
      if (!(out_scale[i] > 0.0) || !std::isfinite(out_scale[i]))
        return fail(c, BQP_E_INVALID, "out_scale must be finite and > 0");
    for (int i = 0; i < 3; ++i) c->mlp->out_scale[i] = out_scale[i];
    c->mlp->packed_fmt = -1;
  }
  return BQP_OK;
}

int32_t bqp_surrogate_train_dev(bqp_ctx* c, const float* d_feat, const float* d_targ, int64_t n, int32_t steps,
                                int32_t batch, double lr, uint64_t seed, float* loss) {
  if (!c) return BQP_E_INVALID;
  bqp_mlp_state* s = c->mlp;
  if (!s || !s->d_params) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (n <= 0 || !d_feat || !d_targ) return fail(c, BQP_E_INVALID, "need n > 0 feature and target rows");
  if (steps < 0 || steps > MAX_LOSS_STEPS) return fail(c, BQP_E_INVALID, "steps must be in 0..65536 per call");
  if (batch < 1 || batch > (1 << 17)) return fail(c, BQP_E_INVALID, "batch must be in 1..131072");
  if (!(lr > 0.0) || !std::isfinite(lr)) return fail(c, BQP_E_INVALID, "lr must be finite and > 0");
  Work w;
  int rc = ensure_work(c, batch, w);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f;
  const unsigned pb = (unsigned)((MLP_NPARAMS + 255) / 256);
  for (int it = 0; it < steps; ++it) {
    const unsigned long long sd = mix64(seed ^ mix64(0x5eedull + (unsigned long long)it));
    hipLaunchKernelGGL(gather_kernel, dim3((batch + 255) / 256), dim3(256), 0, c->stream, d_feat, d_targ,
                       (long long)n, batch, sd, w.X, w.T);
    forward(c->stream, s->d_params, w.X, batch, w);
    loss_backward(c->stream, s->d_params, s->d_grad, w.X, w.T, batch, w, w.lsteps + it);
    ++s->adam_t;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)s->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)s->adam_t)));
    hipLaunchKernelGGL(adam_kernel, dim3(pb), dim3(256), 0, c->stream, s->d_params, s->d_grad, s->d_m, s->d_v,
                       (long long)MLP_NPARAMS, (float)lr, b1, b2, 1e-8f, c1, c2);
  }
  TRY(hipGetLastError());
  TRY(hipStreamSynchronize(c->stream));
  if (loss && steps > 0)
    TRY(hipMemcpy(loss, w.lsteps, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
  s->packed_fmt = -1;
  return BQP_OK;
}

int32_t bqp_surrogate_loss_grad_dev(bqp_ctx* c, const float* d_feat, const float* d_targ, int64_t n, double* loss,
                                    float* grad) {
  if (!c) return BQP_E_INVALID;
  bqp_mlp_state* s = c->mlp;
  if (!s || !s->d_params) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (n <= 0 || n > (1 << 17) || !d_feat || !d_targ) return fail(c, BQP_E_INVALID, "need 0 < n <= 131072 rows");
  Work w;
  int rc = ensure_work(c, (int)n, w);
  if (rc) return rc;
  forward(c->stream, s->d_params, d_feat, (int)n, w);
  loss_backward(c->stream, s->d_params, s->d_grad, d_feat, d_targ, (int)n, w, nullptr);
  TRY(hipGetLastError());
  TRY(hipStreamSynchronize(c->stream));
  if (loss) TRY(hipMemcpy(loss, w.lossd, sizeof(double), hipMemcpyDeviceToHost));
  if (grad) TRY(hipMemcpy(grad, s->d_grad, (size_t)MLP_NPARAMS * sizeof(float), hipMemcpyDeviceToHost));
  return BQP_OK;
}

int32_t bqp_surrogate_pack(bqp_ctx* c, int32_t format, const float* d_calib, int64_t n_calib) {
  if (!c) return BQP_E_INVALID;
  bqp_mlp_state* s = c->mlp;
  if (!s || !s->d_params) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (format != BQP_MLP_BF16 && format != BQP_MLP_FP8) return fail(c, BQP_E_INVALID, "format must be BF16 or FP8");
  if (format == BQP_MLP_FP8 && (n_calib <= 0 || !d_calib))
    return fail(c, BQP_E_INVALID, "FP8 packing needs calibration feature rows");
  std::vector<float> P((size_t)MLP_NPARAMS);
  TRY(hipStreamSynchronize(c->stream));
  TRY(hipMemcpy(P.data(), s->d_params, P.size() * sizeof(float), hipMemcpyDeviceToHost));

  // FP8: activation scales of the inputs of hidden layers 1..3 (= outputs H_0..H_2)
  float amax[3] = {1.f, 1.f, 1.f};
  if (format == BQP_MLP_FP8) {
    const int cap = (int)std::min<int64_t>(n_calib, 65536);
    Work w;
    int rc = ensure_work(c, cap, w);
    if (rc) return rc;
    std::vector<float> part(256);
    for (int l = 0; l < 3; ++l) amax[l] = 0.f;
    for (int64_t off = 0; off < n_calib; off += cap) {
      const int B = (int)std::min<int64_t>(cap, n_calib - off);
      forward(c->stream, s->d_params, d_calib + off * MLP_IN, B, w);
      for (int l = 0; l < 3; ++l) {
        hipLaunchKernelGGL(absmax_kernel, dim3(256), dim3(256), 0, c->stream, w.H[l], (long long)B * MLP_WIDTH,
                           w.amax);
        TRY(hipGetLastError());
        TRY(hipMemcpyAsync(part.data(), w.amax, 256 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        TRY(hipStreamSynchronize(c->stream));
        for (float v : part) amax[l] = std::max(amax[l], v);
      }
    }
    for (int l = 0; l < 3; ++l)
      if (!(amax[l] > 0.f) || !std::isfinite(amax[l])) amax[l] = 1.f;
  }

  const size_t n0 = (size_t)MLP_TILES * 64 * 8;                // layer 0 elements
  const size_t nh = (size_t)MLP_TILES * MLP_CHUNKS * 64 * 8;   // per hidden layer
  const size_t n4 = (size_t)MLP_CHUNKS * 64 * 8;               // output layer
  const size_t hbytes = (format == BQP_MLP_FP8) ? nh : 2 * nh;
  const size_t bytes = 2 * n0 + 3 * hbytes + 2 * n4;
  std::vector<uint8_t> buf(bytes, 0);
  uint16_t* w0 = (uint16_t*)buf.data();
  uint8_t* wh = buf.data() + 2 * n0;
  uint16_t* w4 = (uint16_t*)(buf.data() + 2 * n0 + 3 * hbytes);
  for (int mt = 0; mt < MLP_TILES; ++mt)
    for (int ln = 0; ln < 64; ++ln)
      for (int j = 0; j < 8; ++j) {
        const int row = 16 * mt + (ln & 15), k = 8 * (ln >> 4) + j;
        const float v = k < MLP_IN ? P[mlp_w_off(0) + (size_t)row * MLP_IN + k] : 0.f;
        w0[((size_t)mt * 64 + ln) * 8 + j] = bf16_rne(v);
      }
  float wscale[3] = {1.f, 1.f, 1.f};
  for (int h = 0; h < 3; ++h) {
    const float* W = P.data() + mlp_w_off(h + 1);
    if (format == BQP_MLP_FP8) {
      float m = 0.f;
      for (size_t i = 0; i < (size_t)MLP_WIDTH * MLP_WIDTH; ++i) m = std::max(m, std::fabs(W[i]));
      wscale[h] = m > 0.f ? m / 448.0f : 1.0f;
    }
    for (int mt = 0; mt < MLP_TILES; ++mt)
      for (int kc = 0; kc < MLP_CHUNKS; ++kc)
        for (int ln = 0; ln < 64; ++ln)
          for (int j = 0; j < 8; ++j) {
            const int row = 16 * mt + (ln & 15), k = mlp_perm(kc, ln >> 4, j);
            const float v = W[(size_t)row * MLP_WIDTH + k];
            const size_t e = (((size_t)mt * MLP_CHUNKS + kc) * 64 + ln) * 8 + j;
            if (format == BQP_MLP_FP8)
              wh[(size_t)h * hbytes + e] = e4m3_rne(v / wscale[h]);
            else
              ((uint16_t*)(wh + (size_t)h * hbytes))[e] = bf16_rne(v);
          }
  }
  for (int kc = 0; kc < MLP_CHUNKS; ++kc)
    for (int ln = 0; ln < 64; ++ln)
      for (int j = 0; j < 8; ++j) {
        const int row = ln & 15, k = mlp_perm(kc, ln >> 4, j);
        const float v = row < MLP_OUT ? P[mlp_w_off(4) + (size_t)row * MLP_WIDTH + k] : 0.f;
        w4[((size_t)kc * 64 + ln) * 8 + j] = bf16_rne(v);
      }
  std::vector<float> bias(MLP_BIAS_FLOATS, 0.f);
  for (int l = 0; l < 4; ++l)
    for (int i = 0; i < MLP_WIDTH; ++i) bias[(size_t)l * MLP_WIDTH + i] = P[mlp_b_off(l) + i];
  for (int i = 0; i < MLP_OUT; ++i) bias[4 * MLP_WIDTH + i] = P[mlp_b_off(4) + i];

  (void)hipStreamSynchronize(c->stream);
  bqp_drop_graph(c);
  (void)hipFree(s->d_packed);
  (void)hipFree(s->d_bias);
  s->d_packed = nullptr;
  s->d_bias = nullptr;
  s->packed_fmt = -1;
  TRY(hipMalloc(&s->d_packed, bytes));
  TRY(hipMalloc(&s->d_bias, bias.size() * sizeof(float)));
  TRY(hipMemcpy(s->d_packed, buf.data(), bytes, hipMemcpyHostToDevice));
  TRY(hipMemcpy(s->d_bias, bias.data(), bias.size() * sizeof(float), hipMemcpyHostToDevice));
  MlpPacked& pk = s->packed;
  uint8_t* base = (uint8_t*)s->d_packed;
  pk.w0 = (const uint4*)base;
  for (int h = 0; h < 3; ++h) pk.wh[h] = base + 2 * n0 + (size_t)h * hbytes;
  pk.w4 = (const uint4*)(base + 2 * n0 + 3 * hbytes);
  pk.bias = s->d_bias;
  for (int h = 0; h < 3; ++h) {
    const float sa = amax[h] / 448.0f;  // activation scale of the input of hidden layer h+1
    pk.hscale[h] = format == BQP_MLP_FP8 ? wscale[h] * sa : 1.0f;
    pk.ainv[h] = format == BQP_MLP_FP8 ? 1.0f / sa : 1.0f;
  }
  for (int i = 0; i < 3; ++i) pk.out_scale[i] = (float)s->out_scale[i];
  const double rmid = 0.5 * (s->r_lo + s->r_hi), rhalf = 0.5 * (s->r_hi - s->r_lo);
  pk.r_mid = (float)rmid;
  pk.r_inv_half = (float)(1.0 / rhalf);
  pk.fmt = format;
  s->packed_fmt = format;
  return BQP_OK;
}

int32_t bqp_surrogate_infer_dev(bqp_ctx* c, int32_t format, const float* d_feat, int64_t n, float* d_out,
                                double* kernel_ms) {
  if (!c) return BQP_E_INVALID;
  bqp_mlp_state* s = c->mlp;
  if (!s || !s->d_params) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (n < 0 || (n > 0 && (!d_feat || !d_out))) return fail(c, BQP_E_INVALID, "bad buffers");
  if (format != BQP_MLP_FP32 && format != BQP_MLP_BF16 && format != BQP_MLP_FP8)
    return fail(c, BQP_E_INVALID, "format must be 0 (FP32), 1 (BF16) or 2 (FP8)");
  if (format != BQP_MLP_FP32 && s->packed_fmt != format)
    return fail(c, BQP_E_INVALID, "weights not packed in this format (bqp_surrogate_pack)");
  if (n == 0) {
    if (kernel_ms) *kernel_ms = 0.0;
    return BQP_OK;
  }
  if (!c->ev0) {
    TRY(hipEventCreate(&c->ev0));
    TRY(hipEventCreate(&c->ev1));
  }
  if (format == BQP_MLP_FP32) {
    const int cap = (int)std::min<int64_t>(n, 65536);
    Work w;
    int rc = ensure_work(c, cap, w);
    if (rc) return rc;
    TRY(hipEventRecord(c->ev0, c->stream));
    for (int64_t off = 0; off < n; off += cap) {
      const int B = (int)std::min<int64_t>(cap, n - off);
      forward(c->stream, s->d_params, d_feat + off * MLP_IN, B, w);
      TRY(hipMemcpyAsync(d_out + off * MLP_OUT, w.Y, (size_t)B * MLP_OUT * sizeof(float), hipMemcpyDeviceToDevice,
                         c->stream));
    }
    TRY(hipEventRecord(c->ev1, c->stream));
  } else {
    constexpr int NT = 2;
    const long long tiles = (n + 16 * NT - 1) / (16 * NT);
    const long long want = (tiles + 3) / 4;  // 4 waves per block
    const unsigned blocks = (unsigned)std::max<long long>(1, std::min<long long>(want, (long long)c->cus * 8));
    TRY(hipEventRecord(c->ev0, c->stream));
    if (format == BQP_MLP_FP8)
      hipLaunchKernelGGL((mlp_infer_kernel<MLP_FMT_FP8, NT>), dim3(blocks), dim3(256), 0, c->stream, s->packed,
                         d_feat, (long long)n, d_out);
    else
      hipLaunchKernelGGL((mlp_infer_kernel<MLP_FMT_BF16, NT>), dim3(blocks), dim3(256), 0, c->stream, s->packed,
                         d_feat, (long long)n, d_out);
    TRY(hipEventRecord(c->ev1, c->stream));
  }
  TRY(hipGetLastError());
  TRY(hipEventSynchronize(c->ev1));
  if (kernel_ms) {
    float ms = 0.f;
    TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *kernel_ms = ms;
  }
  return BQP_OK;
}

int32_t bqp_surrogate_set_check(bqp_ctx* c, int32_t check_every, double threshold) {
  if (!c) return BQP_E_INVALID;
  if (!c->mlp) return fail(c, BQP_E_INVALID, "surrogate not initialised (bqp_surrogate_init)");
  if (check_every < 0 || !(threshold >= 0.0)) return fail(c, BQP_E_INVALID, "need check_every >= 0, threshold >= 0");
  c->mlp->check_every = check_every;
  c->mlp->threshold = threshold;
  return BQP_OK;
}

int32_t bqp_surrogate_last_stats(const bqp_ctx* c, bqp_surrogate_stats* out) {
  if (!c || !out) return BQP_E_INVALID;
  *out = c->mlp ? c->mlp->last : bqp_surrogate_stats{};
  return BQP_OK;
}

}  // extern "C"
