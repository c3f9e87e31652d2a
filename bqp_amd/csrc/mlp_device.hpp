// Learned gravity surrogate: the network and its gfx950 MFMA forward pass.
//
// No reference counterpart: the reference evaluates every harmonic term
// exactly (SC/src/gravity.rs:2638).  The surrogate replaces the part of the
// field above degree NLOW (default 4) with a small network,
//
//   features(r_body) (16) -> 4 x [256, SiLU] -> 3  (scaled residual accel)
//
// features: u = r/|r| (3), s = (|r| - r_mid)/r_half, sin/cos(pi u_i),
// sin/cos(2 pi u_i)  (order: ux uy uz s, sin(pi ux) cos(pi ux) sin(pi uy)
// cos(pi uy) sin(pi uz) cos(pi uz), then the same at 2 pi).
//
// Parameters (FP32 master copy, flat): for each layer l = 0..4, W_l
// (out x in, row-major) followed by b_l (out).
//
// Inference uses v_mfma_f32_16x16x32_{bf16,fp8_fp8} in the transposed form
// H_out^T = W . H_in^T: the A operand is a 16 x 32 tile of W (rows = output
// neurons), the B operand a 32 x 16 tile of activations (columns = 16
// samples), the FP32 accumulator a 16 x 16 tile.  Operand layouts (one wave):
//   A: lane l holds row l%16, k = 8(l/16) .. 8(l/16)+7
//   B: lane l holds column l%16, k = 8(l/16) .. 8(l/16)+7
//   C: lane l holds column l%16, rows 4(l/16) .. 4(l/16)+3
// A layer's accumulators therefore leave lane l holding neurons
// 16 mt + 4 q + r (q = l/16) of sample l%16, which is already the B operand
// of the next layer if the k order inside each 32-wide chunk is permuted:
// k-slot j of lane group q in chunk kc is neuron perm(kc, q, j) =
// 32 kc + 4 q + j (j < 4) or 32 kc + 16 + 4 q + (j - 4) (j >= 4).  The host
// packs W with the same permutation, so activations never leave registers
// between layers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace bqp {

// splitmix64 finaliser: the counter-based generator behind the seeded
// training data and minibatch draws
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr int MLP_IN = 16;
constexpr int MLP_WIDTH = 256;
constexpr int MLP_OUT = 3;
constexpr int MLP_LAYERS = 5;  // 4 hidden (SiLU) + linear output
constexpr int MLP_TILES = MLP_WIDTH / 16;   // 16 output tiles per hidden layer
constexpr int MLP_CHUNKS = MLP_WIDTH / 32;  // 8 k-chunks per hidden layer

__host__ __device__ constexpr int mlp_in_dim(int l) { return l == 0 ? MLP_IN : MLP_WIDTH; }
__host__ __device__ constexpr int mlp_out_dim(int l) { return l == MLP_LAYERS - 1 ? MLP_OUT : MLP_WIDTH; }
__host__ __device__ constexpr long mlp_w_off(int l) {
  return l == 0 ? 0 : mlp_w_off(l - 1) + (long)mlp_in_dim(l - 1) * mlp_out_dim(l - 1) + mlp_out_dim(l - 1);
}
__host__ __device__ constexpr long mlp_b_off(int l) { return mlp_w_off(l) + (long)mlp_in_dim(l) * mlp_out_dim(l); }
constexpr long MLP_NPARAMS = mlp_w_off(MLP_LAYERS);  // 202,499
// algorithmic flops of one evaluation (multiply-add = 2)
constexpr double MLP_FLOPS = 2.0 * (MLP_IN * MLP_WIDTH + 3.0 * MLP_WIDTH * MLP_WIDTH + MLP_WIDTH * MLP_OUT);

__host__ __device__ constexpr int mlp_perm(int kc, int q, int j) {
  return 32 * kc + (j < 4 ? 4 * q + j : 16 + 4 * q + (j - 4));
}

enum { MLP_FMT_FP32 = 0, MLP_FMT_BF16 = 1, MLP_FMT_FP8 = 2 };

// Packed inference weights (byte offsets into one buffer).  Layer 0 and the
// output layer are BF16 in both formats; the three 256 x 256 hidden layers
// are BF16 or FP8 E4M3 (OCP, the gfx950 format) with a per-layer scale.
//   layer 0:     [16 mt][64 lane][8]        bf16  (k = feature 8q + j, zero past 16)
//   layers 1-3:  [16 mt][8 kc][64 lane][8]  bf16 or fp8
//   layer 4:     [8 kc][64 lane][8]         bf16  (rows >= 3 zero)
struct MlpPacked {
  const uint4* w0;                    // 16 * 64 uint4
  const void* wh[3];                  // bf16: 16*8*64 uint4; fp8: 16*8*64 uint2
  const uint4* w4;                    // 8 * 64 uint4
  const float* bias;                  // natural order: b0 (256), b1, b2, b3, b4 (3, padded to 16)
  float hscale[3];                    // FP8: dequantisation factor of hidden layer 1..3 (w_scale * a_scale)
  float ainv[3];                      // FP8: 1 / activation scale of the input of hidden layer 1..3
  float out_scale[3];                 // residual (m/s^2) = output * out_scale
  float r_mid, r_inv_half;            // feature radius normalisation (m)
  int fmt;
};

constexpr size_t MLP_BIAS_FLOATS = 4 * MLP_WIDTH + 16;

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float mlp_silu(float z) { return z / (1.0f + __expf(-z)); }

__device__ __forceinline__ void mlp_features(float x, float y, float z, float r_mid, float r_inv_half, float* f) {
  const float r = sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
  const float ir = 1.0f / r;
  const float u[3] = {x * ir, y * ir, z * ir};
  f[0] = u[0];
  f[1] = u[1];
  f[2] = u[2];
  f[3] = (r - r_mid) * r_inv_half;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float s1, c1;
    sincospif(u[i], &s1, &c1);
    f[4 + 2 * i] = s1;
    f[5 + 2 * i] = c1;
    f[10 + 2 * i] = 2.0f * s1 * c1;          // sin(2 pi u)
    f[11 + 2 * i] = fmaf(-2.0f * s1, s1, 1.0f);  // cos(2 pi u)
  }
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
  const __bf16 ha = (__bf16)a, hb = (__bf16)b;
  return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}

__device__ __forceinline__ float clamp448(float v) { return fminf(fmaxf(v, -448.0f), 448.0f); }

// 4 floats -> 4 fp8 (one dword), values clamped to the finite range first
__device__ __forceinline__ uint32_t pack_fp8x4(float a, float b, float c, float d) {
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(clamp448(a), clamp448(b), 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(clamp448(c), clamp448(d), w, true);
  return (uint32_t)w;
}

// Activation operand of one 32-wide k chunk, built from the accumulators of
// output tiles 2kc and 2kc+1 (see perm above).
struct ActBF16 {
  uint4 v[MLP_CHUNKS];
};
struct ActFP8 {
  uint2 v[MLP_CHUNKS];
};

__device__ __forceinline__ void act_store(ActBF16& a, int kc, const f32x4& lo, const f32x4& hi, float) {
  a.v[kc] = make_uint4(pack_bf16x2(lo[0], lo[1]), pack_bf16x2(lo[2], lo[3]), pack_bf16x2(hi[0], hi[1]),
                       pack_bf16x2(hi[2], hi[3]));
}
__device__ __forceinline__ void act_store(ActFP8& a, int kc, const f32x4& lo, const f32x4& hi, float inv) {
  a.v[kc] = make_uint2(pack_fp8x4(lo[0] * inv, lo[1] * inv, lo[2] * inv, lo[3] * inv),
                       pack_fp8x4(hi[0] * inv, hi[1] * inv, hi[2] * inv, hi[3] * inv));
}

__device__ __forceinline__ f32x4 mfma_bf16(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a), as_bf16x8(b), c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_fp8(uint2 a, uint2 b, f32x4 c) {
  const long la = (long)(((unsigned long)a.y << 32) | a.x);
  const long lb = (long)(((unsigned long)b.y << 32) | b.x);
  return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(la, lb, c, 0, 0, 0);
}

// bias (+ optional dequantisation) and SiLU of accumulator tile mt, lane group q
__device__ __forceinline__ f32x4 mlp_epilogue(f32x4 acc, const float* bias, int mt, int q, float scale,
                                              bool silu) {
  const float4 b = *reinterpret_cast<const float4*>(bias + 16 * mt + 4 * q);
  f32x4 z;
  z[0] = fmaf(acc[0], scale, b.x);
  z[1] = fmaf(acc[1], scale, b.y);
  z[2] = fmaf(acc[2], scale, b.z);
  z[3] = fmaf(acc[3], scale, b.w);
  if (silu) {
#pragma unroll
    for (int r = 0; r < 4; ++r) z[r] = mlp_silu(z[r]);
  }
  return z;
}

// One hidden 256 x 256 layer for NT sample tiles: acc = W . act_in, then the
// epilogue writes act_out (format of the next layer's input).
template <int FMT_IN, int NT, class ActIn, class ActOut>
__device__ __forceinline__ void mlp_hidden_layer(const void* w, const float* bias, float scale, float next_inv,
                                                 int lane, const ActIn (&in)[NT], ActOut (&out)[NT]) {
  const int q = lane >> 4;
#pragma unroll
  for (int mp = 0; mp < MLP_TILES / 2; ++mp) {  // pairs of output tiles -> one k chunk of the next layer
    f32x4 acc[NT][2];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t][0] = acc[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < MLP_CHUNKS; ++kc) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int mt = 2 * mp + h;
        if constexpr (FMT_IN == MLP_FMT_FP8) {
          const uint2 a = static_cast<const uint2*>(w)[(mt * MLP_CHUNKS + kc) * 64 + lane];
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[t][h] = mfma_fp8(a, in[t].v[kc], acc[t][h]);
        } else {
          const uint4 a = static_cast<const uint4*>(w)[(mt * MLP_CHUNKS + kc) * 64 + lane];
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[t][h] = mfma_bf16(a, in[t].v[kc], acc[t][h]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const f32x4 lo = mlp_epilogue(acc[t][0], bias, 2 * mp, q, scale, true);
      const f32x4 hi = mlp_epilogue(acc[t][1], bias, 2 * mp + 1, q, scale, true);
      act_store(out[t], mp, lo, hi, next_inv);
    }
    // keep the scheduler from hoisting the next tile pair's weight loads
    // (16 x 16 B per lane) above this one: bounded registers, no spills
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Forward pass of NT tiles of 16 samples held by one wave.  xin[t] is the
// B operand of layer 0 for tile t (lane l: features 8(l/16)..+7 of sample
// l%16 as bf16, zero for k >= 16).  On return, lanes 0..15 hold the three
// outputs of sample 16 t + l in out[t][0..2] (other lanes: unspecified).
// The weight addresses depend only on the lane, so inside a loop over
// sample tiles the compiler would hoist all ~400 weight loads out of the loop
// (and spill them).  Passing the base pointers through an empty asm makes
// them opaque per call, so each load stays next to its MFMA.
template <class T>
__device__ __forceinline__ T* opaque_ptr(T* p) {
  asm volatile("" : "+s"(p));
  return p;
}

template <int FMT, int NT>
__device__ __forceinline__ void mlp_forward_wave(const MlpPacked& PP, int lane, const uint4 (&xin)[NT],
                                                 f32x4 (&out)[NT]) {
  MlpPacked P = PP;
  P.w0 = opaque_ptr(P.w0);
  P.w4 = opaque_ptr(P.w4);
  P.bias = opaque_ptr(P.bias);
#pragma unroll
  for (int h = 0; h < 3; ++h) P.wh[h] = opaque_ptr(P.wh[h]);
  const int q = lane >> 4;
  using ActH = typename std::conditional<FMT == MLP_FMT_FP8, ActFP8, ActBF16>::type;
  ActH h0[NT], h1[NT];
  // layer 0: 16 (padded to 32) -> 256
#pragma unroll
  for (int mp = 0; mp < MLP_TILES / 2; ++mp) {
    const uint4 a0 = P.w0[(2 * mp) * 64 + lane];
    const uint4 a1 = P.w0[(2 * mp + 1) * 64 + lane];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      f32x4 z0 = mfma_bf16(a0, xin[t], f32x4{0.f, 0.f, 0.f, 0.f});
      f32x4 z1 = mfma_bf16(a1, xin[t], f32x4{0.f, 0.f, 0.f, 0.f});
      z0 = mlp_epilogue(z0, P.bias, 2 * mp, q, 1.0f, true);
      z1 = mlp_epilogue(z1, P.bias, 2 * mp + 1, q, 1.0f, true);
      act_store(h0[t], mp, z0, z1, P.ainv[0]);
    }
  }
  // hidden layers 1..3 (ping-pong); the input of the output layer is BF16
  const float one = 1.0f;
  mlp_hidden_layer<FMT, NT>(P.wh[0], P.bias + 1 * MLP_WIDTH, FMT == MLP_FMT_FP8 ? P.hscale[0] : one, P.ainv[1],
                            lane, h0, h1);
  mlp_hidden_layer<FMT, NT>(P.wh[1], P.bias + 2 * MLP_WIDTH, FMT == MLP_FMT_FP8 ? P.hscale[1] : one, P.ainv[2],
                            lane, h1, h0);
  ActBF16 h3[NT];
  mlp_hidden_layer<FMT, NT>(P.wh[2], P.bias + 3 * MLP_WIDTH, FMT == MLP_FMT_FP8 ? P.hscale[2] : one, one, lane,
                            h0, h3);
  // output layer: 256 -> 3 (one 16-row tile)
#pragma unroll
  for (int t = 0; t < NT; ++t) out[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kc = 0; kc < MLP_CHUNKS; ++kc) {
    const uint4 a = P.w4[kc * 64 + lane];
#pragma unroll
    for (int t = 0; t < NT; ++t) out[t] = mfma_bf16(a, h3[t].v[kc], out[t]);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) out[t] = mlp_epilogue(out[t], P.bias + 4 * MLP_WIDTH, 0, q, 1.0f, false);
}

}  // namespace bqp
