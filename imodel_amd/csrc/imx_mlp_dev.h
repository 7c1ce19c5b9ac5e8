// Device-side pieces of the libimx MLP shared by the HIP translation units:
// parameter and fragment layouts, BF16 conversion, GELU, the MFMA wrapper,
// and the inference kernel with LDS-resident weights, templated on how a
// wave obtains its input fragments (stored BF16 rows, or features gathered
// from the model state by the learned-correction coupling).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

#define MLP_IN 48
#define MLP_H 128
#define MLP_OUT 2
#define MLP_WAVES 4  // waves per block, 16 edges each

// parameter offsets in the flat FP32 master vector
#define P_W1 0
#define P_B1 (P_W1 + MLP_IN * MLP_H)
#define P_W2 (P_B1 + MLP_H)
#define P_B2 (P_W2 + MLP_H * MLP_H)
#define P_W3 (P_B2 + MLP_H)
#define P_B3 (P_W3 + MLP_H * MLP_H)
#define P_W4 (P_B3 + MLP_H)
#define P_B4 (P_W4 + MLP_H * MLP_OUT)
#define P_N (P_B4 + MLP_OUT)

// forward fragments: layer l has (N/16) x (K/16) tiles of 64 lanes x 4 bf16
#define FF_1 0
#define FF_2 (FF_1 + MLP_IN * MLP_H)
#define FF_3 (FF_2 + MLP_H * MLP_H)
#define FF_4 (FF_3 + MLP_H * MLP_H)
#define FF_N (FF_4 + MLP_H * 16)
// The inference copy of the same weights (K32 fragments, same offsets and
// size): per output tile of layer 1 a K32 fragment of 64 lanes x 8 bf16
// then a K16 fragment of 64 x 4; layers 2..4 (K = 128) four K32 fragments
// per output tile, tile-major.
// backward fragments (layers 2..4, transposed)
#define FB_2 0
#define FB_3 (FB_2 + MLP_H * MLP_H)
#define FB_4 (FB_3 + MLP_H * MLP_H)
#define FB_N (FB_4 + MLP_H * 16)

__device__ __forceinline__ short f2bf(float x) {
  __hip_bfloat16 b = __float2bfloat16(x);
  return *reinterpret_cast<short *>(&b);
}
__device__ __forceinline__ float bf2f(short s) {
  __hip_bfloat16 b = *reinterpret_cast<__hip_bfloat16 *>(&s);
  return __bfloat162float(b);
}

// Activation: GELU in its tanh form (Hendrycks & Gimpel 2016),
//   g(x) = x/2 (1 + tanh(c (x + a x^3))) = x s(x),
//   s(x) = sigma(2 c (x + a x^3)) = 1 / (1 + exp(-2 c (x + a x^3))),
// c = sqrt(2/pi), a = 0.044715: one exponential and one reciprocal
// (v_exp_f32, v_rcp_f32) per activation. The exponent is written base 2,
// exp(-2 c (x + a x^3)) = 2^(x (k1 + k2 x^2)) with k1 = -2 c log2(e) and
// k2 = a k1, so it costs one multiply and one FMA before v_exp_f32. Large
// |x| saturates s to 0 or 1 (2^u overflows to inf, whose reciprocal is 0).
// g'(x) = s + x s (1 - s) 2 c (1 + 3 a x^2).
#define GELU_C2 1.5957691216057308f  // 2 sqrt(2/pi)
#define GELU_A 0.044715f
#define GELU_K1 (-2.302208198144325f)    // -2 sqrt(2/pi) log2(e)
#define GELU_K2 (-0.1029432395800235f)   // GELU_K1 * 0.044715
__device__ __forceinline__ float gelu_s(float x) {
  const float u = x * __builtin_fmaf(GELU_K2, x * x, GELU_K1);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
__device__ __forceinline__ float gelu_d(float x) {
  const float s = gelu_s(x);
  return s + x * s * (1.0f - s) * GELU_C2 * __builtin_fmaf(3.0f * GELU_A, x * x, 1.0f);
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// 16x16x32 BF16 MFMA (gfx950), twice the contraction of the 16x16x16 form
// at the same issue cost. Lane l holds A row l % 16 and B column l % 16 at
// k = 8 (l / 16) + i, i = 0..7; D is laid out as for the 16x16x16 form.
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                 0);
}

// two floats to a packed BF16 pair (round to nearest even), low half first
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v2f{lo, hi}, v2bf));
}

// Inference with the K32 forward weight fragments (FF_N bf16 = 80 KB)
// staged once per block in LDS; each wave evaluates ET tiles of 16 edges,
// so every weight fragment read from LDS feeds ET MFMAs. Waves loop over
// tiles (grid-stride), so a grid of a few blocks per CU amortises the
// staging.
//
// Activations stay in the MFMA result layout: after a layer, lane l holds
// features 16 m + 4 (l / 16) + i (i < 4) of output tile m for edge l % 16.
// The B operand of K32 chunk c is the concatenation of tiles 2c and 2c+1,
// so its k = 8 (l / 16) + i stands for feature 32 c + kappa(l / 16, i),
//   kappa(q, i) = 4 q + i (i < 4),  16 + 4 q + i - 4 (i >= 4),
// and the packed weights follow the same permutation (mlp_pack), which
// leaves each dot product unchanged.
//
// ld(e0, n, lane, x) fills the input fragments x[kc][t] of edges
// e0 + 16 t + lane % 16 (features 16 kc + 4 (lane / 16) + i), zeros past n;
// x[0], x[1] form K32 chunk 0 of the first layer and x[2] its K16 tail.
#define MLP_FWD_WAVES 4
// GELU of activation tiles z0, z1 (features 16 m .. 16 m + 31) as the K32
// operand of their chunk
__device__ __forceinline__ v8s gelu_chunk(const v4f &z0, const v4f &z1) {
  uint32_t p[4];
  p[0] = f2bf2(gelu(z0[0]), gelu(z0[1]));
  p[1] = f2bf2(gelu(z0[2]), gelu(z0[3]));
  p[2] = f2bf2(gelu(z1[0]), gelu(z1[1]));
  p[3] = f2bf2(gelu(z1[2]), gelu(z1[3]));
  return __builtin_bit_cast(v8s, p);
}

// Hidden layer zn = b + W^T g(z) (K = N = 128), chunk-outer: chunk c of
// the input needs only tiles 2c, 2c+1 of z, so the activation of chunk
// c + 1 can issue while the MFMAs of chunk c run. Each output tile still
// sums its chunks in ascending order.
template <int ET>
__device__ __forceinline__ void layer_pipe(const short *Ws, const float *__restrict__ b, const v4f (&z)[8][ET],
                                           v4f (&zn)[8][ET], int lane) {
#pragma unroll
  for (int nn = 0; nn < 8; ++nn) {
    v4f bias;
#pragma unroll
    for (int i = 0; i < 4; ++i) bias[i] = b[nn * 16 + 4 * (lane / 16) + i];
#pragma unroll
    for (int t = 0; t < ET; ++t) zn[nn][t] = bias;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    v8s a[ET];
#pragma unroll
    for (int t = 0; t < ET; ++t) a[t] = gelu_chunk(z[2 * c][t], z[2 * c + 1][t]);
#pragma unroll
    for (int nn = 0; nn < 8; ++nn) {
      const v8s w = *reinterpret_cast<const v8s *>(Ws + ((nn * 4 + c) * 64 + lane) * 8);
#pragma unroll
      for (int t = 0; t < ET; ++t) zn[nn][t] = mfma32(w, a[t], zn[nn][t]);
    }
  }
}

template <int ET, class Loader>
__global__ __launch_bounds__(64 * MLP_FWD_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
void mlp_fwd_lds(int64_t n, Loader ld, const v8s *__restrict__ Wf, const float *__restrict__ P,
                 float *__restrict__ Y) {
  __shared__ v8s Wv[FF_N / 8];
  for (int i = threadIdx.x; i < FF_N / 8; i += 64 * MLP_FWD_WAVES) Wv[i] = Wf[i];
  __syncthreads();
  const short *Ws = reinterpret_cast<const short *>(Wv);
  const int lane = threadIdx.x & 63;
  const int64_t per = 16 * ET;
  const int64_t ntiles = (n + per - 1) / per;
  for (int64_t tile = (int64_t)blockIdx.x * MLP_FWD_WAVES + (threadIdx.x >> 6); tile < ntiles;
       tile += (int64_t)gridDim.x * MLP_FWD_WAVES) {
    const int64_t e0 = tile * per;
    // keep the weight reads inside the loop: hoisted out of it, the whole
    // 80 KB fragment set would be held in registers
    asm volatile("" ::: "memory");
    v4s x[3][ET];
    ld(e0, n, lane, x);
    v4f z[8][ET], z2[8][ET];
    // layer 1: per output tile one K32 fragment (features 0..31) then one
    // K16 fragment (32..47), 768 bf16. The tail runs as a K32 MFMA with
    // zero upper halves (slots i < 4 of a K32 chunk are the K16 layout):
    // a 16x16x16 MFMA reading a 16x16x32 result as its accumulator right
    // after it gave wrong sums (the compiler inserts no wait between the
    // two shapes), while chains of one shape are exact.
    const v4s zero4 = {0, 0, 0, 0};
#pragma unroll
    for (int nn = 0; nn < 8; ++nn) {
      v4f bias;
#pragma unroll
      for (int i = 0; i < 4; ++i) bias[i] = P[P_B1 + nn * 16 + 4 * (lane / 16) + i];
      const v8s w = *reinterpret_cast<const v8s *>(Ws + FF_1 + nn * 768 + lane * 8);
      const v4s w2 = *reinterpret_cast<const v4s *>(Ws + FF_1 + nn * 768 + 512 + lane * 4);
      const v8s w2p = __builtin_shufflevector(w2, zero4, 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
      for (int t = 0; t < ET; ++t) {
        const v8s x01 = __builtin_shufflevector(x[0][t], x[1][t], 0, 1, 2, 3, 4, 5, 6, 7);
        const v8s x2p = __builtin_shufflevector(x[2][t], zero4, 0, 1, 2, 3, 4, 5, 6, 7);
        z[nn][t] = mfma32(w2p, x2p, mfma32(w, x01, bias));
      }
    }
    layer_pipe<ET>(Ws + FF_2, P + P_B2, z, z2, lane);
    layer_pipe<ET>(Ws + FF_3, P + P_B3, z2, z, lane);
    // output layer: one tile of 16 feature rows, 2 real
    v4f y[ET];
#pragma unroll
    for (int t = 0; t < ET; ++t) {
      y[t] = v4f{0.f, 0.f, 0.f, 0.f};
      if (lane < 16) {
        y[t][0] = P[P_B4];
        y[t][1] = P[P_B4 + 1];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const v8s w = *reinterpret_cast<const v8s *>(Ws + FF_4 + (c * 64 + lane) * 8);
#pragma unroll
      for (int t = 0; t < ET; ++t) y[t] = mfma32(w, gelu_chunk(z[2 * c][t], z[2 * c + 1][t]), y[t]);
    }
    if (lane < 16) {
#pragma unroll
      for (int t = 0; t < ET; ++t) {
        const int64_t e = e0 + 16 * t + lane;
        if (e < n) {
          Y[e * 2] = y[t][0];
          Y[e * 2 + 1] = y[t][1];
        }
      }
    }
  }
}

// input fragments from stored BF16 rows X[n][48]
struct RowLoader {
  const short *X;
  template <int ET>
  __device__ __forceinline__ void operator()(int64_t e0, int64_t n, int lane, v4s (&x)[3][ET]) const {
#pragma unroll
    for (int t = 0; t < ET; ++t) {
      const int64_t e = e0 + 16 * t + (lane % 16);
#pragma unroll
      for (int kc = 0; kc < 3; ++kc) {
        if (e < n) x[kc][t] = *reinterpret_cast<const v4s *>(X + e * MLP_IN + kc * 16 + 4 * (lane / 16));
        else x[kc][t] = v4s{0, 0, 0, 0};
      }
    }
  }
};

// Launch the LDS inference kernel with the K32 fragments: ET = 2 tiles per
// wave, at most two resident blocks per CU (80 KB of LDS each) on 256 CUs.
// The grid is one resident set (512 blocks) for small inputs; from 32768
// tiles on, two sets (1024 blocks) balance the tail better (measured on
// random weights: 7-10 % faster at 1.97 M edges, 15 % slower at 123 k).
template <class Loader>
inline hipError_t mlp_fwd_launch(const short *Ff32, const float *P, int64_t n, Loader ld, float *Y,
                                 hipStream_t st) {
  constexpr int ET = 2;
  const int64_t ntiles = (n + 16 * ET - 1) / (16 * ET);
  int64_t blocks = (ntiles + MLP_FWD_WAVES - 1) / MLP_FWD_WAVES;
  const int64_t cap = ntiles >= 32768 ? 1024 : 512;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  mlp_fwd_lds<ET, Loader><<<(unsigned)blocks, 64 * MLP_FWD_WAVES, 0, st>>>(n, ld, (const v8s *)Ff32, P, Y);
  return hipGetLastError();
}
