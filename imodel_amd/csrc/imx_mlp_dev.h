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

// Activation: GELU with its sigmoid replaced by the hard (piecewise
// linear) sigmoid matching sigma(1.702 x) at 0,
//   g(x) = x s(x),  s(x) = clamp(0.5 + 0.4255 x, 0, 1),
// three full-rate vector instructions (FMA, clamp, multiply). Smooth GELU
// forms need exp and a reciprocal, transcendental instructions that issue
// at a quarter of the vector rate and bound the network, not the MFMAs.
// g'(x) = s + 0.4255 x on 0 < s < 1, else s.
#define GELU_K 0.4255f
__device__ __forceinline__ float gelu_s(float x) {
  return __builtin_amdgcn_fmed3f(0.5f + GELU_K * x, 0.0f, 1.0f);
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
__device__ __forceinline__ float gelu_d(float x) {
  const float s = gelu_s(x);
  return (s > 0.0f && s < 1.0f) ? s + GELU_K * x : s;
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Inference with the forward weight fragments (FF_N bf16 = 80 KB) staged
// once per block in LDS; each wave evaluates ET tiles of 16 edges, so every
// weight fragment read from LDS feeds ET MFMAs. Waves loop over tiles
// (grid-stride), so a grid of a few blocks per CU amortises the staging.
// ld(e0, n, lane, x) fills the input fragments x[kc][t] of edges
// e0 + 16 t + lane % 16 (features 16 kc + 4 (lane / 16) + i), zeros past n.
#define MLP_FWD_WAVES 4
template <int KT, int NT, int ET>
__device__ __forceinline__ void layer_fwd_lds(const v4s *Ws, const float *__restrict__ b,
                                              const v4s (&xin)[KT][ET], v4f (&z)[NT][ET], int lane) {
#pragma unroll
  for (int nn = 0; nn < NT; ++nn) {
    v4f bias;
#pragma unroll
    for (int i = 0; i < 4; ++i) bias[i] = b[nn * 16 + 4 * (lane / 16) + i];
#pragma unroll
    for (int t = 0; t < ET; ++t) z[nn][t] = bias;
#pragma unroll
    for (int kc = 0; kc < KT; ++kc) {
      const v4s w = Ws[(nn * KT + kc) * 64 + lane];
#pragma unroll
      for (int t = 0; t < ET; ++t) z[nn][t] = mfma(w, xin[kc][t], z[nn][t]);
    }
  }
}

template <int NT, int ET>
__device__ __forceinline__ void act_gelu(const v4f (&z)[NT][ET], v4s (&a)[NT][ET]) {
#pragma unroll
  for (int nn = 0; nn < NT; ++nn)
#pragma unroll
    for (int t = 0; t < ET; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) a[nn][t][i] = f2bf(gelu(z[nn][t][i]));
}

template <int ET, class Loader>
__global__ __launch_bounds__(64 * MLP_FWD_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
void mlp_fwd_lds(int64_t n, Loader ld,
                                                                  const v4s *__restrict__ Wf,
                                                                  const float *__restrict__ P,
                                                                  float *__restrict__ Y) {
  __shared__ v4s Ws[FF_N / 4];
  for (int i = threadIdx.x; i < FF_N / 4; i += 64 * MLP_FWD_WAVES) Ws[i] = Wf[i];
  __syncthreads();
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
    v4f z[8][ET];
    v4s a[8][ET];
    layer_fwd_lds<3, 8, ET>(Ws + FF_1 / 4, P + P_B1, x, z, lane);
    act_gelu<8, ET>(z, a);
    layer_fwd_lds<8, 8, ET>(Ws + FF_2 / 4, P + P_B2, a, z, lane);
    act_gelu<8, ET>(z, a);
    layer_fwd_lds<8, 8, ET>(Ws + FF_3 / 4, P + P_B3, a, z, lane);
    act_gelu<8, ET>(z, a);
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
    for (int kc = 0; kc < 8; ++kc) {
      const v4s w = Ws[(FF_4 / 4) + kc * 64 + lane];
#pragma unroll
      for (int t = 0; t < ET; ++t) y[t] = mfma(w, a[kc][t], y[t]);
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

// Launch the LDS inference kernel: ET = 4 tiles per wave, at most two
// resident blocks per CU (80 KB of LDS each) on 256 CUs.
template <class Loader>
inline hipError_t mlp_fwd_launch(const short *Ff, const float *P, int64_t n, Loader ld, float *Y,
                                 hipStream_t st) {
  constexpr int ET = 2;
  const int64_t ntiles = (n + 16 * ET - 1) / (16 * ET);
  int64_t blocks = (ntiles + MLP_FWD_WAVES - 1) / MLP_FWD_WAVES;
  if (blocks > 512) blocks = 512;
  if (blocks < 1) blocks = 1;
  mlp_fwd_lds<ET, Loader><<<(unsigned)blocks, 64 * MLP_FWD_WAVES, 0, st>>>(n, ld, (const v4s *)Ff, P, Y);
  return hipGetLastError();
}
