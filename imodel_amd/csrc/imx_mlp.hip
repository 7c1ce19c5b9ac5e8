// libimx learned correction: per-edge MLP 48 -> 128 -> 128 -> 128 -> 2 with
// GELU (tanh form), BF16 weights and activations on MFMA, FP32 accumulation;
// hand-written inference, training (forward + backward), weight-gradient and
// AdamW kernels.
//
// Register-resident layers. v_mfma_f32_16x16x16_bf16 takes, per lane l,
//   A[m = l%16][k = 4(l/16)+i], B[k = 4(l/16)+i][n = l%16], and returns
//   D[m = 4(l/16)+i][n = l%16]                       (i = 0..3).
// Every layer is evaluated transposed, Z^T(features x edges) = W^T A^T, with
// weights as the A operand and the 16 edges of a wave as the columns. The D
// fragment of output-feature tile t is then exactly the B fragment of input
// chunk t of the next layer, so activations never leave registers between
// layers. The backward data pass (dA^T = W dZ^T) has the same shape with a
// second, transposed weight fragment set. Weight gradients
// dW = A^T(features x edges) . dZ(edges x features) contract over edges and
// read activations/gradients stored feature-major by the training kernel.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "imx_device.h"

#include "imx_mlp_dev.h"

// Y^T tiles (NT x 16 edges) = W^T (N x K) X^T (K x 16), from fragments
// Wf[(n*KT + kc)*64 + lane] and input fragments xin[kc].
template <int KT, int NT>
__device__ __forceinline__ void layer_fwd(const v4s *__restrict__ Wf, const float *__restrict__ b,
                                          const v4s (&xin)[KT], v4f (&z)[NT], int lane) {
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    v4f acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = b[n * 16 + 4 * (lane / 16) + i];
#pragma unroll
    for (int kc = 0; kc < KT; ++kc) acc = mfma(Wf[(n * KT + kc) * 64 + lane], xin[kc], acc);
    z[n] = acc;
  }
}

// load the 3 input fragments of 16 edges (rows of X, 48 bf16 each)
__device__ __forceinline__ void load_x(const short *__restrict__ X, int64_t e0, int64_t n, int lane,
                                       v4s (&x)[3]) {
  const int64_t e = e0 + (lane % 16);
#pragma unroll
  for (int kc = 0; kc < 3; ++kc) {
    if (e < n) x[kc] = *reinterpret_cast<const v4s *>(X + e * MLP_IN + kc * 16 + 4 * (lane / 16));
    else x[kc] = v4s{0, 0, 0, 0};
  }
}

// Training pass over a batch: forward, weighted MSE loss
//   L = sum_e w_e sum_o (y_eo - t_eo)^2,
// backward through every layer; stores (feature-major, bf16) the layer
// inputs A0^T = X^T .. A3^T and the pre-activation gradients dZ1^T .. dZ4^T
// for the weight-gradient kernel, and one loss partial per block.
// Output o with bit o of gmask set takes T[e][o] as dL/dy_eo itself (a loss
// defined outside the network, e.g. through a divergence) and adds nothing
// to the loss partial.
__global__ __launch_bounds__(64 * MLP_WAVES) void mlp_train(
    int64_t n, int64_t npad, const short *__restrict__ X, const float *__restrict__ T,
    const float *__restrict__ w, int gmask, const v4s *__restrict__ Wf, const v4s *__restrict__ Wb,
    const float *__restrict__ P, short *__restrict__ AT, short *__restrict__ DZT,
    float *__restrict__ loss_part) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t e0 = ((int64_t)blockIdx.x * MLP_WAVES + wave) * 16;
  __shared__ float lred[MLP_WAVES];
  float lsum = 0.f;
  if (e0 < npad) {
    const int64_t e = e0 + (lane % 16);
    const bool live = e < n;
    v4s x[3];
    load_x(X, e0, n, lane, x);
    v4f z1[8], z2[8], z3[8];
    v4s a1[8], a2[8], a3[8];
    layer_fwd<3, 8>(Wf + FF_1 / 4, P + P_B1, x, z1, lane);
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) a1[t][i] = f2bf(gelu(z1[t][i]));
    layer_fwd<8, 8>(Wf + FF_2 / 4, P + P_B2, a1, z2, lane);
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) a2[t][i] = f2bf(gelu(z2[t][i]));
    layer_fwd<8, 8>(Wf + FF_3 / 4, P + P_B3, a2, z3, lane);
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) a3[t][i] = f2bf(gelu(z3[t][i]));
    v4f y = {0.f, 0.f, 0.f, 0.f};
    if (lane < 16) {
      y[0] = P[P_B4];
      y[1] = P[P_B4 + 1];
    }
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) y = mfma(Wf[(FF_4 / 4) + kc * 64 + lane], a3[kc], y);
    // dL/dy (features 0,1 live in lanes 0..15, elements 0,1)
    v4s dz4 = {0, 0, 0, 0};
    if (lane < 16 && live) {
      const float we = w[e];
      const float r0 = y[0] - T[e * 2], r1 = y[1] - T[e * 2 + 1];
      lsum = ((gmask & 1) ? 0.f : we * r0 * r0) + ((gmask & 2) ? 0.f : we * r1 * r1);
      dz4[0] = f2bf((gmask & 1) ? T[e * 2] : 2.f * we * r0);
      dz4[1] = f2bf((gmask & 2) ? T[e * 2 + 1] : 2.f * we * r1);
    }
    // stores: feature-major [row][npad]; a wave writes its 16 edge columns
    const int64_t col = e0 + (lane % 16);
    auto store4 = [&](short *base, int row0, v4s v) {
#pragma unroll
      for (int i = 0; i < 4; ++i) base[(int64_t)(row0 + 4 * (lane / 16) + i) * npad + col] = v[i];
    };
    // layer inputs A0 (X, 48 rows), A1..A3 (128 rows each)
#pragma unroll
    for (int kc = 0; kc < 3; ++kc) store4(AT, kc * 16, x[kc]);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      store4(AT + (int64_t)48 * npad, t * 16, a1[t]);
      store4(AT + (int64_t)176 * npad, t * 16, a2[t]);
      store4(AT + (int64_t)304 * npad, t * 16, a3[t]);
    }
    // dZ4 (16 rows, 2 live)
    store4(DZT + (int64_t)384 * npad, 0, dz4);
    // dA3^T = W4 dZ4^T, dZ3 = dA3 * gelu'(z3)
    v4s dz[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(Wb[(FB_4 / 4) + m * 64 + lane], dz4, acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) dz[m][i] = f2bf(acc[i] * gelu_d(z3[m][i]));
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) store4(DZT + (int64_t)256 * npad, t * 16, dz[t]);
    // dA2^T = W3 dZ3^T, dZ2
    v4s dzn[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < 8; ++kc) acc = mfma(Wb[(FB_3 / 4) + (m * 8 + kc) * 64 + lane], dz[kc], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) dzn[m][i] = f2bf(acc[i] * gelu_d(z2[m][i]));
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) store4(DZT + (int64_t)128 * npad, t * 16, dzn[t]);
    // dA1^T = W2 dZ2^T, dZ1
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < 8; ++kc) acc = mfma(Wb[(FB_2 / 4) + (m * 8 + kc) * 64 + lane], dzn[kc], acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) dz[m][i] = f2bf(acc[i] * gelu_d(z1[m][i]));
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) store4(DZT, t * 16, dz[t]);
  }
  // loss partial: wave sum, then block
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (lane == 0) lred[wave] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int q = 0; q < MLP_WAVES; ++q) s += lred[q];
    loss_part[blockIdx.x] = s;
  }
}

// Weight gradients: one wave per (layer, in-tile, out-tile, edge split);
// dW[in][out] = sum_e A^T[in][e] dZ^T[out][e] on MFMA, one FP32 partial
// tile per split (reduced deterministically by the optimiser kernel).
__constant__ int c_wg_layer_tiles[5] = {0, 24, 88, 152, 160};  // prefix of 3x8, 8x8, 8x8, 8x1

__global__ __launch_bounds__(64 * MLP_WAVES) void mlp_wgrad(int64_t npad, int nsplit,
                                                            const short *__restrict__ AT,
                                                            const short *__restrict__ DZT,
                                                            float *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * MLP_WAVES + (threadIdx.x >> 6);
  const int tile = gw / nsplit, split = gw % nsplit;
  if (tile >= 160) return;
  int layer = 0;
  while (tile >= c_wg_layer_tiles[layer + 1]) ++layer;
  const int lt = tile - c_wg_layer_tiles[layer];
  // layer geometry: input rows in AT, dz rows in DZT, out tiles, param offset
  const int a_row = (layer == 0) ? 0 : 48 + (layer - 1) * 128;
  const int dz_row = layer * 128;
  const int n_out_t = (layer == 3) ? 1 : 8;
  const int nout = (layer == 3) ? MLP_OUT : MLP_H;
  const int poff = (layer == 0) ? P_W1 : (layer == 1) ? P_W2 : (layer == 2) ? P_W3 : P_W4;
  const int mt = lt / n_out_t, nt = lt % n_out_t;
  const int64_t per = ((npad / 16 + nsplit - 1) / nsplit) * 16;
  const int64_t k0 = split * per, k1 = (k0 + per < npad) ? k0 + per : npad;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  const short *ar = AT + (int64_t)(a_row + mt * 16 + lane % 16) * npad;
  const short *br = DZT + (int64_t)(dz_row + nt * 16 + lane % 16) * npad;
  for (int64_t k = k0; k < k1; k += 16) {
    const v4s a = *reinterpret_cast<const v4s *>(ar + k + 4 * (lane / 16));
    const v4s b = *reinterpret_cast<const v4s *>(br + k + 4 * (lane / 16));
    acc = mfma(a, b, acc);
  }
  // D[in = mt*16 + 4(l/16)+i][out = nt*16 + l%16]
  const int outc = nt * 16 + lane % 16;
  if (outc < nout)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int inr = mt * 16 + 4 * (lane / 16) + i;
      part[(int64_t)split * P_N + poff + inr * nout + outc] = acc[i];
    }
}

// Bias gradients: db[f] = sum_e dZ^T[f][e], one block per bias entry row;
// written into split 0's partial slots (other splits hold 0 there).
__global__ __launch_bounds__(256) void mlp_bgrad(int64_t npad, const short *__restrict__ DZT,
                                                 float *__restrict__ part) {
  const int r = blockIdx.x;  // 0..385: b1 (128), b2, b3, b4 (2)
  int row, poff;
  if (r < 128) { row = r; poff = P_B1 + r; }
  else if (r < 256) { row = r; poff = P_B2 + (r - 128); }
  else if (r < 384) { row = r; poff = P_B3 + (r - 256); }
  else { row = 384 + (r - 384); poff = P_B4 + (r - 384); }
  float s = 0.f;
  for (int64_t e = threadIdx.x; e < npad; e += 256) s += bf2f(DZT[(int64_t)row * npad + e]);
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[poff] = red[0];
}

// AdamW (decoupled weight decay on the matrices, not the biases), summing
// the split partials; then the BF16 fragment copies are rebuilt.
__global__ __launch_bounds__(256) void mlp_adamw(int nsplit, const float *__restrict__ part,
                                                 float *__restrict__ P, float *__restrict__ M,
                                                 float *__restrict__ V, float *__restrict__ G, float lr,
                                                 float b1, float b2, float eps, float wd, float bc1,
                                                 float bc2, int apply) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P_N) return;
  float g = 0.f;
  for (int s = 0; s < nsplit; ++s) g += part[(int64_t)s * P_N + p];
  G[p] = g;
  if (!apply) return;
  const bool is_bias = (p >= P_B1 && p < P_W2) || (p >= P_B2 && p < P_W3) || (p >= P_B3 && p < P_W4) ||
                       p >= P_B4;
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  float x = P[p];
  if (!is_bias) x -= lr * wd * x;
  x -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  P[p] = x;
}

// fragment copies from the FP32 master (forward: W^T tiles; backward: W tiles)
// k-permutation of a K32 chunk: operand slot i of lane group g holds
// feature kappa(g, i) of the chunk (the inference kernel's B operand is the
// concatenation of two 16-feature activation tiles)
__device__ __forceinline__ int kappa32(int g, int i) { return i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4); }

__global__ __launch_bounds__(256) void mlp_pack(const float *__restrict__ P, short *__restrict__ Ff,
                                                short *__restrict__ Fb, short *__restrict__ Ff32) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  // inference (K32) fragments: index -> (layer, output tile, chunk, lane, i)
  if (q < FF_N) {
    int row, col, N, poff;
    if (q < FF_2) {  // layer 1: per output tile a K32 fragment then a K16 one
      const int nn = q / 768, s = q % 768;
      N = MLP_H;
      poff = P_W1;
      if (s < 512) {
        const int lane = s / 8, i = s % 8;
        row = kappa32(lane / 16, i);
        col = nn * 16 + lane % 16;
      } else {
        const int lane = (s - 512) / 4, i = (s - 512) % 4;
        row = 32 + 4 * (lane / 16) + i;
        col = nn * 16 + lane % 16;
      }
    } else {
      int base;
      if (q < FF_3) { base = FF_2; N = MLP_H; poff = P_W2; }
      else if (q < FF_4) { base = FF_3; N = MLP_H; poff = P_W3; }
      else { base = FF_4; N = MLP_OUT; poff = P_W4; }
      const int r = q - base;
      const int nn = r / 2048, c = (r / 512) % 4, lane = (r / 8) % 64, i = r % 8;
      row = 32 * c + kappa32(lane / 16, i);
      col = nn * 16 + lane % 16;
    }
    Ff32[q] = (col < N) ? f2bf(P[poff + row * N + col]) : (short)0;
  }
  // forward: index -> (layer, n, kc, lane, i)
  if (q < FF_N) {
    int base, K, N, poff, KT;
    int r = q;
    if (r < FF_2) { base = FF_1; K = MLP_IN; N = MLP_H; poff = P_W1; }
    else if (r < FF_3) { base = FF_2; K = MLP_H; N = MLP_H; poff = P_W2; }
    else if (r < FF_4) { base = FF_3; K = MLP_H; N = MLP_H; poff = P_W3; }
    else { base = FF_4; K = MLP_H; N = MLP_OUT; poff = P_W4; }
    r -= base;
    KT = K / 16;
    const int i = r & 3, lane = (r >> 2) & 63, tile = r >> 8;
    const int n = tile / KT, kc = tile % KT;
    const int row = kc * 16 + 4 * (lane / 16) + i;  // k (input feature)
    const int col = n * 16 + lane % 16;             // output feature
    Ff[q] = (col < N) ? f2bf(P[poff + row * N + col]) : (short)0;
  }
  if (q < FB_N) {
    int base, K, N, poff, KT;
    int r = q;
    if (r < FB_3) { base = FB_2; K = MLP_H; N = MLP_H; poff = P_W2; }
    else if (r < FB_4) { base = FB_3; K = MLP_H; N = MLP_H; poff = P_W3; }
    else { base = FB_4; K = MLP_H; N = MLP_OUT; poff = P_W4; }
    r -= base;
    KT = (N == MLP_OUT) ? 1 : N / 16;  // contraction over output features
    const int i = r & 3, lane = (r >> 2) & 63, tile = r >> 8;
    const int m = tile / KT, kc = tile % KT;
    const int row = m * 16 + lane % 16;            // input feature of the layer
    const int col = kc * 16 + 4 * (lane / 16) + i;  // output feature
    (void)K;
    Fb[q] = (col < N) ? f2bf(P[poff + row * N + col]) : (short)0;
  }
}

__global__ void f32_to_bf16(int64_t n, const float *__restrict__ x, short *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// ================================================================= handle ==

namespace imx {

static int mlp_fail(hipError_t e, const char *what) {
  return set_error(IMX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define MTRY(expr, what)                          \
  do {                                            \
    hipError_t _e = (expr);                       \
    if (_e != hipSuccess) return mlp_fail(_e, what); \
  } while (0)

int mlp_reserve(imx_mlp *m, int64_t n) {
  const int64_t npad = ((n + 63) / 64) * 64;
  if (npad <= m->cap) return IMX_OK;
  (void)hipFree(m->X);
  (void)hipFree(m->AT);
  (void)hipFree(m->DZT);
  (void)hipFree(m->T);
  (void)hipFree(m->w);
  (void)hipFree(m->Y);
  (void)hipFree(m->Xf);
  (void)hipFree(m->lossp);
  m->X = m->AT = m->DZT = nullptr;
  m->T = m->w = m->Y = m->Xf = m->lossp = nullptr;
  MTRY(hipMalloc(&m->X, npad * MLP_IN * 2), "mlp scratch");
  MTRY(hipMalloc(&m->AT, npad * 432 * 2), "mlp scratch");
  MTRY(hipMalloc(&m->DZT, npad * 400 * 2), "mlp scratch");
  MTRY(hipMalloc(&m->T, npad * 2 * 4), "mlp scratch");
  MTRY(hipMalloc(&m->w, npad * 4), "mlp scratch");
  MTRY(hipMalloc(&m->Y, npad * 2 * 4), "mlp scratch");
  MTRY(hipMalloc(&m->Xf, npad * MLP_IN * 4), "mlp scratch");
  MTRY(hipMalloc(&m->lossp, (npad / 64 + 1) * 4), "mlp scratch");
  m->cap = npad;
  return IMX_OK;
}

// He-normal init of the matrices, zero biases (SplitMix64 + Box-Muller)
void mlp_initial_params(uint64_t seed, std::vector<float> &P) {
  P.assign(P_N, 0.f);
  uint64_t s = seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto normal = [&]() {
    const double u1 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  auto init = [&](int off, int fan_in, int count, double gain) {
    for (int i = 0; i < count; ++i) P[off + i] = (float)(normal() * gain * std::sqrt(2.0 / fan_in));
  };
  init(P_W1, MLP_IN, MLP_IN * MLP_H, 1.0);
  init(P_W2, MLP_H, MLP_H * MLP_H, 1.0);
  init(P_W3, MLP_H, MLP_H * MLP_H, 1.0);
  init(P_W4, MLP_H, MLP_H * MLP_OUT, 0.1);  // small output layer: starts near "no correction"
}

int mlp_pack_launch(imx_mlp *m) {
  const int n = FF_N > FB_N ? FF_N : FB_N;
  mlp_pack<<<(n + 255) / 256, 256, 0, m->stream>>>(m->P, m->Ff, m->Fb, m->Ff32);
  MTRY(hipGetLastError(), "mlp_pack");
  return IMX_OK;
}

// Forward on device buffers (X bf16 [n][48]) -> Y fp32 [n][2].
int mlp_forward_dev(imx_mlp *m, int64_t n, const short *X, float *Y, hipStream_t st) {
  if (n <= 0) return IMX_OK;
  MTRY(mlp_fwd_launch(m->Ff32, m->P, n, RowLoader{X}, Y, st), "mlp_fwd_lds");
  return IMX_OK;
}

// One training pass on the device batch X (m->X when NULL), m->T, m->w (n edges).
int mlp_train_dev(imx_mlp *m, int64_t n, const short *X, float lr, float wd, int apply, double *loss,
                  int gmask) {
  const int64_t npad = ((n + 63) / 64) * 64;
  const int64_t blocks = npad / 64;
  hipStream_t st = m->stream;
  // zero the padding columns of the stored tensors (rows beyond n in the batch)
  mlp_train<<<(unsigned)blocks, 64 * MLP_WAVES, 0, st>>>(n, npad, X ? X : m->X, m->T, m->w, gmask, (const v4s *)m->Ff,
                                                        (const v4s *)m->Fb, m->P, m->AT, m->DZT, m->lossp);
  MTRY(hipGetLastError(), "mlp_train");
  MTRY(hipMemsetAsync(m->part, 0, (size_t)m->nsplit * P_N * 4, st), "mlp part");
  const int waves = 160 * m->nsplit;
  mlp_wgrad<<<(waves + MLP_WAVES - 1) / MLP_WAVES, 64 * MLP_WAVES, 0, st>>>(npad, m->nsplit, m->AT, m->DZT,
                                                                          m->part);
  mlp_bgrad<<<386, 256, 0, st>>>(npad, m->DZT, m->part);
  MTRY(hipGetLastError(), "mlp_wgrad");
  m->t += apply ? 1 : 0;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - powf(b1, (float)(m->t > 0 ? m->t : 1));
  const float bc2 = 1.f - powf(b2, (float)(m->t > 0 ? m->t : 1));
  mlp_adamw<<<(P_N + 255) / 256, 256, 0, st>>>(m->nsplit, m->part, m->P, m->M, m->V, m->G, lr, b1, b2,
                                                1e-8f, wd, bc1, bc2, apply);
  MTRY(hipGetLastError(), "mlp_adamw");
  if (apply) {
    int rc = mlp_pack_launch(m);
    if (rc) return rc;
  }
  if (loss) {
    std::vector<float> lp(blocks);
    MTRY(hipMemcpyAsync(lp.data(), m->lossp, blocks * 4, hipMemcpyDeviceToHost, st), "loss");
    MTRY(hipStreamSynchronize(st), "loss");
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = s;
  }
  return IMX_OK;
}

}  // namespace imx

extern "C" {

int imx_mlp_create(uint64_t seed, imx_mlp **out) {
  if (!out) return imx::set_error(IMX_EINVAL, "imx_mlp_create: NULL out");
  *out = nullptr;
  imx_mlp *m = new (std::nothrow) imx_mlp();
  if (!m) return imx::set_error(IMX_ENOMEM, "imx_mlp_create: out of memory");
  std::vector<float> P;
  imx::mlp_initial_params(seed, P);
  int rc = IMX_OK;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&m->P, P_N * 4) != hipSuccess || hipMalloc(&m->M, P_N * 4) != hipSuccess ||
      hipMalloc(&m->V, P_N * 4) != hipSuccess || hipMalloc(&m->G, P_N * 4) != hipSuccess ||
      hipMalloc(&m->part, (size_t)m->nsplit * P_N * 4) != hipSuccess ||
      hipMalloc(&m->Ff, FF_N * 2) != hipSuccess || hipMalloc(&m->Fb, FB_N * 2) != hipSuccess ||
      hipMalloc(&m->Ff32, FF_N * 2) != hipSuccess) {
    rc = imx::set_error(IMX_EDEVICE, "imx_mlp_create: device allocation failed");
  } else {
    (void)hipMemcpy(m->P, P.data(), P_N * 4, hipMemcpyHostToDevice);
    // the fills run on the null stream, which m->stream does not wait for
    (void)hipMemsetAsync(m->M, 0, P_N * 4, nullptr);
    (void)hipMemsetAsync(m->V, 0, P_N * 4, nullptr);
    (void)hipMemsetAsync(m->G, 0, P_N * 4, nullptr);
    (void)hipStreamSynchronize(nullptr);
    rc = imx::mlp_pack_launch(m);
    if (!rc && hipStreamSynchronize(m->stream) != hipSuccess)
      rc = imx::set_error(IMX_EDEVICE, "imx_mlp_create: pack failed");
  }
  if (rc) {
    imx_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return IMX_OK;
}

int imx_mlp_nparams(void) { return P_N; }

int imx_mlp_get_params(imx_mlp *m, float *p) {
  if (!m || !p) return imx::set_error(IMX_EINVAL, "imx_mlp_get_params: NULL argument");
  if (hipMemcpy(p, m->P, P_N * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return imx::set_error(IMX_EDEVICE, "imx_mlp_get_params: copy failed");
  return IMX_OK;
}

int imx_mlp_set_params(imx_mlp *m, const float *p) {
  if (!m || !p) return imx::set_error(IMX_EINVAL, "imx_mlp_set_params: NULL argument");
  if (hipMemcpy(m->P, p, P_N * 4, hipMemcpyHostToDevice) != hipSuccess)
    return imx::set_error(IMX_EDEVICE, "imx_mlp_set_params: copy failed");
  int rc = imx::mlp_pack_launch(m);
  if (rc) return rc;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return imx::set_error(IMX_EDEVICE, "imx_mlp_set_params");
  return IMX_OK;
}

int imx_mlp_init(imx_mlp *m, uint64_t seed) {
  if (!m) return imx::set_error(IMX_EINVAL, "imx_mlp_init: NULL handle");
  std::vector<float> P;
  imx::mlp_initial_params(seed, P);
  if (hipStreamSynchronize(m->stream) != hipSuccess ||
      hipMemcpy(m->P, P.data(), P_N * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(m->M, 0, P_N * 4, nullptr) != hipSuccess ||
      hipMemsetAsync(m->V, 0, P_N * 4, nullptr) != hipSuccess ||
      hipMemsetAsync(m->G, 0, P_N * 4, nullptr) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess)
    return imx::set_error(IMX_EDEVICE, "imx_mlp_init: device copy failed");
  m->t = 0;
  int rc = imx::mlp_pack_launch(m);
  if (rc) return rc;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return imx::set_error(IMX_EDEVICE, "imx_mlp_init");
  return IMX_OK;
}

// host-array entry points (tests, gradient check, small batches)
int imx_mlp_forward(imx_mlp *m, int64_t n, const float *X, float *Y) {
  if (!m || !X || !Y || n < 0) return imx::set_error(IMX_EINVAL, "imx_mlp_forward: bad argument");
  int rc = imx::mlp_reserve(m, n);
  if (rc) return rc;
  if (hipMemcpyAsync(m->Xf, X, n * MLP_IN * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    return imx::set_error(IMX_EDEVICE, "imx_mlp_forward: upload");
  f32_to_bf16<<<(unsigned)((n * MLP_IN + 255) / 256), 256, 0, m->stream>>>(n * MLP_IN, m->Xf, m->X);
  rc = imx::mlp_forward_dev(m, n, m->X, m->Y, m->stream);
  if (rc) return rc;
  if (hipMemcpyAsync(Y, m->Y, n * 2 * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess)
    return imx::set_error(IMX_EDEVICE, "imx_mlp_forward: download");
  return IMX_OK;
}

// Gradient of L = sum_e w_e |y_e - t_e|^2 (no update when apply == 0);
// grad receives the P_N parameter gradients (may be NULL).
int imx_mlp_train(imx_mlp *m, int64_t n, const float *X, const float *T, const float *w, float lr,
                  float wd, int32_t apply, double *loss, float *grad, int32_t gmask) {
  if (!m || !X || !T || !w || n <= 0 || gmask < 0 || gmask > 3)
    return imx::set_error(IMX_EINVAL, "imx_mlp_train: bad argument");
  int rc = imx::mlp_reserve(m, n);
  if (rc) return rc;
  hipStream_t st = m->stream;
  if (hipMemcpyAsync(m->Xf, X, n * MLP_IN * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(m->T, T, n * 2 * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(m->w, w, n * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    return imx::set_error(IMX_EDEVICE, "imx_mlp_train: upload");
  f32_to_bf16<<<(unsigned)((n * MLP_IN + 255) / 256), 256, 0, st>>>(n * MLP_IN, m->Xf, m->X);
  rc = imx::mlp_train_dev(m, n, nullptr, lr, wd, apply, loss, gmask);
  if (rc) return rc;
  if (grad && (hipMemcpyAsync(grad, m->G, P_N * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
               hipStreamSynchronize(st) != hipSuccess))
    return imx::set_error(IMX_EDEVICE, "imx_mlp_train: download");
  if (hipStreamSynchronize(st) != hipSuccess) return imx::set_error(IMX_EDEVICE, "imx_mlp_train");
  return IMX_OK;
}

// Inference timing on n synthetic rows already resident (for throughput).
int imx_mlp_forward_timed(imx_mlp *m, int64_t n, int32_t reps, double *ms) {
  if (!m || !ms || n <= 0 || reps < 1) return imx::set_error(IMX_EINVAL, "imx_mlp_forward_timed: bad argument");
  int rc = imx::mlp_reserve(m, n);
  if (rc) return rc;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemsetAsync(m->X, 0, n * MLP_IN * 2, m->stream);
  rc = imx::mlp_forward_dev(m, n, m->X, m->Y, m->stream);
  (void)hipEventRecord(e0, m->stream);
  for (int r = 0; r < reps && !rc; ++r) rc = imx::mlp_forward_dev(m, n, m->X, m->Y, m->stream);
  (void)hipEventRecord(e1, m->stream);
  (void)hipEventSynchronize(e1);
  float f = 0.f;
  (void)hipEventElapsedTime(&f, e0, e1);
  *ms = f / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

void imx_mlp_destroy(imx_mlp *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  float *fs[] = {m->P, m->M, m->V, m->G, m->part, m->lossp, m->T, m->w, m->Y, m->Xf};
  for (float *p : fs) (void)hipFree(p);
  short *ss[] = {m->Ff, m->Fb, m->Ff32, m->X, m->AT, m->DZT};
  for (short *p : ss) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

}  // extern "C"
