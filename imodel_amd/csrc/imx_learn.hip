// libimx learned correction coupled to the shallow-water step.
//
// A coarse FAST handle advances with the plain RK4 step; the per-edge MLP
// (imx_mlp.hip) then predicts, from the 2-ring of every edge, the correction
// (du_e, dh) that maps the plain coarse step toward the restricted state of
// a finer run. Features of edge e (48, normalised by the caller's mean/std):
//   0-10   s_k u_k at e and the other ring edges of its two cells (TRiSK
//          orientation signs s_k), state before the step
//   11-21  the plain step's increment of the same
//   22-31  h - h_ref at the edge's 2 cells, the 2 triangle apexes (left,
//          right of v1->v2) and the other ring cells of both cells
//   32-41  the plain step's increment of the same
//   42-47  z and sqrt(1-z^2) of the edge midpoint, l_e, l~_e, A_c1, A_c2
//          relative to their means
// Output: (du_e, dh_e); the cell gets the mean dh_e of its edges. The
// correction is skipped (plain step kept) when its RMS exceeds a bound.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "imx_device.h"

#define NF 48
#define NSE 11  // stencil edges
#define NSC 10  // stencil cells
#define LTPB 256

struct imx_learn {
  imx_swm *s = nullptr;
  imx_mlp *m = nullptr;
  int nv = 0, NE = 0, nvalid = 0;
  imx::Alloc A;
  int32_t *st_e = nullptr;  // [NSE][NE], -1 pad / dummy
  float *st_s = nullptr;    // [NSE][NE]
  int32_t *st_c = nullptr;  // [NSC][NE]
  float *geo = nullptr;     // [6][NE]
  void *hn = nullptr, *un = nullptr;  // state before the step (T)
  float *raw = nullptr;     // [NE][NF]
  float *mu = nullptr, *sig = nullptr;
  float tsu = 1.f, tsh = 1.f;
  float *Y = nullptr;       // [NE][2]
  float *wv = nullptr;      // [NE]
  double *rms = nullptr;    // [1]
  int *fb = nullptr;        // [1] fallback flag of the last step
  double bound = 1e300;
};

__device__ __forceinline__ short lf2bf(float x) {
  __hip_bfloat16 b = __float2bfloat16(x);
  return *reinterpret_cast<short *>(&b);
}

template <typename T>
__global__ __launch_bounds__(LTPB) void learn_feat(int NE, const int32_t *__restrict__ st_e,
                                                   const float *__restrict__ st_s,
                                                   const int32_t *__restrict__ st_c,
                                                   const float *__restrict__ geo,
                                                   const T *__restrict__ un, const T *__restrict__ hn,
                                                   const T *__restrict__ up, const T *__restrict__ hp,
                                                   float *__restrict__ raw) {
  const int e = blockIdx.x * LTPB + threadIdx.x;
  if (e >= NE) return;
  float *f = raw + (int64_t)e * NF;
  if (st_e[e] < 0) {
    for (int j = 0; j < NF; ++j) f[j] = 0.f;
    return;
  }
#pragma unroll
  for (int k = 0; k < NSE; ++k) {
    const int ek = st_e[k * NE + e];
    float a = 0.f, d = 0.f;
    if (ek >= 0) {
      const float s = st_s[k * NE + e];
      a = s * (float)un[ek];
      d = s * (float)(up[ek] - un[ek]);
    }
    f[k] = a;
    f[NSE + k] = d;
  }
#pragma unroll
  for (int k = 0; k < NSC; ++k) {
    const int c = st_c[k * NE + e];
    float a = 0.f, d = 0.f;
    if (c >= 0) {
      a = (float)hn[c];
      d = (float)(hp[c] - hn[c]);
    }
    f[2 * NSE + k] = a;
    f[2 * NSE + NSC + k] = d;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) f[2 * NSE + 2 * NSC + k] = geo[k * NE + e];
}

__global__ __launch_bounds__(LTPB) void learn_norm(int64_t n, const float *__restrict__ raw,
                                                   const float *__restrict__ mu,
                                                   const float *__restrict__ sig, short *__restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * LTPB + threadIdx.x;
  if (i >= n * NF) return;
  const int j = (int)(i % NF);
  X[i] = lf2bf((raw[i] - mu[j]) / sig[j]);
}

// weighted RMS of the velocity output (one block)
__global__ __launch_bounds__(1024) void learn_rms(int NE, const float *__restrict__ Y,
                                                  const float *__restrict__ w, double *__restrict__ out) {
  double s = 0.0;
  for (int e = threadIdx.x; e < NE; e += 1024) s += (double)w[e] * (double)Y[2 * e] * (double)Y[2 * e];
  __shared__ double red[1024];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sqrt(red[0]);
}

template <typename T>
__global__ __launch_bounds__(LTPB) void learn_apply_u(int NE, const int32_t *__restrict__ st_e,
                                                      const float *__restrict__ Y, float inv_tsu,
                                                      const double *__restrict__ rms, double bound,
                                                      T *u, int *__restrict__ fb) {
  const int e = blockIdx.x * LTPB + threadIdx.x;
  if (e >= NE) return;
  const bool ok = rms[0] <= bound;
  if (e == 0) fb[0] = ok ? 0 : 1;
  if (!ok || st_e[e] < 0) return;
  u[e] = (T)((float)u[e] + Y[2 * e] * inv_tsu);
}

template <typename T>
__global__ __launch_bounds__(LTPB) void learn_apply_h(int nv, const int32_t *__restrict__ c_nnb,
                                                      const int32_t *__restrict__ c_ed,
                                                      const float *__restrict__ Y, float inv_tsh,
                                                      const double *__restrict__ rms, double bound, T *h) {
  const int c = blockIdx.x * LTPB + threadIdx.x;
  if (c >= nv || !(rms[0] <= bound)) return;
  const int n = c_nnb[c];
  float s = 0.f;
  for (int j = 0; j < n; ++j) s += Y[2 * c_ed[j * nv + c] + 1];
  h[c] = (T)((float)h[c] + (s / (float)n) * inv_tsh);
}

namespace imx {

static int lfail(hipError_t e, const char *what) {
  return set_error(IMX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define LTRY(expr, what)                        \
  do {                                          \
    hipError_t _e = (expr);                     \
    if (_e != hipSuccess) return imx::lfail(_e, what); \
  } while (0)

// save x_n, one plain step, raw features of (x_n, x_plain)
static int learn_sample(imx_learn *L) {
  imx_swm *s = L->s;
  const size_t es = s->single ? 4 : 8;
  LTRY(hipMemcpyAsync(L->hn, s->h0, es * s->nv, hipMemcpyDeviceToDevice, s->stream), "learn save");
  LTRY(hipMemcpyAsync(L->un, s->u0, es * s->NE, hipMemcpyDeviceToDevice, s->stream), "learn save");
  int rc = imx_swm_step(s, 1);
  if (rc) return rc;
  const int g = (L->NE + LTPB - 1) / LTPB;
  if (s->single)
    learn_feat<float><<<g, LTPB, 0, s->stream>>>(L->NE, L->st_e, L->st_s, L->st_c, L->geo,
                                                (const float *)L->un, (const float *)L->hn,
                                                (const float *)s->u0, (const float *)s->h0, L->raw);
  else
    learn_feat<double><<<g, LTPB, 0, s->stream>>>(L->NE, L->st_e, L->st_s, L->st_c, L->geo,
                                                 (const double *)L->un, (const double *)L->hn,
                                                 (const double *)s->u0, (const double *)s->h0, L->raw);
  LTRY(hipGetLastError(), "learn_feat");
  return IMX_OK;
}

static int learn_normalise(imx_learn *L) {
  int rc = mlp_reserve(L->m, L->NE);
  if (rc) return rc;
  const int64_t n = (int64_t)L->NE * NF;
  learn_norm<<<(unsigned)((n + LTPB - 1) / LTPB), LTPB, 0, L->s->stream>>>(L->NE, L->raw, L->mu, L->sig,
                                                                           L->m->X);
  LTRY(hipGetLastError(), "learn_norm");
  return IMX_OK;
}

static int learn_correct(imx_learn *L, double bound) {
  imx_swm *s = L->s;
  int rc = learn_normalise(L);
  if (rc) return rc;
  rc = mlp_forward_dev(L->m, L->NE, L->m->X, L->Y, s->stream);
  if (rc) return rc;
  learn_rms<<<1, 1024, 0, s->stream>>>(L->NE, L->Y, L->wv, L->rms);
  const int g = (L->NE + LTPB - 1) / LTPB, gc = (s->nv + LTPB - 1) / LTPB;
  if (s->single) {
    learn_apply_u<float><<<g, LTPB, 0, s->stream>>>(L->NE, L->st_e, L->Y, 1.f / L->tsu, L->rms, bound,
                                                   (float *)s->u0, L->fb);
    learn_apply_h<float><<<gc, LTPB, 0, s->stream>>>(s->nv, s->c_nnb, s->c_ed, L->Y, 1.f / L->tsh, L->rms,
                                                    bound, (float *)s->h0);
  } else {
    learn_apply_u<double><<<g, LTPB, 0, s->stream>>>(L->NE, L->st_e, L->Y, 1.f / L->tsu, L->rms, bound,
                                                    (double *)s->u0, L->fb);
    learn_apply_h<double><<<gc, LTPB, 0, s->stream>>>(s->nv, s->c_nnb, s->c_ed, L->Y, 1.f / L->tsh, L->rms,
                                                     bound, (double *)s->h0);
  }
  LTRY(hipGetLastError(), "learn_apply");
  return IMX_OK;
}

}  // namespace imx

extern "C" {

int imx_learn_create(imx_swm *s, imx_mlp *m, const imx_mesh_desc *d, imx_learn **out) {
  if (!s || !m || !d || !out) return imx::set_error(IMX_EINVAL, "imx_learn_create: NULL argument");
  *out = nullptr;
  if (!s->fast) return imx::set_error(IMX_ENOTSUP, "imx_learn_create: needs an FP32/FP64 fused handle");
  if (d->nv != s->nv || d->ne != s->ne || d->nt != s->nt)
    return imx::set_error(IMX_EINVAL, "imx_learn_create: mesh does not match the handle");
  imx_learn *L = new (std::nothrow) imx_learn();
  if (!L) return imx::set_error(IMX_ENOMEM, "imx_learn_create: out of memory");
  L->s = s;
  L->m = m;
  L->nv = s->nv;
  L->NE = s->NE;
  const int NE = s->NE, MX = IMX_MAXNB;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new;
  std::vector<int32_t> st_e((size_t)NSE * NE, -1), st_c((size_t)NSC * NE, -1);
  std::vector<float> st_s((size_t)NSE * NE, 0.f), geo((size_t)6 * NE, 0.f), wv(NE, 0.f);
  double ml = 0, md = 0, ma = 0;
  for (int eo = 0; eo < d->ne; ++eo) {
    ml += d->ed_leng[eo];
    md += d->edhx_leng[eo];
  }
  for (int c = 0; c < d->nv; ++c) ma += d->hx_areag[c];
  ml /= d->ne;
  md /= d->ne;
  ma /= d->nv;
  for (int eo = 0; eo < d->ne; ++eo) {
    const int e = en[eo];
    int cell[2] = {d->ed_v[2 * eo] - 1, d->ed_v[2 * eo + 1] - 1};
    int slot[2] = {-1, -1};
    for (int q = 0; q < 2; ++q)
      for (int j = 0; j < d->v_nnb[cell[q]]; ++j)
        if (d->v_ed[MX * cell[q] + j] == eo + 1) slot[q] = j;
    st_e[e] = e;
    st_s[e] = 1.f;
    int k = 1;
    for (int q = 0; q < 2; ++q) {
      const int c = cell[q], n = d->v_nnb[c], i = slot[q];
      for (int t = 1; t < 6; ++t, ++k) {
        if (t >= n) continue;  // pentagon: one slot stays padded
        const int j = (i + t) % n;
        st_e[(size_t)k * NE + e] = en[d->v_ed[MX * c + j] - 1];
        st_s[(size_t)k * NE + e] =
            (float)(d->hx_ttgout[MX * c + j] * d->hx_tnrccw[MX * c + i] < 0 ? -1 : 1);
      }
    }
    // cells: c1, c2, left apex, right apex, other ring cells of c1 and c2
    const double *p1 = d->v_p + 3 * cell[0], *p2 = d->v_p + 3 * cell[1];
    double nrm[3];
    imx::cross3(p1, p2, nrm);
    int apex[2] = {-1, -1};
    for (int q = 0; q < 2; ++q) {
      const int t = d->ed_sh[2 * eo + q] - 1;
      for (int l = 0; l < 3; ++l) {
        const int v = d->tr_v[3 * t + l] - 1;
        if (v != cell[0] && v != cell[1]) {
          if (imx::dot3(nrm, d->v_p + 3 * v) > 0) apex[0] = v;
          else apex[1] = v;
        }
      }
    }
    int cl[NSC];
    for (int q = 0; q < NSC; ++q) cl[q] = -1;
    cl[0] = cell[0];
    cl[1] = cell[1];
    cl[2] = apex[0];
    cl[3] = apex[1];
    int kc = 4;
    for (int q = 0; q < 2; ++q) {
      const int c = cell[q], n = d->v_nnb[c], i = slot[q];
      for (int t = 2; t <= 4; ++t, ++kc) {
        if (t > n - 2) continue;
        cl[kc] = d->v_nb[MX * c + (i + t) % n] - 1;
      }
    }
    for (int q = 0; q < NSC; ++q) st_c[(size_t)q * NE + e] = cl[q] >= 0 ? cn[cl[q]] : -1;
    const double z = d->ed_c[3 * eo + 2];
    geo[e] = (float)z;
    geo[(size_t)NE + e] = (float)std::sqrt(std::max(0.0, 1 - z * z));
    geo[(size_t)2 * NE + e] = (float)(d->ed_leng[eo] / ml);
    geo[(size_t)3 * NE + e] = (float)(d->edhx_leng[eo] / md);
    geo[(size_t)4 * NE + e] = (float)(d->hx_areag[cell[0]] / ma);
    geo[(size_t)5 * NE + e] = (float)(d->hx_areag[cell[1]] / ma);
    wv[e] = 1.f / (float)d->ne;
  }
  L->nvalid = d->ne;
  imx::Alloc &A = L->A;
  L->st_e = A.put(st_e);
  L->st_s = A.put(st_s);
  L->st_c = A.put(st_c);
  L->geo = A.put(geo);
  L->wv = A.put(wv);
  const size_t es = s->single ? 4 : 8;
  L->hn = A.get<char>(es * s->nv);
  L->un = A.get<char>(es * NE);
  L->raw = A.get<float>((size_t)NE * NF);
  L->mu = A.put(std::vector<float>(NF, 0.f));
  L->sig = A.put(std::vector<float>(NF, 1.f));
  L->Y = A.get<float>((size_t)2 * NE);
  L->rms = A.get<double>(1);
  L->fb = A.get<int>(1);
  if (A.err != hipSuccess) {
    int rc = imx::lfail(A.err, "imx_learn_create");
    imx_learn_destroy(L);
    return rc;
  }
  *out = L;
  return IMX_OK;
}

int imx_learn_sample(imx_learn *L, float *raw) {
  if (!L) return imx::set_error(IMX_EINVAL, "imx_learn_sample: NULL handle");
  int rc = imx::learn_sample(L);
  if (rc) return rc;
  if (raw) {  // caller edge order
    std::vector<float> b((size_t)L->NE * NF);
    LTRY(hipMemcpyAsync(b.data(), L->raw, b.size() * 4, hipMemcpyDeviceToHost, L->s->stream), "learn raw");
    LTRY(hipStreamSynchronize(L->s->stream), "learn raw");
    for (int eo = 0; eo < L->s->ne; ++eo)
      std::memcpy(raw + (size_t)eo * NF, b.data() + (size_t)L->s->lay.edge_new[eo] * NF, NF * 4);
  }
  return IMX_OK;
}

int imx_learn_set_scaling(imx_learn *L, const float *mu, const float *sig, float tscale_u, float tscale_h) {
  if (!L || !mu || !sig || !(tscale_u > 0) || !(tscale_h > 0))
    return imx::set_error(IMX_EINVAL, "imx_learn_set_scaling: bad argument");
  LTRY(hipMemcpy(L->mu, mu, NF * 4, hipMemcpyHostToDevice), "learn scaling");
  LTRY(hipMemcpy(L->sig, sig, NF * 4, hipMemcpyHostToDevice), "learn scaling");
  L->tsu = tscale_u;
  L->tsh = tscale_h;
  return IMX_OK;
}

// Train on the features of the last imx_learn_sample against the targets
// du[ne], dh[nv] (caller numbering): `epochs` AdamW steps on all edges.
int imx_learn_train(imx_learn *L, const double *du, const double *dh, int32_t epochs, float lr, float wd,
                    double *loss) {
  if (!L || !du || !dh || epochs < 1) return imx::set_error(IMX_EINVAL, "imx_learn_train: bad argument");
  imx_swm *s = L->s;
  int rc = imx::learn_normalise(L);
  if (rc) return rc;
  const int NE = L->NE;
  std::vector<float> T((size_t)2 * NE, 0.f), w(NE, 0.f);
  std::vector<double> dhi(s->nv);
  for (int c = 0; c < s->nv; ++c) dhi[s->lay.cell_new[c]] = dh[c];
  std::vector<int32_t> ev((size_t)2 * s->NE);
  LTRY(hipMemcpy(ev.data(), s->e_v, ev.size() * 4, hipMemcpyDeviceToHost), "learn train");
  for (int eo = 0; eo < s->ne; ++eo) {
    const int e = s->lay.edge_new[eo];
    T[2 * e] = (float)(du[eo] * L->tsu);
    T[2 * e + 1] = (float)(0.5 * (dhi[ev[e]] + dhi[ev[(size_t)s->NE + e]]) * L->tsh);
    w[e] = 1.f / (float)s->ne;
  }
  LTRY(hipMemcpyAsync(L->m->T, T.data(), T.size() * 4, hipMemcpyHostToDevice, s->stream), "learn train");
  LTRY(hipMemcpyAsync(L->m->w, w.data(), w.size() * 4, hipMemcpyHostToDevice, s->stream), "learn train");
  LTRY(hipStreamSynchronize(s->stream), "learn train");
  for (int k = 0; k < epochs; ++k) {
    rc = imx::mlp_train_dev(L->m, NE, lr, wd, 1, (k == epochs - 1) ? loss : nullptr);
    if (rc) return rc;
  }
  LTRY(hipStreamSynchronize(L->m->stream), "learn train");
  return IMX_OK;
}

// nsteps corrected steps: plain RK4, then the MLP correction unless its RMS
// exceeds `bound` (scaled units); *nfallback counts the plain-only steps.
int imx_learn_step(imx_learn *L, int32_t nsteps, double bound, int32_t *nfallback) {
  if (!L || nsteps < 0) return imx::set_error(IMX_EINVAL, "imx_learn_step: bad argument");
  int nfb = 0;
  for (int k = 0; k < nsteps; ++k) {
    int rc = imx::learn_sample(L);
    if (rc) return rc;
    rc = imx::learn_correct(L, bound);
    if (rc) return rc;
    if (nfallback) {
      int f = 0;
      LTRY(hipMemcpyAsync(&f, L->fb, 4, hipMemcpyDeviceToHost, L->s->stream), "learn step");
      LTRY(hipStreamSynchronize(L->s->stream), "learn step");
      nfb += f;
    }
  }
  LTRY(hipStreamSynchronize(L->s->stream), "learn step");
  if (nfallback) *nfallback = nfb;
  return IMX_OK;
}

int imx_learn_last_rms(imx_learn *L, double *rms) {
  if (!L || !rms) return imx::set_error(IMX_EINVAL, "imx_learn_last_rms: bad argument");
  LTRY(hipMemcpy(rms, L->rms, 8, hipMemcpyDeviceToHost), "learn rms");
  return IMX_OK;
}

void imx_learn_destroy(imx_learn *L) {
  if (!L) return;
  L->A.release();
  delete L;
}

}  // extern "C"
