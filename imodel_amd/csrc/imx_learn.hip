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
// Outputs per edge: du_e, added to the edge velocity, and a mass flux F_e
// whose discrete divergence D F is added to the cell depths,
//   (D F)_c = -kappa/A_c sum_{e in c} ttgout_ce l~_e F_e,
// so the h correction moves mass between neighbours and the total mass
// sum_c A_c h_c is unchanged. Training fits du to its target directly and
// fits D F to the dh target through D (the gradient D^T r is handed to the
// MLP backward pass). The correction is skipped (plain step kept) when its
// RMS exceeds a bound. The corrected step (save, RK4 step, features, MLP,
// correction) is one hipGraph launch.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "imx_device.h"
#include "imx_mlp_dev.h"

#define NF 48
#define NSE 11  // stencil edges
#define NSC 10  // stencil cells
#define LTPB 256
#define LPART 1024  // blocks of the correction-norm pass
#define SREC 32    // words per edge stencil record (128 bytes)

struct imx_learn {
  imx_swm *s = nullptr;
  imx_mlp *m = nullptr;
  int nv = 0, NE = 0;
  imx::Alloc A;
  int32_t *stn = nullptr;   // [NE][SREC] stencil record per edge slot (see learn_featx)
  float *dc = nullptr;      // [6][nv] coefficient of F_{c_ed[j][c]} in (D F)_c
  float *de = nullptr;      // [2][NE] coefficient of F_e in (D F) at e_v[0/1][e]
  void *hn = nullptr, *un = nullptr;  // state before the step (T)
  float *raw = nullptr;     // [NE][NF]
  float *mu = nullptr, *sig = nullptr;
  float tsu = 1.f, tsh = 1.f;
  float *Y = nullptr;       // [NE][2]
  float *wv = nullptr;      // [NE] 1/ne on real edges
  float *dhc = nullptr;     // [nv] (D F)_c, or its residual in training
  float *th = nullptr;      // [nv] scaled dh target
  double *part = nullptr;   // [LPART] partial sums of the correction norm
  double *rms = nullptr;    // [1] correction RMS of the last step
  double *bound = nullptr;  // [2] RMS bound, limiter mode (IMX_LIMIT_*)
  double *lh = nullptr;     // [1] h loss of the last training epoch
  int *nfb = nullptr;       // [1] limited steps since the counter was cleared
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const short *gX = nullptr;  // MLP input buffer the graph was captured with
  uint64_t gen = 0;           // the step handle's generation at capture
  // device sample store: normalised features, scaled targets
  int cap = 0, count = 0;
  int64_t xstride = 0;      // shorts per stored feature matrix
  short *Xb = nullptr;      // [cap][xstride]
  float *tub = nullptr;     // [cap][NE] scaled du target, internal order
  float *thb = nullptr;     // [cap][nv] scaled dh target
};


template <typename T>
__global__ __launch_bounds__(LTPB) void learn_save(int nv, int NE, const T *__restrict__ h,
                                                   const T *__restrict__ u, T *__restrict__ hn,
                                                   T *__restrict__ un) {
  const int i = blockIdx.x * LTPB + threadIdx.x;
  if (i < NE) un[i] = u[i];
  if (i < nv) hn[i] = h[i];
}

// Features of (saved state un/hn, plain step up/hp), one thread per edge
// slot: the 128-byte stencil record is read with vector loads and all 42
// state gathers are issued before any is used, so a thread keeps its
// loads in flight together (one feature per thread would chain a record
// read and a gather for every feature). Record words: 0-10 ring edges
// (id << 1 | sign bit, -1 pad; word 0 is the edge itself, -1 for a dummy
// slot), 11-20 stencil cells (-1 pad), 21-26 geometry features (FP32
// bits). Dummy slots get all-zero raw features.
template <typename T>
__global__ __launch_bounds__(LTPB) void learn_featx(int NE, const int32_t *__restrict__ stn,
                                                    const T *__restrict__ un, const T *__restrict__ hn,
                                                    const T *__restrict__ up, const T *__restrict__ hp,
                                                    const float *__restrict__ mu,
                                                    const float *__restrict__ sig,
                                                    float *__restrict__ raw, short *__restrict__ X) {
  const int e = blockIdx.x * LTPB + threadIdx.x;
  if (e >= NE) return;
  int32_t r[28];
  const int4 *rv = reinterpret_cast<const int4 *>(stn + (int64_t)e * SREC);
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int4 v = rv[q];
    r[4 * q] = v.x;
    r[4 * q + 1] = v.y;
    r[4 * q + 2] = v.z;
    r[4 * q + 3] = v.w;
  }
  const bool real = r[0] >= 0;
  float f[NF];
  T a[NSE], b[NSE], c0[NSC], c1[NSC];
#pragma unroll
  for (int k = 0; k < NSE; ++k) {
    const int w = r[k];
    const int ek = (real && w >= 0) ? (w >> 1) : 0;
    a[k] = un[ek];
    b[k] = up[ek];
  }
#pragma unroll
  for (int k = 0; k < NSC; ++k) {
    const int c = (real && r[NSE + k] >= 0) ? r[NSE + k] : 0;
    c0[k] = hn[c];
    c1[k] = hp[c];
  }
#pragma unroll
  for (int k = 0; k < NSE; ++k) {
    const int w = r[k];
    const bool ok = real && w >= 0;
    const float sg = (w & 1) ? -1.f : 1.f;
    f[k] = ok ? sg * (float)a[k] : 0.f;
    f[NSE + k] = ok ? sg * (float)(b[k] - a[k]) : 0.f;
  }
#pragma unroll
  for (int k = 0; k < NSC; ++k) {
    const bool ok = real && r[NSE + k] >= 0;
    f[2 * NSE + k] = ok ? (float)c0[k] : 0.f;
    f[2 * NSE + NSC + k] = ok ? (float)(c1[k] - c0[k]) : 0.f;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) f[2 * NSE + 2 * NSC + k] = real ? __int_as_float(r[NSE + NSC + k]) : 0.f;
  const int64_t o = (int64_t)e * NF;
  if (raw) {
    float4 *rw = reinterpret_cast<float4 *>(raw + o);
#pragma unroll
    for (int q = 0; q < NF / 4; ++q) rw[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
  }
  if (X) {
    v4s *xw = reinterpret_cast<v4s *>(X + o);
#pragma unroll
    for (int q = 0; q < NF / 4; ++q) {
      v4s x;
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = f2bf((f[4 * q + i] - mu[4 * q + i]) / sig[4 * q + i]);
      xw[q] = x;
    }
  }
}

// out_c = (D F)_c - t_c (t may be NULL), F = the flux output Y[.][1]
__global__ __launch_bounds__(LTPB) void learn_div(int nv, const int32_t *__restrict__ c_nnb,
                                                  const int32_t *__restrict__ c_ed,
                                                  const float *__restrict__ dc, const float *__restrict__ Y,
                                                  const float *__restrict__ t, float *__restrict__ out) {
  const int c = blockIdx.x * LTPB + threadIdx.x;
  if (c >= nv) return;
  const int n = c_nnb[c];
  float v = 0.f;
  for (int j = 0; j < n; ++j) v += dc[j * nv + c] * Y[2 * c_ed[j * nv + c] + 1];
  out[c] = t ? v - t[c] : v;
}

// dL/dF_e = scale * (D^T r)_e into T[e][1] (the MLP's gradient input)
__global__ __launch_bounds__(LTPB) void learn_divT(int NE, const int32_t *__restrict__ e_v,
                                                   const float *__restrict__ de, const float *__restrict__ r,
                                                   float scale, float *__restrict__ T) {
  const int e = blockIdx.x * LTPB + threadIdx.x;
  if (e >= NE) return;
  const int c1 = e_v[e], c2 = e_v[NE + e];
  T[2 * e + 1] = (c1 < 0) ? 0.f : scale * (de[e] * r[c1] + de[NE + e] * r[c2]);
}

// T[e] = (scaled du target, scale * (D^T r)_e) for a stored sample's step
__global__ __launch_bounds__(LTPB) void learn_targets(int NE, const int32_t *__restrict__ e_v,
                                                      const float *__restrict__ de,
                                                      const float *__restrict__ r, float scale,
                                                      const float *__restrict__ tu, float *__restrict__ T) {
  const int e = blockIdx.x * LTPB + threadIdx.x;
  if (e >= NE) return;
  const int c1 = e_v[e], c2 = e_v[NE + e];
  T[2 * e] = tu[e];
  T[2 * e + 1] = (c1 < 0) ? 0.f : scale * (de[e] * r[c1] + de[NE + e] * r[c2]);
}

// out = sqrt(sum_e w_e a_e^2 + sb sum_c b_c^2) (one block; a strided by sa)
__global__ __launch_bounds__(1024) void learn_norm2(int na, const float *__restrict__ a, int sa,
                                                    const float *__restrict__ w, int nb,
                                                    const float *__restrict__ b, double sb, int take_sqrt,
                                                    double *__restrict__ out) {
  double s = 0.0;
  if (a)
    for (int e = threadIdx.x; e < na; e += 1024) s += (double)w[e] * (double)a[sa * e] * (double)a[sa * e];
  double t = 0.0;
  if (b)
    for (int c = threadIdx.x; c < nb; c += 1024) t += (double)b[c] * (double)b[c];
  s += sb * t;
  __shared__ double red[1024];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(red[0]) : red[0];
}

// (D F)_c of the flux output into dhc, and per-block partial sums of the
// correction's squared norm sum_e w_e Y0^2 + (1/nv) sum_c (D F)_c^2
// (grid-stride over max(nv, NE), LPART blocks)
__global__ __launch_bounds__(LTPB) void learn_div_part(int nv, int NE, const int32_t *__restrict__ c_nnb,
                                                       const int32_t *__restrict__ c_ed,
                                                       const float *__restrict__ dc,
                                                       const float *__restrict__ Y,
                                                       const float *__restrict__ w,
                                                       float *__restrict__ dhc, double *__restrict__ part) {
  double s = 0.0, t = 0.0;
  const int nmax = nv > NE ? nv : NE;
  for (int i = blockIdx.x * LTPB + threadIdx.x; i < nmax; i += gridDim.x * LTPB) {
    if (i < nv) {
      const int n = c_nnb[i];
      float v = 0.f;
      for (int j = 0; j < n; ++j) v += dc[j * nv + i] * Y[2 * c_ed[j * nv + i] + 1];
      dhc[i] = v;
      t += (double)v * (double)v;
    }
    if (i < NE) s += (double)w[i] * (double)Y[2 * i] * (double)Y[2 * i];
  }
  __shared__ double red[LTPB];
  red[threadIdx.x] = s + t / nv;
  __syncthreads();
  for (int o = LTPB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// every block sums the LPART partials (fixed order: the same RMS in every
// block), then applies u += Y0/tsu and h += (D F)/tsh unless RMS > bound
template <typename T>
__global__ __launch_bounds__(LTPB) void learn_apply(int nv, int NE, const float *__restrict__ wv,
                                                    const float *__restrict__ Y,
                                                    const float *__restrict__ dhc, float inv_tsu,
                                                    float inv_tsh, const double *__restrict__ part,
                                                    const double *__restrict__ bound, T *u, T *h,
                                                    double *__restrict__ rms, int *__restrict__ nfb) {
  __shared__ double red[LTPB];
  double p = 0.0;
  for (int k = threadIdx.x; k < LPART; k += LTPB) p += part[k];
  red[threadIdx.x] = p;
  __syncthreads();
  for (int o = LTPB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  // limiter (bound[0] the RMS bound, bound[1] the mode): a correction whose
  // RMS exceeds the bound is dropped and the plain step kept (mode 0), or
  // scaled down to the bound (mode 1); a non-finite one (or bound 0) always
  // leaves the plain step untouched
  const double r = sqrt(red[0]);
  const bool ok = r <= bound[0];
  const bool scale = bound[1] > 0.5;
  const float sc = ok ? 1.f : ((scale && r < 1e300) ? (float)(bound[0] / r) : 0.f);
  const int i = blockIdx.x * LTPB + threadIdx.x;
  if (i == 0) {
    rms[0] = r;
    if (!ok) nfb[0] = nfb[0] + 1;  // one thread of the grid counts
  }
  if (!(sc > 0.f)) return;
  if (i < NE && wv[i] > 0.f) u[i] = (T)((float)u[i] + sc * (Y[2 * i] * inv_tsu));
  if (i < nv) h[i] = (T)((float)h[i] + sc * (dhc[i] * inv_tsh));
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

static void learn_drop_graph(imx_learn *L) {
  if (L->gexec) (void)hipGraphExecDestroy(L->gexec);
  if (L->graph) (void)hipGraphDestroy(L->graph);
  L->gexec = nullptr;
  L->graph = nullptr;
}

// enqueue: save x_n, one plain step
template <typename T>
static void learn_enqueue_sample(imx_learn *L) {
  imx_swm *s = L->s;
  const int g = (L->NE + LTPB - 1) / LTPB;
  learn_save<T><<<g, LTPB, 0, s->stream>>>(s->nv, L->NE, (const T *)s->h0, (const T *)s->u0, (T *)L->hn,
                                           (T *)L->un);
  swm_launch_step(s);
}

// enqueue: features of (saved state, current state) into raw and/or X
template <typename T>
static void learn_enqueue_featx_t(imx_learn *L, hipStream_t st, float *raw, short *X) {
  imx_swm *s = L->s;
  learn_featx<T><<<(unsigned)((L->NE + LTPB - 1) / LTPB), LTPB, 0, st>>>(
      L->NE, L->stn, (const T *)L->un, (const T *)L->hn, (const T *)s->u0, (const T *)s->h0, L->mu, L->sig,
      raw, X);
}

static void learn_enqueue_featx(imx_learn *L, hipStream_t st, float *raw, short *X) {
  if (L->s->single) learn_enqueue_featx_t<float>(L, st, raw, X);
  else learn_enqueue_featx_t<double>(L, st, raw, X);
}

static void learn_free_store(imx_learn *L) {
  (void)hipFree(L->Xb);
  (void)hipFree(L->tub);
  (void)hipFree(L->thb);
  L->Xb = nullptr;
  L->tub = L->thb = nullptr;
  L->cap = L->count = 0;
}

// enqueue: features -> MLP -> correction (with the RMS bound) on the state
template <typename T>
static int learn_enqueue_correct(imx_learn *L) {
  imx_swm *s = L->s;
  hipStream_t st = s->stream;
  learn_enqueue_featx_t<T>(L, st, nullptr, L->m->X);
  int rc = mlp_forward_dev(L->m, L->NE, L->m->X, L->Y, st);
  if (rc) return rc;
  const int nmax = s->nv > L->NE ? s->nv : L->NE;
  learn_div_part<<<LPART, LTPB, 0, st>>>(s->nv, L->NE, s->c_nnb, s->c_ed, L->dc, L->Y, L->wv, L->dhc,
                                         L->part);
  learn_apply<T><<<(nmax + LTPB - 1) / LTPB, LTPB, 0, st>>>(s->nv, L->NE, L->wv, L->Y, L->dhc,
                                                           1.f / L->tsu, 1.f / L->tsh, L->part, L->bound,
                                                           (T *)s->u0, (T *)s->h0, L->rms, L->nfb);
  LTRY(hipGetLastError(), "learn_apply");
  return IMX_OK;
}

static int learn_sample(imx_learn *L) {
  if (L->s->single) learn_enqueue_sample<float>(L);
  else learn_enqueue_sample<double>(L);
  LTRY(hipGetLastError(), "learn sample");
  return IMX_OK;
}

static int learn_ensure_graph(imx_learn *L) {
  int rc = mlp_reserve(L->m, L->NE);
  if (rc) return rc;
  // recapture when the MLP scratch moved or the handle's step changed
  // (scheme, tuning, topography, h_ref)
  if (L->gexec && L->gX == L->m->X && L->gen == L->s->generation) return IMX_OK;
  learn_drop_graph(L);
  hipStream_t st = L->s->stream;
  LTRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "learn graph capture");
  if (L->s->single) {
    learn_enqueue_sample<float>(L);
    rc = learn_enqueue_correct<float>(L);
  } else {
    learn_enqueue_sample<double>(L);
    rc = learn_enqueue_correct<double>(L);
  }
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(st, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  LTRY(e, "learn graph capture end");
  L->graph = gr;
  LTRY(hipGraphInstantiate(&L->gexec, L->graph, nullptr, nullptr, 0), "learn graph instantiate");
  L->gX = L->m->X;
  L->gen = L->s->generation;
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
  const int NE = s->NE, nv = s->nv, MX = IMX_MAXNB;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new;
  std::vector<int32_t> stn((size_t)SREC * NE, 0);
  for (int e = 0; e < NE; ++e)
    for (int k = 0; k < NSE + NSC; ++k) stn[(size_t)e * SREC + k] = -1;
  std::vector<float> wv(NE, 0.f);
  std::vector<float> dc((size_t)6 * nv, 0.f), de((size_t)2 * NE, 0.f);
  double ml = 0, md = 0, ma = 0;
  for (int eo = 0; eo < d->ne; ++eo) {
    ml += d->ed_leng[eo];
    md += d->edhx_leng[eo];
  }
  for (int c = 0; c < d->nv; ++c) ma += d->hx_areag[c];
  ml /= d->ne;
  md /= d->ne;
  ma /= d->nv;
  // (D F)_c coefficients, scaled by kappa = mean area / mean dual length
  const double kappa = ma / md;
  for (int co = 0; co < d->nv; ++co) {
    const int c = cn[co];
    for (int j = 0; j < d->v_nnb[co]; ++j) {
      const int eo = d->v_ed[MX * co + j] - 1, e = en[eo];
      const double sg = d->hx_ttgout[MX * co + j] < 0 ? -1.0 : 1.0;
      const float coef = (float)(-kappa * sg * d->edhx_leng[eo] / d->hx_areag[co]);
      dc[(size_t)j * nv + c] = coef;
      de[(size_t)((d->ed_v[2 * eo] - 1 == co) ? 0 : 1) * NE + e] = coef;
    }
  }
  for (int eo = 0; eo < d->ne; ++eo) {
    const int e = en[eo];
    int cell[2] = {d->ed_v[2 * eo] - 1, d->ed_v[2 * eo + 1] - 1};
    int slot[2] = {-1, -1};
    for (int q = 0; q < 2; ++q)
      for (int j = 0; j < d->v_nnb[cell[q]]; ++j)
        if (d->v_ed[MX * cell[q] + j] == eo + 1) slot[q] = j;
    int32_t *rec = stn.data() + (size_t)e * SREC;
    rec[0] = e << 1;
    int k = 1;
    for (int q = 0; q < 2; ++q) {
      const int c = cell[q], n = d->v_nnb[c], i = slot[q];
      for (int t = 1; t < 6; ++t, ++k) {
        if (t >= n) continue;  // pentagon: one slot stays padded
        const int j = (i + t) % n;
        const int neg = d->hx_ttgout[MX * c + j] * d->hx_tnrccw[MX * c + i] < 0 ? 1 : 0;
        rec[k] = (en[d->v_ed[MX * c + j] - 1] << 1) | neg;
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
    for (int q = 0; q < NSC; ++q) rec[NSE + q] = cl[q] >= 0 ? cn[cl[q]] : -1;
    const double z = d->ed_c[3 * eo + 2];
    const float geo[6] = {(float)z, (float)std::sqrt(std::max(0.0, 1 - z * z)),
                          (float)(d->ed_leng[eo] / ml), (float)(d->edhx_leng[eo] / md),
                          (float)(d->hx_areag[cell[0]] / ma), (float)(d->hx_areag[cell[1]] / ma)};
    std::memcpy(rec + NSE + NSC, geo, sizeof(geo));
    wv[e] = 1.f / (float)d->ne;
  }
  imx::Alloc &A = L->A;
  L->stn = A.put(stn);
  L->wv = A.put(wv);
  L->dc = A.put(dc);
  L->de = A.put(de);
  const size_t es = s->single ? 4 : 8;
  L->hn = A.get<char>(es * nv);
  L->un = A.get<char>(es * NE);
  L->raw = A.get<float>((size_t)NE * NF);
  L->mu = A.put(std::vector<float>(NF, 0.f));
  L->sig = A.put(std::vector<float>(NF, 1.f));
  L->Y = A.get<float>((size_t)2 * NE);
  L->dhc = A.get<float>(nv);
  L->th = A.get<float>(nv);
  L->part = A.get<double>(LPART);
  L->rms = A.get<double>(1);
  L->bound = A.put(std::vector<double>{1e300, (double)IMX_LIMIT_FALLBACK});
  L->lh = A.get<double>(1);
  L->nfb = A.get<int>(1);
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
  if (raw) imx::learn_enqueue_featx(L, L->s->stream, L->raw, nullptr);
  LTRY(hipGetLastError(), "learn_featx");
  LTRY(hipStreamSynchronize(L->s->stream), "learn sample");
  if (raw) {  // caller edge order
    std::vector<float> b((size_t)L->NE * NF);
    LTRY(hipMemcpy(b.data(), L->raw, b.size() * 4, hipMemcpyDeviceToHost), "learn raw");
    for (int eo = 0; eo < L->s->ne; ++eo)
      std::memcpy(raw + (size_t)eo * NF, b.data() + (size_t)L->s->lay.edge_new[eo] * NF, NF * 4);
  }
  return IMX_OK;
}

int imx_learn_set_scaling(imx_learn *L, const float *mu, const float *sig, float tscale_u, float tscale_h) {
  if (!L || !mu || !sig || !(tscale_u > 0) || !(tscale_h > 0))
    return imx::set_error(IMX_EINVAL, "imx_learn_set_scaling: bad argument");
  for (int j = 0; j < NF; ++j)
    if (!(sig[j] > 0)) return imx::set_error(IMX_EINVAL, "imx_learn_set_scaling: sig must be positive");
  LTRY(hipMemcpy(L->mu, mu, NF * 4, hipMemcpyHostToDevice), "learn scaling");
  LTRY(hipMemcpy(L->sig, sig, NF * 4, hipMemcpyHostToDevice), "learn scaling");
  L->tsu = tscale_u;
  L->tsh = tscale_h;
  imx::learn_drop_graph(L);  // the scales are launch arguments of the graph
  return IMX_OK;
}

// Train on the features of the last imx_learn_sample against the targets
// du[ne], dh[nv] (caller numbering): `epochs` AdamW steps on all edges.
// loss[0] = mean_e (du_e - t_e)^2, loss[1] = mean_c ((D F)_c - t_c)^2 in
// the scaled units, before the last step.
int imx_learn_train(imx_learn *L, const double *du, const double *dh, int32_t epochs, float lr, float wd,
                    double *loss) {
  if (!L || !du || !dh || epochs < 1) return imx::set_error(IMX_EINVAL, "imx_learn_train: bad argument");
  imx_swm *s = L->s;
  imx_mlp *m = L->m;
  int rc = imx::mlp_reserve(m, L->NE);
  if (rc) return rc;
  const int NE = L->NE, nv = s->nv;
  imx::learn_enqueue_featx(L, s->stream, nullptr, m->X);
  LTRY(hipGetLastError(), "learn_featx");
  LTRY(hipStreamSynchronize(s->stream), "learn train");
  std::vector<float> T((size_t)2 * NE, 0.f), w(NE, 0.f), th(nv);
  for (int c = 0; c < nv; ++c) th[s->lay.cell_new[c]] = (float)(dh[c] * L->tsh);
  for (int eo = 0; eo < s->ne; ++eo) {
    const int e = s->lay.edge_new[eo];
    T[2 * e] = (float)(du[eo] * L->tsu);
    w[e] = 1.f / (float)s->ne;
  }
  hipStream_t st = m->stream;
  LTRY(hipMemcpyAsync(m->T, T.data(), T.size() * 4, hipMemcpyHostToDevice, st), "learn train");
  LTRY(hipMemcpyAsync(m->w, w.data(), w.size() * 4, hipMemcpyHostToDevice, st), "learn train");
  LTRY(hipMemcpyAsync(L->th, th.data(), th.size() * 4, hipMemcpyHostToDevice, st), "learn train");
  const int g = (NE + LTPB - 1) / LTPB, gc = (nv + LTPB - 1) / LTPB;
  double lu = 0.0;
  for (int k = 0; k < epochs; ++k) {
    const bool last = (k == epochs - 1);
    rc = imx::mlp_forward_dev(m, NE, m->X, L->Y, st);
    if (rc) return rc;
    learn_div<<<gc, LTPB, 0, st>>>(nv, s->c_nnb, s->c_ed, L->dc, L->Y, L->th, L->dhc);
    if (last) learn_norm2<<<1, 1024, 0, st>>>(0, nullptr, 1, nullptr, nv, L->dhc, 1.0 / nv, 0, L->lh);
    learn_divT<<<g, LTPB, 0, st>>>(NE, s->e_v, L->de, L->dhc, 2.f / (float)nv, m->T);
    LTRY(hipGetLastError(), "learn train");
    rc = imx::mlp_train_dev(m, NE, nullptr, lr, wd, 1, last ? &lu : nullptr, 2);
    if (rc) return rc;
  }
  double lh = 0.0;
  LTRY(hipMemcpyAsync(&lh, L->lh, 8, hipMemcpyDeviceToHost, st), "learn train");
  LTRY(hipStreamSynchronize(st), "learn train");
  if (loss) {
    loss[0] = lu;
    loss[1] = lh;
  }
  return IMX_OK;
}

// nsteps corrected steps: plain RK4, then the MLP correction, scaled down
// to RMS `bound` (scaled units) where it exceeds it; *nlimited counts the
// scaled steps.
int imx_learn_set_limiter(imx_learn *L, int32_t mode) {
  if (!L || (mode != IMX_LIMIT_FALLBACK && mode != IMX_LIMIT_SCALE))
    return imx::set_error(IMX_EINVAL, "imx_learn_set_limiter: bad argument");
  const double m = (double)mode;
  LTRY(hipMemcpyAsync(L->bound + 1, &m, 8, hipMemcpyHostToDevice, L->s->stream), "imx_learn_set_limiter");
  LTRY(hipStreamSynchronize(L->s->stream), "imx_learn_set_limiter");
  return IMX_OK;
}

int imx_learn_step(imx_learn *L, int32_t nsteps, double bound, int32_t *nlimited) {
  if (!L || nsteps < 0) return imx::set_error(IMX_EINVAL, "imx_learn_step: bad argument");
  imx_swm *s = L->s;
  int rc = imx::learn_ensure_graph(L);
  if (rc) return rc;
  LTRY(hipMemcpyAsync(L->bound, &bound, 8, hipMemcpyHostToDevice, s->stream), "learn step");
  LTRY(hipMemsetAsync(L->nfb, 0, 4, s->stream), "learn step");
  for (int k = 0; k < nsteps; ++k) LTRY(hipGraphLaunch(L->gexec, s->stream), "learn step");
  int nfb = 0;
  LTRY(hipMemcpyAsync(&nfb, L->nfb, 4, hipMemcpyDeviceToHost, s->stream), "learn step");
  rc = imx::swm_check_flag(s, "imx_learn_step");
  if (nlimited) *nlimited = nfb;
  return rc;
}

// Device sample store (capacity `cap` samples; existing samples dropped).
int imx_learn_reserve(imx_learn *L, int32_t cap) {
  if (!L || cap < 1) return imx::set_error(IMX_EINVAL, "imx_learn_reserve: bad argument");
  imx::learn_free_store(L);
  const int64_t npad = ((int64_t)L->NE + 63) / 64 * 64;
  L->xstride = npad * NF;
  LTRY(hipMalloc(&L->Xb, (size_t)cap * L->xstride * 2), "imx_learn_reserve");
  LTRY(hipMalloc(&L->tub, (size_t)cap * L->NE * 4), "imx_learn_reserve");
  LTRY(hipMalloc(&L->thb, (size_t)cap * L->nv * 4), "imx_learn_reserve");
  L->cap = cap;
  return IMX_OK;
}

// Store the last sample's normalised features with its targets du[ne],
// dh[nv] (caller numbering); the oldest sample is replaced when full.
int imx_learn_store(imx_learn *L, const double *du, const double *dh, int32_t *slot) {
  if (!L || !du || !dh) return imx::set_error(IMX_EINVAL, "imx_learn_store: bad argument");
  if (L->cap < 1) return imx::set_error(IMX_ESTATE, "imx_learn_store: no store (imx_learn_reserve)");
  imx_swm *s = L->s;
  const int k = L->count % L->cap;
  imx::learn_enqueue_featx(L, s->stream, nullptr, L->Xb + (size_t)k * L->xstride);
  LTRY(hipGetLastError(), "learn_featx");
  std::vector<float> tu(L->NE, 0.f), th(s->nv);
  for (int c = 0; c < s->nv; ++c) th[s->lay.cell_new[c]] = (float)(dh[c] * L->tsh);
  for (int eo = 0; eo < s->ne; ++eo) tu[s->lay.edge_new[eo]] = (float)(du[eo] * L->tsu);
  LTRY(hipMemcpyAsync(L->tub + (size_t)k * L->NE, tu.data(), tu.size() * 4, hipMemcpyHostToDevice,
                      s->stream), "imx_learn_store");
  LTRY(hipMemcpyAsync(L->thb + (size_t)k * s->nv, th.data(), th.size() * 4, hipMemcpyHostToDevice,
                      s->stream), "imx_learn_store");
  LTRY(hipStreamSynchronize(s->stream), "imx_learn_store");
  ++L->count;
  if (slot) *slot = k;
  return IMX_OK;
}

// nsteps AdamW steps, each on one stored sample drawn uniformly (SplitMix64
// from `seed`); the learning rate goes linearly from lr0 to lr1. loss[2]:
// mean scaled (u, h) losses over the last min(nsteps, 16) steps.
int imx_learn_fit(imx_learn *L, int32_t nsteps, float lr0, float lr1, float wd, uint64_t seed,
                  double loss[2]) {
  if (!L || nsteps < 1) return imx::set_error(IMX_EINVAL, "imx_learn_fit: bad argument");
  const int n = L->count < L->cap ? L->count : L->cap;
  if (n < 1) return imx::set_error(IMX_ESTATE, "imx_learn_fit: the store is empty");
  imx_swm *s = L->s;
  imx_mlp *m = L->m;
  int rc = imx::mlp_reserve(m, L->NE);
  if (rc) return rc;
  const int NE = L->NE, nv = s->nv;
  hipStream_t st = m->stream;
  LTRY(hipStreamSynchronize(s->stream), "imx_learn_fit");
  LTRY(hipMemcpyAsync(m->w, L->wv, (size_t)NE * 4, hipMemcpyDeviceToDevice, st), "imx_learn_fit");
  const int g = (NE + LTPB - 1) / LTPB, gc = (nv + LTPB - 1) / LTPB;
  const int tail = nsteps < 16 ? nsteps : 16;
  double su = 0.0, sh = 0.0;
  uint64_t x = seed;
  for (int k = 0; k < nsteps; ++k) {
    x += 0x9E3779B97F4A7C15ull;  // SplitMix64
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const int j = (int)(z % (uint64_t)n);
    const short *X = L->Xb + (size_t)j * L->xstride;
    const float lr = nsteps > 1 ? lr0 + (lr1 - lr0) * (float)k / (float)(nsteps - 1) : lr0;
    const bool last = k >= nsteps - tail;
    rc = imx::mlp_forward_dev(m, NE, X, L->Y, st);
    if (rc) return rc;
    learn_div<<<gc, LTPB, 0, st>>>(nv, s->c_nnb, s->c_ed, L->dc, L->Y, L->thb + (size_t)j * nv, L->dhc);
    if (last) learn_norm2<<<1, 1024, 0, st>>>(0, nullptr, 1, nullptr, nv, L->dhc, 1.0 / nv, 0, L->lh);
    learn_targets<<<g, LTPB, 0, st>>>(NE, s->e_v, L->de, L->dhc, 2.f / (float)nv,
                                      L->tub + (size_t)j * NE, m->T);
    LTRY(hipGetLastError(), "imx_learn_fit");
    double lu = 0.0;
    rc = imx::mlp_train_dev(m, NE, X, lr, wd, 1, last ? &lu : nullptr, 2);
    if (rc) return rc;
    if (last) {
      double lh = 0.0;
      LTRY(hipMemcpyAsync(&lh, L->lh, 8, hipMemcpyDeviceToHost, st), "imx_learn_fit");
      LTRY(hipStreamSynchronize(st), "imx_learn_fit");
      su += lu;
      sh += lh;
    }
  }
  LTRY(hipStreamSynchronize(st), "imx_learn_fit");
  if (loss) {
    loss[0] = su / tail;
    loss[1] = sh / tail;
  }
  return IMX_OK;
}

int imx_learn_last_rms(imx_learn *L, double *rms) {
  if (!L || !rms) return imx::set_error(IMX_EINVAL, "imx_learn_last_rms: bad argument");
  LTRY(hipMemcpy(rms, L->rms, 8, hipMemcpyDeviceToHost), "learn rms");
  return IMX_OK;
}

void imx_learn_destroy(imx_learn *L) {
  if (!L) return;
  imx::learn_drop_graph(L);
  imx::learn_free_store(L);
  L->A.release();
  delete L;
}

}  // extern "C"
