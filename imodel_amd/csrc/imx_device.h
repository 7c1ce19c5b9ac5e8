// Device-side handle definitions shared by the HIP translation units of
// libimx (the step, the MLP and the learned-correction coupling).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "imx_internal.h"

namespace imx {

struct Alloc {
  std::vector<void *> ptrs;
  size_t bytes = 0;
  hipError_t err = hipSuccess;
  template <typename X>
  X *get(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(X));
    if (e != hipSuccess) {
      err = e;
      return nullptr;
    }
    // dummy slots read as zeros. The fill runs on the null stream, which the
    // handles' non-blocking streams do not wait for: it must be complete
    // before a copy or kernel on such a stream writes the buffer, or it can
    // land afterwards and wipe what they wrote.
    e = hipMemsetAsync(p, 0, n * sizeof(X), nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) err = e;
    ptrs.push_back(p);
    bytes += n * sizeof(X);
    return (X *)p;
  }
  template <typename X>
  X *put(const std::vector<X> &h) {
    X *p = get<X>(h.size());
    if (p && !h.empty()) {
      hipError_t e = hipMemcpy(p, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice);
      if (e != hipSuccess) err = e;
    }
    return p;
  }
  void release() {
    for (void *p : ptrs) (void)hipFree(p);
    ptrs.clear();
  }
};

}  // namespace imx

struct imx_swm {
  imx_swm_opts o{};
  int nv = 0, ne = 0, nt = 0;  // caller counts
  int NE = 0, NT = 0;          // internal edge/triangle slots (3 nv, 2 nv)
  imx::Layout lay;             // caller -> internal maps
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  // bumped whenever the step's kernel sequence or arguments change (scheme,
  // tuning, topography, h_ref): graphs captured over the step by other
  // objects (imx_learn) compare it with the value they captured at
  uint64_t generation = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  imx::Alloc A;
  bool fast = false;
  bool single = false;  // FAST with T=float
  double href = 0.0;
  bool href_set = false;

  // geometry / REF tables (always present), internal numbering
  int32_t *c_nnb = nullptr, *c_ed = nullptr;
  int8_t *c_ttg = nullptr, *c_tnr = nullptr;
  double *c_area = nullptr, *c_trskw = nullptr;
  // the TRiSK terms of each edge, edge-major [2 MX][NE] (term ic MX + j of
  // cell ic of the edge): ring edge k (-1 past the cell's edges) and weight
  int32_t *e_tk = nullptr;
  double *e_tw = nullptr, *e_tdl = nullptr;  // weight, dual length of k
  int8_t *e_ts = nullptr;                    // sign of the term (HTC)
  int8_t *hc_ts = nullptr;                   // the same for HC
  // analysis fields of the last LTE evaluation (imx_swm_lte): q_hx [nv],
  // vhq_tr [3][NT], vh_hx and vhq_hx [3][nv]; the ring triangles and the
  // tr -> hx weights of each cell [MX][nv]
  double *f_lte_qhx = nullptr, *f_lte_vtr = nullptr, *f_lte_vhh = nullptr, *f_lte_vhq = nullptr;
  int32_t *lte_ctr = nullptr;
  double *lte_cw = nullptr;
  bool lte_done = false;
  // fused handles with diffusion: K at cells and triangles and K zeta at
  // triangles (per stage), typed T
  bool f_diff = false;
  void *x_kh = nullptr, *x_kt = nullptr, *x_cz = nullptr;
  // fused handles with hyperdiffusion: per edge its triangles [2][NE] and the
  // factors 1/(|e| a) and s/(|e*| a); per stage K div u, lapu, K div lapu and
  // K curl lapu (typed T)
  bool f_hyper = false;
  int32_t *x_hy_esh = nullptr;
  void *x_hy_ea = nullptr, *x_hy_eb = nullptr, *x_hy_kd = nullptr, *x_hy_lap = nullptr, *x_hy_d = nullptr,
       *x_hy_z = nullptr;
  // fused handles with APVM: per cell slot and per triangle slot the Perot
  // weight with the projection folded in, as (x, y, z, 0) records ([MX][nv],
  // [3][NT]), per triangle slot the triangle across the edge (-1 padding),
  // and per stage the velocity at cells and grad q at triangles [N][4]
  // (typed T)
  bool f_apvm = false;
  // fused handles with pxt16: per cell slot the q_e weights of ring
  // triangles j-1 and j [2][MX][nv], per edge the folded normal of the
  // pered term [NE][4]; per stage A = P(uh) and B = P(uh q) at cells
  // [2][nv][4] and q_e at edges (typed T). The cell Perot weights are x_ap_wc.
  // pxt16's methods on the fused kernels, each on its own: barycentric
  // interpolation, the pered Coriolis term, the Perot kinetic energy
  bool f_bary = false, f_pered = false, f_perhx = false;
  // fused handles with gass18 (1 HTC, 2 HC): per triangle slot the areas of
  // (circumcentre, edge m) and their sum [4][NT] and the area [NT]; per
  // edge slot its two triangles [2][NE]; per cell slot the outer edge of
  // ring triangle j [MX][nv]; per stage E and H at triangles and q at edges
  // (typed T)
  int f_gass = 0;
  bool f_glink = false;              // gass mode: the linking-edge PV of the gass Coriolis term
  bool f_gtrsk = false;              // gass mode: the edge PV from trsk interpolation
  void *x_g_w = nullptr, *x_g_a = nullptr, *x_g_E = nullptr, *x_g_H = nullptr, *x_g_q = nullptr;
  int32_t *x_g_et = nullptr, *x_g_out = nullptr;
  void *x_g_qw = nullptr;              // gass mode with bary (HTC): the two triangles' weights per edge
  bool f_gqw = false;
  void *x_px_wq = nullptr, *x_px_n = nullptr, *x_px_ab = nullptr, *x_px_qe = nullptr;
  // pxt16 with HC: h at each Voronoi edge midpoint from the barycentric
  // weights [3][NE] (typed T) of the cells x_px_hv [3][NE] of the triangle
  // holding it, and the per-stage result [NE]
  bool f_pxhc = false;
  bool f_nopv = false;  // fused level model (trsk10)
  // Gassmann's kinetic energy on the fused kernels: sub-area weights [3][NT],
  // per-stage K_tr [NT] (typed T) and gasscoef
  bool f_gk = false;
  int f_melv = 0;  // Melvin's kinetic energy on the fused kernels: 1 HTC, 2 HC
  // dtred's Coriolis term on the fused kernels: triangle weights [3][NT] and
  // per-stage vhq [NT] (V4 records), edge interpolation weights [2][NE],
  // normals [NE] (V4), the two triangles of each edge slot [2][NE]
  bool f_dtred = false;
  void *x_dt_w = nullptr, *x_dt_vhq = nullptr, *x_dt_ew = nullptr, *x_dt_n = nullptr;
  int32_t *x_dt_et = nullptr;
  double gk_coef = 1.0;
  void *x_gk_w = nullptr, *x_gk_k = nullptr;
  void *x_px_hb = nullptr, *x_px_he = nullptr;
  int32_t *x_px_hv = nullptr;
  int32_t *x_ap_tnb = nullptr;
  void *x_ap_wc = nullptr, *x_ap_wt = nullptr;
  void *x_ap_v = nullptr, *x_ap_g = nullptr;
  int32_t *e_v = nullptr, *e_sh = nullptr;
  int8_t *e_loc = nullptr;
  double *e_leng = nullptr, *e_dleng = nullptr;
  int32_t *t_ed = nullptr, *t_v = nullptr;
  int8_t *t_tg = nullptr;
  double *t_area = nullptr, *t_f = nullptr, *t_trhx = nullptr;
  double *bt = nullptr;  // [nv] bottom topography (internal order, FP64; zeros when none)
  void *x_gb = nullptr;   // g b typed T for the fused kernels: x_gb_buf, or nullptr when b = 0
  void *x_gb_buf = nullptr;
  bool has_topo = false;
  double topo_sb = 0.0, topo_sb2 = 0.0;  // sum A b, sum A b^2 (unit-sphere areas)
  int *flag = nullptr;
  double *part = nullptr, *diag = nullptr;
  int nparts = 0;

  // REF state and fields
  double *rh = nullptr, *ru = nullptr, *rhn = nullptr, *run = nullptr;
  double *massf = nullptr, *momf = nullptr;
  double *f_h_ed = nullptr, *f_uh = nullptr, *f_eta = nullptr, *f_h_tr = nullptr,
         *f_q_tr = nullptr, *f_q_ed = nullptr, *f_divuh = nullptr, *f_ke = nullptr,
         *f_ghbk = nullptr, *f_grad = nullptr, *f_uhq = nullptr, *f_divu = nullptr,
         *f_zeta = nullptr;
  // REF scheme options (imx_swm_set_scheme): APVM, diffusion, hyperdiffusion
  imx_scheme sc{};
  bool sc_apvm = false, sc_diff = false, sc_hyper = false;
  int sc_wrap = 0;                   // IMX_WRAP_*
  // the methods the wrapper (or, with IMX_WRAP_NONE, the caller) selects
  int sc_sint = IMX_SINT_TRSK, sc_cor = IMX_COR_TRSK, sc_rec = IMX_REC_TRSK;
  double sc_gcoef = 1.0;
  bool sc_nopv = false;
  // Coriolis parameter (imx_swm_set_coriolis): f = 2 Omega sin(lat) per
  // triangle slot, kept on the host to restore it; f_lin: eta = f alone
  std::vector<double> f_var;
  bool f_lin = false;
  double *t_ka = nullptr;            // [3][NT] Gassmann K: areas of (circumcentre, edge ends)
  int32_t *c_tr = nullptr;           // [MX][nv] ring triangles of each cell
  double *c_hxtr = nullptr;          // [MX][nv] hxtr_area of each ring triangle
  double *f_ketr = nullptr, *f_eta_ed = nullptr, *f_vhq = nullptr, *f_vuh = nullptr;
  bool sc_hc = false;                // HC staggering
  int8_t *hc_c = nullptr;            // [MX][nv] HC: hx%nr per cell slot (div, Perot, Coriolis)
  int8_t *hc_t = nullptr;            // [3][NT] HC: vorticity signs
  int8_t *hc_g = nullptr;            // [NE] HC: grad_ed sign
  int8_t *hc_tt = nullptr;           // [NE] HC: grad_ed_tg sign
  double *hc_pc = nullptr;           // [3][NE] HC: Voronoi edge midpoints
  double *hc_hb = nullptr;           // [3][NE] HC + pxt16: barycentric weights of h at edges
  int32_t *hc_hv = nullptr;          // [3][NE] HC + pxt16: their cells
  double *t_w = nullptr;             // [3][NT] pxt16: barycentric coords; gass18: edge areas
  double *e_dw = nullptr;            // [2][NE] pxt16: distances edge -> its triangles
  double *e_nr = nullptr;            // [3][NE] pxt16: unit normal ed%nr
  int32_t *e_ed3 = nullptr;          // [2][MX][NE] gass18: linking edge per TRiSK term
  int8_t *e_nrs = nullptr;           // sign of ed%nr . (tr(sh2)%c - tr(sh1)%c)
  double *e_c = nullptr, *t_c = nullptr, *c_p = nullptr;  // [3][n] points
  double *dif_hx = nullptr, *dif_tr = nullptr;            // coefficients
  double *f_qg_ed = nullptr, *f_qg_tr = nullptr, *f_vhx = nullptr;  // APVM
  double *f_lapu = nullptr, *f_lap2 = nullptr, *f_divl = nullptr, *f_vortl = nullptr;

  // FAST tables and state (typed by precision; void* holds T*)
  int32_t *x_c_nb = nullptr, *x_c_bitsA = nullptr, *x_c_bitsB = nullptr;
  void *x_c_R = nullptr, *x_c_ika = nullptr;
  double *x_c_iaa = nullptr;
  // [MX][nv] dual length of each cell's edge slots, typed T: both cells of
  // an edge read the same rounded value, so the mass fluxes stay
  // antisymmetric and the total mass exact in FP32 as well
  void *x_c_dl = nullptr;
  void *x_c_kw = nullptr, *x_e_il = nullptr;
  int32_t *x_e_oth = nullptr;
  int32_t *x_t_ab = nullptr, *x_t_eb = nullptr;
  void *x_t_ceta = nullptr, *x_t_kap = nullptr, *x_t_f = nullptr;
  void *h0 = nullptr, *hA = nullptr, *hB = nullptr, *u0 = nullptr, *uA = nullptr, *uB = nullptr;
  void *acch = nullptr, *accu = nullptr, *q = nullptr, *phi = nullptr, *Pown = nullptr,
       *Poth = nullptr;
  double *dh_out = nullptr;
  void *du_out = nullptr;
  // ensembles: per-cell / per-edge-slot time step (nullptr: scalar o.dt)
  double *dt_c = nullptr;
  void *dt_e = nullptr;
  // launch configuration per fused kernel (tri, cell, edge): block size and
  // XCD remap, chosen by imx_swm_tune
  int kbs[3] = {256, 256, 256};
  int kremap[3] = {1, 1, 1};
};

// MLP handle: FP32 master parameters with AdamW moments, BF16 fragment
// copies of the weights, and the scratch of the training pass.
struct imx_mlp {
  float *P = nullptr, *M = nullptr, *V = nullptr, *G = nullptr, *part = nullptr, *lossp = nullptr;
  short *Ff = nullptr, *Fb = nullptr;  // K16 fragments of the training kernels
  short *Ff32 = nullptr;                // K32 fragments of the inference kernel
  short *X = nullptr, *AT = nullptr, *DZT = nullptr;
  float *T = nullptr, *w = nullptr, *Y = nullptr, *Xf = nullptr;
  int64_t cap = 0;  // edges the scratch holds
  int nsplit = 16;
  int64_t t = 0;    // optimiser step
  hipStream_t stream = nullptr;
};

namespace imx {
int mlp_reserve(imx_mlp *m, int64_t n);
int mlp_forward_dev(imx_mlp *m, int64_t n, const short *X, float *Y, hipStream_t st);
int mlp_train_dev(imx_mlp *m, int64_t n, const short *X, float lr, float wd, int apply, double *loss,
                  int gmask);
// one fused RK4 step of a FAST handle, enqueued on its stream (no graph)
void swm_launch_step(imx_swm *s);
// waits for the handle's stream; IMX_ESTATE if a step met a vanishing depth
int swm_check_flag(imx_swm *s, const char *what);
}  // namespace imx
