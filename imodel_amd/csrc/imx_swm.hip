// libimx device side: the TRiSK/HTC shallow-water step on MI355X (gfx950).
//
// Two kernel families share one handle:
//  * REF  (IMX_PREC_FP64_REF): one kernel per reference operator, FP64, the
//    reference's expression order and no FMA contraction (the library is
//    built with -ffp-contract=off). Bit-for-bit with the reference Fortran
//    built the same way. Used for parity and for FP64 callers.
//  * FAST (IMX_PREC_FP32 / IMX_PREC_FP64): three fused passes per RK stage,
//      triangle pass  q_t = (zeta_t + f_t) / h_t
//      cell pass      mass-flux divergence (FP64, compensated), kinetic
//                     energy, Bernoulli potential, and the cell's TRiSK
//                     partial sums for each of its edges (a 6x6 local
//                     product using weights rebuilt from the 6 kite areas)
//      edge pass      grad of the Bernoulli potential + the two Coriolis
//                     partials, RK update
//    with the RK4 state resident in HBM and one step captured in a hipGraph.
// Tables are structure-of-arrays in an internal numbering along a Hilbert
// curve; blocks are remapped so each XCD walks a contiguous stretch of it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "imx_internal.h"

#define TPB 256
#define NXCD 8
#define MX IMX_MAXNB

// --------------------------------------------------------------- helpers ---
static inline int grid_for(int n) {
  int nb = (n + TPB - 1) / TPB;
  if (nb < 1) nb = 1;
  return ((nb + NXCD - 1) / NXCD) * NXCD;
}

// Blocks are dispatched round-robin over the 8 XCDs. Give XCD x the
// contiguous logical range [x*per, (x+1)*per) so that SFC-neighbouring work
// shares one L2. The grid is always a multiple of NXCD blocks.
__device__ __forceinline__ int xcd_index() {
  const int per = gridDim.x / NXCD;
  const int lb = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  return lb * TPB + threadIdx.x;
}

// Same with a compile-time block size; remap = 0 keeps the hardware order
// (a choice the autotuner makes per kernel).
template <int BS>
__device__ __forceinline__ int ent_index(int remap) {
  if (!remap) return blockIdx.x * BS + threadIdx.x;
  const int per = gridDim.x / NXCD;
  const int lb = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  return lb * BS + threadIdx.x;
}

static inline int grid_bs(int n, int bs) {
  int nb = (n + bs - 1) / bs;
  if (nb < 1) nb = 1;
  return ((nb + NXCD - 1) / NXCD) * NXCD;
}

// epsilon(1.)/1e5 of the reference zero guard (swm_operators.f90:1603)
#define HTR_GUARD (1.1920928955078125e-07 / 100000.0)

// ============================================================ REF kernels ==
// scalar_hx2ed + scalar_elem_product (swm_operators.f90:88-114, :1567-1573):
// the average of the two cells, or (HC with pxt16) barycentric weights hb of
// the cells hv of the triangle holding the dual edge midpoint
__global__ __launch_bounds__(TPB) void ref_edge_flux(int ne, const int32_t *__restrict__ e_v,
                                                     const double *__restrict__ hb,
                                                     const int32_t *__restrict__ hv,
                                                     const double *__restrict__ h,
                                                     const double *__restrict__ u,
                                                     double *__restrict__ h_ed,
                                                     double *__restrict__ uh) {
  const int l = xcd_index();
  if (l >= ne || e_v[l] < 0) return;
  double he;
  if (hb) {
    he = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) he = he + hb[i * ne + l] * h[hv[i * ne + l]];
  } else {
    he = (h[e_v[l]] + h[e_v[ne + l]]) * 0.5;
  }
  h_ed[l] = he;
  uh[l] = u[l] * he;
}

// vort_tr + rel_vort_tr + scalar_hx2tr + scalar_elem_divide
// (swm_operators.f90:551-611, :780-827, :258-263, :1579-1615)
// h at triangles: trsk area weights (hmode 0), barycentric coordinates t_w
// (1, pxt16), or edge values weighted by the edge's sub-triangle areas t_w
// (2, gass18: scalar_edtr2trcc, swm_operators.f90:286-346)
__global__ __launch_bounds__(TPB) void ref_tri(int nt, const int32_t *__restrict__ t_ed,
                                               const int8_t *__restrict__ t_tg,
                                               const int32_t *__restrict__ t_v,
                                               const double *__restrict__ t_area,
                                               const double *__restrict__ t_f,
                                               const double *__restrict__ t_trhx,
                                               const double *__restrict__ e_leng,
                                               const double *__restrict__ u,
                                               const double *__restrict__ h, int hmode,
                                               const double *__restrict__ t_w,
                                               const double *__restrict__ h_ed, double erad,
                                               int use_pv, int flin, double *__restrict__ eta,
                                               double *__restrict__ zeta,
                                               double *__restrict__ h_tr,
                                               double *__restrict__ q_tr, int *__restrict__ flag) {
  const int k = xcd_index();
  if (k >= nt || t_ed[k] < 0) return;
  double e = 0.0;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const int ed = t_ed[l * nt + k];
    e = e + u[ed] * e_leng[ed] * (double)t_tg[l * nt + k];
  }
  const double z = e / t_area[k] / erad;
  zeta[k] = z;
  // absolute vorticity; the linearised f-sphere keeps f alone
  const double et = flin ? t_f[k] : z + t_f[k];
  eta[k] = et;
  double s = 0.0, ht;
  if (hmode == 1) {
#pragma unroll
    for (int i = 0; i < 3; ++i) s = s + t_w[i * nt + k] * h[t_v[i * nt + k]];
    ht = s;
  } else if (hmode == 2) {
#pragma unroll
    for (int i = 0; i < 3; ++i) s = s + t_w[i * nt + k] * h_ed[t_ed[i * nt + k]];
    ht = s / t_area[k];
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) s = s + t_trhx[i * nt + k] * h[t_v[i * nt + k]];
    ht = s / t_area[k];
  }
  h_tr[k] = ht;
  double q = 0.0;
  if (use_pv) {
    if (fabs(ht) < HTR_GUARD) flag[0] = 1;
    else q = et / ht;
  }
  q_tr[k] = q;
}

// PV at edges: scalar_tr2ed of q, simple average (qmode 0, trsk10,
// swm_operators.f90:152-153) or distance-weighted (1, pxt16 on HTC, :163-168);
// or (2, gass18) eta and h averaged over the rhombus with the triangle areas
// (:155-162) and divided (swm.f90:2469-2473, zero guard as the reference's)
// The vorticity at edges (eta_ed, scalar_tr2ed of eta with the same
// interpolation, swm.f90:2465) is written when eta_ed is not NULL (the level
// model's Coriolis term reads it).
__global__ __launch_bounds__(TPB) void ref_edge_q(int ne, int qmode, const int32_t *__restrict__ e_sh,
                                                  const double *__restrict__ e_dw,
                                                  const double *__restrict__ t_area,
                                                  const double *__restrict__ eta,
                                                  const double *__restrict__ h_tr,
                                                  const double *__restrict__ q_tr,
                                                  double *__restrict__ q_ed, double *__restrict__ eta_ed,
                                                  int *__restrict__ flag) {
  const int l = xcd_index();
  if (l >= ne || e_sh[l] < 0) return;
  const int k1 = e_sh[l], k2 = e_sh[ne + l];
  if (qmode == 1) {
    const double d1 = e_dw[l], d2 = e_dw[ne + l];
    q_ed[l] = (d2 * q_tr[k1] + d1 * q_tr[k2]) / (d1 + d2);
    if (eta_ed) eta_ed[l] = (d2 * eta[k1] + d1 * eta[k2]) / (d1 + d2);
  } else if (qmode >= 2) {
    // gass18: rhombus averages with the triangle areas, or (HC, qmode 3)
    // the simple averages scalar_tr2ed takes on HC
    double er, hr;
    if (qmode == 2) {
      const double a1 = t_area[k1], a2 = t_area[k2];
      er = (a1 * eta[k1] + a2 * eta[k2]) / (a1 + a2);
      hr = (a1 * h_tr[k1] + a2 * h_tr[k2]) / (a1 + a2);
    } else {
      er = 0.5 * (eta[k1] + eta[k2]);
      hr = 0.5 * (h_tr[k1] + h_tr[k2]);
    }
    if (eta_ed) eta_ed[l] = er;
    if (fabs(hr) < HTR_GUARD) {
      flag[0] = 1;
      q_ed[l] = 0.0;
    } else {
      q_ed[l] = er / hr;
    }
  } else {
    q_ed[l] = 0.5 * (q_tr[k1] + q_tr[k2]);
    if (eta_ed) eta_ed[l] = 0.5 * (eta[k1] + eta[k2]);
  }
}

// kinetic_energy_tr (swm_operators.f90:1074-1130, geodesic areas): K at
// triangles, sum over the edges of the area of (circumcentre, edge ends)
// times u^2, over the triangle's area (Gassmann's kinetic energy)
__global__ __launch_bounds__(TPB) void ref_ketr(int nt, const int32_t *__restrict__ t_ed,
                                                const double *__restrict__ t_ka,
                                                const double *__restrict__ t_area,
                                                const double *__restrict__ u, double *__restrict__ ketr) {
  const int k = xcd_index();
  if (k >= nt || t_ed[k] < 0) return;
  double s = 0.0;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const double uu = u[t_ed[l * nt + k]];
    s = s + t_ka[l * nt + k] * (uu * uu);
  }
  ketr[k] = s / t_area[k];
}

// div_hx(uh), div_hx(u), kinetic_energy_hx, g(h+b)+K
// (swm_operators.f90:613-661, :1132-1217; swm.f90:2533-2536)
__global__ __launch_bounds__(TPB) void ref_cell(int nv, const int32_t *__restrict__ c_nnb,
                                                const int32_t *__restrict__ c_ed,
                                                const int8_t *__restrict__ c_ttg,
                                                const double *__restrict__ c_area,
                                                const double *__restrict__ e_leng,
                                                const double *__restrict__ e_dleng,
                                                const double *__restrict__ u,
                                                const double *__restrict__ uh,
                                                const double *__restrict__ h,
                                                const double *__restrict__ bt, int kmode,
                                                const double *__restrict__ vhx,
                                                const double *__restrict__ vuh,
                                                const double *__restrict__ ketr,
                                                const int32_t *__restrict__ c_tr,
                                                const double *__restrict__ c_hxtr, double gcoef,
                                                double erad, double grav, double *__restrict__ divuh,
                                                double *__restrict__ masseq,
                                                double *__restrict__ divu,
                                                double *__restrict__ ke,
                                                double *__restrict__ ghbk) {
  const int i = xcd_index();
  if (i >= nv) return;
  const int n = c_nnb[i];
  double s = 0.0, sd = 0.0, k = 0.0, ca = 0.0;
  for (int j = 0; j < n; ++j) {
    const int l = c_ed[j * nv + i];
    const double sg = (double)c_ttg[j * nv + i];
    s = s + sg * uh[l] * e_dleng[l];
    sd = sd + sg * u[l] * e_dleng[l];
    const double ed_area = 0.25 * e_leng[l] * e_dleng[l];
    ca = ca + ed_area;
    // melv (HTC): sum of ed_area uh u
    k = (kmode == IMX_REC_MELV) ? k + ed_area * uh[l] * u[l] : k + ed_area * (u[l] * u[l]);
  }
  const double dv = s / c_area[i] / erad;
  divuh[i] = dv;
  masseq[i] = -dv;
  divu[i] = sd / c_area[i] / erad;
  double kk = k / ca;
  if (kmode == IMX_REC_PERHX) {  // perhx (pxt16): |v|^2 / 2 of the Perot vector
    const double v0 = vhx[i], v1 = vhx[nv + i], v2 = vhx[2 * nv + i];
    double dd = 0.0;
    dd = dd + v0 * v0;
    dd = dd + v1 * v1;
    dd = dd + v2 * v2;
    kk = dd / 2.0;
  } else if (kmode == IMX_REC_MELV) {
    if (vuh) {  // HC: Perot vector of uh dotted with that of u, over 2 h
      double dd = 0.0;
      dd = dd + vuh[i] * vhx[i];
      dd = dd + vuh[nv + i] * vhx[nv + i];
      dd = dd + vuh[2 * nv + i] * vhx[2 * nv + i];
      kk = dd / h[i] / 2.0;
    } else {
      kk = k / h[i] / ca;
    }
  } else if (kmode == IMX_REC_GASS) {
    // Gassmann: the trsk K blended with the ring triangles' K
    kk = kk * gcoef;
    for (int j = 0; j < n; ++j) kk = kk + (1.0 - gcoef) * c_hxtr[j * nv + i] * ketr[c_tr[j * nv + i]];
  }
  ke[i] = kk;
  ghbk[i] = grav * (h[i] + bt[i]) + kk;
}

// grad_ed (HTC) + coriolis_ed (trsk, PV form) -> momeq
// (swm_operators.f90:535-542, :1283-1309; swm.f90:2542, :2567)
__global__ __launch_bounds__(TPB) void ref_edge_mom(
    int ne, int nv, const int32_t *__restrict__ e_v, const int8_t *__restrict__ e_loc,
    const double *__restrict__ e_leng, const double *__restrict__ e_dleng,
    const int32_t *__restrict__ c_nnb, const int32_t *__restrict__ c_ed,
    const int8_t *__restrict__ c_ttg, const int8_t *__restrict__ c_tnr,
    const int32_t *__restrict__ e_tk, const double *__restrict__ e_tw, const double *__restrict__ e_tdl,
    const int8_t *__restrict__ e_ts, const double *__restrict__ ghbk,
    const double *__restrict__ uh, const double *__restrict__ q_ed,
    const int32_t *__restrict__ e_ed3, const int8_t *__restrict__ e_gs, double erad,
    int zero_mom, double *__restrict__ grad, double *__restrict__ uhq,
    double *__restrict__ momeq) {
  const int l = xcd_index();
  if (l >= ne || e_v[l] < 0) return;
  // grad_ed: HC orients the difference by edhx%nr (e_gs)
  const double gs = e_gs ? (double)e_gs[l] : 1.0;
  const double g = gs * (ghbk[e_v[ne + l]] - ghbk[e_v[l]]) / e_leng[l] / erad;
  grad[l] = g;
  double s = 0.0;
  const double ql = q_ed[l];
  // the terms in the reference's order (cell 1 then cell 2, each cell's
  // edges in ring order), ring edges and weights from the edge-major term
  // tables (coalesced) instead of the cell-major weight table
  // the sign of a term is dsign(1, ttgout(j) tnrccw(ei)) (HTC) or
  // dsign(1, hx%nr(j) hx%nr(ei)) (HC), tabulated per term
#pragma unroll
  for (int t = 0; t < 2 * MX; ++t) {
    const int k = e_tk[t * ne + l];
    if (k < 0) continue;
    const double signcor = (double)e_ts[t * ne + l];
    // gass18: q of the rhombus linking edges l and k where there is one
    const int e3 = e_ed3 ? e_ed3[t * ne + l] : -1;
    const double qtmp = e3 >= 0 ? q_ed[e3] : 0.5 * (ql + q_ed[k]);
    s = s + uh[k] * qtmp * e_tdl[t * ne + l] * e_tw[t * ne + l] * signcor;
  }
  const double p = s / e_leng[l];
  uhq[l] = p;
  momeq[l] = zero_mom ? 0.0 : (-g) - p;
}

// ---- scheme options of the REF path (APVM, diffusion, hyperdiffusion) ----
// Points are stored component-major: x of entity i at [i], y at [n + i],
// z at [2n + i].

// grad_tr2ed (swm_operators.f90:1427-1463): gradient of q across each edge,
// oriented by the primal edge's normal
__global__ __launch_bounds__(TPB) void ref_qgrad_ed(int ne, const int32_t *__restrict__ e_sh,
                                                    const int8_t *__restrict__ e_nrs,
                                                    const double *__restrict__ e_dleng,
                                                    const double *__restrict__ q_tr, double erad,
                                                    double *__restrict__ qg) {
  const int l = xcd_index();
  if (l >= ne || e_sh[l] < 0) return;
  qg[l] = (double)e_nrs[l] * (q_tr[e_sh[ne + l]] - q_tr[e_sh[l]]) / e_dleng[l] / erad;
}

// proj_vec_sphere (smeshpack.f90:4419-4433)
__device__ __forceinline__ void proj_sphere(double &v0, double &v1, double &v2, double p0, double p1,
                                            double p2) {
  double vp = 0.0;
  vp = vp + v0 * p0;
  vp = vp + v1 * p1;
  vp = vp + v2 * p2;
  double pn = 0.0;
  pn = pn + p0 * p0;
  pn = pn + p1 * p1;
  pn = pn + p2 * p2;
  pn = sqrt(pn);
  v0 = v0 - vp * p0 / pn;
  v1 = v1 - vp * p1 / pn;
  v2 = v2 - vp * p2 / pn;
}

// vector_edtr2tr_perot (swm_operators.f90:408-452) of the edge field qg:
// the vector at each triangle's circumcentre (tr%nr = -tr%tg)
__global__ __launch_bounds__(TPB) void ref_perot_tr(int nt, int ne, const int32_t *__restrict__ t_ed,
                                                    const int8_t *__restrict__ t_tg,
                                                    const double *__restrict__ e_c,
                                                    const double *__restrict__ t_c,
                                                    const double *__restrict__ e_leng,
                                                    const double *__restrict__ t_area,
                                                    const double *__restrict__ qg,
                                                    double *__restrict__ vt) {
  const int k = xcd_index();
  if (k >= nt || t_ed[k] < 0) return;
  const double p0 = t_c[k], p1 = t_c[nt + k], p2 = t_c[2 * nt + k];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const int ed = t_ed[l * nt + k];
    const double a = -(double)t_tg[l * nt + k] * qg[ed];
    const double len = e_leng[ed];
    v0 = v0 + a * (e_c[ed] - p0) * len;
    v1 = v1 + a * (e_c[ne + ed] - p1) * len;
    v2 = v2 + a * (e_c[2 * ne + ed] - p2) * len;
  }
  const double ar = t_area[k];
  v0 = v0 / ar;
  v1 = v1 / ar;
  v2 = v2 / ar;
  proj_sphere(v0, v1, v2, p0, p1, p2);
  vt[k] = v0;
  vt[nt + k] = v1;
  vt[2 * nt + k] = v2;
}

// vector_edhx2hx_perot (swm_operators.f90:454-509, HTC): velocity vector at
// each cell node
__global__ __launch_bounds__(TPB) void ref_perot_hx(int nv, int ne, const int32_t *__restrict__ c_nnb,
                                                    const int32_t *__restrict__ c_ed,
                                                    const int8_t *__restrict__ c_ttg,
                                                    const double *__restrict__ e_c,
                                                    const double *__restrict__ c_p,
                                                    const double *__restrict__ e_dleng,
                                                    const double *__restrict__ c_area,
                                                    const double *__restrict__ u,
                                                    double *__restrict__ vh) {
  const int i = xcd_index();
  if (i >= nv) return;
  const double p0 = c_p[i], p1 = c_p[nv + i], p2 = c_p[2 * nv + i];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  const int n = c_nnb[i];
  for (int j = 0; j < n; ++j) {
    const int ed = c_ed[j * nv + i];
    const double a = (double)c_ttg[j * nv + i] * u[ed];
    const double len = e_dleng[ed];
    v0 = v0 + a * (e_c[ed] - p0) * len;
    v1 = v1 + a * (e_c[ne + ed] - p1) * len;
    v2 = v2 + a * (e_c[2 * ne + ed] - p2) * len;
  }
  const double ar = c_area[i];
  v0 = v0 / ar;
  v1 = v1 / ar;
  v2 = v2 / ar;
  proj_sphere(v0, v1, v2, p0, p1, p2);
  vh[i] = v0;
  vh[nv + i] = v1;
  vh[2 * nv + i] = v2;
}

// Analysis fields the tendency adds under the local truncation error
// harness (test_lterror = 1, swm.f90:2500-2503): vhq_tr =
// vector_edhx2tr_perot(uh) q_tr (swm_operators.f90:348-406, then
// vector_elem_product :1617-1649). Signs tr%tg (HTC) or tr%tg s_e (HC), edge
// points ed%c (HTC) or edhx%c (HC); the result is cross(tr%c, v) / area.
__global__ __launch_bounds__(TPB) void ref_lte_tr(int nt, int ne, const int32_t *__restrict__ t_ed,
                                                  const int8_t *__restrict__ tsg,
                                                  const double *__restrict__ epc,
                                                  const double *__restrict__ t_c,
                                                  const double *__restrict__ e_leng,
                                                  const double *__restrict__ t_area,
                                                  const double *__restrict__ uh,
                                                  const double *__restrict__ q_tr,
                                                  double *__restrict__ vhq) {
  const int k = xcd_index();
  if (k >= nt || t_ed[k] < 0) return;
  const double p0 = t_c[k], p1 = t_c[nt + k], p2 = t_c[2 * nt + k];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const int ed = t_ed[l * nt + k];
    const double a = (double)tsg[l * nt + k] * uh[ed];
    const double len = e_leng[ed];
    v0 = v0 + a * (epc[ed] - p0) * len;
    v1 = v1 + a * (epc[ne + ed] - p1) * len;
    v2 = v2 + a * (epc[2 * ne + ed] - p2) * len;
  }
  const double ar = t_area[k];
  const double c0 = (p1 * v2 - p2 * v1) / ar;
  const double c1 = (p2 * v0 - p0 * v2) / ar;
  const double c2 = (p0 * v1 - p1 * v0) / ar;
  const double q = q_tr[k];
  vhq[k] = c0 * q;
  vhq[nt + k] = c1 * q;
  vhq[2 * nt + k] = c2 * q;
}

// ... and at cells (swm.f90:2546-2551): vh_hx = v_hx h, q_hx =
// scalar_tr2hx(q_tr) (swm_operators.f90:178-227: weights w over the cell's
// ring triangles, hxtr_area or Wachspress coordinates), vhq_hx = vh_hx q_hx
__global__ __launch_bounds__(TPB) void ref_lte_cell(int nv, const int32_t *__restrict__ c_nnb,
                                                    const int32_t *__restrict__ c_tr,
                                                    const double *__restrict__ c_w,
                                                    const double *__restrict__ vh,
                                                    const double *__restrict__ h,
                                                    const double *__restrict__ q_tr,
                                                    double *__restrict__ vhh, double *__restrict__ qhx,
                                                    double *__restrict__ vhq) {
  const int i = xcd_index();
  if (i >= nv) return;
  const double hh = h[i];
  const double a0 = vh[i] * hh, a1 = vh[nv + i] * hh, a2 = vh[2 * nv + i] * hh;
  double q = 0.0;
  const int n = c_nnb[i];
  for (int j = 0; j < n; ++j) q = q + c_w[j * nv + i] * q_tr[c_tr[j * nv + i]];
  vhh[i] = a0;
  vhh[nv + i] = a1;
  vhh[2 * nv + i] = a2;
  qhx[i] = q;
  vhq[i] = a0 * q;
  vhq[nv + i] = a1 * q;
  vhq[2 * nv + i] = a2 * q;
}

// apvm (swm_operators.f90:1497-1501): q_ed -= pvspar dt (qg_t1 + qg_t2) . (v_c1 + v_c2)
__global__ __launch_bounds__(TPB) void ref_apvm_q(int ne, int nt, int nv, const int32_t *__restrict__ e_sh,
                                                  const int32_t *__restrict__ e_v,
                                                  const double *__restrict__ qgt,
                                                  const double *__restrict__ vh, double pvspar,
                                                  double dt, double *__restrict__ q_ed) {
  const int l = xcd_index();
  if (l >= ne || e_sh[l] < 0) return;
  const int k1 = e_sh[l], k2 = e_sh[ne + l], c1 = e_v[l], c2 = e_v[ne + l];
  double dp = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q)
    dp = dp + (qgt[q * nt + k1] + qgt[q * nt + k2]) * (vh[q * nv + c1] + vh[q * nv + c2]);
  q_ed[l] = q_ed[l] - pvspar * dt * dp;
}

// diffusion_ed (swm_operators.f90:855-895, HTC): grad_ed(K_hx div) -
// grad_ed_tg(K_tr vort); mode 1 adds the result to momeq, mode 2 subtracts it
__global__ __launch_bounds__(TPB) void ref_lap(int ne, const int32_t *__restrict__ e_v,
                                               const int32_t *__restrict__ e_sh,
                                               const int8_t *__restrict__ e_nrs,
                                               const int8_t *__restrict__ e_gs,
                                               const double *__restrict__ e_leng,
                                               const double *__restrict__ e_dleng,
                                               const double *__restrict__ kh,
                                               const double *__restrict__ kt,
                                               const double *__restrict__ div,
                                               const double *__restrict__ vort, double erad, int mode,
                                               double *__restrict__ out, double *__restrict__ momeq) {
  const int l = xcd_index();
  if (l >= ne || e_v[l] < 0) return;
  const int c1 = e_v[l], c2 = e_v[ne + l], k1 = e_sh[l], k2 = e_sh[ne + l];
  const double gs = e_gs ? (double)e_gs[l] : 1.0;
  const double g = gs * (kh[c2] * div[c2] - kh[c1] * div[c1]) / e_leng[l] / erad;
  const double gt = (double)e_nrs[l] * (kt[k2] * vort[k2] - kt[k1] * vort[k1]);
  const double gv = gt / (e_dleng[l] * erad);
  const double lap = g - gv;
  out[l] = lap;
  if (mode == 1) momeq[l] = momeq[l] + lap;
  else if (mode == 2) momeq[l] = momeq[l] - lap;
}

// div_hx (swm_operators.f90:613-661, HTC) of an edge field
__global__ __launch_bounds__(TPB) void ref_div(int nv, const int32_t *__restrict__ c_nnb,
                                               const int32_t *__restrict__ c_ed,
                                               const int8_t *__restrict__ c_ttg,
                                               const double *__restrict__ c_area,
                                               const double *__restrict__ e_dleng,
                                               const double *__restrict__ f, double erad,
                                               double *__restrict__ out) {
  const int i = xcd_index();
  if (i >= nv) return;
  const int n = c_nnb[i];
  double s = 0.0;
  for (int j = 0; j < n; ++j) {
    const int l = c_ed[j * nv + i];
    s = s + (double)c_ttg[j * nv + i] * f[l] * e_dleng[l];
  }
  out[i] = s / c_area[i] / erad;
}

// rel_vort_tr (swm_operators.f90:780-827, HTC) of an edge field
__global__ __launch_bounds__(TPB) void ref_rvort(int nt, const int32_t *__restrict__ t_ed,
                                                 const int8_t *__restrict__ t_tg,
                                                 const double *__restrict__ t_area,
                                                 const double *__restrict__ e_leng,
                                                 const double *__restrict__ f, double erad,
                                                 double *__restrict__ out) {
  const int k = xcd_index();
  if (k >= nt || t_ed[k] < 0) return;
  double z = 0.0;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const int ed = t_ed[l * nt + k];
    z = z + f[ed] * e_leng[ed] * (double)t_tg[l * nt + k];
  }
  out[k] = z / t_area[k] / erad;
}

// grad_ed + coriolis_ed pered (pxt16, swm_operators.f90:1330-1368): Perot
// reconstruction of uh q at the edge's two cells, averaged, projected at the
// edge point, then (HTC) minus its ed%nr component or (HC) its edhx%tg
// component -> momeq
// coriolis_ed dtred (swm_operators.f90:1310-1329): the Perot vector of
// uh q at the edge's two triangles (vhq_tr) interpolated to the edge, with
// the distances to the circumcentres and projected on the tangent plane
// (HTC, -component along ed%nr), or averaged (HC, component along edhx%tg)
__global__ __launch_bounds__(TPB) void ref_edge_mom_dtred(
    int ne, int nt, const int32_t *__restrict__ e_v, const int32_t *__restrict__ e_sh,
    const double *__restrict__ e_leng, const double *__restrict__ e_dw, const double *__restrict__ e_c,
    const double *__restrict__ e_nr, const double *__restrict__ vhq, const double *__restrict__ ghbk,
    const int8_t *__restrict__ e_gs, int hc, double erad, int zero_mom, double *__restrict__ grad,
    double *__restrict__ uhq, double *__restrict__ momeq) {
  const int l = xcd_index();
  if (l >= ne || e_v[l] < 0) return;
  const double gs = e_gs ? (double)e_gs[l] : 1.0;
  const double g = gs * (ghbk[e_v[ne + l]] - ghbk[e_v[l]]) / e_leng[l] / erad;
  grad[l] = g;
  const int k1 = e_sh[l], k2 = e_sh[ne + l];
  double v0, v1, v2, p;
  if (!hc) {
    const double d1 = e_dw[l], d2 = e_dw[ne + l];
    v0 = (d2 * vhq[k1] + d1 * vhq[k2]) / (d1 + d2);
    v1 = (d2 * vhq[nt + k1] + d1 * vhq[nt + k2]) / (d1 + d2);
    v2 = (d2 * vhq[2 * nt + k1] + d1 * vhq[2 * nt + k2]) / (d1 + d2);
    proj_sphere(v0, v1, v2, e_c[l], e_c[ne + l], e_c[2 * ne + l]);
    double d = 0.0;
    d = d + v0 * e_nr[l];
    d = d + v1 * e_nr[ne + l];
    d = d + v2 * e_nr[2 * ne + l];
    p = -d;
  } else {
    v0 = 0.5 * vhq[k1] + 0.5 * vhq[k2];
    v1 = 0.5 * vhq[nt + k1] + 0.5 * vhq[nt + k2];
    v2 = 0.5 * vhq[2 * nt + k1] + 0.5 * vhq[2 * nt + k2];
    double d = 0.0;
    d = d + v0 * e_nr[l];
    d = d + v1 * e_nr[ne + l];
    d = d + v2 * e_nr[2 * ne + l];
    p = d;
  }
  uhq[l] = p;
  momeq[l] = zero_mom ? 0.0 : (-g) - p;
}

__global__ __launch_bounds__(TPB) void ref_edge_mom_pered(
    int ne, int nv, const int32_t *__restrict__ e_v, const double *__restrict__ e_leng,
    const double *__restrict__ e_dleng, const double *__restrict__ e_c,
    const double *__restrict__ e_nr, const int32_t *__restrict__ c_nnb,
    const int32_t *__restrict__ c_ed, const int8_t *__restrict__ c_ttg,
    const double *__restrict__ c_p, const double *__restrict__ c_area,
    const double *__restrict__ ghbk, const double *__restrict__ uh, const double *__restrict__ q_ed,
    const int8_t *__restrict__ e_gs, double psign, double erad, int zero_mom, double *__restrict__ grad,
    double *__restrict__ uhq, double *__restrict__ momeq) {
  const int l = xcd_index();
  if (l >= ne || e_v[l] < 0) return;
  const double gs = e_gs ? (double)e_gs[l] : 1.0;
  const double g = gs * (ghbk[e_v[ne + l]] - ghbk[e_v[l]]) / e_leng[l] / erad;
  grad[l] = g;
  const double ql = q_ed[l];
  double w[2][3];
  for (int ic = 0; ic < 2; ++ic) {
    const int node = e_v[ic * ne + l];
    const double p0 = c_p[node], p1 = c_p[nv + node], p2 = c_p[2 * nv + node];
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    const int n = c_nnb[node];
    for (int j = 0; j < n; ++j) {
      const int l2 = c_ed[j * nv + node];
      const double qtmp = 0.5 * (ql + q_ed[l2]);
      const double a = (double)c_ttg[j * nv + node] * uh[l2] * qtmp;
      const double len = e_dleng[l2];
      v0 = v0 + a * (e_c[l2] - p0) * len;
      v1 = v1 + a * (e_c[ne + l2] - p1) * len;
      v2 = v2 + a * (e_c[2 * ne + l2] - p2) * len;
    }
    const double ar = c_area[node];
    v0 = v0 / ar;
    v1 = v1 / ar;
    v2 = v2 / ar;
    proj_sphere(v0, v1, v2, p0, p1, p2);
    w[ic][0] = v0;
    w[ic][1] = v1;
    w[ic][2] = v2;
  }
  double v0 = 0.5 * w[0][0] + 0.5 * w[1][0];
  double v1 = 0.5 * w[0][1] + 0.5 * w[1][1];
  double v2 = 0.5 * w[0][2] + 0.5 * w[1][2];
  proj_sphere(v0, v1, v2, e_c[l], e_c[ne + l], e_c[2 * ne + l]);
  double dp = 0.0;
  dp = dp + v0 * e_nr[l];
  dp = dp + v1 * e_nr[ne + l];
  dp = dp + v2 * e_nr[2 * ne + l];
  // HTC: minus the ed%nr component; HC: the edhx%tg component (e_nr holds
  // the vector, psign the sign)
  const double p = psign < 0.0 ? -dp : dp;
  uhq[l] = p;
  momeq[l] = zero_mom ? 0.0 : (-g) - p;
}

// ode_rk4 combinations (swm.f90:2647-2678)
__global__ __launch_bounds__(TPB) void ref_rk_stage(int n, const double *__restrict__ y,
                                                    const double *__restrict__ k, double dt,
                                                    int half, double *__restrict__ yn) {
  const int i = xcd_index();
  if (i >= n) return;
  yn[i] = half ? y[i] + dt * k[i] / 2.0 : y[i] + dt * k[i];
}

__global__ __launch_bounds__(TPB) void ref_rk_final(int n, double *__restrict__ y,
                                                    const double *__restrict__ k0,
                                                    const double *__restrict__ k1,
                                                    const double *__restrict__ k2,
                                                    const double *__restrict__ k3, double dt) {
  const int i = xcd_index();
  if (i >= n) return;
  y[i] = y[i] + dt * (k0[i] + 2.0 * k1[i] + 2.0 * k2[i] + k3[i]) / 6.0;
}

// =========================================================== FAST kernels ==
// Internal numbering: edge of owning cell c is 3c+k, triangle of owning cell
// c is 2c+k (k < 3, k < 2). A cell rebuilds the ids of its ring edges and
// triangles from the neighbour ids it holds plus per-slot bits:
//   bitsA: nnb (0-2) | ttgout<0 (3-8) | tnrccw<0 (9-14) | cell owns edge j (15-20)
//          | diffusion: sign of the K zeta jump of owned edge j is -1 (21-26)
//   bitsB: k of edge j (2 bits at 2j) | owner of triangle j: 0 cell, 1 nb_j,
//          2 nb_{j+1} (2 bits at 12+2j) | k of triangle j (bit 24+j)
// Pentagons carry a padded 6th slot repeating slot 0, masked out of all sums,
// so every lane issues the same independent gathers and no load sits behind
// a branch.
#define BA_TTG 3
#define BA_TNR 9
#define BA_OWN 15
#define BA_FZ 21
#define BB_TSEL 12
#define BB_KT 24
#define OTH_MASK 0x3FFFFFFF
#define OTH_V2 30

// Points and vectors of the fused APVM pass, one 16 B (FP32) or 32 B (FP64)
// record each, so a gather is one vector load instead of three.
template <typename T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};

template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_tri(int remap, int NT, const int32_t *__restrict__ t_ab,
                                                const int32_t *__restrict__ t_eb,
                                                const T *__restrict__ t_ceta,
                                                const T *__restrict__ t_kap,
                                                const T *__restrict__ t_f,
                                                const T *__restrict__ u,
                                                const T *__restrict__ h, T href, int use_pv, int flin, int nopv,
                                                const T *__restrict__ kt, T *__restrict__ cz,
                                                int gmode, const T *__restrict__ g_w,
                                                const T *__restrict__ g_a, T *__restrict__ gE,
                                                T *__restrict__ gH, const T *__restrict__ gk_w,
                                                T *__restrict__ gk_k, const V4<T> *__restrict__ dt_w,
                                                V4<T> *__restrict__ dt_vhq, const T *__restrict__ dt_he,
                                                T *__restrict__ q,
                                                int *__restrict__ flag) {
  const int t = ent_index<BS>(remap);
  if (t >= NT) return;
  const int a = t_ab[t];
  if (a < 0) return;
  const int o = t >> 1;
  const int b = t_ab[NT + t];
  const int eb = t_eb[t];
  // edges (o,a), (a,b), (b,o): bit 0 of each 3-bit field = first vertex owns it
  const int e0 = 3 * ((eb & 1) ? o : a) + ((eb >> 1) & 3);
  const int e1 = 3 * (((eb >> 3) & 1) ? a : b) + ((eb >> 4) & 3);
  const int e2 = 3 * (((eb >> 6) & 1) ? b : o) + ((eb >> 7) & 3);
  T eta = t_f[t];
  if (!flin) {  // the linearised f-sphere keeps f alone
    eta += u[e0] * t_ceta[t];
    eta += u[e1] * t_ceta[NT + t];
    eta += u[e2] * t_ceta[2 * NT + t];
  }
  if (kt) {  // diffusion: K zeta (relative vorticity) at the triangle
    T z = u[e0] * t_ceta[t];
    z += u[e1] * t_ceta[NT + t];
    z += u[e2] * t_ceta[2 * NT + t];
    cz[t] = kt[t] * z;
  }
  if (gk_w) {
    // kinetic_energy_tr (Gassmann's kinetic energy): sum of the
    // (circumcentre, edge) sub-areas times u^2, over A_t
    const T u0 = u[e0], u1 = u[e1], u2 = u[e2];
    T k = gk_w[t] * (u0 * u0);
    k += gk_w[NT + t] * (u1 * u1);
    k += gk_w[2 * NT + t] * (u2 * u2);
    gk_k[t] = k;
  }
  if (gmode) {
    // gass18: h at each edge from its two vertices, h_t = sum w_m h_e / A_t
    // with w_m the area of (circumcentre, edge m) (scalar_edtr2trcc); the
    // rhombus PV at edges takes E = alpha eta and H = alpha h_t of both
    // triangles, alpha = A_t with HTC and 1 with HC (scalar_tr2ed)
    const T ho = h[o], ha = h[a], hb = h[b];
    T hs = g_w[t] * ((T)0.5 * (ho + ha));
    hs += g_w[NT + t] * ((T)0.5 * (ha + hb));
    hs += g_w[2 * NT + t] * ((T)0.5 * (hb + ho));
    const T Hs = href * g_w[3 * NT + t] + hs;  // h = href + h'
    const T At = g_a[t];
    if (use_pv && fabs((double)Hs / (double)At) < HTR_GUARD) flag[0] = 1;
    if (gmode == 1) {
      gE[t] = At * eta;
      gH[t] = Hs;
    } else {
      gE[t] = eta;
      gH[t] = Hs / At;
    }
    if (q) {  // APVM's gradient takes q at the triangle, eta / h_t
      T qq = (T)0;
      if (use_pv && !(fabs((double)Hs / (double)At) < HTR_GUARD)) qq = eta / (Hs / At);
      q[t] = qq;
    }
    return;
  }
  T hsum = t_kap[t] * h[o];
  hsum += t_kap[NT + t] * h[a];
  hsum += t_kap[2 * NT + t] * h[b];
  const T ht = href + hsum;
  T qq = (T)0;
  if (use_pv) {
    if (fabs((double)ht) < HTR_GUARD) flag[0] = 1;
    else qq = eta / ht;
  }
  // the level model's Coriolis term reads the absolute vorticity, the PV
  // still guards the depth (scalar_elem_divide runs either way)
  q[t] = nopv ? eta : qq;
  if (dt_w) {
    // dtred: q times the Perot reconstruction of uh at the circumcentre
    // (vector_edhx2tr_perot, swm.f90:2501-2503), linear in uh: the host
    // folds signs, lengths, the cross product with the circumcentre and
    // 1/A_t into one vector per edge slot; uh takes h at the edge as the
    // mean of its two cells, or with bary on HC the barycentric depth
    // fast_hc_he wrote (dt_he)
    T uh0, uh1, uh2;
    if (dt_he) {
      uh0 = u[e0] * dt_he[e0];
      uh1 = u[e1] * dt_he[e1];
      uh2 = u[e2] * dt_he[e2];
    } else {
      const T ho = h[o], ha = h[a], hb = h[b];
      uh0 = u[e0] * (href + (T)0.5 * (ho + ha));
      uh1 = u[e1] * (href + (T)0.5 * (ha + hb));
      uh2 = u[e2] * (href + (T)0.5 * (hb + ho));
    }
    const V4<T> w0 = dt_w[t], w1 = dt_w[NT + t], w2 = dt_w[2 * NT + t];
    dt_vhq[t] = V4<T>{qq * (uh0 * w0.x + uh1 * w1.x + uh2 * w2.x), qq * (uh0 * w0.y + uh1 * w1.y + uh2 * w2.y),
                      qq * (uh0 * w0.z + uh1 * w1.z + uh2 * w2.z), (T)0};
  }
}

// pxt16 with HC: h at the Voronoi edge midpoint, barycentric in the cells
// hv of the triangle that holds it (scalar_hx2ed, swm_operators.f90:88-114),
// one value per edge slot that both end cells read, so the mass flux still
// cancels exactly between them
template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_hc_he(int remap, int NE, const int32_t *__restrict__ e_oth,
                                                  const T *__restrict__ hb, const int32_t *__restrict__ hv,
                                                  const T *__restrict__ hs, T href, T *__restrict__ he) {
  const int e = ent_index<BS>(remap);
  if (e >= NE || e_oth[e] < 0) return;
  T s = hb[e] * hs[hv[e]];
  s += hb[NE + e] * hs[hv[NE + e]];
  s += hb[2 * NE + e] * hs[hv[2 * NE + e]];
  he[e] = href + s;
}

// gass18's PV at every edge slot from the rhombus of its two triangles:
// q = (E1 + E2) / (H1 + H2), or with trsk interpolation (qtr non-null) the
// mean of the two triangles' PV; with APVM (apv non-null) minus
// pvspar dt (grad q(t1) + grad q(t2)) . (v(c1) + v(c2)) (apvm,
// swm_operators.f90:1497-1501), from fast_apvm_pre's vectors
template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_gass_q(int remap, int NE, const int32_t *__restrict__ e_t,
                                                   const T *__restrict__ gE, const T *__restrict__ gH,
                                                   const int32_t *__restrict__ e_oth,
                                                   const V4<T> *__restrict__ apv,
                                                   const V4<T> *__restrict__ apg, T pvsdt,
                                                   const T *__restrict__ qtr, const T *__restrict__ qw,
                                                   T *__restrict__ gq) {
  const int l = ent_index<BS>(remap);
  if (l >= NE) return;
  const int t1 = e_t[l];
  if (t1 < 0) return;
  const int t2 = e_t[NE + l];
  T qe;
  if (qtr) {  // gass Coriolis term with trsk interpolation: the plain average (scalar_tr2ed)
    // or with bary (HTC) the weights of the distances to the circumcentres
    qe = qw ? qw[l] * qtr[t1] + qw[NE + l] * qtr[t2] : (T)0.5 * (qtr[t1] + qtr[t2]);
  } else {
    const T H = gH[t1] + gH[t2];
    qe = (H != (T)0) ? (gE[t1] + gE[t2]) / H : (T)0;
  }
  if (apv) {
    const int c1 = l / 3, c2 = e_oth[l] & OTH_MASK;
    const V4<T> g1 = apg[t1], g2 = apg[t2], v1 = apv[c1], v2 = apv[c2];
    const T d = (g1.x + g2.x) * (v1.x + v2.x) + (g1.y + g2.y) * (v1.y + v2.y) + (g1.z + g2.z) * (v1.z + v2.z);
    qe -= pvsdt * d;
  }
  gq[l] = qe;
}

// Neumaier's compensated summation step.
__device__ __forceinline__ void kadd(double &s, double &c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

// APVM on the fused path, one thread per cell after fast_tri: the Perot
// velocity at the cell (vector_edhx2hx_perot) and, for the cell's owned
// triangles, the Perot reconstruction of grad q (grad_tr2ed then
// vector_edtr2tr_perot). Both are linear in u and in q differences, and the
// tangent-plane projection is linear too, so the host folds geometry, signs,
// lengths, areas and the projection into one vector weight per cell slot
// (wc) and per triangle slot (wt): v = sum_j wc_j u_j and
// grad q = sum_l wt_l (q(across l) - q(t)). tnb < 0 marks padding triangles.
template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_apvm_pre(
    int remap, int nv, int NT, const int32_t *__restrict__ c_nb, const int32_t *__restrict__ c_bitsA,
    const int32_t *__restrict__ c_bitsB, const V4<T> *__restrict__ wc, const int32_t *__restrict__ tnb,
    const V4<T> *__restrict__ wt, const T *__restrict__ us, const T *__restrict__ q, V4<T> *__restrict__ vc,
    V4<T> *__restrict__ qg) {
  const int c = ent_index<BS>(remap);
  if (c >= nv) return;
  const int A = c_bitsA[c];
  const int B = c_bitsB[c];
  T v0 = (T)0, v1 = (T)0, v2 = (T)0;
#pragma unroll
  for (int j = 0; j < MX; ++j) {  // a pentagon's padded slot has a zero weight
    const int nbj = c_nb[j * nv + c];
    const int eo = ((A >> (BA_OWN + j)) & 1) ? c : nbj;
    const T uj = us[3 * eo + ((B >> (2 * j)) & 3)];
    const V4<T> w = wc[j * nv + c];
    v0 += w.x * uj;
    v1 += w.y * uj;
    v2 += w.z * uj;
  }
  vc[c] = V4<T>{v0, v1, v2, (T)0};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = 2 * c + k;
    const int n0 = tnb[t];
    if (n0 < 0) continue;
    const T qt = q[t];
    const T d0 = q[n0] - qt, d1 = q[tnb[NT + t]] - qt, d2 = q[tnb[2 * NT + t]] - qt;
    const V4<T> a0 = wt[t], a1 = wt[NT + t], a2 = wt[2 * NT + t];
    qg[t] = V4<T>{a0.x * d0 + a1.x * d1 + a2.x * d2, a0.y * d0 + a1.y * d1 + a2.y * d2,
                  a0.z * d0 + a1.z * d1 + a2.z * d2, (T)0};
  }
}

// Hyperdiffusion on the fused path: lap_lapu = diffusion_ed of the
// divergence and vorticity of lapu = diffusion_ed(div u, zeta)
// (swm_operators.f90:897-975), in three passes after fast_tri (which holds
// K zeta in cz). hyper_a: Kd = K div u at cells.
template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_hyper_a(int remap, int nv, const int32_t *__restrict__ c_nb,
                                                    const int32_t *__restrict__ c_bitsA,
                                                    const int32_t *__restrict__ c_bitsB,
                                                    const T *__restrict__ c_dl, const double *__restrict__ c_iaa,
                                                    const T *__restrict__ kh, const T *__restrict__ f,
                                                    T *__restrict__ out) {
  const int c = ent_index<BS>(remap);
  if (c >= nv) return;
  const int A = c_bitsA[c];
  const int B = c_bitsB[c];
  const bool hex = ((A & 7) == MX);
  double du = 0.0;
#pragma unroll
  for (int j = 0; j < MX; ++j) {
    const int eo = ((A >> (BA_OWN + j)) & 1) ? c : c_nb[j * nv + c];
    double g = (double)f[3 * eo + ((B >> (2 * j)) & 3)] * (double)c_dl[j * nv + c];
    if ((A >> (BA_TTG + j)) & 1) g = -g;
    if ((j < MX - 1) || hex) du += g;
  }
  out[c] = kh[c] * (T)(du * c_iaa[c]);
}

// hyper_b: lapu at each edge, grad(Kd) - grad_tg(K zeta):
// (Kd(v2) - Kd(v1)) / (|e| a) - s (Kz(sh2) - Kz(sh1)) / (|e*| a), with the
// end cells in the edge's gradient orientation (e_oth) and s folded into eb.
template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_hyper_b(int remap, int NE, const int32_t *__restrict__ e_oth,
                                                    const int32_t *__restrict__ e_sh, const T *__restrict__ ea,
                                                    const T *__restrict__ eb, const T *__restrict__ kd,
                                                    const T *__restrict__ cz, T *__restrict__ lap) {
  const int e = ent_index<BS>(remap);
  if (e >= NE) return;
  const int oth = e_oth[e];
  if (oth < 0) return;
  const int o = e / 3;
  const int c2 = oth & OTH_MASK;
  const bool own_v2 = (oth >> OTH_V2) & 1;
  const int v1 = own_v2 ? c2 : o, v2 = own_v2 ? o : c2;
  lap[e] = (kd[v2] - kd[v1]) * ea[e] - (cz[e_sh[NE + e]] - cz[e_sh[e]]) * eb[e];
}

// hyper_c: K div(lapu) at cells and, for the cell's owned triangles,
// K curl(lapu) (the vorticity coefficients of fast_tri)
template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_hyper_c(int remap, int nv, int NT, const int32_t *__restrict__ c_nb,
                                                    const int32_t *__restrict__ c_bitsA,
                                                    const int32_t *__restrict__ c_bitsB,
                                                    const T *__restrict__ c_dl, const double *__restrict__ c_iaa,
                                                    const T *__restrict__ kh, const T *__restrict__ kt,
                                                    const int32_t *__restrict__ t_ab,
                                                    const int32_t *__restrict__ t_eb,
                                                    const T *__restrict__ t_ceta, const T *__restrict__ lap,
                                                    T *__restrict__ hd, T *__restrict__ hz) {
  const int c = ent_index<BS>(remap);
  if (c >= nv) return;
  const int A = c_bitsA[c];
  const int B = c_bitsB[c];
  const bool hex = ((A & 7) == MX);
  double du = 0.0;
#pragma unroll
  for (int j = 0; j < MX; ++j) {
    const int eo = ((A >> (BA_OWN + j)) & 1) ? c : c_nb[j * nv + c];
    double g = (double)lap[3 * eo + ((B >> (2 * j)) & 3)] * (double)c_dl[j * nv + c];
    if ((A >> (BA_TTG + j)) & 1) g = -g;
    if ((j < MX - 1) || hex) du += g;
  }
  hd[c] = kh[c] * (T)(du * c_iaa[c]);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = 2 * c + k;
    const int ta = t_ab[t];
    if (ta < 0) continue;
    const int tb = t_ab[NT + t];
    const int eb = t_eb[t];
    const int e0 = 3 * ((eb & 1) ? c : ta) + ((eb >> 1) & 3);
    const int e1 = 3 * (((eb >> 3) & 1) ? ta : tb) + ((eb >> 4) & 3);
    const int e2 = 3 * (((eb >> 6) & 1) ? tb : c) + ((eb >> 7) & 3);
    T z = lap[e0] * t_ceta[t];
    z += lap[e1] * t_ceta[NT + t];
    z += lap[e2] * t_ceta[2 * NT + t];
    hz[t] = kt[t] * z;
  }
}

// OPT selects the code compiled in, and with it the register footprint:
// 0 the plain scheme, 1 with diffusion and hyperdiffusion, 2 with every
// option (APVM and pxt16 too), 3 gass18, 4 the level model (with
// diffusion), 5 gass18 with the Perot or Melvin kinetic energy. Pointers of
// options left out are ignored.
template <typename T, int BS, int OPT>
__global__ __launch_bounds__(BS) void fast_cell(
    int remap, int nv, const int32_t *__restrict__ c_nb, const int32_t *__restrict__ c_bitsA,
    const int32_t *__restrict__ c_bitsB, const T *__restrict__ c_R,
    const double *__restrict__ c_iaa, const T *__restrict__ c_ika,
    const T *__restrict__ c_dl, const T *__restrict__ c_kw, const T *__restrict__ q,
    const T *__restrict__ kh, const T *__restrict__ cz, const T *__restrict__ hd,
    const T *__restrict__ hz, const void *tang, const V4<T> *__restrict__ pxw, const V4<T> *__restrict__ pkw,
    const T *__restrict__ pxq, V4<T> *__restrict__ pxab, T *__restrict__ pxqe, const V4<T> *__restrict__ apv,
    const V4<T> *__restrict__ apg, T pvsdt, const T *__restrict__ gq, const int32_t *__restrict__ gout,
    const T *__restrict__ hce, int glink, const T *__restrict__ gk_k, T gk_c, int melv,
    const V4<T> *__restrict__ mw, int dtred, const T *__restrict__ hs, const T *__restrict__ us, double href, const T *__restrict__ gb, T grav,
    T inv_a, T *__restrict__ phi, T *__restrict__ Pown, T *__restrict__ Poth, int stage, double dt_all,
    const double *__restrict__ dtc, const T *h0, T *hout, T *__restrict__ acch,
    double *__restrict__ dh_out) {
  // the level model (OPT 4): the TRiSK sum takes u and eta at edges in
  // place of uh and q
  constexpr bool kDiff = OPT == 1 || OPT == 2 || OPT == 4, kAll = OPT == 2, kGass = OPT == 3 || OPT == 5,
                 kNopv = OPT == 4, kKe = OPT == 2 || OPT == 5;
  const int c = ent_index<BS>(remap);
  if (c >= nv) return;
  const int A = c_bitsA[c];
  const int B = c_bitsB[c];
  const bool hex = ((A & 7) == MX);
  int nb[MX];
#pragma unroll
  for (int j = 0; j < MX; ++j) nb[j] = c_nb[j * nv + c];
  T R[MX];
#pragma unroll
  for (int j = 0; j < MX; ++j) R[j] = c_R[j * nv + c];
  int ed[MX], tr[MX];
#pragma unroll
  for (int j = 0; j < MX; ++j) {
    const int eo = ((A >> (BA_OWN + j)) & 1) ? c : nb[j];
    ed[j] = 3 * eo + ((B >> (2 * j)) & 3);
    const int ts = (B >> (BB_TSEL + 2 * j)) & 3;
    const int to = (ts == 0) ? c : (ts == 1) ? nb[j] : nb[(j + 1) % MX];
    tr[j] = 2 * to + ((B >> (BB_KT + j)) & 1);
  }
  const T hc = hs[c];
  T ue[MX], hn[MX], qt[MX], kw[MX];
  double dl[MX];
#pragma unroll
  for (int j = 0; j < MX; ++j) {
    ue[j] = us[ed[j]];
    hn[j] = hs[nb[j]];
    qt[j] = kGass ? (T)0 : q[tr[j]];
    dl[j] = (double)c_dl[j * nv + c];
    kw[j] = c_kw[j * nv + c];
  }
  // h at the cell's edges: the average of the two cells, or (pxt16 with HC)
  // the barycentric value of the edge's own triangle
  double hE[MX];
#pragma unroll
  for (int j = 0; j < MX; ++j)
    hE[j] = ((kAll || kGass) && hce) ? (double)hce[ed[j]] : href + 0.5 * ((double)hc + (double)hn[j]);
  double dv = 0.0, dvc = 0.0;
  T ke = (T)0;
  T G[MX];
#pragma unroll
  for (int j = 0; j < MX; ++j) {
    const bool valid = (j < MX - 1) || hex;
    const double he = hE[j];
    double g = (double)ue[j] * he * dl[j];
    if ((A >> (BA_TTG + j)) & 1) g = -g;
    if (!valid) g = 0.0;
    kadd(dv, dvc, g);
    if (kNopv) {  // coriolis_ed noPV (swm_operators.f90:1297-1300): u, not uh
      T gu = ue[j] * (T)dl[j];
      if ((A >> (BA_TTG + j)) & 1) gu = -gu;
      G[j] = valid ? gu : (T)0;
    } else {
      G[j] = (T)g;
    }
    ke += valid ? kw[j] * ue[j] * ue[j] : (T)0;
  }
  const double dh = -(dv + dvc) * c_iaa[c];
  T K = ke * c_ika[c];
  if ((kAll || kGass) && gk_k) {
    // Gassmann's kinetic energy (kinetic_energy_hx, swm_operators.f90:
    // 1198-1207): gasscoef K_trsk + (1 - gasscoef) sum_j hxtr_j K_tr(t_j)
    T kk = K * gk_c;
#pragma unroll
    for (int j = 0; j < MX; ++j)
      if ((j < MX - 1) || hex) kk += ((T)1 - gk_c) * R[j] * gk_k[tr[j]];
    K = kk;
  }
  if (kKe && melv) {
    // Melvin's kinetic energy (kinetic_energy_hx, swm_operators.f90:
    // 1174-1197): HTC, sum_j kw_j u_j (u_j h_e) / (h sum_j kw_j); HC, the
    // Perot vectors of uh and u dotted, over 2 h
    const double hcell = href + (double)hc;
    if (mw) {
      T a0 = (T)0, a1 = (T)0, a2 = (T)0, b0 = (T)0, b1 = (T)0, b2 = (T)0;
#pragma unroll
      for (int j = 0; j < MX; ++j) {  // a pentagon's padded slot has a zero weight
        const V4<T> w = mw[j * nv + c];
        const T uh = ue[j] * (T)hE[j];
        a0 += w.x * uh;
        a1 += w.y * uh;
        a2 += w.z * uh;
        b0 += w.x * ue[j];
        b1 += w.y * ue[j];
        b2 += w.z * ue[j];
      }
      K = (T)((double)(a0 * b0 + a1 * b1 + a2 * b2) / hcell / 2.0);
    } else {
      T km = (T)0;
#pragma unroll
      for (int j = 0; j < MX; ++j)
        if ((j < MX - 1) || hex) km += kw[j] * ue[j] * (ue[j] * (T)hE[j]);
      K = (T)((double)(km * c_ika[c]) / hcell);
    }
  }
  if (kKe && pkw) {  // perhx (pxt16): |v|^2 / 2 of the cell's Perot velocity
    T v0 = (T)0, v1 = (T)0, v2 = (T)0;
#pragma unroll
    for (int j = 0; j < MX; ++j) {  // a pentagon's padded slot has a zero weight
      const V4<T> w = pkw[j * nv + c];
      v0 += w.x * ue[j];
      v1 += w.y * ue[j];
      v2 += w.z * ue[j];
    }
    K = (v0 * v0 + v1 * v1 + v2 * v2) / (T)2;
  }
  T Kd = (T)0;
  if (kDiff && kh) {  // diffusion, normal part: grad(K div u) folds into the potential
    double du = 0.0;
#pragma unroll
    for (int j = 0; j < MX; ++j) {
      double g = (double)ue[j] * dl[j];
      if ((A >> (BA_TTG + j)) & 1) g = -g;
      if ((j < MX - 1) || hex) du += g;
    }
    Kd = kh[c] * (T)(du * c_iaa[c]);
  }
  if (kDiff && hd) Kd -= hd[c];  // hyperdiffusion: - grad(K div lapu)
  T pot = grav * hc;
  if (gb) pot += gb[c];  // bottom topography: g (h + b)
  phi[c] = (pot + K - Kd) * inv_a;

  // PV at the cell's edges: edge j lies between ring triangles j-1 and j
  T qe[MX], qo[MX];
  if (kGass) {
    // gass18: the rhombus PV of each ring edge and of the outer edge of each
    // ring triangle (the edge that links ring edges j and j+1)
#pragma unroll
    for (int j = 0; j < MX; ++j) {
      qe[j] = gq[ed[j]];
      qo[j] = glink ? gq[gout[j * nv + c]] : (T)0;
    }
  } else if (kAll && pxq) {  // pxt16: weighted by the distances to the circumcentres
#pragma unroll
    for (int j = 0; j < MX; ++j) {
      const int jp = (j == 0) ? (hex ? MX - 1 : MX - 2) : j - 1;
      qe[j] = pxq[j * nv + c] * qt[jp] + pxq[(MX + j) * nv + c] * qt[j];
    }
  } else {
    qe[0] = (T)0.5 * ((hex ? qt[MX - 1] : qt[MX - 2]) + qt[0]);
#pragma unroll
    for (int j = 1; j < MX; ++j) qe[j] = (T)0.5 * (qt[j - 1] + qt[j]);
  }
  if (kAll && apv) {
    // APVM: q_e -= pvspar dt (grad q(t_{j-1}) + grad q(t_j)) . (v(c) + v(nb_j))
    const V4<T> vs = apv[c];
    V4<T> g[MX];
#pragma unroll
    for (int j = 0; j < MX; ++j) g[j] = apg[tr[j]];
#pragma unroll
    for (int j = 0; j < MX; ++j) {
      const int jp = (j == 0) ? (hex ? MX - 1 : MX - 2) : j - 1;
      const V4<T> vn = apv[nb[j]];
      const T d = (g[jp].x + g[j].x) * (vs.x + vn.x) + (g[jp].y + g[j].y) * (vs.y + vn.y) +
                  (g[jp].z + g[j].z) * (vs.z + vn.z);
      qe[j] -= pvsdt * d;
    }
  }
  if (kAll && (pxw || dtred)) {
    // the Coriolis term is fast_edge's (pered, dtred); this cell stores the
    // diffusion part of the partials of the edges it owns
    if (pxw) {
      // pxt16's pered Coriolis term: the Perot reconstruction at this cell
      // of uh (q_e + q_l) / 2 for an edge l is (q_l A + B) / 2 with
      // A = P(uh) and B = P(uh q); fast_edge combines both cells' A and B
      // with q_l
      T a0 = (T)0, a1 = (T)0, a2 = (T)0, b0 = (T)0, b1 = (T)0, b2 = (T)0;
#pragma unroll
      for (int j = 0; j < MX; ++j) {
        const V4<T> w = pxw[j * nv + c];
        const T uh = ue[j] * (T)hE[j];
        const T uq = uh * qe[j];
        a0 += w.x * uh;
        a1 += w.y * uh;
        a2 += w.z * uh;
        b0 += w.x * uq;
        b1 += w.y * uq;
        b2 += w.z * uq;
      }
      pxab[c] = V4<T>{a0, a1, a2, (T)0};
      pxab[nv + c] = V4<T>{b0, b1, b2, (T)0};
    }
#pragma unroll
    for (int i = 0; i < MX; ++i) {
      if ((i < MX - 1 || hex) && ((A >> (BA_OWN + i)) & 1)) {
        T P = (T)0;
        if (kDiff && tang) {  // diffusion's tangential part, as below
          const int ip = (i == 0) ? (hex ? MX - 1 : MX - 2) : i - 1;
          T fz = (T)4 * kw[i] * (T)(1.0 / (dl[i] * dl[i])) * inv_a;
          if ((A >> (BA_FZ + i)) & 1) fz = -fz;
          T jump = (T)0;
          if (kh) jump = cz[tr[i]] - cz[tr[ip]];
          if (hz) jump -= hz[tr[i]] - hz[tr[ip]];
          P = fz * jump;
        }
        Pown[ed[i]] = P;
        if (pxw) pxqe[ed[i]] = qe[i];
      }
    }
  } else {
  // TRiSK weights from prefix sums of the kite areas:
  // w(i,j) = S_j - S_i - 1/2 (j>i), tot - S_i + S_j - 1/2 (j<i), w(i,i) = 0
  T S[MX + 1];
  S[0] = (T)0;
#pragma unroll
  for (int j = 0; j < MX; ++j) S[j + 1] = S[j] + R[j];
  const T tot = S[MX];
#pragma unroll
  for (int i = 0; i < MX; ++i) {
    T acc = (T)0;
#pragma unroll
    for (int j = 0; j < MX; ++j) {
      if (j != i) {
        const T w = ((j > i) ? (S[j] - S[i]) : (tot - S[i] + S[j])) - (T)0.5;
        T qs = qe[i] + qe[j];
        if (kGass && glink) {
          // coriolis_ed gass (swm_operators.f90:1378-1403): the PV of the
          // edge linking i and j (gethxedgeconnection): the outer edge of
          // the ring triangle between neighbours, the ring edge between
          // ring edges two apart, else the trsk average
          const int n = hex ? MX : MX - 1;
          const int dj = (j - i + n) % n;
          if (dj == 1) qs = (T)2 * qo[i];
          else if (dj == n - 1) qs = (T)2 * qo[j];
          else if (dj == 2) qs = (T)2 * qe[(i + 1) % n];
          else if (dj == n - 2) qs = (T)2 * qe[(j + 1) % n];
        }
        acc += G[j] * qs * w;
      }
    }
    T P = (T)0.5 * acc;
    if ((A >> (BA_TNR + i)) & 1) P = -P;
    if (i < MX - 1 || hex) {
      if ((A >> (BA_OWN + i)) & 1) {
        // diffusion, tangential part (owner only): the jump of K zeta
        // between the edge's ring triangles i-1 and i, times |e| / (|e*| a)
        // = 4 kw / (|e*|^2 a)
        if (kDiff && tang) {  // non-null with diffusion or hyperdiffusion
          const int ip = (i == 0) ? (hex ? MX - 1 : MX - 2) : i - 1;
          T fz = (T)4 * kw[i] * (T)(1.0 / (dl[i] * dl[i])) * inv_a;
          if ((A >> (BA_FZ + i)) & 1) fz = -fz;
          T jump = (T)0;
          if (kh) jump = cz[tr[i]] - cz[tr[ip]];
          if (hz) jump -= hz[tr[i]] - hz[tr[ip]];  // hyperdiffusion: + grad_tg(K curl lapu)
          P += fz * jump;
        }
        Pown[ed[i]] = P;
      } else {
        Poth[ed[i]] = P;
      }
    }
  }
  }

  const double dt = dtc ? dtc[c] : dt_all;  // ensembles: per-instance step
  if (stage < 0) {
    dh_out[c] = dh;
  } else if (stage < 3) {
    const double a = (stage == 2) ? dt : 0.5 * dt;
    const double b = (stage == 0) ? dt / 6.0 : dt / 3.0;
    const double y0 = (double)h0[c];
    hout[c] = (T)(y0 + a * dh);
    acch[c] = (stage == 0) ? (T)(b * dh) : (T)((double)acch[c] + b * dh);
  } else {
    hout[c] = (T)((double)h0[c] + ((double)acch[c] + (dt / 6.0) * dh));
  }
}

template <typename T, int BS>
__global__ __launch_bounds__(BS) void fast_edge(int remap, int NE, const int32_t *__restrict__ e_oth,
                                                 const T *__restrict__ e_il,
                                                 const T *__restrict__ phi,
                                                 const T *__restrict__ Pown,
                                                 const T *__restrict__ Poth,
                                                 const V4<T> *__restrict__ pxab,
                                                 const T *__restrict__ pxqe,
                                                 const V4<T> *__restrict__ pxn,
                                                 const int32_t *__restrict__ dt_et,
                                                 const T *__restrict__ dt_ew,
                                                 const V4<T> *__restrict__ dt_n,
                                                 const V4<T> *__restrict__ dt_vhq, int zero_mom,
                                                 int stage, T dt_all, const T *__restrict__ dte,
                                                 const T *u0, T *uout, T *__restrict__ accu,
                                                 T *__restrict__ du_out) {
  const int e = ent_index<BS>(remap);
  if (e >= NE) return;
  const int oth = e_oth[e];
  if (oth < 0) return;
  const int o = e / 3;
  const int c2 = oth & OTH_MASK;
  const bool own_v2 = (oth >> OTH_V2) & 1;
  const int v1 = own_v2 ? c2 : o, v2 = own_v2 ? o : c2;
  T du;
  if (pxab) {
    // pered (pxt16): the tangent-plane average of both cells' Perot vectors
    // of uh (q_e + q_l) / 2, dotted with the edge normal (minus the ed%nr
    // component; the projection and the sign are folded into pxn)
    const int nv = NE / 3;
    const T ql = pxqe[e];
    const V4<T> A1 = pxab[v1], A2 = pxab[v2], B1 = pxab[nv + v1], B2 = pxab[nv + v2], n = pxn[e];
    const T uhq = (T)0.25 * (n.x * (ql * (A1.x + A2.x) + (B1.x + B2.x)) + n.y * (ql * (A1.y + A2.y) + (B1.y + B2.y)) +
                             n.z * (ql * (A1.z + A2.z) + (B1.z + B2.z)));
    du = -((phi[v2] - phi[v1]) + Pown[e]) * e_il[e] - uhq;
  } else if (dt_vhq) {
    // dtred (coriolis_ed, swm_operators.f90:1310-1329): the two triangles'
    // vhq interpolated to the edge (distance weights, HTC; the mean, HC),
    // dotted with the edge normal (projection and sign folded into dt_n)
    const V4<T> a = dt_vhq[dt_et[e]], b = dt_vhq[dt_et[NE + e]], n = dt_n[e];
    const T w1 = dt_ew[e], w2 = dt_ew[NE + e];
    const T uhq = n.x * (w1 * a.x + w2 * b.x) + n.y * (w1 * a.y + w2 * b.y) + n.z * (w1 * a.z + w2 * b.z);
    du = -((phi[v2] - phi[v1]) + Pown[e]) * e_il[e] - uhq;
  } else {
    du = -((phi[v2] - phi[v1]) + (Pown[e] + Poth[e])) * e_il[e];
  }
  if (zero_mom) du = (T)0;
  const T dt = dte ? dte[e] : dt_all;
  if (stage < 0) {
    du_out[e] = du;
  } else if (stage < 3) {
    const T a = (stage == 2) ? dt : (T)0.5 * dt;
    const T b = (stage == 0) ? dt / (T)6 : dt / (T)3;
    uout[e] = u0[e] + a * du;
    accu[e] = (stage == 0) ? b * du : accu[e] + b * du;
  } else {
    uout[e] = u0[e] + (accu[e] + (dt / (T)6) * du);
  }
}

// ============================================================ diagnostics ==
// FP64 reductions of the current state (h = href + hs, u) with wavefront
// shuffles + LDS, one partial per block, then a one-block final pass.
#define NPART 8

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ void block_store(double (&v)[NPART], double *part) {
  __shared__ double red[TPB / 64][NPART];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NPART; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[wid][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < NPART) {
    double s = 0.0;
    for (int w = 0; w < TPB / 64; ++w) s += red[w][threadIdx.x];
    part[(size_t)blockIdx.x * NPART + threadIdx.x] = s;
  }
}

// slot 0: sum A h, 1: sum A, 2: sum A g h (h/2 + b), 3: sum A divu^2, 7: bad
template <typename T>
__global__ __launch_bounds__(TPB) void diag_cell(int off, int cnt, int nv, const int32_t *__restrict__ c_nnb,
                                                 const int32_t *__restrict__ c_ed,
                                                 const int8_t *__restrict__ c_ttg,
                                                 const double *__restrict__ c_area,
                                                 const double *__restrict__ e_dleng,
                                                 const T *__restrict__ hs,
                                                 const T *__restrict__ us, double href,
                                                 const double *__restrict__ bt, double grav, double erad,
                                                 double *__restrict__ part) {
  double v[NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int i = off + blockIdx.x * TPB + threadIdx.x;
  if (i < off + cnt) {
    const double h = href + (double)hs[i];
    const double A = c_area[i];
    double s = 0.0;
    for (int j = 0; j < c_nnb[i]; ++j) {
      const int l = c_ed[j * nv + i];
      s = s + (double)c_ttg[j * nv + i] * (double)us[l] * e_dleng[l];
    }
    const double dvu = s / A / erad;
    v[0] = A * h;
    v[1] = A;
    v[2] = A * grav * h * (h * 0.5 + bt[i]);
    v[3] = A * (dvu * dvu);
    v[7] = isfinite(h) ? 0.0 : 1.0;
  }
  block_store(v, part);
}

// slot 4: sum 1/2 l d h_e u^2, 7: bad
template <typename T>
__global__ __launch_bounds__(TPB) void diag_edge(int off, int cnt, int ne, const int32_t *__restrict__ e_v,
                                                 const double *__restrict__ e_leng,
                                                 const double *__restrict__ e_dleng,
                                                 const T *__restrict__ hs,
                                                 const T *__restrict__ us, double href,
                                                 double *__restrict__ part) {
  double v[NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int l = off + blockIdx.x * TPB + threadIdx.x;
  if (l < off + cnt && e_v[l] >= 0) {
    const double he = (href + (double)hs[e_v[l]] + href + (double)hs[e_v[ne + l]]) * 0.5;
    const double u = (double)us[l];
    v[4] = 0.5 * e_leng[l] * e_dleng[l] * he * (u * u);
    v[7] = isfinite(u) ? 0.0 : 1.0;
  }
  block_store(v, part);
}

// slot 5: sum 1/2 A_t a^2 h_t q_t^2 (potential enstrophy)
template <typename T>
__global__ __launch_bounds__(TPB) void diag_tri(int off, int cnt, int nt, const int32_t *__restrict__ t_ed,
                                                const int8_t *__restrict__ t_tg,
                                                const int32_t *__restrict__ t_v,
                                                const double *__restrict__ t_area,
                                                const double *__restrict__ t_f,
                                                const double *__restrict__ t_trhx,
                                                const double *__restrict__ e_leng,
                                                const T *__restrict__ hs,
                                                const T *__restrict__ us, double href,
                                                double erad, int flin, double *__restrict__ part) {
  double v[NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int k = off + blockIdx.x * TPB + threadIdx.x;
  if (k < off + cnt && t_ed[k] >= 0) {
    double e = 0.0, s = 0.0;
    for (int l = 0; l < 3; ++l) {
      const int ed = t_ed[l * nt + k];
      e += (double)us[ed] * e_leng[ed] * (double)t_tg[l * nt + k];
      s += t_trhx[l * nt + k] * (href + (double)hs[t_v[l * nt + k]]);
    }
    const double eta = flin ? t_f[k] : e / t_area[k] / erad + t_f[k];
    const double ht = s / t_area[k];
    const double qq = eta / ht;
    v[5] = 0.5 * t_area[k] * erad * erad * ht * qq * qq;
  }
  block_store(v, part);
}

__global__ __launch_bounds__(TPB) void diag_final(int nparts, const double *__restrict__ part,
                                                  double *__restrict__ out) {
  double v[NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nparts; b += TPB)
#pragma unroll
    for (int k = 0; k < NPART; ++k) v[k] += part[(size_t)b * NPART + k];
  __shared__ double red[TPB / 64][NPART];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NPART; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[wid][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < NPART) {
    double s = 0.0;
    for (int w = 0; w < TPB / 64; ++w) s += red[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// ======================================================= bandwidth probe ==
// Streaming copy, used to measure the attainable HBM bandwidth and to scale
// the memory counters of a profiler (bytes moved are known exactly). Each
// lane issues UNR independent 16-byte loads before its stores, so every
// wave keeps UNR x 1 KiB of reads in flight; NT marks the accesses
// non-temporal (streamed past the caches).
typedef float probe4 __attribute__((ext_vector_type(4)));
template <int UNR, bool NT>
__global__ __launch_bounds__(TPB) void copy_probe(const probe4 *__restrict__ a, probe4 *__restrict__ b,
                                                  size_t n) {
  const size_t stride = (size_t)gridDim.x * TPB;
  size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
  for (; i + (UNR - 1) * stride < n; i += UNR * stride) {
    probe4 x[UNR];
#pragma unroll
    for (int k = 0; k < UNR; ++k) x[k] = NT ? __builtin_nontemporal_load(&a[i + k * stride]) : a[i + k * stride];
#pragma unroll
    for (int k = 0; k < UNR; ++k) {
      if (NT) __builtin_nontemporal_store(x[k], &b[i + k * stride]);
      else b[i + k * stride] = x[k];
    }
  }
  for (; i < n; i += stride) b[i] = a[i];
}

// =============================================================== handle ====
#include "imx_device.h"

namespace imx {

static int hip_fail(hipError_t e, const char *what) {
  return set_error(IMX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPTRY(expr, what)                                 \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return imx::hip_fail(_e, what); \
  } while (0)

template <typename T>
static std::vector<T> cvt(const std::vector<double> &a) {
  return std::vector<T>(a.begin(), a.end());
}

// --------------------------------------------------------------- launches --
static void launch_ref_tendency(imx_swm *s, const double *h, const double *u, double *masseq,
                                double *momeq) {
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  hipStream_t st = s->stream;
  // the method triple: scalar interpolation, Coriolis, kinetic energy
  const int sint = s->sc_sint, cor = s->sc_cor, rec = s->sc_rec;
  // the staggering's signs and Perot edge points (HC: hx%nr, edhx%c, ...)
  const bool hc = s->sc_hc;
  const int8_t *csg = hc ? s->hc_c : s->c_ttg, *ctn = hc ? s->hc_c : s->c_tnr;
  const int8_t *tsg = hc ? s->hc_t : s->t_tg, *egs = hc ? s->hc_g : nullptr;
  const int8_t *ett = hc ? s->hc_tt : s->e_nrs;
  const double *epc = hc ? s->hc_pc : s->e_c;
  const int qmode = sint == IMX_SINT_GASS ? (hc ? 3 : 2) : (!hc && sint == IMX_SINT_BARY ? 1 : 0);
  const bool hbary = hc && sint == IMX_SINT_BARY;
  ref_edge_flux<<<grid_for(NE), TPB, 0, st>>>(NE, s->e_v, hbary ? s->hc_hb : nullptr, s->hc_hv, h, u,
                                              s->f_h_ed, s->f_uh);
  ref_tri<<<grid_for(NT), TPB, 0, st>>>(NT, s->t_ed, tsg, s->t_v, s->t_area, s->t_f,
                                        s->t_trhx, s->e_leng, u, h, sint, s->t_w, s->f_h_ed,
                                        s->o.erad, s->o.testcase >= 2, s->f_lin ? 1 : 0, s->f_eta, s->f_zeta,
                                        s->f_h_tr, s->f_q_tr, s->flag);
  ref_edge_q<<<grid_for(NE), TPB, 0, st>>>(NE, qmode, s->e_sh, s->e_dw, s->t_area, s->f_eta, s->f_h_tr,
                                           s->f_q_tr, s->f_q_ed, s->sc_nopv ? s->f_eta_ed : nullptr,
                                           s->flag);
  // Perot velocity at cells: APVM, the pxt16 and Melvin kinetic energies
  // (vector_edhx2hx_perot, swm.f90:2516)
  if (s->sc_apvm || rec == IMX_REC_PERHX || (hc && rec == IMX_REC_MELV))
    ref_perot_hx<<<grid_for(nv), TPB, 0, st>>>(nv, NE, s->c_nnb, s->c_ed, csg, epc, s->c_p,
                                               s->e_dleng, s->c_area, u, s->f_vhx);
  if (hc && rec == IMX_REC_MELV)  // ... and of uh
    ref_perot_hx<<<grid_for(nv), TPB, 0, st>>>(nv, NE, s->c_nnb, s->c_ed, csg, epc, s->c_p,
                                               s->e_dleng, s->c_area, s->f_uh, s->f_vuh);
  if (rec == IMX_REC_GASS)  // kinetic_energy_tr (swm.f90:2521-2523)
    ref_ketr<<<grid_for(NT), TPB, 0, st>>>(NT, s->t_ed, s->t_ka, s->t_area, u, s->f_ketr);
  if (cor == IMX_COR_DTRED)  // vhq_tr (swm.f90:2501-2503)
    ref_lte_tr<<<grid_for(NT), TPB, 0, st>>>(NT, NE, s->t_ed, tsg, epc, s->t_c, s->e_leng, s->t_area, s->f_uh,
                                             s->f_q_tr, s->f_vhq);
  if (s->sc_apvm) {
    // grad q to edges and triangles, then the APVM PV at edges
    // (tendency, swm.f90:2481-2492, :2556-2559)
    ref_qgrad_ed<<<grid_for(NE), TPB, 0, st>>>(NE, s->e_sh, s->e_nrs, s->e_dleng, s->f_q_tr, s->o.erad,
                                               s->f_qg_ed);
    ref_perot_tr<<<grid_for(NT), TPB, 0, st>>>(NT, NE, s->t_ed, s->t_tg, s->e_c, s->t_c, s->e_leng,
                                               s->t_area, s->f_qg_ed, s->f_qg_tr);
    ref_apvm_q<<<grid_for(NE), TPB, 0, st>>>(NE, NT, nv, s->e_sh, s->e_v, s->f_qg_tr, s->f_vhx,
                                             s->sc.pvspar, s->o.dt, s->f_q_ed);
  }
  ref_cell<<<grid_for(nv), TPB, 0, st>>>(nv, s->c_nnb, s->c_ed, csg, s->c_area, s->e_leng,
                                         s->e_dleng, u, s->f_uh, h, s->bt, rec, s->f_vhx,
                                         (hc && rec == IMX_REC_MELV) ? s->f_vuh : nullptr, s->f_ketr, s->c_tr,
                                         s->c_hxtr, s->sc_gcoef, s->o.erad, s->o.grav, s->f_divuh, masseq,
                                         s->f_divu, s->f_ke, s->f_ghbk);
  if (cor == IMX_COR_DTRED)
    ref_edge_mom_dtred<<<grid_for(NE), TPB, 0, st>>>(NE, NT, s->e_v, s->e_sh, s->e_leng, s->e_dw, s->e_c, s->e_nr,
                                                     s->f_vhq, s->f_ghbk, egs, hc ? 1 : 0, s->o.erad,
                                                     s->o.testcase <= 1, s->f_grad, s->f_uhq, momeq);
  else if (cor == IMX_COR_PERED)
    ref_edge_mom_pered<<<grid_for(NE), TPB, 0, st>>>(NE, nv, s->e_v, s->e_leng, s->e_dleng, epc,
                                                     s->e_nr, s->c_nnb, s->c_ed, csg, s->c_p,
                                                     s->c_area, s->f_ghbk, s->f_uh, s->f_q_ed, egs,
                                                     hc ? 1.0 : -1.0, s->o.erad, s->o.testcase <= 1,
                                                     s->f_grad, s->f_uhq, momeq);
  else
    ref_edge_mom<<<grid_for(NE), TPB, 0, st>>>(NE, nv, s->e_v, s->e_loc, s->e_leng, s->e_dleng,
                                               s->c_nnb, s->c_ed, csg, ctn, s->e_tk, s->e_tw, s->e_tdl,
                                               hc ? s->hc_ts : s->e_ts, s->f_ghbk,
                                               s->sc_nopv ? u : s->f_uh,             // level model:
                                               s->sc_nopv ? s->f_eta_ed : s->f_q_ed,  // u and eta
                                               cor == IMX_COR_GASS ? s->e_ed3 : nullptr, egs, s->o.erad,
                                               s->o.testcase <= 1, s->f_grad, s->f_uhq, momeq);
  if (s->sc_diff || s->sc_hyper) {
    // diffusion_ed of u (momeq + lapu with K2), then with K4 diffusion_ed of
    // lapu's divergence and vorticity (momeq - lap_lapu) (swm.f90:2575-2590);
    // testcase <= 1 leaves momeq zero
    const int live = s->o.testcase > 1;
    ref_lap<<<grid_for(NE), TPB, 0, st>>>(NE, s->e_v, s->e_sh, ett, egs, s->e_leng, s->e_dleng, s->dif_hx,
                                          s->dif_tr, s->f_divu, s->f_zeta, s->o.erad,
                                          (live && s->sc_diff) ? 1 : 0, s->f_lapu, momeq);
    if (s->sc_hyper) {
      ref_div<<<grid_for(nv), TPB, 0, st>>>(nv, s->c_nnb, s->c_ed, csg, s->c_area, s->e_dleng,
                                            s->f_lapu, s->o.erad, s->f_divl);
      ref_rvort<<<grid_for(NT), TPB, 0, st>>>(NT, s->t_ed, tsg, s->t_area, s->e_leng, s->f_lapu,
                                              s->o.erad, s->f_vortl);
      ref_lap<<<grid_for(NE), TPB, 0, st>>>(NE, s->e_v, s->e_sh, ett, egs, s->e_leng, s->e_dleng,
                                            s->dif_hx, s->dif_tr, s->f_divl, s->f_vortl, s->o.erad,
                                            live ? 2 : 0, s->f_lap2, momeq);
    }
  }
}

static void launch_ref_step(imx_swm *s) {
  const int nv = s->nv, NE = s->NE;
  const double dt = s->o.dt;
  hipStream_t st = s->stream;
  double *m[4], *f[4];
  for (int k = 0; k < 4; ++k) {
    m[k] = s->massf + (size_t)k * nv;
    f[k] = s->momf + (size_t)k * NE;
  }
  launch_ref_tendency(s, s->rh, s->ru, m[0], f[0]);
  for (int k = 0; k < 3; ++k) {
    const int half = (k < 2);
    ref_rk_stage<<<grid_for(NE), TPB, 0, st>>>(NE, s->ru, f[k], dt, half, s->run);
    ref_rk_stage<<<grid_for(nv), TPB, 0, st>>>(nv, s->rh, m[k], dt, half, s->rhn);
    launch_ref_tendency(s, s->rhn, s->run, m[k + 1], f[k + 1]);
  }
  ref_rk_final<<<grid_for(NE), TPB, 0, st>>>(NE, s->ru, f[0], f[1], f[2], f[3], dt);
  ref_rk_final<<<grid_for(nv), TPB, 0, st>>>(nv, s->rh, m[0], m[1], m[2], m[3], dt);
}

// Block-size dispatch of the fused kernels (the autotuner's search space).
#define IMX_BS_SWITCH(bs, CALL) \
  switch (bs) {                 \
    case 64: CALL(64); break;    \
    case 128: CALL(128); break;  \
    case 512: CALL(512); break;  \
    default: CALL(256); break;   \
  }

// One FAST stage: stage<0 evaluates the tendency of (hs, us) into dh_out/du_out.
template <typename T>
static void launch_fast_stage(imx_swm *s, int stage, const T *hs, const T *us, T *hout,
                              T *uout, hipEvent_t *evs) {
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  hipStream_t st = s->stream;
  const T href = (T)s->href;
  if (evs) (void)hipEventRecord(evs[0], st);
  if (s->f_pxhc)  // before fast_tri, whose dtred term reads these depths
    fast_hc_he<T, 256><<<grid_bs(NE, 256), 256, 0, st>>>(s->kremap[2], NE, s->x_e_oth, (const T *)s->x_px_hb,
                                                         s->x_px_hv, hs, href, (T *)s->x_px_he);
#define IMX_TRI(B)                                                                          \
  fast_tri<T, B><<<grid_bs(NT, B), B, 0, st>>>(                                            \
      s->kremap[0], NT, s->x_t_ab, s->x_t_eb, (const T *)s->x_t_ceta, (const T *)s->x_t_kap, \
      (const T *)s->x_t_f, us, hs, href, s->o.testcase >= 2, s->f_lin ? 1 : 0, s->f_nopv ? 1 : 0,  \
      (s->f_diff || s->f_hyper) ? (const T *)s->x_kt : nullptr,                                    \
      (T *)s->x_cz, s->f_gtrsk ? 0 : s->f_gass, (const T *)s->x_g_w, (const T *)s->x_g_a, (T *)s->x_g_E, (T *)s->x_g_H, \
      s->f_gk ? (const T *)s->x_gk_w : nullptr, (T *)s->x_gk_k,                                     \
      s->f_dtred ? (const V4<T> *)s->x_dt_w : nullptr, (V4<T> *)s->x_dt_vhq,                        \
      (s->f_dtred && s->f_pxhc) ? (const T *)s->x_px_he : nullptr,                                  \
      (s->f_gass && !s->f_apvm && !s->f_gtrsk) ? nullptr : (T *)s->q, s->flag)
  IMX_BS_SWITCH(s->kbs[0], IMX_TRI)
#undef IMX_TRI
  if (s->f_hyper) {
    fast_hyper_a<T, 256><<<grid_bs(nv, 256), 256, 0, st>>>(s->kremap[1], nv, s->x_c_nb, s->x_c_bitsA, s->x_c_bitsB,
                                                           (const T *)s->x_c_dl, s->x_c_iaa, (const T *)s->x_kh,
                                                           us, (T *)s->x_hy_kd);
    fast_hyper_b<T, 256><<<grid_bs(NE, 256), 256, 0, st>>>(s->kremap[2], NE, s->x_e_oth, s->x_hy_esh,
                                                           (const T *)s->x_hy_ea, (const T *)s->x_hy_eb,
                                                           (const T *)s->x_hy_kd, (const T *)s->x_cz,
                                                           (T *)s->x_hy_lap);
    fast_hyper_c<T, 256><<<grid_bs(nv, 256), 256, 0, st>>>(
        s->kremap[1], nv, NT, s->x_c_nb, s->x_c_bitsA, s->x_c_bitsB, (const T *)s->x_c_dl, s->x_c_iaa,
        (const T *)s->x_kh, (const T *)s->x_kt, s->x_t_ab, s->x_t_eb, (const T *)s->x_t_ceta,
        (const T *)s->x_hy_lap, (T *)s->x_hy_d, (T *)s->x_hy_z);
  }
  if (s->f_apvm)
    fast_apvm_pre<T, 256><<<grid_bs(nv, 256), 256, 0, st>>>(
        s->kremap[1], nv, NT, s->x_c_nb, s->x_c_bitsA, s->x_c_bitsB, (const V4<T> *)s->x_ap_wc, s->x_ap_tnb,
        (const V4<T> *)s->x_ap_wt, us, (const T *)s->q, (V4<T> *)s->x_ap_v, (V4<T> *)s->x_ap_g);
  if (s->f_gass)  // after fast_apvm_pre, whose vectors correct the rhombus PV
    fast_gass_q<T, 256><<<grid_bs(NE, 256), 256, 0, st>>>(
        s->kremap[2], NE, s->x_g_et, (const T *)s->x_g_E, (const T *)s->x_g_H, s->x_e_oth,
        s->f_apvm ? (const V4<T> *)s->x_ap_v : nullptr, (const V4<T> *)s->x_ap_g, (T)(s->sc.pvspar * s->o.dt),
        s->f_gtrsk ? (const T *)s->q : nullptr, s->f_gqw ? (const T *)s->x_g_qw : nullptr, (T *)s->x_g_q);
  if (evs) (void)hipEventRecord(evs[1], st);
#define IMX_CELL(B)                                                                           \
  fast_cell<T, B, OPT><<<grid_bs(nv, B), B, 0, st>>>(                                        \
      s->kremap[1], nv, s->x_c_nb, s->x_c_bitsA, s->x_c_bitsB, (const T *)s->x_c_R, s->x_c_iaa, \
      (const T *)s->x_c_ika, (const T *)s->x_c_dl, (const T *)s->x_c_kw, (const T *)s->q,       \
      s->f_diff ? (const T *)s->x_kh : nullptr, (const T *)s->x_cz,                              \
      s->f_hyper ? (const T *)s->x_hy_d : nullptr, s->f_hyper ? (const T *)s->x_hy_z : nullptr,    \
      (s->f_diff || s->f_hyper) ? s->x_kh : nullptr,                                               \
      s->f_pered ? (const V4<T> *)s->x_ap_wc : nullptr, s->f_perhx ? (const V4<T> *)s->x_ap_wc : nullptr, \
      s->f_bary ? (const T *)s->x_px_wq : nullptr,                                                 \
      (V4<T> *)s->x_px_ab, (T *)s->x_px_qe,                                                        \
      s->f_apvm ? (const V4<T> *)s->x_ap_v : nullptr, (const V4<T> *)s->x_ap_g, (T)(s->sc.pvspar * s->o.dt), \
      (const T *)s->x_g_q, s->x_g_out, s->f_pxhc ? (const T *)s->x_px_he : nullptr, s->f_glink ? 1 : 0, \
      s->f_gk ? (const T *)s->x_gk_k : nullptr, (T)s->gk_coef, s->f_melv,                          \
      s->f_melv == 2 ? (const V4<T> *)s->x_ap_wc : nullptr, s->f_dtred ? 1 : 0, hs, us, s->href,     \
      (const T *)s->x_gb,                                                                          \
      (T)s->o.grav, (T)(1.0 / s->o.erad), (T *)s->phi, (T *)s->Pown, (T *)s->Poth, stage,       \
      s->o.dt, s->dt_c, (const T *)s->h0, hout, (T *)s->acch, s->dh_out)
  if (s->f_gass && (s->f_melv || s->f_perhx)) {
    constexpr int OPT = 5;
    IMX_BS_SWITCH(s->kbs[1], IMX_CELL)
  } else if (s->f_gass) {
    constexpr int OPT = 3;
    IMX_BS_SWITCH(s->kbs[1], IMX_CELL)
  } else if (s->f_nopv) {
    constexpr int OPT = 4;
    IMX_BS_SWITCH(s->kbs[1], IMX_CELL)
  } else if (s->f_apvm || s->f_bary || s->f_pered || s->f_perhx || s->f_gk || s->f_melv || s->f_dtred) {
    constexpr int OPT = 2;
    IMX_BS_SWITCH(s->kbs[1], IMX_CELL)
  } else if (s->f_diff || s->f_hyper) {
    constexpr int OPT = 1;
    IMX_BS_SWITCH(s->kbs[1], IMX_CELL)
  } else {
    constexpr int OPT = 0;
    IMX_BS_SWITCH(s->kbs[1], IMX_CELL)
  }
#undef IMX_CELL
  if (evs) (void)hipEventRecord(evs[2], st);
#define IMX_EDGE(B)                                                                              \
  fast_edge<T, B><<<grid_bs(NE, B), B, 0, st>>>(                                                \
      s->kremap[2], NE, s->x_e_oth, (const T *)s->x_e_il, (const T *)s->phi, (const T *)s->Pown, \
      (const T *)s->Poth, s->f_pered ? (const V4<T> *)s->x_px_ab : nullptr, (const T *)s->x_px_qe, \
      (const V4<T> *)s->x_px_n, s->x_dt_et, (const T *)s->x_dt_ew, (const V4<T> *)s->x_dt_n,      \
      s->f_dtred ? (const V4<T> *)s->x_dt_vhq : nullptr, s->o.testcase <= 1, stage, (T)s->o.dt,   \
      (const T *)s->dt_e, (const T *)s->u0, uout, (T *)s->accu, (T *)s->du_out)
  IMX_BS_SWITCH(s->kbs[2], IMX_EDGE)
#undef IMX_EDGE
  if (evs) (void)hipEventRecord(evs[3], st);
}

template <typename T>
static void launch_fast_step(imx_swm *s, std::vector<hipEvent_t> *evs) {
  T *h0 = (T *)s->h0, *hA = (T *)s->hA, *hB = (T *)s->hB;
  T *u0 = (T *)s->u0, *uA = (T *)s->uA, *uB = (T *)s->uB;
  const T *hin[4] = {h0, hA, hB, hA};
  const T *uin[4] = {u0, uA, uB, uA};
  T *hout[4] = {hA, hB, hA, h0};
  T *uout[4] = {uA, uB, uA, u0};
  for (int st = 0; st < 4; ++st)
    launch_fast_stage<T>(s, st, hin[st], uin[st], hout[st], uout[st],
                         evs ? &(*evs)[4 * st] : nullptr);
}

static void launch_step(imx_swm *s) {
  if (!s->fast) launch_ref_step(s);
  else if (s->single) launch_fast_step<float>(s, nullptr);
  else launch_fast_step<double>(s, nullptr);
}

void swm_launch_step(imx_swm *s) { launch_step(s); }

// Diagnostics of the cells [c0, c1) and the edge/triangle slots they own.
template <typename T>
static void launch_diag(imx_swm *s, const T *hs, const T *us, double href, int c0, int c1, double *part,
                        double *diag) {
  hipStream_t st = s->stream;
  const int nc = c1 - c0, ne = 3 * nc, nt = 2 * nc;
  const int gv = (nc + TPB - 1) / TPB, ge = (ne + TPB - 1) / TPB, gt = (nt + TPB - 1) / TPB;
  diag_cell<T><<<gv, TPB, 0, st>>>(c0, nc, s->nv, s->c_nnb, s->c_ed, s->sc_hc ? s->hc_c : s->c_ttg, s->c_area,
                                   s->e_dleng, hs, us, href, s->bt, s->o.grav, s->o.erad, part);
  diag_edge<T><<<ge, TPB, 0, st>>>(3 * c0, ne, s->NE, s->e_v, s->e_leng, s->e_dleng, hs, us, href,
                                   part + (size_t)gv * NPART);
  diag_tri<T><<<gt, TPB, 0, st>>>(2 * c0, nt, s->NT, s->t_ed, s->sc_hc ? s->hc_t : s->t_tg, s->t_v, s->t_area, s->t_f,
                                  s->t_trhx, s->e_leng, hs, us, href, s->o.erad, s->f_lin ? 1 : 0,
                                  part + (size_t)(gv + ge) * NPART);
  diag_final<<<1, TPB, 0, st>>>(gv + ge + gt, part, diag);
}

// ------------------------------------------------------------ create -------
// Tables of the fused kernels in the ownership numbering (see fast_cell).
template <typename T>
static int build_fast(imx_swm *s, const imx_mesh_desc &d, const double *trhx, const double *hxtr) {
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  const double a = s->o.erad;
  std::vector<int32_t> c_nb((size_t)MX * nv, 0), bitsA(nv, 0), bitsB(nv, 0);
  std::vector<double> c_R((size_t)MX * nv, 0.0), c_iaa(nv), c_ika(nv);
  // per-slot dual length and K weight of each cell's edges, slot-major so a
  // cell reads them with coalesced loads instead of gathering them by edge
  std::vector<double> c_dl((size_t)MX * nv, 0.0), c_kw((size_t)MX * nv, 0.0);
  for (int co = 0; co < nv; ++co) {
    const int c = cn[co];
    const int n = d.v_nnb[co];
    if (n < MX - 1) return set_error(IMX_ENOTSUP, "fused kernels need cells with 5 or 6 edges");
    int A = n, B = 0;
    double ksum = 0.0;
    for (int j = 0; j < MX; ++j) {
      const int js = (j < n) ? j : 0;  // padded slot of a pentagon repeats slot 0
      const int nbo = d.v_nb[MX * co + js] - 1;
      const int nbi = cn[nbo];
      c_nb[(size_t)j * nv + c] = nbi;
      const int eo = d.v_ed[MX * co + js] - 1;
      const int ei = en[eo];
      const int owner = ei / 3;
      c_dl[(size_t)j * nv + c] = d.edhx_leng[eo];
      c_kw[(size_t)j * nv + c] = 0.25 * d.ed_leng[eo] * d.edhx_leng[eo];
      if (owner != c && owner != nbi) return set_error(IMX_EINVAL, "layout: edge owner is not an end cell");
      if (owner == c) A |= 1 << (BA_OWN + j);
      B |= (ei % 3) << (2 * j);
      const int ti = tn[d.v_tr[MX * co + js] - 1];
      const int town = ti / 2;
      const int nbn = cn[d.v_nb[MX * co + (js + 1) % n] - 1];
      const int ts = (town == c) ? 0 : (town == nbi) ? 1 : (town == nbn) ? 2 : -1;
      if (ts < 0) return set_error(IMX_EINVAL, "layout: triangle owner is not a vertex");
      B |= ts << (BB_TSEL + 2 * j);
      B |= (ti % 2) << (BB_KT + j);
      if (j < n) {
        if (d.hx_ttgout[MX * co + j] < 0) A |= 1 << (BA_TTG + j);
        if (d.hx_tnrccw[MX * co + j] < 0) A |= 1 << (BA_TNR + j);
        c_R[(size_t)j * nv + c] = hxtr[MX * co + j];
        ksum += 0.25 * d.ed_leng[eo] * d.edhx_leng[eo];
      }
    }
    bitsA[c] = A;
    bitsB[c] = B;
    c_iaa[c] = 1.0 / (d.hx_areag[co] * a);
    c_ika[c] = 1.0 / ksum;
  }
  std::vector<int32_t> e_oth(NE, -1);
  std::vector<double> e_il(NE, 0.0);
  for (int eo = 0; eo < d.ne; ++eo) {
    const int e = en[eo];
    const int o = e / 3;
    const int v1 = cn[d.ed_v[2 * eo] - 1], v2 = cn[d.ed_v[2 * eo + 1] - 1];
    if (o == v1) e_oth[e] = v2;
    else e_oth[e] = v1 | (1 << OTH_V2);
    e_il[e] = 1.0 / d.ed_leng[eo];
  }
  std::vector<int32_t> t_ab((size_t)2 * NT, -1), t_eb(NT, 0);
  std::vector<double> t_ceta((size_t)3 * NT, 0.0), t_kap((size_t)3 * NT, 0.0), t_f(NT, 0.0);
  for (int ko = 0; ko < d.nt; ++ko) {
    const int t = tn[ko];
    const int o = t / 2;
    int r = -1;  // rotate the CCW vertex list so the owner comes first
    for (int l = 0; l < 3; ++l)
      if (cn[d.tr_v[3 * ko + l] - 1] == o) r = l;
    if (r < 0) return set_error(IMX_EINVAL, "layout: triangle owner not found");
    int vv[3];
    for (int m = 0; m < 3; ++m) vv[m] = cn[d.tr_v[3 * ko + (r + m) % 3] - 1];
    t_ab[t] = vv[1];
    t_ab[(size_t)NT + t] = vv[2];
    int eb = 0;
    for (int m = 0; m < 3; ++m) {
      const int lr = (r + m) % 3;  // reference edge l runs tr%v(l) -> tr%v(l+1)
      const int eo = d.tr_ed[3 * ko + lr] - 1;
      const int ei = en[eo];
      const int first = vv[m], second = vv[(m + 1) % 3];
      const int owner = ei / 3;
      if (owner != first && owner != second) return set_error(IMX_EINVAL, "layout: bad triangle edge");
      eb |= ((owner == first) ? 1 : 0) << (3 * m);
      eb |= (ei % 3) << (3 * m + 1);
      t_ceta[(size_t)m * NT + t] = d.ed_leng[eo] * (double)d.tr_tg[3 * ko + lr] / (d.tr_areag[ko] * a);
      t_kap[(size_t)m * NT + t] = trhx[3 * ko + (r + m) % 3] / d.tr_areag[ko];
    }
    t_eb[t] = eb;
    t_f[t] = 2.0 * s->o.omega * std::sin(d.tr_clat[ko]);
  }
  Alloc &A = s->A;
  s->x_c_nb = A.put(c_nb);
  s->x_c_bitsA = A.put(bitsA);
  s->x_c_bitsB = A.put(bitsB);
  s->x_c_R = A.put(cvt<T>(c_R));
  s->x_c_iaa = A.put(c_iaa);
  s->x_c_ika = A.put(cvt<T>(c_ika));
  s->x_e_oth = A.put(e_oth);
  s->x_c_dl = A.put(cvt<T>(c_dl));
  s->x_c_kw = A.put(cvt<T>(c_kw));
  s->x_e_il = A.put(cvt<T>(e_il));
  s->x_t_ab = A.put(t_ab);
  s->x_t_eb = A.put(t_eb);
  s->x_t_ceta = A.put(cvt<T>(t_ceta));
  s->x_t_kap = A.put(cvt<T>(t_kap));
  s->x_t_f = A.put(cvt<T>(t_f));
  s->h0 = A.get<T>(nv);
  s->hA = A.get<T>(nv);
  s->hB = A.get<T>(nv);
  s->u0 = A.get<T>(NE);
  s->uA = A.get<T>(NE);
  s->uB = A.get<T>(NE);
  s->acch = A.get<T>(nv);
  s->accu = A.get<T>(NE);
  s->q = A.get<T>(NT);
  s->phi = A.get<T>(nv);
  s->Pown = A.get<T>(NE);
  s->Poth = A.get<T>(NE);
  s->dh_out = A.get<double>(nv);
  s->du_out = A.get<T>(NE);
  if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_create (fast tables)");
  return IMX_OK;
}

// Geometry and buffers of the scheme options and the LTE analysis, built
// once per handle: internal numbering, points component-major.
static int scheme_tables(imx_swm *s, const imx_mesh_desc *d) {
  if (s->e_nrs) return IMX_OK;
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  Alloc &A = s->A;
  // geometry of the options, internal numbering, points component-major
  std::vector<int8_t> nrs(NE, 1);
  std::vector<double> ec((size_t)3 * NE, 0.0), tc((size_t)3 * NT, 0.0), cp((size_t)3 * nv, 0.0);
  for (int eo = 0; eo < d->ne; ++eo) {
    const int e = en[eo];
    const double *a = d->v_p + 3 * (d->ed_v[2 * eo] - 1), *b = d->v_p + 3 * (d->ed_v[2 * eo + 1] - 1);
    double nr[3];
    cross3(a, b, nr);  // ed%nr up to its positive norm (smeshpack.f90:725-726)
    const double *t1 = d->tr_c + 3 * (d->ed_sh[2 * eo] - 1), *t2 = d->tr_c + 3 * (d->ed_sh[2 * eo + 1] - 1);
    double dp = 0.0;
    for (int q = 0; q < 3; ++q) dp = dp + nr[q] * (t2[q] - t1[q]);
    nrs[e] = (int8_t)(std::signbit(dp) ? -1 : 1);
    for (int q = 0; q < 3; ++q) ec[(size_t)q * NE + e] = d->ed_c[3 * eo + q];
  }
  for (int ko = 0; ko < d->nt; ++ko)
    for (int q = 0; q < 3; ++q) tc[(size_t)q * NT + tn[ko]] = d->tr_c[3 * ko + q];
  for (int co = 0; co < nv; ++co)
    for (int q = 0; q < 3; ++q) cp[(size_t)q * nv + cn[co]] = d->v_p[3 * co + q];
  s->e_nrs = A.put(nrs);
  s->e_c = A.put(ec);
  s->t_c = A.put(tc);
  s->c_p = A.put(cp);
  s->f_qg_ed = A.get<double>(NE);
  s->f_qg_tr = A.get<double>((size_t)3 * NT);
  s->f_vhx = A.get<double>((size_t)3 * nv);
  s->f_lapu = A.get<double>(NE);
  s->f_lap2 = A.get<double>(NE);
  s->f_divl = A.get<double>(nv);
  s->f_vortl = A.get<double>(NT);
  s->dif_hx = A.get<double>(nv);
  s->dif_tr = A.get<double>(NT);
  s->t_w = A.get<double>((size_t)3 * NT);
  s->e_dw = A.get<double>((size_t)2 * NE);
  s->e_nr = A.get<double>((size_t)3 * NE);
  s->e_ed3 = A.get<int32_t>((size_t)2 * MX * NE);
  s->hc_c = A.get<int8_t>((size_t)MX * nv);
  s->hc_t = A.get<int8_t>((size_t)3 * NT);
  s->hc_g = A.get<int8_t>(NE);
  s->hc_tt = A.get<int8_t>(NE);
  s->hc_pc = A.get<double>((size_t)3 * NE);
  s->hc_hb = A.get<double>((size_t)3 * NE);
  s->hc_hv = A.get<int32_t>((size_t)3 * NE);
  s->t_ka = A.get<double>((size_t)3 * NT);
  s->c_tr = A.get<int32_t>((size_t)MX * nv);
  s->c_hxtr = A.get<double>((size_t)MX * nv);
  s->f_ketr = A.get<double>(NT);
  s->f_eta_ed = A.get<double>(NE);
  s->f_vhq = A.get<double>((size_t)3 * NT);
  s->f_vuh = A.get<double>((size_t)3 * nv);
  if (A.err != hipSuccess) return hip_fail(A.err, "scheme tables");
  return IMX_OK;
}

// A component-major [3][N] device vector field (kind 0 cells, 2 triangles)
// to caller order, x y z per entry.
static int get_vector(imx_swm *s, const double *src, int kind, double *out) {
  const int n = kind == 0 ? s->nv : s->nt, ni = kind == 0 ? s->nv : s->NT;
  const std::vector<int32_t> &perm = kind == 0 ? s->lay.cell_new : s->lay.tri_new;
  std::vector<double> b((size_t)3 * ni);
  HIPTRY(hipMemcpy(b.data(), src, b.size() * sizeof(double), hipMemcpyDeviceToHost), "get_field");
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < 3; ++q) out[3 * i + q] = b[(size_t)q * ni + perm[i]];
  return IMX_OK;
}

// K2 diffusion on the fused kernels (Klemp form, HTC, TRiSK): the normal
// part grad(K div u) enters the cell potential phi, and the tangential part
// -grad_tg(K zeta) enters the owner's partial P of each edge as
// fz (Kz[t_i] - Kz[t_{i-1}]) with fz = e (+-1) |e| / (|e*| a), where e is the
// sign of ed%nr.(t2 - t1) and +-1 says whether ring triangle i is sh(2);
// the sign goes into bits BA_FZ of the cell's bitsA.
// fast_edge then forms -(dphi + P) / |e|, i.e.
// + (Kd(v2) - Kd(v1)) / (|e| a) - e (Kz(sh2) - Kz(sh1)) / (|e*| a).
template <typename T>
static int fast_diffusion(imx_swm *s, const imx_mesh_desc &d, const imx_scheme &sc, bool bary) {
  s->f_diff = false;
  s->f_hyper = false;
  if (sc.k2 == 0.0 && sc.k4 == 0.0) return IMX_OK;
  const int nv = s->nv, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  // HC orients grad_ed_tg by edhx%tg (the hc_tt table) instead of ed%nr
  std::vector<int8_t> htt;
  if (s->sc_hc) {
    htt.resize(s->NE);
    HIPTRY(hipMemcpy(htt.data(), s->hc_tt, htt.size(), hipMemcpyDeviceToHost), "set_scheme");
  }
  std::vector<double> trhx_v, hxtr_v, trskw_v;
  const double *trhx = d.tr_trhx_area;
  if (!trhx) {
    trhx_v.assign((size_t)3 * d.nt, 0.0);
    hxtr_v.assign((size_t)MX * d.nv, 0.0);
    trskw_v.assign((size_t)MX * MX * d.nv, 0.0);
    gridprop(d, trhx_v.data(), hxtr_v.data(), trskw_v.data());
    trhx = trhx_v.data();
  }
  std::vector<double> hx, tr;
  // with K4 > 0 both terms take the square roots of the hyperdiffusion
  // coefficients, as the reference sets them up
  diffusion_coef(d, trhx, bary, sc.k2, sc.diffus, sc.k4, sc.hyperdiffus, hx, tr);
  std::vector<double> kh(nv, 0.0), kt(NT, 0.0);
  for (int co = 0; co < nv; ++co) kh[cn[co]] = hx[co];
  for (int ko = 0; ko < d.nt; ++ko) kt[tn[ko]] = tr[ko];
  std::vector<int32_t> bitsA(nv);
  HIPTRY(hipMemcpy(bitsA.data(), s->x_c_bitsA, nv * sizeof(int32_t), hipMemcpyDeviceToHost), "set_scheme");
  for (int co = 0; co < nv; ++co) {
    const int c = cn[co], n = d.v_nnb[co];
    bitsA[c] &= ~(((1 << MX) - 1) << BA_FZ);
    for (int j = 0; j < n; ++j) {
      const int eo = d.v_ed[MX * co + j] - 1;
      if (en[eo] / 3 != c) continue;  // the owner cell writes the edge's P
      const int ti = d.v_tr[MX * co + j] - 1, tp = d.v_tr[MX * co + (j + n - 1) % n] - 1;
      const int s1 = d.ed_sh[2 * eo] - 1, s2 = d.ed_sh[2 * eo + 1] - 1;
      double side;
      if (ti == s2 && tp == s1) side = 1.0;
      else if (ti == s1 && tp == s2) side = -1.0;
      else return set_error(IMX_EINVAL, "imx_swm_set_scheme: cell ring triangles do not flank the edge");
      const double *pa = d.v_p + 3 * (d.ed_v[2 * eo] - 1), *pb = d.v_p + 3 * (d.ed_v[2 * eo + 1] - 1);
      double nr[3];
      cross3(pa, pb, nr);
      const double *t1 = d.tr_c + 3 * s1, *t2 = d.tr_c + 3 * s2;
      double dp = 0.0;
      for (int q = 0; q < 3; ++q) dp = dp + nr[q] * (t2[q] - t1[q]);
      const double e = s->sc_hc ? (double)htt[en[eo]] : (std::signbit(dp) ? -1.0 : 1.0);
      if (e * side < 0.0) bitsA[c] |= 1 << (BA_FZ + j);
    }
  }
  Alloc &A = s->A;
  if (!s->x_kh) {
    s->x_kh = A.get<T>(nv);
    s->x_kt = A.get<T>(NT);
    s->x_cz = A.get<T>(NT);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
  }
  const std::vector<T> khT = cvt<T>(kh), ktT = cvt<T>(kt);
  HIPTRY(hipMemcpy(s->x_kh, khT.data(), khT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_kt, ktT.data(), ktT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_c_bitsA, bitsA.data(), nv * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  if (sc.k4 > 0.0) {
    // lapu at edges: the triangles either side and the two factors
    const int NE = s->NE;
    const double a = s->o.erad;
    std::vector<int32_t> esh((size_t)2 * NE, 0);
    std::vector<double> ea(NE, 0.0), eb(NE, 0.0);
    for (int eo = 0; eo < d.ne; ++eo) {
      const int e = en[eo];
      const int s1 = d.ed_sh[2 * eo] - 1, s2 = d.ed_sh[2 * eo + 1] - 1;
      esh[e] = tn[s1];
      esh[(size_t)NE + e] = tn[s2];
      double sg;
      if (s->sc_hc) {
        sg = (double)htt[e];
      } else {
        const double *pa = d.v_p + 3 * (d.ed_v[2 * eo] - 1), *pb = d.v_p + 3 * (d.ed_v[2 * eo + 1] - 1);
        double nr[3];
        cross3(pa, pb, nr);
        double dp = 0.0;
        for (int q = 0; q < 3; ++q) dp = dp + nr[q] * (d.tr_c[3 * s2 + q] - d.tr_c[3 * s1 + q]);
        sg = std::signbit(dp) ? -1.0 : 1.0;
      }
      ea[e] = 1.0 / (d.ed_leng[eo] * a);
      eb[e] = sg / (d.edhx_leng[eo] * a);
    }
    if (!s->x_hy_esh) {
      s->x_hy_esh = A.get<int32_t>((size_t)2 * NE);
      s->x_hy_ea = A.get<T>(NE);
      s->x_hy_eb = A.get<T>(NE);
      s->x_hy_kd = A.get<T>(nv);
      s->x_hy_lap = A.get<T>(NE);
      s->x_hy_d = A.get<T>(nv);
      s->x_hy_z = A.get<T>(NT);
      if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
    }
    const std::vector<T> eaT = cvt<T>(ea), ebT = cvt<T>(eb);
    HIPTRY(hipMemcpy(s->x_hy_esh, esh.data(), esh.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
    HIPTRY(hipMemcpy(s->x_hy_ea, eaT.data(), eaT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
    HIPTRY(hipMemcpy(s->x_hy_eb, ebT.data(), ebT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  }
  s->f_diff = sc.k2 > 0.0;
  s->f_hyper = sc.k4 > 0.0;
  return IMX_OK;
}

// HC sign tables and Voronoi-midpoint points (the REF kernels' HC branches
// and, for fused handles, the diagnostics), built from the geometry.
// hbary: barycentric h at the dual edge midpoints (bary interpolation);
// tg: e_nr takes edhx%tg (pered and dtred Coriolis)
static int hc_tables(imx_swm *s, const imx_mesh_desc *d, bool hbary, bool tg) {
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  // HC signs from the geometry: s_e = sign(ed%tg . edhx%nr) turns the
  // HTC indicators into hx%nr = ttgout s_e and the vorticity signs
  // tr%tg s_e (smeshpack.f90:1085-1116); grad_ed and grad_ed_tg orient
  // by edhx%nr and edhx%tg (swm_operators.f90:531-534, :755-760)
  std::vector<int8_t> sgn(d->ne), hcc((size_t)MX * nv, 1), hct((size_t)3 * NT, 1), hg(NE, 1), htt(NE, 1);
  std::vector<double> pc((size_t)3 * NE, 0.0), tgv((size_t)3 * NE, 0.0), hb((size_t)3 * NE, 0.0);
  std::vector<int32_t> hv((size_t)3 * NE, 0);
  for (int eo = 0; eo < d->ne; ++eo) {
    const int e = en[eo];
    const double *v1 = d->v_p + 3 * (d->ed_v[2 * eo] - 1), *v2 = d->v_p + 3 * (d->ed_v[2 * eo + 1] - 1);
    const double *t1 = d->tr_c + 3 * (d->ed_sh[2 * eo] - 1), *t2 = d->tr_c + 3 * (d->ed_sh[2 * eo + 1] - 1);
    const double *hxc = d->edhx_c + 3 * eo;
    double enr[3], etg[3], hnr[3], htg[3];
    cross3(v1, v2, enr);
    double n = norm3(enr);
    for (int q = 0; q < 3; ++q) enr[q] = enr[q] / n;
    cross3(enr, d->ed_c + 3 * eo, etg);
    cross3(t1, t2, hnr);
    n = norm3(hnr);
    for (int q = 0; q < 3; ++q) hnr[q] = hnr[q] / n;
    cross3(hnr, hxc, htg);
    n = norm3(htg);
    for (int q = 0; q < 3; ++q) htg[q] = htg[q] / n;
    sgn[eo] = std::signbit(dot3(etg, hnr)) ? -1 : 1;
    double dv[3] = {v2[0] - v1[0], v2[1] - v1[1], v2[2] - v1[2]};
    double dt3[3] = {t2[0] - t1[0], t2[1] - t1[1], t2[2] - t1[2]};
    hg[e] = std::signbit(dot3(hnr, dv)) ? -1 : 1;
    htt[e] = std::signbit(dot3(htg, dt3)) ? -1 : 1;
    for (int q = 0; q < 3; ++q) {
      pc[(size_t)q * NE + e] = hxc[q];
      tgv[(size_t)q * NE + e] = htg[q];
    }
    if (hbary) {  // h at the dual edge midpoint (initialize_gridprop, swm.f90:1007-1018)
      const int kt = insidetrmesh(*d, hxc, d->ed_sh[2 * eo] - 1) ? d->ed_sh[2 * eo] - 1 : d->ed_sh[2 * eo + 1] - 1;
      double b[3];
      bary_at(*d, hxc, kt, b);
      for (int i = 0; i < 3; ++i) {
        hb[(size_t)i * NE + e] = b[i];
        hv[(size_t)i * NE + e] = cn[d->tr_v[3 * kt + i] - 1];
      }
    }
  }
  for (int co = 0; co < nv; ++co)
    for (int j = 0; j < d->v_nnb[co]; ++j)
      hcc[(size_t)j * nv + cn[co]] = (int8_t)((d->hx_ttgout[MX * co + j] < 0 ? -1 : 1) * sgn[d->v_ed[MX * co + j] - 1]);
  for (int ko = 0; ko < d->nt; ++ko)
    for (int l = 0; l < 3; ++l)
      hct[(size_t)l * NT + tn[ko]] = (int8_t)((d->tr_tg[3 * ko + l] < 0 ? -1 : 1) * sgn[d->tr_ed[3 * ko + l] - 1]);
  // the HC sign of each TRiSK term: hx%nr(j) hx%nr(ei) of its cell
  std::vector<int8_t> hts((size_t)2 * MX * NE, 1);
  for (int eo = 0; eo < d->ne; ++eo)
    for (int ic = 0; ic < 2; ++ic) {
      const int co = d->ed_v[2 * eo + ic] - 1, c = cn[co];
      int ei = 0;
      for (int j = 0; j < d->v_nnb[co]; ++j)
        if (d->v_ed[MX * co + j] == eo + 1) ei = j;
      for (int j = 0; j < d->v_nnb[co]; ++j)
        hts[((size_t)ic * MX + j) * NE + en[eo]] = (int8_t)(hcc[(size_t)j * nv + c] * hcc[(size_t)ei * nv + c]);
    }
  if (s->hc_ts)  // the REF TRiSK term table (fused handles have none)
    HIPTRY(hipMemcpy(s->hc_ts, hts.data(), hts.size(), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->hc_c, hcc.data(), hcc.size(), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->hc_t, hct.data(), hct.size(), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->hc_g, hg.data(), hg.size(), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->hc_tt, htt.data(), htt.size(), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->hc_pc, pc.data(), pc.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->hc_hb, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->hc_hv, hv.data(), hv.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  // pered and dtred on HC take the edhx%tg component
  if (tg)
    HIPTRY(hipMemcpy(s->e_nr, tgv.data(), tgv.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  return IMX_OK;
}

// The staggering's signs in the fused tables: the flux and TRiSK sign bits
// of each cell slot (hx%ttgout / hx%tnrccw with HTC; hx%nr = ttgout s_e for
// both with HC, whose TRiSK term sign is hx%nr(j) hx%nr(i)), the vorticity
// coefficients (tr%tg, or tr%tg s_e), and the orientation of each edge's
// potential difference (HC: sign of edhx%nr . (v2 - v1), applied by swapping
// the edge's end cells). HC reads the tables hc_tables built.
template <typename T>
static int fast_stagger(imx_swm *s, const imx_mesh_desc &d, bool hc) {
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  const double a = s->o.erad;
  std::vector<int8_t> hcc, hct, hg;
  if (hc) {
    hcc.resize((size_t)MX * nv);
    hct.resize((size_t)3 * NT);
    hg.resize(NE);
    HIPTRY(hipMemcpy(hcc.data(), s->hc_c, hcc.size(), hipMemcpyDeviceToHost), "set_scheme");
    HIPTRY(hipMemcpy(hct.data(), s->hc_t, hct.size(), hipMemcpyDeviceToHost), "set_scheme");
    HIPTRY(hipMemcpy(hg.data(), s->hc_g, hg.size(), hipMemcpyDeviceToHost), "set_scheme");
  }
  std::vector<int32_t> bitsA(nv), e_oth(NE, -1);
  HIPTRY(hipMemcpy(bitsA.data(), s->x_c_bitsA, nv * sizeof(int32_t), hipMemcpyDeviceToHost), "set_scheme");
  for (int co = 0; co < nv; ++co) {
    const int c = cn[co];
    bitsA[c] &= ~((((1 << MX) - 1) << BA_TTG) | (((1 << MX) - 1) << BA_TNR));
    for (int j = 0; j < d.v_nnb[co]; ++j) {
      const int sf = hc ? hcc[(size_t)j * nv + c] : d.hx_ttgout[MX * co + j];
      const int sn = hc ? hcc[(size_t)j * nv + c] : d.hx_tnrccw[MX * co + j];
      if (sf < 0) bitsA[c] |= 1 << (BA_TTG + j);
      if (sn < 0) bitsA[c] |= 1 << (BA_TNR + j);
    }
  }
  for (int eo = 0; eo < d.ne; ++eo) {
    const int e = en[eo];
    const int o = e / 3;
    int v1 = cn[d.ed_v[2 * eo] - 1], v2 = cn[d.ed_v[2 * eo + 1] - 1];
    if (hc && hg[e] < 0) std::swap(v1, v2);
    e_oth[e] = (o == v1) ? v2 : (v1 | (1 << OTH_V2));
  }
  std::vector<double> t_ceta((size_t)3 * NT, 0.0);
  for (int ko = 0; ko < d.nt; ++ko) {
    const int t = tn[ko];
    int r = -1;
    for (int l = 0; l < 3; ++l)
      if (cn[d.tr_v[3 * ko + l] - 1] == t / 2) r = l;
    if (r < 0) return set_error(IMX_EINVAL, "imx_swm_set_scheme: triangle owner not found");
    for (int m = 0; m < 3; ++m) {
      const int lr = (r + m) % 3;
      const int eo = d.tr_ed[3 * ko + lr] - 1;
      const double sg = hc ? (double)hct[(size_t)lr * NT + t] : (double)d.tr_tg[3 * ko + lr];
      t_ceta[(size_t)m * NT + t] = d.ed_leng[eo] * sg / (d.tr_areag[ko] * a);
    }
  }
  const std::vector<T> tcT = cvt<T>(t_ceta);
  HIPTRY(hipMemcpy(s->x_c_bitsA, bitsA.data(), nv * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_e_oth, e_oth.data(), NE * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_t_ceta, tcT.data(), tcT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  return IMX_OK;
}

static void put_projected(double *dst, const double w[3], const double *p) {
  const double wp = w[0] * p[0] + w[1] * p[1] + w[2] * p[2];
  for (int q = 0; q < 3; ++q) dst[q] = w[q] - wp * p[q];
}

// The cell Perot weights of the fused kernels (x_ap_wc, [MX][nv] records):
// v = sum_j wc_j u_j is vector_edhx2hx_perot of u at the cell node,
// projected, with hx%nr and the Voronoi midpoints under HC. APVM and pxt16
// both read them.
template <typename T>
static int fast_cell_weights(imx_swm *s, const imx_mesh_desc &d) {
  const int nv = s->nv;
  const std::vector<int32_t> &cn = s->lay.cell_new;
  std::vector<double> wc((size_t)4 * MX * nv, 0.0);
  // HC: hx%nr signs and the Voronoi edge midpoints
  std::vector<int8_t> hcc;
  if (s->sc_hc) {
    hcc.resize((size_t)MX * nv);
    HIPTRY(hipMemcpy(hcc.data(), s->hc_c, hcc.size(), hipMemcpyDeviceToHost), "set_scheme");
  }
  for (int co = 0; co < nv; ++co) {
    const double *p = d.v_p + 3 * co;
    for (int j = 0; j < d.v_nnb[co]; ++j) {
      const int eo = d.v_ed[MX * co + j] - 1;
      // ttgout u (ed%c - v%p) edhx%leng / hx%areag (swm_operators.f90:474-505)
      const double sg = s->sc_hc ? (double)hcc[(size_t)j * nv + cn[co]] : (double)d.hx_ttgout[MX * co + j];
      const double *pe = s->sc_hc ? d.edhx_c + 3 * eo : d.ed_c + 3 * eo;
      const double f = sg * d.edhx_leng[eo] / d.hx_areag[co];
      double w[3];
      for (int q = 0; q < 3; ++q) w[q] = f * (pe[q] - p[q]);
      put_projected(&wc[4 * ((size_t)j * nv + cn[co])], w, p);
    }
  }
  if (!s->x_ap_wc) {
    s->x_ap_wc = s->A.get<T>((size_t)4 * MX * nv);
    if (s->A.err != hipSuccess) return hip_fail(s->A.err, "imx_swm_set_scheme");
  }
  const std::vector<T> wcT = cvt<T>(wc);
  HIPTRY(hipMemcpy(s->x_ap_wc, wcT.data(), wcT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  return IMX_OK;
}

// pxt16's three methods on the fused kernels, each on its own (bary: the
// scalar interpolation; pered: the Coriolis term; perhx: the kinetic
// energy). bary: h at triangles from the barycentric coordinates of the
// circumcentre (t_kap), q at edges weighted by the distances to the two
// circumcentres (HC: the plain average) and, with HC, h at the Voronoi edge
// midpoints (fast_hc_he); switching it off restores the kite-area t_kap.
// pered: its normals and per-stage buffers. pered and perhx: the cell's
// Perot weights.
template <typename T>
static int fast_px(imx_swm *s, const imx_mesh_desc &d, bool bary, bool pered, bool perhx) {
  const bool px = bary;
  s->f_pered = false;
  s->f_perhx = false;
  if (!bary && !s->f_bary) {
    s->f_pxhc = false;
    if (!pered && !perhx) return IMX_OK;
  }
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  std::vector<double> trhx_v, hxtr_v, trskw_v;
  const double *trhx = d.tr_trhx_area;
  if (!px && s->f_bary && !trhx) {  // restoring the kite weights
    trhx_v.assign((size_t)3 * d.nt, 0.0);
    hxtr_v.assign((size_t)MX * d.nv, 0.0);
    trskw_v.assign((size_t)MX * MX * d.nv, 0.0);
    gridprop(d, trhx_v.data(), hxtr_v.data(), trskw_v.data());
    trhx = trhx_v.data();
  }
  std::vector<double> tk((size_t)3 * NT, 0.0);
  for (int ko = 0; ko < d.nt && (bary || s->f_bary); ++ko) {
    const int t = tn[ko];
    int r = -1;
    for (int l = 0; l < 3; ++l)
      if (cn[d.tr_v[3 * ko + l] - 1] == t / 2) r = l;
    if (r < 0) return set_error(IMX_EINVAL, "imx_swm_set_scheme: triangle owner not found");
    double b[3];
    if (px) bary_circumcentre(d, ko, b);
    for (int m = 0; m < 3; ++m)
      tk[(size_t)m * NT + t] = px ? b[(r + m) % 3] : trhx[3 * ko + (r + m) % 3] / d.tr_areag[ko];
  }
  if (bary || s->f_bary) {
    const std::vector<T> tkT = cvt<T>(tk);
    HIPTRY(hipMemcpy(s->x_t_kap, tkT.data(), tkT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  }
  s->f_bary = bary;
  s->f_pxhc = false;
  if (!bary && !pered && !perhx) return IMX_OK;
  if (pered || perhx) {
    const int rc = fast_cell_weights<T>(s, d);
    if (rc) return rc;
  }
  // q_e = (d2 q(sh1) + d1 q(sh2)) / (d1 + d2), d_k the arc from ed%c to the
  // circumcentre of sh(k), as weights of ring triangles j-1 and j; with HC
  // the plain average (scalar_tr2ed, swm_operators.f90:152-168)
  const bool hc = s->sc_hc;
  std::vector<double> wq((size_t)2 * MX * nv, 0.0), pn((size_t)4 * NE, 0.0);
  for (int co = 0; co < nv; ++co) {
    const int c = cn[co], n = d.v_nnb[co];
    for (int j = 0; j < n && bary; ++j) {
      if (hc) {
        wq[(size_t)j * nv + c] = 0.5;
        wq[(size_t)(MX + j) * nv + c] = 0.5;
        continue;
      }
      const int eo = d.v_ed[MX * co + j] - 1;
      const int tp = d.v_tr[MX * co + (j + n - 1) % n] - 1;
      const int s1 = d.ed_sh[2 * eo] - 1;
      const double d1 = arclen(d.ed_c + 3 * eo, d.tr_c + 3 * s1);
      const double d2 = arclen(d.ed_c + 3 * eo, d.tr_c + 3 * (d.ed_sh[2 * eo + 1] - 1));
      const double w1 = d2 / (d1 + d2), w2 = d1 / (d1 + d2);  // of sh1, sh2
      wq[(size_t)j * nv + c] = (tp == s1) ? w1 : w2;
      wq[(size_t)(MX + j) * nv + c] = (tp == s1) ? w2 : w1;
    }
  }
  // pered: HTC, minus the ed%nr component of the tangent-plane average at
  // ed%c; HC, the edhx%tg component of the average at edhx%c
  // (coriolis_ed, swm_operators.f90:1330-1376)
  const bool hcb = hc && bary;
  std::vector<double> hb(hcb ? (size_t)3 * NE : 0, 0.0);
  std::vector<int32_t> hv(hcb ? (size_t)3 * NE : 0, 0);
  for (int eo = 0; eo < d.ne && (pered || hcb); ++eo) {
    const int e = en[eo];
    if (hc) {
      const double *t1 = d.tr_c + 3 * (d.ed_sh[2 * eo] - 1), *t2 = d.tr_c + 3 * (d.ed_sh[2 * eo + 1] - 1);
      const double *hxc = d.edhx_c + 3 * eo;
      double hnr[3], htg[3];
      cross3(t1, t2, hnr);
      double nn = norm3(hnr);
      for (int q = 0; q < 3; ++q) hnr[q] = hnr[q] / nn;
      cross3(hnr, hxc, htg);
      nn = norm3(htg);
      for (int q = 0; q < 3; ++q) htg[q] = htg[q] / nn;
      put_projected(&pn[(size_t)4 * e], htg, hxc);
      if (!hcb) continue;
      // h at the dual edge midpoint: barycentric in the triangle that holds
      // it (initialize_gridprop, swm.f90:1007-1018)
      const int kt = insidetrmesh(d, hxc, d.ed_sh[2 * eo] - 1) ? d.ed_sh[2 * eo] - 1 : d.ed_sh[2 * eo + 1] - 1;
      double b[3];
      bary_at(d, hxc, kt, b);
      for (int i = 0; i < 3; ++i) {
        hb[(size_t)i * NE + e] = b[i];
        hv[(size_t)i * NE + e] = cn[d.tr_v[3 * kt + i] - 1];
      }
    } else {
      double nr[3];
      cross3(d.v_p + 3 * (d.ed_v[2 * eo] - 1), d.v_p + 3 * (d.ed_v[2 * eo + 1] - 1), nr);
      const double nn = norm3(nr);
      for (int q = 0; q < 3; ++q) nr[q] = -nr[q] / nn;
      put_projected(&pn[(size_t)4 * e], nr, d.ed_c + 3 * eo);
    }
  }
  Alloc &A = s->A;
  if (!s->x_px_wq) {
    s->x_px_wq = A.get<T>((size_t)2 * MX * nv);
    s->x_px_n = A.get<T>((size_t)4 * NE);
    s->x_px_ab = A.get<T>((size_t)4 * 2 * nv);
    s->x_px_qe = A.get<T>(NE);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
  }
  if (hcb && !s->x_px_hb) {
    s->x_px_hb = A.get<T>((size_t)3 * NE);
    s->x_px_hv = A.get<int32_t>((size_t)3 * NE);
    s->x_px_he = A.get<T>(NE);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
  }
  const std::vector<T> wqT = cvt<T>(wq), pnT = cvt<T>(pn);
  HIPTRY(hipMemcpy(s->x_px_wq, wqT.data(), wqT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_px_n, pnT.data(), pnT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  if (hcb) {
    const std::vector<T> hbT = cvt<T>(hb);
    HIPTRY(hipMemcpy(s->x_px_hb, hbT.data(), hbT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
    HIPTRY(hipMemcpy(s->x_px_hv, hv.data(), hv.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  }
  s->f_pxhc = hcb;
  s->f_pered = pered;
  s->f_perhx = perhx;
  return IMX_OK;
}

// APVM on the fused kernels: the geometry fast_apvm_pre reads (see there),
// in the ownership numbering, typed T.
template <typename T>
static int fast_apvm(imx_swm *s, const imx_mesh_desc &d, const imx_scheme &sc) {
  if (!sc.apvm) {
    s->f_apvm = false;
    return IMX_OK;
  }
  const int nv = s->nv, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &tn = s->lay.tri_new;
  const double a = s->o.erad;
  // weights as (x, y, z, 0) records: wt [3][NT]
  std::vector<double> wt((size_t)4 * 3 * NT, 0.0);
  std::vector<int32_t> tnb((size_t)3 * NT, -1);
  int rc = fast_cell_weights<T>(s, d);
  if (rc) return rc;
  for (int ko = 0; ko < d.nt; ++ko) {
    const int t = tn[ko];
    const double *p = d.tr_c + 3 * ko;
    int r = -1;  // slot m of the fused kernels is reference edge (r + m) % 3
    for (int l = 0; l < 3; ++l)
      if (cn[d.tr_v[3 * ko + l] - 1] == t / 2) r = l;
    if (r < 0) return set_error(IMX_EINVAL, "imx_swm_set_scheme: triangle owner not found");
    for (int m = 0; m < 3; ++m) {
      const int lr = (r + m) % 3;
      const int eo = d.tr_ed[3 * ko + lr] - 1;
      const int s1 = d.ed_sh[2 * eo] - 1, s2 = d.ed_sh[2 * eo + 1] - 1;
      if (s1 != ko && s2 != ko) return set_error(IMX_EINVAL, "imx_swm_set_scheme: triangle not on its edge");
      const double *pa = d.v_p + 3 * (d.ed_v[2 * eo] - 1), *pb = d.v_p + 3 * (d.ed_v[2 * eo + 1] - 1);
      double nr[3];
      cross3(pa, pb, nr);
      double dp = 0.0;
      for (int q = 0; q < 3; ++q) dp = dp + nr[q] * (d.tr_c[3 * s2 + q] - d.tr_c[3 * s1 + q]);
      const double nrs = std::signbit(dp) ? -1.0 : 1.0;
      const double side = (ko == s1) ? 1.0 : -1.0;  // q(across) - q(t) = +-(q(sh2) - q(sh1))
      tnb[(size_t)m * NT + t] = tn[ko == s1 ? s2 : s1];
      // grad q at the edge: ed%nr-sign (q(sh2) - q(sh1)) / (|e*| a); its Perot
      // term: -tr%tg |e| (ed%c - tr%c) / tr%areag (swm_operators.f90:1427-1462)
      const double f = -(double)d.tr_tg[3 * ko + lr] * d.ed_leng[eo] / d.tr_areag[ko] * nrs /
                       (d.edhx_leng[eo] * a) * side;
      double w[3];
      for (int q = 0; q < 3; ++q) w[q] = f * (d.ed_c[3 * eo + q] - p[q]);
      put_projected(&wt[4 * ((size_t)m * NT + t)], w, p);
    }
  }
  Alloc &A = s->A;
  if (!s->x_ap_tnb) {
    s->x_ap_tnb = A.get<int32_t>((size_t)3 * NT);
    s->x_ap_wt = A.get<T>((size_t)4 * 3 * NT);
    s->x_ap_v = A.get<T>((size_t)4 * nv);
    s->x_ap_g = A.get<T>((size_t)4 * NT);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
  }
  const std::vector<T> wtT = cvt<T>(wt);
  HIPTRY(hipMemcpy(s->x_ap_tnb, tnb.data(), tnb.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_ap_wt, wtT.data(), wtT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  s->f_apvm = true;
  return IMX_OK;
}

static int validate_desc(const imx_mesh_desc &d) {
  if (d.nv <= 0 || d.ne <= 0 || d.nt <= 0) return set_error(IMX_EINVAL, "mesh: empty mesh");
  if (d.nv >= (1 << OTH_V2)) return set_error(IMX_EINVAL, "mesh: too many cells (max 2^30 - 1)");
  if (!d.v_nnb || !d.v_nb || !d.v_ed || !d.v_tr || !d.hx_ttgout || !d.hx_tnrccw || !d.v_p ||
      !d.hx_areag || !d.ed_v || !d.ed_sh || !d.ed_leng || !d.edhx_leng || !d.ed_c || !d.edhx_c ||
      !d.tr_v || !d.tr_ed || !d.tr_tg || !d.tr_areag || !d.tr_c || !d.tr_clat)
    return set_error(IMX_EINVAL, "mesh: a required array is NULL");
  for (int i = 0; i < d.nv; ++i) {
    const int n = d.v_nnb[i];
    if (n < 3 || n > MX) return set_error(IMX_EINVAL, "mesh: cell with nnb outside 3..6");
    for (int j = 0; j < n; ++j) {
      const int e = d.v_ed[MX * i + j], t = d.v_tr[MX * i + j], c = d.v_nb[MX * i + j];
      if (e < 1 || e > d.ne || t < 1 || t > d.nt || c < 1 || c > d.nv)
        return set_error(IMX_EINVAL, "mesh: cell list index out of range");
    }
  }
  for (int l = 0; l < d.ne; ++l)
    for (int k = 0; k < 2; ++k) {
      const int c = d.ed_v[2 * l + k], t = d.ed_sh[2 * l + k];
      if (c < 1 || c > d.nv || t < 1 || t > d.nt)
        return set_error(IMX_EINVAL, "mesh: edge index out of range");
    }
  for (int k = 0; k < d.nt; ++k)
    for (int l = 0; l < 3; ++l) {
      const int e = d.tr_ed[3 * k + l], c = d.tr_v[3 * k + l];
      if (e < 1 || e > d.ne || c < 1 || c > d.nv)
        return set_error(IMX_EINVAL, "mesh: triangle index out of range");
    }
  // every edge must appear in the rings of both its end cells
  for (int l = 0; l < d.ne; ++l)
    for (int k = 0; k < 2; ++k) {
      const int c = d.ed_v[2 * l + k] - 1;
      bool found = false;
      for (int j = 0; j < d.v_nnb[c]; ++j) found |= (d.v_ed[MX * c + j] == l + 1);
      if (!found) return set_error(IMX_EINVAL, "mesh: edge missing from the ring of its cell");
    }
  return IMX_OK;
}

static int create(const imx_mesh_desc &d, const imx_swm_opts &o, imx_swm *s,
                  const int32_t *group = nullptr, const double *dt_cell = nullptr,
                  const double *dt_edge = nullptr, const Layout *pre = nullptr) {
  int rc = validate_desc(d);
  if (rc) return rc;
  if (o.precision < IMX_PREC_FP64_REF || o.precision > IMX_PREC_FP32)
    return set_error(IMX_EINVAL, "opts: unknown precision");
  if (!(o.erad > 0) || !(o.grav > 0)) return set_error(IMX_EINVAL, "opts: erad and grav must be > 0");
  s->o = o;
  s->nv = d.nv;
  s->ne = d.ne;
  s->nt = d.nt;
  s->fast = o.precision != IMX_PREC_FP64_REF;
  s->single = o.precision == IMX_PREC_FP32;
  if (o.h_ref != 0.0) {
    s->href = o.h_ref;
    s->href_set = true;
  }

  // gridprop: caller's arrays or computed here (reference formulas, FP64)
  std::vector<double> trhx_v, hxtr_v, trskw_v;
  const double *trhx = d.tr_trhx_area, *hxtr = d.hx_hxtr_area, *trskw = d.hx_trskw;
  if (!trhx || !hxtr || !trskw) {
    trhx_v.assign((size_t)3 * d.nt, 0.0);
    hxtr_v.assign((size_t)MX * d.nv, 0.0);
    trskw_v.assign((size_t)MX * MX * d.nv, 0.0);
    gridprop(d, trhx_v.data(), hxtr_v.data(), trskw_v.data());
    trhx = trhx_v.data();
    hxtr = hxtr_v.data();
    trskw = trskw_v.data();
  }

  std::string err;
  if (pre) s->lay = *pre;
  else if (!build_layout(d, o.use_sfc != 0, s->lay, err, group))
    return set_error(IMX_EINVAL, "layout: " + err);
  s->NE = s->lay.NE;
  s->NT = s->lay.NT;
  const int nv = d.nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;

  HIPTRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
  HIPTRY(hipEventCreate(&s->ev0), "hipEventCreate");
  HIPTRY(hipEventCreate(&s->ev1), "hipEventCreate");

  // geometry / REF tables in the internal numbering (slot-major [j][entity])
  std::vector<int32_t> c_nnb(nv), c_ed((size_t)MX * nv, 0);
  std::vector<int8_t> c_ttg((size_t)MX * nv, 0), c_tnr((size_t)MX * nv, 0);
  std::vector<double> c_area(nv), c_trskw;
  if (!s->fast) c_trskw.assign((size_t)MX * MX * nv, 0.0);
  for (int co = 0; co < nv; ++co) {
    const int c = cn[co];
    const int n = d.v_nnb[co];
    c_nnb[c] = n;
    c_area[c] = d.hx_areag[co];
    for (int j = 0; j < n; ++j) {
      c_ed[(size_t)j * nv + c] = en[d.v_ed[MX * co + j] - 1];
      c_ttg[(size_t)j * nv + c] = (int8_t)(d.hx_ttgout[MX * co + j] < 0 ? -1 : 1);
      c_tnr[(size_t)j * nv + c] = (int8_t)(d.hx_tnrccw[MX * co + j] < 0 ? -1 : 1);
      if (!s->fast)
        for (int k = 0; k < n; ++k)
          c_trskw[((size_t)j * MX + k) * nv + c] = trskw[MX * MX * co + MX * j + k];
    }
  }
  std::vector<int32_t> e_v((size_t)2 * NE, -1), e_sh((size_t)2 * NE, -1);
  std::vector<int8_t> e_loc((size_t)2 * NE, 0);
  std::vector<double> e_leng(NE, 1.0), e_dleng(NE, 0.0);
  for (int eo = 0; eo < d.ne; ++eo) {
    const int e = en[eo];
    for (int k = 0; k < 2; ++k) {
      const int co = d.ed_v[2 * eo + k] - 1;
      e_v[(size_t)k * NE + e] = cn[co];
      e_sh[(size_t)k * NE + e] = tn[d.ed_sh[2 * eo + k] - 1];
      int loc = -1;  // getedindexonhx (smeshpack.f90:4099-4125)
      for (int j = 0; j < d.v_nnb[co]; ++j)
        if (d.v_ed[MX * co + j] == eo + 1) {
          loc = j;
          break;
        }
      e_loc[(size_t)k * NE + e] = (int8_t)loc;
    }
    e_leng[e] = d.ed_leng[eo];
    e_dleng[e] = d.edhx_leng[eo];
  }
  std::vector<int32_t> t_ed((size_t)3 * NT, -1), t_v((size_t)3 * NT, 0);
  std::vector<int8_t> t_tg((size_t)3 * NT, 0);
  std::vector<double> t_area(NT, 1.0), t_f(NT, 0.0), t_trhx((size_t)3 * NT, 0.0);
  for (int ko = 0; ko < d.nt; ++ko) {
    const int k = tn[ko];
    for (int l = 0; l < 3; ++l) {
      t_ed[(size_t)l * NT + k] = en[d.tr_ed[3 * ko + l] - 1];
      t_v[(size_t)l * NT + k] = cn[d.tr_v[3 * ko + l] - 1];
      t_tg[(size_t)l * NT + k] = (int8_t)(d.tr_tg[3 * ko + l] < 0 ? -1 : 1);
      t_trhx[(size_t)l * NT + k] = trhx[3 * ko + l];
    }
    t_area[k] = d.tr_areag[ko];
    // 2.*Omega*dsin(lat) of vort_tr (swm_operators.f90:597), libm sin on the host
    t_f[k] = 2.0 * o.omega * std::sin(d.tr_clat[ko]);
  }
  Alloc &A = s->A;
  s->c_nnb = A.put(c_nnb);
  s->c_ed = A.put(c_ed);
  s->c_ttg = A.put(c_ttg);
  s->c_tnr = A.put(c_tnr);
  s->c_area = A.put(c_area);
  s->e_v = A.put(e_v);
  s->e_sh = A.put(e_sh);
  s->e_loc = A.put(e_loc);
  s->e_leng = A.put(e_leng);
  s->e_dleng = A.put(e_dleng);
  s->t_ed = A.put(t_ed);
  s->t_v = A.put(t_v);
  s->t_tg = A.put(t_tg);
  s->t_area = A.put(t_area);
  s->t_f = A.put(t_f);
  s->f_var = t_f;
  s->t_trhx = A.put(t_trhx);
  s->bt = A.put(std::vector<double>(nv, 0.0));
  s->flag = A.put(std::vector<int>(1, 0));
  s->nparts = (nv + TPB - 1) / TPB + (NE + TPB - 1) / TPB + (NT + TPB - 1) / TPB;
  s->part = A.get<double>((size_t)s->nparts * NPART);
  s->diag = A.get<double>(NPART);
  if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_create (geometry)");

  if (!s->fast) {
    s->c_trskw = A.put(c_trskw);
    {  // TRiSK terms per edge, edge-major
      std::vector<int32_t> tk((size_t)2 * MX * NE, -1);
      std::vector<double> tw((size_t)2 * MX * NE, 0.0), tdl((size_t)2 * MX * NE, 0.0);
      std::vector<int8_t> ts((size_t)2 * MX * NE, 1);
      for (int e = 0; e < NE; ++e) {
        if (e_v[e] < 0) continue;
        for (int ic = 0; ic < 2; ++ic) {
          const int node = e_v[(size_t)ic * NE + e], ei = e_loc[(size_t)ic * NE + e];
          for (int j = 0; j < c_nnb[node]; ++j) {
            const size_t t = ((size_t)ic * MX + j) * NE + e;
            const int k = c_ed[(size_t)j * nv + node];
            tk[t] = k;
            tw[t] = c_trskw[((size_t)ei * MX + j) * nv + node];
            tdl[t] = e_dleng[k];
            ts[t] = (int8_t)(c_ttg[(size_t)j * nv + node] * c_tnr[(size_t)ei * nv + node]);
          }
        }
      }
      s->e_tk = A.put(tk);
      s->e_tw = A.put(tw);
      s->e_tdl = A.put(tdl);
      s->e_ts = A.put(ts);
      s->hc_ts = A.get<int8_t>((size_t)2 * MX * NE);
    }
    s->rh = A.get<double>(nv);
    s->ru = A.get<double>(NE);
    s->rhn = A.get<double>(nv);
    s->run = A.get<double>(NE);
    s->massf = A.get<double>((size_t)4 * nv);
    s->momf = A.get<double>((size_t)4 * NE);
    s->f_h_ed = A.get<double>(NE);
    s->f_uh = A.get<double>(NE);
    s->f_eta = A.get<double>(NT);
    s->f_h_tr = A.get<double>(NT);
    s->f_q_tr = A.get<double>(NT);
    s->f_q_ed = A.get<double>(NE);
    s->f_divuh = A.get<double>(nv);
    s->f_ke = A.get<double>(nv);
    s->f_ghbk = A.get<double>(nv);
    s->f_grad = A.get<double>(NE);
    s->f_uhq = A.get<double>(NE);
    s->f_divu = A.get<double>(nv);
    s->f_zeta = A.get<double>(NT);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_create (ref fields)");
  } else {
    rc = s->single ? build_fast<float>(s, d, trhx, hxtr) : build_fast<double>(s, d, trhx, hxtr);
    if (rc) return rc;
    if (dt_cell && dt_edge) {  // per-instance time steps of an ensemble
      std::vector<double> dc(nv), de(NE, 0.0);
      for (int co = 0; co < nv; ++co) dc[cn[co]] = dt_cell[co];
      for (int eo = 0; eo < d.ne; ++eo) de[en[eo]] = dt_edge[eo];
      s->dt_c = A.put(dc);
      if (s->single) s->dt_e = A.put(cvt<float>(de));
      else s->dt_e = A.put(de);
      if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_create (time steps)");
    }
  }
  HIPTRY(hipDeviceSynchronize(), "imx_swm_create");
  return IMX_OK;
}

static int check_flag(imx_swm *s, const char *what) {
  int f = 0;
  HIPTRY(hipMemcpyAsync(&f, s->flag, sizeof(int), hipMemcpyDeviceToHost, s->stream), what);
  HIPTRY(hipStreamSynchronize(s->stream), what);
  if (f) {
    int z = 0;
    HIPTRY(hipMemcpy(s->flag, &z, sizeof(int), hipMemcpyHostToDevice), what);
    return set_error(IMX_ESTATE, std::string(what) +
                                     ": division by zero, |h_tr| < epsilon/1e5 "
                                     "(swm_operators.f90:1603)");
  }
  return IMX_OK;
}

template <typename T>
static int upload_state(imx_swm *s, const double *h, const double *u, T *hd, T *ud) {
  std::vector<T> hb(s->nv, (T)0), ub(s->NE, (T)0);
  for (int i = 0; i < s->nv; ++i) hb[s->lay.cell_new[i]] = (T)(h[i] - s->href);
  for (int l = 0; l < s->ne; ++l) ub[s->lay.edge_new[l]] = (T)u[l];
  HIPTRY(hipMemcpyAsync(hd, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, s->stream),
         "upload h");
  HIPTRY(hipMemcpyAsync(ud, ub.data(), ub.size() * sizeof(T), hipMemcpyHostToDevice, s->stream),
         "upload u");
  HIPTRY(hipStreamSynchronize(s->stream), "upload");
  return IMX_OK;
}

template <typename T>
static int download_state(imx_swm *s, double *h, double *u, const T *hd, const T *ud) {
  std::vector<T> hb(s->nv), ub(s->NE);
  HIPTRY(hipMemcpyAsync(hb.data(), hd, hb.size() * sizeof(T), hipMemcpyDeviceToHost, s->stream),
         "download h");
  HIPTRY(hipMemcpyAsync(ub.data(), ud, ub.size() * sizeof(T), hipMemcpyDeviceToHost, s->stream),
         "download u");
  HIPTRY(hipStreamSynchronize(s->stream), "download");
  if (h)
    for (int i = 0; i < s->nv; ++i) h[i] = (double)hb[s->lay.cell_new[i]] + s->href;
  if (u)
    for (int l = 0; l < s->ne; ++l) u[l] = (double)ub[s->lay.edge_new[l]];
  return IMX_OK;
}

static void drop_graph(imx_swm *s);

static void set_href(imx_swm *s, const double *h) {
  if (s->href_set || !s->fast) return;
  double sa = 0.0, sh = 0.0;
  std::vector<double> area(s->nv);
  // area-weighted mean of the first state set (cell areas live on the device)
  (void)hipMemcpy(area.data(), s->c_area, s->nv * sizeof(double), hipMemcpyDeviceToHost);
  for (int i = 0; i < s->nv; ++i) {
    const double Ar = area[s->lay.cell_new[i]];
    sa += Ar;
    sh += Ar * h[i];
  }
  s->href = sh / sa;
  s->href_set = true;
  drop_graph(s);  // a step captured before holds href = 0
}

int swm_stencil_rings(const imx_swm *s) {
  return 1 + (s->sc_apvm ? 1 : 0) + (s->sc_hyper ? 1 : 0) + (s->sc_cor == IMX_COR_GASS ? 1 : 0) +
         (s->sc_rec == IMX_REC_GASS ? 1 : 0);
}

static void drop_graph(imx_swm *s) {
  ++s->generation;  // graphs captured elsewhere over this handle's step are stale too
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
}

static int ensure_graph(imx_swm *s) {
  if (s->gexec) return IMX_OK;
  HIPTRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "graph capture");
  launch_step(s);
  HIPTRY(hipStreamEndCapture(s->stream, &s->graph), "graph capture end");
  HIPTRY(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0), "graph instantiate");
  return IMX_OK;
}

static int run_steps(imx_swm *s, int nsteps) {
  if (s->o.use_graph) {
    int rc = ensure_graph(s);
    if (rc) return rc;
    for (int k = 0; k < nsteps; ++k) HIPTRY(hipGraphLaunch(s->gexec, s->stream), "hipGraphLaunch");
  } else {
    for (int k = 0; k < nsteps; ++k) launch_step(s);
    HIPTRY(hipGetLastError(), "step launch");
  }
  return IMX_OK;
}

// Diagnostics vector of the cells [c0, c1) (IMX_DIAG_* order, FP64).
static void launch_diag_state(imx_swm *s, int c0, int c1, double *part, double *diag) {
  if (!s->fast) launch_diag<double>(s, s->rh, s->ru, 0.0, c0, c1, part, diag);
  else if (s->single)
    launch_diag<float>(s, (const float *)s->h0, (const float *)s->u0, s->href, c0, c1, part, diag);
  else launch_diag<double>(s, (const double *)s->h0, (const double *)s->u0, s->href, c0, c1, part, diag);
}

static int diag_range(imx_swm *s, int c0, int c1, double *out) {
  launch_diag_state(s, c0, c1, s->part, s->diag);
  double v[NPART];
  HIPTRY(hipMemcpyAsync(v, s->diag, sizeof(v), hipMemcpyDeviceToHost, s->stream), "diagnostics");
  HIPTRY(hipStreamSynchronize(s->stream), "diagnostics");
  const double mass = v[0] / v[1];
  const double pe = v[2];
  const double ke = v[4];
  // unavailable potential energy: the fluid at rest at the mean height m of
  // h + b (calc_energies, src/swm.f90:2706-2715):
  // sum A g (m - b)((m - b)/2 + b) = g/2 (m^2 sum A - sum A b^2)
  double upe;
  if (s->has_topo) {
    const double m = (v[0] + s->topo_sb) / v[1];
    upe = 0.5 * s->o.grav * (m * m * v[1] - s->topo_sb2);
  } else {
    upe = v[1] * s->o.grav * mass * (mass * 0.5);
  }
  out[IMX_DIAG_MASS] = mass;
  out[IMX_DIAG_PENERGY] = pe;
  out[IMX_DIAG_KENERGY] = ke;
  out[IMX_DIAG_TENERGY] = pe + ke;
  out[IMX_DIAG_AENERGY] = (pe + ke) - upe;
  out[IMX_DIAG_PENSTROPHY] = v[5];
  out[IMX_DIAG_RMSDIV] = std::sqrt(v[3] / v[1]);
  out[IMX_DIAG_BADFLAG] = v[7] > 0 ? 1.0 : 0.0;
  return IMX_OK;
}

// Autotuner: for each fused kernel, time every (block size, XCD remap)
// candidate over `nsteps` steps with the other kernels fixed, keep the
// fastest. The state is restored afterwards.
static int tune(imx_swm *s, int nsteps, int32_t *out) {
  if (!s->fast) return set_error(IMX_ENOTSUP, "imx_swm_tune: FAST precisions only");
  if (nsteps < 1) nsteps = 1;
  const size_t es = s->single ? sizeof(float) : sizeof(double);
  void *sh = nullptr, *su = nullptr;
  HIPTRY(hipMalloc(&sh, es * s->nv), "tune");
  HIPTRY(hipMalloc(&su, es * s->NE), "tune");
  HIPTRY(hipMemcpyAsync(sh, s->h0, es * s->nv, hipMemcpyDeviceToDevice, s->stream), "tune");
  HIPTRY(hipMemcpyAsync(su, s->u0, es * s->NE, hipMemcpyDeviceToDevice, s->stream), "tune");
  std::vector<hipEvent_t> ev(16);
  for (auto &e : ev) HIPTRY(hipEventCreate(&e), "event");
  const int bss[4] = {64, 128, 256, 512};
  for (int k = 0; k < 3; ++k) {
    double best = 1e30;
    int bb = s->kbs[k], br = s->kremap[k];
    for (int bi = 0; bi < 4; ++bi)
      for (int r = 0; r < 2; ++r) {
        s->kbs[k] = bss[bi];
        s->kremap[k] = r;
        double acc = 0.0;
        for (int it = 0; it <= nsteps; ++it) {  // first pass warms up
          if (s->single) launch_fast_step<float>(s, &ev);
          else launch_fast_step<double>(s, &ev);
          HIPTRY(hipEventSynchronize(ev[15]), "event");
          if (it == 0) continue;
          for (int st = 0; st < 4; ++st) {
            float f = 0.f;
            HIPTRY(hipEventElapsedTime(&f, ev[4 * st + k], ev[4 * st + k + 1]), "event");
            acc += f;
          }
        }
        if (acc < best) {
          best = acc;
          bb = bss[bi];
          br = r;
        }
      }
    s->kbs[k] = bb;
    s->kremap[k] = br;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  HIPTRY(hipMemcpyAsync(s->h0, sh, es * s->nv, hipMemcpyDeviceToDevice, s->stream), "tune");
  HIPTRY(hipMemcpyAsync(s->u0, su, es * s->NE, hipMemcpyDeviceToDevice, s->stream), "tune");
  HIPTRY(hipStreamSynchronize(s->stream), "tune");
  (void)hipFree(sh);
  (void)hipFree(su);
  drop_graph(s);
  if (out)
    for (int k = 0; k < 3; ++k) {
      out[2 * k] = s->kbs[k];
      out[2 * k + 1] = s->kremap[k];
    }
  return check_flag(s, "imx_swm_tune");
}

int swm_check_flag(imx_swm *s, const char *what) { return check_flag(s, what); }

}  // namespace imx

// ================================================================== C ABI ==
extern "C" {

void imx_swm_default_opts(imx_swm_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->precision = IMX_PREC_FP32;
  o->testcase = 2;
  o->use_sfc = 1;
  o->use_graph = 1;
  o->erad = 6.37122e6;
  o->grav = 9.80616;
  o->omega = 7.292e-5;
  o->dt = 900.0;
  o->h_ref = 0.0;
}

int imx_swm_create(const imx_mesh_desc *d, const imx_swm_opts *o, imx_swm **out) {
  if (!d || !o || !out) return imx::set_error(IMX_EINVAL, "imx_swm_create: NULL argument");
  *out = nullptr;
  imx_swm *s = new (std::nothrow) imx_swm();
  if (!s) return imx::set_error(IMX_ENOMEM, "imx_swm_create: out of memory");
  int rc;
  try {
    rc = imx::create(*d, *o, s);
  } catch (const std::bad_alloc &) {
    rc = imx::set_error(IMX_ENOMEM, "imx_swm_create: host out of memory");
  }
  if (rc) {
    imx_swm_destroy(s);
    return rc;
  }
  *out = s;
  return IMX_OK;
}

int imx_swm_set_state(imx_swm *s, const double *h, const double *u) {
  if (!s || !h || !u) return imx::set_error(IMX_EINVAL, "imx_swm_set_state: NULL argument");
  if (!s->fast) return imx::upload_state<double>(s, h, u, s->rh, s->ru);
  imx::set_href(s, h);
  if (s->single) return imx::upload_state<float>(s, h, u, (float *)s->h0, (float *)s->u0);
  return imx::upload_state<double>(s, h, u, (double *)s->h0, (double *)s->u0);
}

int imx_swm_set_topography(imx_swm *s, const double *bt) {
  if (!s) return imx::set_error(IMX_EINVAL, "imx_swm_set_topography: NULL handle");
  std::vector<double> b(s->nv, 0.0);
  double sb = 0.0, sb2 = 0.0;
  bool any = false;
  if (bt) {
    std::vector<double> area(s->nv);
    HIPTRY(hipMemcpy(area.data(), s->c_area, s->nv * sizeof(double), hipMemcpyDeviceToHost), "topography");
    for (int i = 0; i < s->nv; ++i) {
      if (!std::isfinite(bt[i])) return imx::set_error(IMX_EINVAL, "imx_swm_set_topography: non-finite b");
      const int k = s->lay.cell_new[i];
      b[k] = bt[i];
      any = any || bt[i] != 0.0;
    }
    for (int k = 0; k < s->nv; ++k) {  // internal order: the same sums for every caller numbering
      sb += area[k] * b[k];
      sb2 += area[k] * (b[k] * b[k]);
    }
  }
  HIPTRY(hipMemcpyAsync(s->bt, b.data(), s->nv * sizeof(double), hipMemcpyHostToDevice, s->stream), "topography");
  if (s->fast) {
    if (any && !s->x_gb_buf) {
      s->x_gb_buf = s->single ? (void *)s->A.get<float>(s->nv) : (void *)s->A.get<double>(s->nv);
      if (s->A.err != hipSuccess) return imx::set_error(IMX_ENOMEM, "imx_swm_set_topography: out of device memory");
    }
    if (any) {
      const double g = s->o.grav;
      if (s->single) {
        std::vector<float> gb(s->nv);
        for (int k = 0; k < s->nv; ++k) gb[k] = (float)(g * b[k]);
        HIPTRY(hipMemcpyAsync(s->x_gb_buf, gb.data(), gb.size() * sizeof(float), hipMemcpyHostToDevice, s->stream),
               "topography");
        HIPTRY(hipStreamSynchronize(s->stream), "topography");
      } else {
        std::vector<double> gb(s->nv);
        for (int k = 0; k < s->nv; ++k) gb[k] = g * b[k];
        HIPTRY(hipMemcpyAsync(s->x_gb_buf, gb.data(), gb.size() * sizeof(double), hipMemcpyHostToDevice, s->stream),
               "topography");
        HIPTRY(hipStreamSynchronize(s->stream), "topography");
      }
    }
  }
  HIPTRY(hipStreamSynchronize(s->stream), "topography");
  s->has_topo = any;
  s->topo_sb = sb;
  s->topo_sb2 = sb2;
  s->x_gb = (s->fast && any) ? s->x_gb_buf : nullptr;
  imx::drop_graph(s);  // the captured step holds the previous topography pointer
  return IMX_OK;
}

int imx_swm_set_coriolis(imx_swm *s, int32_t fsphere, double fcte) {
  if (!s) return imx::set_error(IMX_EINVAL, "imx_swm_set_coriolis: NULL handle");
  if (fsphere < IMX_F_VARIABLE || fsphere > IMX_F_CONSTANT || !std::isfinite(fcte))
    return imx::set_error(IMX_EINVAL, "imx_swm_set_coriolis: bad fsphere or fcte");
  std::vector<double> f(s->f_var);
  if (fsphere != IMX_F_VARIABLE) std::fill(f.begin(), f.end(), fcte);
  HIPTRY(hipMemcpy(s->t_f, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice), "imx_swm_set_coriolis");
  if (s->fast) {
    if (s->single) {
      std::vector<float> g(f.begin(), f.end());
      HIPTRY(hipMemcpy(s->x_t_f, g.data(), g.size() * sizeof(float), hipMemcpyHostToDevice), "imx_swm_set_coriolis");
    } else {
      HIPTRY(hipMemcpy(s->x_t_f, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice), "imx_swm_set_coriolis");
    }
  }
  s->f_lin = fsphere == IMX_F_LINEAR;
  imx::drop_graph(s);  // the captured step holds the previous flag
  return IMX_OK;
}

int imx_swm_get_state(imx_swm *s, double *h, double *u) {
  if (!s) return imx::set_error(IMX_EINVAL, "imx_swm_get_state: NULL handle");
  if (!s->fast) return imx::download_state<double>(s, h, u, s->rh, s->ru);
  if (s->single) return imx::download_state<float>(s, h, u, (float *)s->h0, (float *)s->u0);
  return imx::download_state<double>(s, h, u, (double *)s->h0, (double *)s->u0);
}

int imx_swm_tendency(imx_swm *s, const double *h, const double *u, double *masseq,
                     double *momeq) {
  if (!s || !h || !u || !masseq || !momeq)
    return imx::set_error(IMX_EINVAL, "imx_swm_tendency: NULL argument");
  int rc;
  std::vector<double> mb(s->nv), fb(s->NE);
  if (!s->fast) {
    if ((rc = imx::upload_state<double>(s, h, u, s->rhn, s->run))) return rc;
    imx::launch_ref_tendency(s, s->rhn, s->run, s->massf, s->momf);
    if ((rc = imx::check_flag(s, "imx_swm_tendency"))) return rc;
    HIPTRY(hipMemcpy(mb.data(), s->massf, s->nv * sizeof(double), hipMemcpyDeviceToHost), "tendency");
    HIPTRY(hipMemcpy(fb.data(), s->momf, s->NE * sizeof(double), hipMemcpyDeviceToHost), "tendency");
  } else {
    imx::set_href(s, h);
    if (s->single) {
      if ((rc = imx::upload_state<float>(s, h, u, (float *)s->hA, (float *)s->uA))) return rc;
      imx::launch_fast_stage<float>(s, -1, (const float *)s->hA, (const float *)s->uA, nullptr,
                                    nullptr, nullptr);
      if ((rc = imx::check_flag(s, "imx_swm_tendency"))) return rc;
      std::vector<float> ff(s->NE);
      HIPTRY(hipMemcpy(ff.data(), s->du_out, s->NE * sizeof(float), hipMemcpyDeviceToHost), "tendency");
      for (int l = 0; l < s->NE; ++l) fb[l] = ff[l];
    } else {
      if ((rc = imx::upload_state<double>(s, h, u, (double *)s->hA, (double *)s->uA))) return rc;
      imx::launch_fast_stage<double>(s, -1, (const double *)s->hA, (const double *)s->uA, nullptr,
                                     nullptr, nullptr);
      if ((rc = imx::check_flag(s, "imx_swm_tendency"))) return rc;
      HIPTRY(hipMemcpy(fb.data(), s->du_out, s->NE * sizeof(double), hipMemcpyDeviceToHost), "tendency");
    }
    HIPTRY(hipMemcpy(mb.data(), s->dh_out, s->nv * sizeof(double), hipMemcpyDeviceToHost), "tendency");
  }
  for (int i = 0; i < s->nv; ++i) masseq[i] = mb[s->lay.cell_new[i]];
  for (int l = 0; l < s->ne; ++l) momeq[l] = fb[s->lay.edge_new[l]];
  return IMX_OK;
}

int imx_swm_get_field(imx_swm *s, int32_t which, double *out) {
  if (!s || !out) return imx::set_error(IMX_EINVAL, "imx_swm_get_field: NULL argument");
  if (s->fast) {
    // the fused passes keep only q on triangles (from the last RK stage)
    if (which != IMX_FIELD_Q_TR)
      return imx::set_error(IMX_ENOTSUP,
                            "imx_swm_get_field: fused handles keep only IMX_FIELD_Q_TR; use IMX_PREC_FP64_REF");
    std::vector<double> b(s->NT);
    if (s->single) {
      std::vector<float> f(s->NT);
      HIPTRY(hipMemcpy(f.data(), s->q, s->NT * sizeof(float), hipMemcpyDeviceToHost), "get_field");
      for (int t = 0; t < s->NT; ++t) b[t] = f[t];
    } else {
      HIPTRY(hipMemcpy(b.data(), s->q, s->NT * sizeof(double), hipMemcpyDeviceToHost), "get_field");
    }
    for (int t = 0; t < s->nt; ++t) out[t] = b[s->lay.tri_new[t]];
    return IMX_OK;
  }
  double *src = nullptr;
  int kind = 0;  // 0 cell, 1 edge, 2 tri
  switch (which) {
    case IMX_FIELD_H_ED: src = s->f_h_ed; kind = 1; break;
    case IMX_FIELD_UH: src = s->f_uh; kind = 1; break;
    case IMX_FIELD_ETA: src = s->f_eta; kind = 2; break;
    case IMX_FIELD_H_TR: src = s->f_h_tr; kind = 2; break;
    case IMX_FIELD_Q_TR: src = s->f_q_tr; kind = 2; break;
    case IMX_FIELD_Q_ED: src = s->f_q_ed; kind = 1; break;
    case IMX_FIELD_DIVUH: src = s->f_divuh; kind = 0; break;
    case IMX_FIELD_KE_HX: src = s->f_ke; kind = 0; break;
    case IMX_FIELD_GHBK: src = s->f_ghbk; kind = 0; break;
    case IMX_FIELD_GRAD_GHBK: src = s->f_grad; kind = 1; break;
    case IMX_FIELD_UHQ_PERP: src = s->f_uhq; kind = 1; break;
    case IMX_FIELD_DIVU: src = s->f_divu; kind = 0; break;
    case IMX_FIELD_ZETA: src = s->f_zeta; kind = 2; break;
    case IMX_FIELD_Q_GRAD_ED: src = s->sc_apvm ? s->f_qg_ed : nullptr; kind = 1; break;
    case IMX_FIELD_LAPU: src = (s->sc_diff || s->sc_hyper) ? s->f_lapu : nullptr; kind = 1; break;
    case IMX_FIELD_LAP_LAPU: src = s->sc_hyper ? s->f_lap2 : nullptr; kind = 1; break;
    case IMX_FIELD_Q_HX:
    case IMX_FIELD_VHQ_TR:
    case IMX_FIELD_VH_HX:
    case IMX_FIELD_VHQ_HX: {
      if (!s->lte_done)
        return imx::set_error(IMX_EINVAL, "imx_swm_get_field: the analysis fields come from imx_swm_lte");
      if (which == IMX_FIELD_Q_HX) {
        src = s->f_lte_qhx;
        kind = 0;
        break;
      }
      const double *v = which == IMX_FIELD_VHQ_TR ? s->f_lte_vtr : which == IMX_FIELD_VH_HX ? s->f_lte_vhh : s->f_lte_vhq;
      return imx::get_vector(s, v, which == IMX_FIELD_VHQ_TR ? 2 : 0, out);
    }
    case IMX_FIELD_Q_GRAD_TR:
      if (!s->sc_apvm) return imx::set_error(IMX_EINVAL, "imx_swm_get_field: the field's scheme option is off");
      return imx::get_vector(s, s->f_qg_tr, 2, out);
    default: return imx::set_error(IMX_EINVAL, "imx_swm_get_field: unknown field");
  }
  if (!src) return imx::set_error(IMX_EINVAL, "imx_swm_get_field: the field's scheme option is off");
  const int n = kind == 0 ? s->nv : kind == 1 ? s->ne : s->nt;
  const int ni = kind == 0 ? s->nv : kind == 1 ? s->NE : s->NT;
  const std::vector<int32_t> &perm =
      kind == 0 ? s->lay.cell_new : kind == 1 ? s->lay.edge_new : s->lay.tri_new;
  std::vector<double> b(ni);
  HIPTRY(hipMemcpy(b.data(), src, ni * sizeof(double), hipMemcpyDeviceToHost), "get_field");
  for (int i = 0; i < n; ++i) out[i] = b[perm[i]];
  return IMX_OK;
}

int imx_swm_lte(imx_swm *s, const imx_mesh *m, double *err, double *mim) {
  return imx_swm_lte_hol(s, m, 0.0, err, mim);
}

int imx_swm_lte_hol(imx_swm *s, const imx_mesh *m, double hollgw, double *err, double *mim) {
  if (!s || !m || !err || !mim) return imx::set_error(IMX_EINVAL, "imx_swm_lte: NULL argument");
  if (s->fast) return imx::set_error(IMX_ENOTSUP, "imx_swm_lte: the LTE analysis runs on FP64_REF handles");
  const int tc = s->o.testcase;
  if (tc != 2 && !(tc >= 32 && tc <= 35))
    return imx::set_error(IMX_ENOTSUP, "imx_swm_lte: the closed forms are those of test cases 2 and 32-35 "
                                       "(the handle's testcase)");
  if (m->nv != s->nv || m->ne != s->ne || m->nt != s->nt)
    return imx::set_error(IMX_EINVAL, "imx_swm_lte: not the handle's mesh");
  imx_mesh_desc d;
  int rc = imx_mesh_desc_of(m, &d);
  if (rc) return rc;
  if ((rc = imx::scheme_tables(s, &d))) return rc;
  const int nv = s->nv, ne = s->ne, nt = s->nt, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &tn = s->lay.tri_new;
  imx::Alloc &A = s->A;
  if (!s->f_lte_qhx) {
    s->f_lte_qhx = A.get<double>(nv);
    s->f_lte_vtr = A.get<double>((size_t)3 * NT);
    s->f_lte_vhh = A.get<double>((size_t)3 * nv);
    s->f_lte_vhq = A.get<double>((size_t)3 * nv);
    s->lte_ctr = A.get<int32_t>((size_t)MX * nv);
    s->lte_cw = A.get<double>((size_t)MX * nv);
    if (A.err != hipSuccess) return imx::hip_fail(A.err, "imx_swm_lte");
  }
  // scalar_tr2hx weights: Wachspress coordinates with pxt16's barycentric
  // interpolation (initialize_gridprop, swm.f90:1001-1005), else hxtr_area
  std::vector<int32_t> ctr((size_t)MX * nv, 0);
  std::vector<double> cw((size_t)MX * nv, 0.0), hxtr_v, trhx_v, trskw_v;
  const double *hxtr = d.hx_hxtr_area;
  if (!hxtr && s->sc_sint != IMX_SINT_BARY) {
    trhx_v.assign((size_t)3 * nt, 0.0);
    hxtr_v.assign((size_t)MX * nv, 0.0);
    trskw_v.assign((size_t)MX * MX * nv, 0.0);
    imx::gridprop(d, trhx_v.data(), hxtr_v.data(), trskw_v.data());
    hxtr = hxtr_v.data();
  }
  for (int co = 0; co < nv; ++co) {
    double w[MX];
    if (s->sc_sint == IMX_SINT_BARY) imx::wachspress_node(d, co, w);
    for (int j = 0; j < d.v_nnb[co]; ++j) {
      ctr[(size_t)j * nv + cn[co]] = tn[d.v_tr[MX * co + j] - 1];
      cw[(size_t)j * nv + cn[co]] = s->sc_sint == IMX_SINT_BARY ? w[j] : hxtr[MX * co + j];
    }
  }
  HIPTRY(hipMemcpy(s->lte_ctr, ctr.data(), ctr.size() * sizeof(int32_t), hipMemcpyHostToDevice), "imx_swm_lte");
  HIPTRY(hipMemcpy(s->lte_cw, cw.data(), cw.size() * sizeof(double), hipMemcpyHostToDevice), "imx_swm_lte");
  // the test case as initialize_fields sets it (state, bottom and Coriolis
  // term), resident like the reference's globals
  std::vector<double> h(nv), u(ne), bt(nv), mass(nv), mom(ne);
  rc = imx_ic_testcase(m, tc, s->sc_hc ? IMX_STAG_HC : IMX_STAG_HTC, s->o.erad, hollgw, h.data(), u.data(),
                       bt.data());
  if (rc) return rc;
  int32_t fsphere;
  double fcte;
  if ((rc = imx_ic_testcase_coriolis(tc, &fsphere, &fcte)) || (rc = imx_swm_set_coriolis(s, fsphere, fcte)) ||
      (rc = imx_swm_set_topography(s, bt.data())))
    return rc;
  if ((rc = imx_swm_set_state(s, h.data(), u.data()))) return rc;
  if ((rc = imx_swm_tendency(s, h.data(), u.data(), mass.data(), mom.data()))) return rc;
  // the analysis fields, on the tendency's input (rhn, run) and fields
  const bool hc = s->sc_hc;
  hipStream_t st = s->stream;
  ref_perot_hx<<<grid_for(nv), TPB, 0, st>>>(nv, s->NE, s->c_nnb, s->c_ed, hc ? s->hc_c : s->c_ttg,
                                                       hc ? s->hc_pc : s->e_c, s->c_p, s->e_dleng, s->c_area,
                                                       s->run, s->f_vhx);
  ref_lte_tr<<<grid_for(NT), TPB, 0, st>>>(NT, s->NE, s->t_ed, hc ? s->hc_t : s->t_tg,
                                                     hc ? s->hc_pc : s->e_c, s->t_c, s->e_leng, s->t_area,
                                                     s->f_uh, s->f_q_tr, s->f_lte_vtr);
  ref_lte_cell<<<grid_for(nv), TPB, 0, st>>>(nv, s->c_nnb, s->lte_ctr, s->lte_cw, s->f_vhx, s->rhn,
                                                       s->f_q_tr, s->f_lte_vhh, s->f_lte_qhx, s->f_lte_vhq);
  HIPTRY(hipGetLastError(), "imx_swm_lte launch");
  HIPTRY(hipStreamSynchronize(st), "imx_swm_lte");
  s->lte_done = true;
  // fields in caller numbering
  auto fld = [&](int which, size_t n, std::vector<double> &v) {
    v.assign(n, 0.0);
    return imx_swm_get_field(s, which, v.data());
  };
  std::vector<double> h_ed, uh, divuh, ke, eta, h_tr, q_tr, q_hx, q_ed, grad, vhq_tr, qg_tr, vh_hx, vhq_hx, uhq, lapu;
  if ((rc = fld(IMX_FIELD_H_ED, ne, h_ed)) || (rc = fld(IMX_FIELD_UH, ne, uh)) ||
      (rc = fld(IMX_FIELD_DIVUH, nv, divuh)) || (rc = fld(IMX_FIELD_KE_HX, nv, ke)) ||
      (rc = fld(IMX_FIELD_ETA, nt, eta)) || (rc = fld(IMX_FIELD_H_TR, nt, h_tr)) ||
      (rc = fld(IMX_FIELD_Q_TR, nt, q_tr)) || (rc = fld(IMX_FIELD_Q_HX, nv, q_hx)) ||
      (rc = fld(IMX_FIELD_Q_ED, ne, q_ed)) || (rc = fld(IMX_FIELD_GRAD_GHBK, ne, grad)) ||
      (rc = fld(IMX_FIELD_VHQ_TR, (size_t)3 * nt, vhq_tr)) || (rc = fld(IMX_FIELD_VH_HX, (size_t)3 * nv, vh_hx)) ||
      (rc = fld(IMX_FIELD_VHQ_HX, (size_t)3 * nv, vhq_hx)) || (rc = fld(IMX_FIELD_UHQ_PERP, ne, uhq)))
    return rc;
  if (s->sc_apvm && (rc = fld(IMX_FIELD_Q_GRAD_TR, (size_t)3 * nt, qg_tr))) return rc;
  if ((s->sc_diff || s->sc_hyper) && (rc = fld(IMX_FIELD_LAPU, ne, lapu))) return rc;
  imx::LteFields f{h.data(), u.data(), h_ed.data(), uh.data(), divuh.data(), ke.data(), eta.data(),
                   h_tr.data(), q_tr.data(), q_hx.data(), q_ed.data(), grad.data(), vhq_tr.data(),
                   qg_tr.empty() ? nullptr : qg_tr.data(), vh_hx.data(), vhq_hx.data(), uhq.data(),
                   lapu.empty() ? nullptr : lapu.data(), mass.data(), mom.data(), bt.data()};
  return imx::lte_analysis(*m, tc, hollgw, hc, s->sc_rec == IMX_REC_PERHX, s->o.erad, s->o.grav, f, err, mim);
}

}  // extern "C"

namespace imx {
// The method triple each wrapper stands for (swm_data.f90:443-517): trsk10 =
// (trsk, trsk, trsk), pxt16 = (bary, pered, perhx), gass18 = (gass, gass,
// trsk) for (scalar interpolation, Coriolis, kinetic energy)
static void wrapper_methods(int wrap, int &sint, int &cor, int &rec) {
  sint = IMX_SINT_TRSK, cor = IMX_COR_TRSK, rec = IMX_REC_TRSK;
  if (wrap == IMX_WRAP_PXT16) sint = IMX_SINT_BARY, cor = IMX_COR_PERED, rec = IMX_REC_PERHX;
  if (wrap == IMX_WRAP_GASS18) sint = IMX_SINT_GASS, cor = IMX_COR_GASS;
}

// area of the triangle (circumcentre of k, ends of its edge i): the
// Gassmann h_tr weights and kinetic_energy_tr's areas (swm_operators.f90:1112-1117)
static double gass_subarea(const imx_mesh_desc &d, int ko, int i) {
  const int eo = d.tr_ed[3 * ko + i] - 1;
  return sphtriarea(d.tr_c + 3 * ko, d.v_p + 3 * (d.ed_v[2 * eo] - 1), d.v_p + 3 * (d.ed_v[2 * eo + 1] - 1));
}

// dtred's Coriolis term on the fused kernels (on = false switches it off):
// per triangle slot, the vector weight of each edge slot in vhq = q sum_m
// uh_m W_m, W_m = tg_m leng_m (c x (e_m - c)) / A_t (vector_edhx2tr_perot
// with the circumcentre c and the edge points e_m: ed%c, HTC; edhx%c, HC);
// per edge slot, its two triangles, their interpolation weights and the
// normal the result is dotted with (HTC: minus ed%nr projected at ed%c;
// HC: edhx%tg), typed T
template <typename T>
static int fast_dtred(imx_swm *s, const imx_mesh_desc &d, bool on) {
  s->f_dtred = false;
  if (!on) return IMX_OK;
  const int NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  const bool hc = s->sc_hc;
  std::vector<int8_t> hct;
  if (hc) {  // the HC vorticity signs tr%tg s_e, in the layout of ref_tri
    hct.resize((size_t)3 * NT);
    HIPTRY(hipMemcpy(hct.data(), s->hc_t, hct.size(), hipMemcpyDeviceToHost), "set_scheme");
  }
  std::vector<double> w((size_t)4 * 3 * NT, 0.0), ew((size_t)2 * NE, 0.0), nn((size_t)4 * NE, 0.0);
  std::vector<int32_t> et((size_t)2 * NE, 0);
  for (int ko = 0; ko < d.nt; ++ko) {
    const int t = tn[ko];
    int r = -1;  // slot m of the fused kernels is reference edge (r + m) % 3
    for (int l = 0; l < 3; ++l)
      if (cn[d.tr_v[3 * ko + l] - 1] == t / 2) r = l;
    if (r < 0) return set_error(IMX_EINVAL, "imx_swm_set_scheme: triangle owner not found");
    const double *p = d.tr_c + 3 * ko;
    for (int m = 0; m < 3; ++m) {
      const int lr = (r + m) % 3;
      const int eo = d.tr_ed[3 * ko + lr] - 1;
      const double sg = hc ? (double)hct[(size_t)lr * NT + t] : (double)d.tr_tg[3 * ko + lr];
      const double *pe = hc ? d.edhx_c + 3 * eo : d.ed_c + 3 * eo;
      const double f = sg * d.ed_leng[eo] / d.tr_areag[ko];
      const double v[3] = {pe[0] - p[0], pe[1] - p[1], pe[2] - p[2]};
      double *o = &w[4 * ((size_t)m * NT + t)];
      o[0] = f * (p[1] * v[2] - p[2] * v[1]);
      o[1] = f * (p[2] * v[0] - p[0] * v[2]);
      o[2] = f * (p[0] * v[1] - p[1] * v[0]);
    }
  }
  for (int eo = 0; eo < d.ne; ++eo) {
    const int e = en[eo];
    const int k1 = d.ed_sh[2 * eo] - 1, k2 = d.ed_sh[2 * eo + 1] - 1;
    et[e] = tn[k1];
    et[(size_t)NE + e] = tn[k2];
    if (hc) {
      const double *t1 = d.tr_c + 3 * k1, *t2 = d.tr_c + 3 * k2, *hxc = d.edhx_c + 3 * eo;
      double hnr[3], htg[3];
      cross3(t1, t2, hnr);
      double n = norm3(hnr);
      for (int q = 0; q < 3; ++q) hnr[q] = hnr[q] / n;
      cross3(hnr, hxc, htg);
      n = norm3(htg);
      for (int q = 0; q < 3; ++q) nn[(size_t)4 * e + q] = htg[q] / n;
      ew[e] = 0.5;
      ew[(size_t)NE + e] = 0.5;
    } else {
      const double d1 = arclen(d.ed_c + 3 * eo, d.tr_c + 3 * k1);
      const double d2 = arclen(d.ed_c + 3 * eo, d.tr_c + 3 * k2);
      ew[e] = d2 / (d1 + d2);
      ew[(size_t)NE + e] = d1 / (d1 + d2);
      double nr[3];
      cross3(d.v_p + 3 * (d.ed_v[2 * eo] - 1), d.v_p + 3 * (d.ed_v[2 * eo + 1] - 1), nr);
      const double n = norm3(nr);
      for (int q = 0; q < 3; ++q) nr[q] = -nr[q] / n;
      put_projected(&nn[(size_t)4 * e], nr, d.ed_c + 3 * eo);
    }
  }
  Alloc &A = s->A;
  if (!s->x_dt_w) {
    s->x_dt_w = A.get<T>((size_t)4 * 3 * NT);
    s->x_dt_vhq = A.get<T>((size_t)4 * NT);
    s->x_dt_ew = A.get<T>((size_t)2 * NE);
    s->x_dt_n = A.get<T>((size_t)4 * NE);
    s->x_dt_et = A.get<int32_t>((size_t)2 * NE);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
  }
  const std::vector<T> wT = cvt<T>(w), ewT = cvt<T>(ew), nT = cvt<T>(nn);
  HIPTRY(hipMemcpy(s->x_dt_w, wT.data(), wT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_dt_ew, ewT.data(), ewT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_dt_n, nT.data(), nT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_dt_et, et.data(), et.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  s->f_dtred = true;
  return IMX_OK;
}

// Gassmann's kinetic energy on the fused kernels (on = false switches it
// off): the (circumcentre, edge) sub-areas over A_t in fast_tri's slot order,
// and the per-stage K_tr, typed T
template <typename T>
static int fast_gassk(imx_swm *s, const imx_mesh_desc &d, bool on, double gcoef) {
  s->f_gk = false;
  if (!on) return IMX_OK;
  const int NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &tn = s->lay.tri_new;
  std::vector<double> kw((size_t)3 * NT, 0.0);
  for (int ko = 0; ko < d.nt; ++ko) {
    const int t = tn[ko];
    int r = -1;  // slot m of the fused kernels is reference edge (r + m) % 3
    for (int l = 0; l < 3; ++l)
      if (cn[d.tr_v[3 * ko + l] - 1] == t / 2) r = l;
    if (r < 0) return set_error(IMX_EINVAL, "imx_swm_set_scheme: triangle owner not found");
    for (int m = 0; m < 3; ++m) kw[(size_t)m * NT + t] = gass_subarea(d, ko, (r + m) % 3) / d.tr_areag[ko];
  }
  Alloc &A = s->A;
  if (!s->x_gk_w) {
    s->x_gk_w = A.get<T>((size_t)3 * NT);
    s->x_gk_k = A.get<T>(NT);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
  }
  const std::vector<T> kwT = cvt<T>(kw);
  HIPTRY(hipMemcpy(s->x_gk_w, kwT.data(), kwT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  s->gk_coef = gcoef;
  s->f_gk = true;
  return IMX_OK;
}

// gass18 on the fused kernels (mode 1 HTC, 2 HC; 0 switches it off): the
// sub-triangle areas in fast_tri's slot order with their sum, the triangle
// areas, the two triangles of each edge slot and the outer edge of each
// ring triangle of each cell, in the ownership numbering, typed T
template <typename T>
static int fast_gass18(imx_swm *s, const imx_mesh_desc &d, int mode, bool qwd) {
  if (!mode) {
    s->f_gass = 0;
    s->f_gqw = false;
    return IMX_OK;
  }
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  std::vector<double> gw((size_t)4 * NT, 0.0), ga(NT, 1.0);
  for (int ko = 0; ko < d.nt; ++ko) {
    const int t = tn[ko];
    int r = -1;  // slot m of the fused kernels is reference edge (r + m) % 3
    for (int l = 0; l < 3; ++l)
      if (cn[d.tr_v[3 * ko + l] - 1] == t / 2) r = l;
    if (r < 0) return set_error(IMX_EINVAL, "imx_swm_set_scheme: triangle owner not found");
    double sw = 0.0;
    for (int m = 0; m < 3; ++m) {
      const double w = gass_subarea(d, ko, (r + m) % 3);
      gw[(size_t)m * NT + t] = w;
      sw = sw + w;
    }
    gw[(size_t)3 * NT + t] = sw;
    ga[t] = d.tr_areag[ko];
  }
  std::vector<int32_t> et((size_t)2 * NE, -1), out((size_t)MX * nv, 0);
  for (int eo = 0; eo < d.ne; ++eo) {
    et[en[eo]] = tn[d.ed_sh[2 * eo] - 1];
    et[(size_t)NE + en[eo]] = tn[d.ed_sh[2 * eo + 1] - 1];
  }
  for (int co = 0; co < nv; ++co) {
    const int n = d.v_nnb[co];
    int first = 0;
    for (int j = 0; j < n; ++j) {
      // ring triangle j lies between ring edges j and j+1: its third edge
      const int ko = d.v_tr[MX * co + j] - 1;
      const int e1 = d.v_ed[MX * co + j] - 1, e2 = d.v_ed[MX * co + (j + 1) % n] - 1;
      int e3 = -1;
      for (int l = 0; l < 3; ++l) {
        const int e = d.tr_ed[3 * ko + l] - 1;
        if (e != e1 && e != e2) e3 = e;
      }
      if (e3 < 0) return set_error(IMX_EINVAL, "imx_swm_set_scheme: ring triangle without an outer edge");
      out[(size_t)j * nv + cn[co]] = en[e3];
      if (j == 0) first = en[e3];
    }
    for (int j = n; j < MX; ++j) out[(size_t)j * nv + cn[co]] = first;  // a pentagon's padded slot
  }
  Alloc &A = s->A;
  if (!s->x_g_w) {
    s->x_g_w = A.get<T>((size_t)4 * NT);
    s->x_g_a = A.get<T>(NT);
    s->x_g_E = A.get<T>(NT);
    s->x_g_H = A.get<T>(NT);
    s->x_g_q = A.get<T>(NE);
    s->x_g_et = A.get<int32_t>((size_t)2 * NE);
    s->x_g_out = A.get<int32_t>((size_t)MX * nv);
    if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
  }
  s->f_gqw = false;
  if (qwd && (!d.ed_c || !d.tr_c)) return set_error(IMX_EINVAL, "imx_swm_set_scheme: bary needs ed_c and tr_c");
  if (qwd) {
    // the gass Coriolis term over bary-interpolated PV (HTC): q_e =
    // (d2 q(sh1) + d1 q(sh2)) / (d1 + d2), d_k the arc from ed%c to the
    // circumcentre of sh(k) (scalar_tr2ed, swm_operators.f90:165-169)
    std::vector<double> qw((size_t)2 * NE, 0.0);
    for (int eo = 0; eo < d.ne; ++eo) {
      const double d1 = arclen(d.ed_c + 3 * eo, d.tr_c + 3 * (d.ed_sh[2 * eo] - 1));
      const double d2 = arclen(d.ed_c + 3 * eo, d.tr_c + 3 * (d.ed_sh[2 * eo + 1] - 1));
      qw[en[eo]] = d2 / (d1 + d2);
      qw[(size_t)NE + en[eo]] = d1 / (d1 + d2);
    }
    if (!s->x_g_qw) {
      s->x_g_qw = A.get<T>((size_t)2 * NE);
      if (A.err != hipSuccess) return hip_fail(A.err, "imx_swm_set_scheme");
    }
    const std::vector<T> qwT = cvt<T>(qw);
    HIPTRY(hipMemcpy(s->x_g_qw, qwT.data(), qwT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
    s->f_gqw = true;
  }
  const std::vector<T> gwT = cvt<T>(gw), gaT = cvt<T>(ga);
  HIPTRY(hipMemcpy(s->x_g_w, gwT.data(), gwT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_g_a, gaT.data(), gaT.size() * sizeof(T), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_g_et, et.data(), et.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  HIPTRY(hipMemcpy(s->x_g_out, out.data(), out.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  s->f_gass = mode;
  return IMX_OK;
}
}  // namespace imx

extern "C" {

int imx_swm_set_scheme(imx_swm *s, const imx_mesh_desc *d, const imx_scheme *sc) {
  if (!s || !d || !sc) return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: NULL argument");
  if (d->nv != s->nv || d->ne != s->ne || d->nt != s->nt)
    return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: not the handle's mesh");
  if (!(sc->k2 >= 0.0) || !(sc->k4 >= 0.0) || !std::isfinite(sc->k2) || !std::isfinite(sc->k4) ||
      !std::isfinite(sc->pvspar) || sc->diffus < IMX_COEF_CONST || sc->diffus > IMX_COEF_DIAM ||
      sc->hyperdiffus < IMX_COEF_CONST || sc->hyperdiffus > IMX_COEF_DIAM ||
      sc->wrap < IMX_WRAP_TRSK10 || sc->wrap > IMX_WRAP_NONE || (sc->nopv != 0 && sc->nopv != 1))
    return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: bad option value");
  int sint, cor, rec;
  double gcoef = 1.0;
  if (sc->wrap == IMX_WRAP_NONE) {
    if (sc->sinterp < IMX_SINT_TRSK || sc->sinterp > IMX_SINT_GASS || sc->coriolis < IMX_COR_TRSK ||
        sc->coriolis > IMX_COR_DTRED || sc->recon < IMX_REC_TRSK || sc->recon > IMX_REC_MELV ||
        (sc->recon == IMX_REC_GASS && !std::isfinite(sc->gasscoef)))
      return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: bad method value (sinterp, coriolis, recon, gasscoef)");
    sint = sc->sinterp, cor = sc->coriolis, rec = sc->recon;
    if (rec == IMX_REC_GASS) gcoef = sc->gasscoef;
  } else {
    imx::wrapper_methods(sc->wrap, sint, cor, rec);
  }
  if (sc->nopv && (cor == IMX_COR_PERED || cor == IMX_COR_DTRED))
    return imx::set_error(IMX_EINVAL,
                          "imx_swm_set_scheme: nopv with pered or dtred Coriolis (the reference stops: these "
                          "reconstructions need the PV)");
  if (sint == IMX_SINT_GASS && (sc->k2 > 0.0 || sc->k4 > 0.0))
    return imx::set_error(IMX_EINVAL,
                          "imx_swm_set_scheme: gass interpolation with diffusion (the reference stops: "
                          "scalar_hx2tr needs edge values for Gassmann's scheme)");
  const bool plain = sint == IMX_SINT_TRSK && cor == IMX_COR_TRSK && rec == IMX_REC_TRSK;
  if (!plain && (!d->ed_c || !d->tr_c || !d->v_p))
    return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: methods other than trsk need ed_c, tr_c and v_p");
  if (sc->stag != IMX_STAG_HTC && sc->stag != IMX_STAG_HC)
    return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: bad option value (stag)");
  if (sc->stag == IMX_STAG_HC && (!d->ed_c || !d->tr_c || !d->v_p || !d->edhx_c))
    return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: HC needs ed_c, edhx_c, tr_c and v_p");
  if ((sc->k2 > 0.0 || sc->k4 > 0.0) && (!d->v_tr || !d->tr_c || !d->v_nb))
    return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: diffusion needs v_tr, v_nb and tr_c");
  if ((sc->apvm || rec == IMX_REC_GASS) && (!d->ed_c || !d->tr_c || !d->v_p || !d->v_tr))
    return imx::set_error(IMX_EINVAL, "imx_swm_set_scheme: APVM and gass K need ed_c, tr_c, v_p and v_tr");
  const bool hc = sc->stag == IMX_STAG_HC;
  if (s->fast) {
    // the fused kernels take every combination of the scalar interpolation
    // trsk or bary, the Coriolis term trsk, pered or dtred and the kinetic
    // energy trsk, perhx, gass (gasscoef; dubos is gass with 0) or melv;
    // gass18's interpolation, Coriolis term or both (trsk's or bary's for
    // the other) with any of these kinetic energies; the level model with trsk10's methods
    struct Methods {
      bool gass = false, bary = false, pered = false, perhx = false, gk = false, melv = false, dtred = false;
      bool glink = false, gtrsk = false;  // gass Coriolis term; gass PV from trsk interpolation
    } fm;
    bool ok = false;
    if ((sint == IMX_SINT_GASS && (cor == IMX_COR_GASS || cor == IMX_COR_TRSK)) ||
        (sint == IMX_SINT_TRSK && cor == IMX_COR_GASS) || (sint == IMX_SINT_BARY && cor == IMX_COR_GASS)) {
      // gass18's rhombus PV with the trsk Coriolis sum, or the gass
      // Coriolis sum over the trsk average of the triangles' PV; fast_cell's
      // gass mode has no diffusion
      ok = sc->k2 <= 0.0 && sc->k4 <= 0.0;
      fm.gass = true;
      fm.glink = cor == IMX_COR_GASS;
      fm.gtrsk = sint != IMX_SINT_GASS;
      fm.bary = sint == IMX_SINT_BARY;
      fm.gk = rec == IMX_REC_GASS;
      fm.perhx = rec == IMX_REC_PERHX;
      fm.melv = rec == IMX_REC_MELV;
    } else if ((sint == IMX_SINT_TRSK || sint == IMX_SINT_BARY) &&
               (cor == IMX_COR_TRSK || cor == IMX_COR_PERED || cor == IMX_COR_DTRED)) {
      ok = true;
      fm.bary = sint == IMX_SINT_BARY;
      fm.pered = cor == IMX_COR_PERED;
      fm.dtred = cor == IMX_COR_DTRED;
      fm.perhx = rec == IMX_REC_PERHX;
      fm.gk = rec == IMX_REC_GASS;
      fm.melv = rec == IMX_REC_MELV;
    }
    const bool trsk3 = ok && !fm.gass && !fm.bary && !fm.pered && !fm.dtred && !fm.perhx && !fm.gk && !fm.melv;
    if (!ok || (sc->nopv && (!trsk3 || sc->apvm)))
      return imx::set_error(IMX_ENOTSUP,
                            "imx_swm_set_scheme: fused handles take the scalar interpolation trsk or bary, the "
                            "Coriolis term trsk, pered or dtred and any kinetic "
                            "energy, with APVM, diffusion and hyperdiffusion; gass18's interpolation, "
                            "Coriolis term or both (trsk's or bary's for the other) with any kinetic energy "
                            "and APVM; "
                            "HTC or HC; and the level model "
                            "with trsk10's methods (diffusion and hyperdiffusion allowed); other method "
                            "triples and the level "
                            "model with other options run on FP64_REF handles");
    int fw = fm.gass ? IMX_WRAP_GASS18 : (fm.bary && fm.pered && fm.perhx) ? IMX_WRAP_PXT16 : IMX_WRAP_TRSK10;
    // the tables of each option, in turn; every step rewrites the tables
    // and flags of its option whether it is on or off
    auto apply = [&](const imx_scheme &q, bool qhc, const Methods &m, double qgc) -> int {
      int rc;
      if (qhc && ((rc = imx::scheme_tables(s, d)) || (rc = imx::hc_tables(s, d, false, false)))) return rc;
      rc = s->single ? imx::fast_stagger<float>(s, *d, qhc) : imx::fast_stagger<double>(s, *d, qhc);
      if (rc) return rc;
      s->sc_hc = qhc;
      rc = s->single ? imx::fast_diffusion<float>(s, *d, q, m.bary) : imx::fast_diffusion<double>(s, *d, q, m.bary);
      if (rc) return rc;
      rc = s->single ? imx::fast_apvm<float>(s, *d, q) : imx::fast_apvm<double>(s, *d, q);
      if (rc) return rc;
      rc = s->single ? imx::fast_px<float>(s, *d, m.bary, m.pered, m.perhx)
                     : imx::fast_px<double>(s, *d, m.bary, m.pered, m.perhx);
      if (rc) return rc;
      rc = s->single ? imx::fast_gassk<float>(s, *d, m.gk, qgc) : imx::fast_gassk<double>(s, *d, m.gk, qgc);
      if (rc) return rc;
      // Melvin's kinetic energy: HC takes the Perot weights of the cell
      s->f_melv = 0;
      if (m.melv && qhc) {
        rc = s->single ? imx::fast_cell_weights<float>(s, *d) : imx::fast_cell_weights<double>(s, *d);
        if (rc) return rc;
      }
      s->f_melv = m.melv ? (qhc ? 2 : 1) : 0;
      rc = s->single ? imx::fast_dtred<float>(s, *d, m.dtred) : imx::fast_dtred<double>(s, *d, m.dtred);
      if (rc) return rc;
      const int gmode = m.gass ? (qhc ? 2 : 1) : 0;
      s->f_glink = m.glink;
      s->f_gtrsk = m.gtrsk;
      const bool qwd = m.gass && m.bary && !qhc;  // HC: the plain average
      return s->single ? imx::fast_gass18<float>(s, *d, gmode, qwd) : imx::fast_gass18<double>(s, *d, gmode, qwd);
    };
    imx_scheme r = *sc;
    r.wrap = fw;
    int rc = apply(r, hc, fm, gcoef);
    s->f_nopv = rc == IMX_OK && r.nopv != 0;
    std::string why;
    if (rc) {
      // a failure part way (device memory) would leave a mix of the old and
      // new options: put the handle back on the default scheme instead
      why = imx_last_error();
      const imx_scheme def{};
      (void)apply(def, false, Methods{}, 1.0);
      r = def, fw = IMX_WRAP_TRSK10, fm = Methods{}, gcoef = 1.0;
      imx::wrapper_methods(fw, sint, cor, rec);
    }
    s->sc_wrap = fw;
    s->sc_sint = sint, s->sc_cor = cor, s->sc_rec = rec, s->sc_gcoef = fm.gk ? gcoef : 1.0, s->sc_nopv = s->f_nopv;
    // the option flags size the Jacobian's stencil superset
    s->sc_apvm = r.apvm != 0;
    s->sc_diff = r.k2 != 0.0;
    s->sc_hyper = r.k4 != 0.0;
    s->sc = r;
    imx::drop_graph(s);
    if (rc) return imx::set_error(rc, why + " (the handle is back on the default scheme)");
    return IMX_OK;
  }
  const int nv = s->nv, NE = s->NE, NT = s->NT;
  const std::vector<int32_t> &cn = s->lay.cell_new, &en = s->lay.edge_new, &tn = s->lay.tri_new;
  if (sc->apvm || sc->k2 > 0.0 || sc->k4 > 0.0 || !plain || sc->nopv || hc) {
    const int rc = imx::scheme_tables(s, d);
    if (rc) return rc;
  }
  const bool tg = cor == IMX_COR_PERED || cor == IMX_COR_DTRED;
  if (hc) {
    const int rc = imx::hc_tables(s, d, sint == IMX_SINT_BARY, tg);
    if (rc) return rc;
  }
  // intersection areas from the mesh, or computed here when it has none
  std::vector<double> trhx_v, hxtr_v, trskw_v;
  const double *trhx = d->tr_trhx_area, *hxtr = d->hx_hxtr_area;
  auto areas = [&]() {
    if (trhx && hxtr) return;
    trhx_v.assign((size_t)3 * d->nt, 0.0);
    hxtr_v.assign((size_t)MX * d->nv, 0.0);
    trskw_v.assign((size_t)MX * MX * d->nv, 0.0);
    imx::gridprop(*d, trhx_v.data(), hxtr_v.data(), trskw_v.data());
    trhx = trhx_v.data();
    hxtr = hxtr_v.data();
  };
  if (sint == IMX_SINT_BARY) {
    // barycentric coordinates of the circumcentres (initialize_gridprop,
    // swm.f90:1021-1026)
    std::vector<double> tw((size_t)3 * NT, 0.0);
    for (int ko = 0; ko < d->nt; ++ko) {
      double b[3];
      imx::bary_circumcentre(*d, ko, b);
      for (int i = 0; i < 3; ++i) tw[(size_t)i * NT + tn[ko]] = b[i];
    }
    HIPTRY(hipMemcpy(s->t_w, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  } else if (sint == IMX_SINT_GASS) {
    // areas of the sub-triangles (circumcentre, edge)
    std::vector<double> tw((size_t)3 * NT, 0.0);
    for (int ko = 0; ko < d->nt; ++ko)
      for (int i = 0; i < 3; ++i) tw[(size_t)i * NT + tn[ko]] = imx::gass_subarea(*d, ko, i);
    HIPTRY(hipMemcpy(s->t_w, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  }
  if (sint == IMX_SINT_BARY || cor == IMX_COR_DTRED || (!hc && tg)) {
    // edge-to-circumcentre distances (bary q at edges, dtred) and unit edge
    // normals ed%nr (pered and dtred, HTC)
    std::vector<double> dw((size_t)2 * NE, 1.0), nr((size_t)3 * NE, 0.0);
    for (int eo = 0; eo < d->ne; ++eo) {
      const int e = en[eo];
      for (int k = 0; k < 2; ++k)
        dw[(size_t)k * NE + e] = imx::arclen(d->ed_c + 3 * eo, d->tr_c + 3 * (d->ed_sh[2 * eo + k] - 1));
      double c[3];
      imx::cross3(d->v_p + 3 * (d->ed_v[2 * eo] - 1), d->v_p + 3 * (d->ed_v[2 * eo + 1] - 1), c);
      const double n = imx::norm3(c);
      for (int q = 0; q < 3; ++q) nr[(size_t)q * NE + e] = c[q] / n;
    }
    HIPTRY(hipMemcpy(s->e_dw, dw.data(), dw.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
    if (!hc && tg)  // HC holds edhx%tg there (set above)
      HIPTRY(hipMemcpy(s->e_nr, nr.data(), nr.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  }
  if (cor == IMX_COR_GASS) {
    // the linking edge of each TRiSK term (gethxedgeconnection)
    std::vector<int32_t> e3((size_t)2 * MX * NE, -1);
    for (int eo = 0; eo < d->ne; ++eo)
      for (int ic = 0; ic < 2; ++ic) {
        const int co = d->ed_v[2 * eo + ic] - 1;
        for (int j = 0; j < d->v_nnb[co]; ++j) {
          const int c3 = imx::hxedgeconn(*d, eo, d->v_ed[MX * co + j] - 1);
          e3[((size_t)ic * MX + j) * NE + en[eo]] = c3 >= 0 ? en[c3] : -1;
        }
      }
    HIPTRY(hipMemcpy(s->e_ed3, e3.data(), e3.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
  }
  if (rec == IMX_REC_GASS) {
    // kinetic_energy_tr's areas and each cell's ring triangles with their
    // hxtr_area (kinetic_energy_hx, swm_operators.f90:1198-1207)
    areas();
    std::vector<double> ka((size_t)3 * NT, 0.0), hw((size_t)MX * nv, 0.0);
    std::vector<int32_t> ct((size_t)MX * nv, 0);
    for (int ko = 0; ko < d->nt; ++ko)
      for (int i = 0; i < 3; ++i) ka[(size_t)i * NT + tn[ko]] = imx::gass_subarea(*d, ko, i);
    for (int co = 0; co < nv; ++co)
      for (int j = 0; j < d->v_nnb[co]; ++j) {
        ct[(size_t)j * nv + cn[co]] = tn[d->v_tr[MX * co + j] - 1];
        hw[(size_t)j * nv + cn[co]] = hxtr[MX * co + j];
      }
    HIPTRY(hipMemcpy(s->t_ka, ka.data(), ka.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
    HIPTRY(hipMemcpy(s->c_tr, ct.data(), ct.size() * sizeof(int32_t), hipMemcpyHostToDevice), "set_scheme");
    HIPTRY(hipMemcpy(s->c_hxtr, hw.data(), hw.size() * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  }
  if (sc->k2 > 0.0 || sc->k4 > 0.0) {
    areas();
    std::vector<double> hx, tr;
    imx::diffusion_coef(*d, trhx, sint == IMX_SINT_BARY, sc->k2, sc->diffus, sc->k4, sc->hyperdiffus, hx, tr);
    std::vector<double> hi(nv, 0.0), ti(NT, 0.0);
    for (int co = 0; co < nv; ++co) hi[cn[co]] = hx[co];
    for (int ko = 0; ko < d->nt; ++ko) ti[tn[ko]] = tr[ko];
    HIPTRY(hipMemcpy(s->dif_hx, hi.data(), nv * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
    HIPTRY(hipMemcpy(s->dif_tr, ti.data(), NT * sizeof(double), hipMemcpyHostToDevice), "set_scheme");
  }
  s->sc = *sc;
  s->sc_apvm = sc->apvm != 0;
  s->sc_diff = sc->k2 != 0.0;
  s->sc_hyper = sc->k4 != 0.0;
  s->sc_wrap = sc->wrap;
  s->sc_sint = sint, s->sc_cor = cor, s->sc_rec = rec, s->sc_gcoef = gcoef, s->sc_nopv = sc->nopv != 0;
  s->sc_hc = hc;
  imx::drop_graph(s);  // the captured step holds the previous kernel sequence
  return IMX_OK;
}

int imx_swm_step(imx_swm *s, int32_t nsteps) {
  if (!s) return imx::set_error(IMX_EINVAL, "imx_swm_step: NULL handle");
  if (nsteps < 0) return imx::set_error(IMX_EINVAL, "imx_swm_step: nsteps < 0");
  int rc = imx::run_steps(s, nsteps);
  if (rc) return rc;
  return imx::check_flag(s, "imx_swm_step");
}

int imx_swm_step_timed(imx_swm *s, int32_t nsteps, double *ms) {
  if (!s || !ms) return imx::set_error(IMX_EINVAL, "imx_swm_step_timed: NULL argument");
  if (nsteps < 0) return imx::set_error(IMX_EINVAL, "imx_swm_step_timed: nsteps < 0");
  if (s->o.use_graph) {
    int rc = imx::ensure_graph(s);  // instantiation stays outside the timed region
    if (rc) return rc;
  }
  HIPTRY(hipEventRecord(s->ev0, s->stream), "event");
  int rc = imx::run_steps(s, nsteps);
  if (rc) return rc;
  HIPTRY(hipEventRecord(s->ev1, s->stream), "event");
  HIPTRY(hipEventSynchronize(s->ev1), "event");
  float f = 0.f;
  HIPTRY(hipEventElapsedTime(&f, s->ev0, s->ev1), "event");
  *ms = f;
  return imx::check_flag(s, "imx_swm_step_timed");
}

static const char *kFastNames[3] = {"fast_tri", "fast_cell", "fast_edge"};

int imx_swm_profile(imx_swm *s, int32_t nsteps, double *ms, int32_t *nk) {
  if (!s || !ms || !nk) return imx::set_error(IMX_EINVAL, "imx_swm_profile: NULL argument");
  if (!s->fast) return imx::set_error(IMX_ENOTSUP, "imx_swm_profile: FAST precisions only");
  std::vector<hipEvent_t> ev(16);
  for (auto &e : ev) HIPTRY(hipEventCreate(&e), "event");
  double acc[3] = {0, 0, 0};
  for (int k = 0; k < nsteps; ++k) {
    if (s->single) imx::launch_fast_step<float>(s, &ev);
    else imx::launch_fast_step<double>(s, &ev);
    HIPTRY(hipEventSynchronize(ev[15]), "event");
    for (int st = 0; st < 4; ++st)
      for (int j = 0; j < 3; ++j) {
        float f = 0.f;
        HIPTRY(hipEventElapsedTime(&f, ev[4 * st + j], ev[4 * st + j + 1]), "event");
        acc[j] += f;
      }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  for (int j = 0; j < 3; ++j) ms[j] = acc[j] / (4.0 * (nsteps > 0 ? nsteps : 1));
  *nk = 3;
  return imx::check_flag(s, "imx_swm_profile");
}

const char *imx_swm_kernel_name(imx_swm *s, int32_t k) {
  (void)s;
  if (k < 0 || k > 2) return "";
  return kFastNames[k];
}

int imx_swm_diagnostics(imx_swm *s, double out[IMX_NDIAG]) {
  if (!s || !out) return imx::set_error(IMX_EINVAL, "imx_swm_diagnostics: NULL argument");
  return imx::diag_range(s, 0, s->nv, out);
}

// ---- diagnostics without allocation -------------------------------------
// The block partials of the diagnostics reductions live in a file-scope
// device array (one per GPU, held until the process ends) instead of the
// handle's buffers; its device address is looked up once per device with
// hipGetSymbolAddress. Calls that overlap on one GPU share the array and race.
#define IMX_DIAG_SCRATCH_DOUBLES (1 << 19)  // 4 MiB: partials of up to 65,535 blocks
static __device__ __align__(256) double scratch[IMX_DIAG_SCRATCH_DOUBLES];

namespace imx {
static std::mutex g_scratch_mu;
static double *g_scratch_dev[64] = {};

static int scratch_address(double **out) {
  int dev = 0;
  HIPTRY(hipGetDevice(&dev), "diagnostics scratch");
  if (dev < 0 || dev >= 64) return set_error(IMX_EDEVICE, "diagnostics scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    HIPTRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "diagnostics scratch");
    g_scratch_dev[dev] = (double *)p;
  }
  *out = g_scratch_dev[dev];
  return IMX_OK;
}
}  // namespace imx

// The host finish of diag_range, on the device: mass, energies (with the
// topography's sums), potential enstrophy, RMS divergence, bad-state flag.
__global__ void diag_finish(const double *__restrict__ v, double grav, int has_topo, double sb, double sb2,
                            double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double mass = v[0] / v[1];
  const double pe = v[2];
  const double ke = v[4];
  double upe;
  if (has_topo) {
    const double m = (v[0] + sb) / v[1];
    upe = 0.5 * grav * (m * m * v[1] - sb2);
  } else {
    upe = v[1] * grav * mass * (mass * 0.5);
  }
  out[IMX_DIAG_MASS] = mass;
  out[IMX_DIAG_PENERGY] = pe;
  out[IMX_DIAG_KENERGY] = ke;
  out[IMX_DIAG_TENERGY] = pe + ke;
  out[IMX_DIAG_AENERGY] = (pe + ke) - upe;
  out[IMX_DIAG_PENSTROPHY] = v[5];
  out[IMX_DIAG_RMSDIV] = sqrt(v[3] / v[1]);
  out[IMX_DIAG_BADFLAG] = v[7] > 0 ? 1.0 : 0.0;
}

int imx_swm_diagnostics_dev(imx_swm *s, double *out_dev) {
  if (!s || !out_dev) return imx::set_error(IMX_EINVAL, "imx_swm_diagnostics_dev: NULL argument");
  const size_t need = (size_t)(s->nparts + 1) * NPART;
  if (need > (size_t)IMX_DIAG_SCRATCH_DOUBLES)
    return imx::set_error(IMX_EINVAL, "imx_swm_diagnostics_dev: the mesh needs " + std::to_string(need) +
                                          " scratch doubles, more than the " +
                                          std::to_string(IMX_DIAG_SCRATCH_DOUBLES) + " it holds (about " +
                                          std::to_string(IMX_DIAG_SCRATCH_DOUBLES / 8 * 256 / 6) + " cells)");
  double *sc = nullptr;
  int rc = imx::scratch_address(&sc);
  if (rc) return rc;
  // every block writes all of its partials, so nothing needs clearing
  double *part = sc, *diag = sc + (size_t)s->nparts * NPART;
  imx::launch_diag_state(s, 0, s->nv, part, diag);
  diag_finish<<<1, 64, 0, s->stream>>>(diag, s->o.grav, s->has_topo ? 1 : 0, s->topo_sb, s->topo_sb2, out_dev);
  HIPTRY(hipGetLastError(), "imx_swm_diagnostics_dev");
  return IMX_OK;
}

int imx_swm_info(imx_swm *s, int64_t out[8]) {
  if (!s || !out) return imx::set_error(IMX_EINVAL, "imx_swm_info: NULL argument");
  out[0] = s->nv;
  out[1] = s->ne;
  out[2] = s->nt;
  out[3] = (int64_t)s->A.bytes;
  out[4] = s->o.precision;
  out[5] = s->o.use_sfc;
  out[6] = s->o.use_graph;
  out[7] = 0;
  return IMX_OK;
}

int imx_copy_probe(int64_t bytes, int32_t reps, double *gbs) {
  if (!gbs || bytes < 16 || reps < 1) return imx::set_error(IMX_EINVAL, "imx_copy_probe: bad argument");
  const size_t n = (size_t)bytes / 16;
  probe4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = IMX_OK;
  float ms = 0.f;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    rc = imx::set_error(IMX_ENOMEM, "imx_copy_probe: hipMalloc failed");
  } else if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
             hipEventCreate(&e1) != hipSuccess) {
    rc = imx::set_error(IMX_EDEVICE, "imx_copy_probe: setup failed");
  } else {
    // the best of a few launch shapes: 8192 or 32768 blocks (32 or 128 per
    // CU), cached or non-temporal accesses, 4 loads in flight per lane
    double best = 0.0;
    for (int shape = 0; shape < 4 && rc == IMX_OK; ++shape) {
      const int blocks = (shape & 1) ? 32768 : 8192;
      auto launch = [&]() {
        if (shape < 2) copy_probe<4, false><<<blocks, TPB>>>(a, b, n);
        else copy_probe<4, true><<<blocks, TPB>>>(a, b, n);
      };
      launch();
      (void)hipEventRecord(e0, nullptr);
      for (int r = 0; r < reps; ++r) launch();
      (void)hipEventRecord(e1, nullptr);
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        rc = imx::set_error(IMX_EDEVICE, "imx_copy_probe: kernel failed");
      else
        best = std::max(best, 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9);
    }
    *gbs = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

void imx_swm_destroy(imx_swm *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  imx::drop_graph(s);
  s->A.release();
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int imx_swm_tune(imx_swm *s, int32_t nsteps, int32_t out[6]) {
  if (!s) return imx::set_error(IMX_EINVAL, "imx_swm_tune: NULL handle");
  return imx::tune(s, nsteps, out);
}

}  // extern "C"

// ------------------------------------------------------------- ensembles --
// n independent instances as one disjoint-union mesh: cells ordered by
// (instance, Hilbert key), so every instance owns contiguous cell, edge and
// triangle ranges; per-cell/per-edge time steps let each instance run at its
// own dt while one launch per fused kernel advances all of them.
struct imx_ens {
  imx_swm *s = nullptr;
  int n = 0;
  std::vector<int32_t> nv, ne, nt, cv0, ce0, c0;  // counts, union caller offsets, internal cell offset
  std::vector<double> area;                       // union cell areas (caller order)
};

namespace imx {

template <typename T>
static int ens_upload(imx_ens *E, int i, const double *h, const double *u) {
  imx_swm *s = E->s;
  const int nc = E->nv[i], c0 = E->c0[i];
  std::vector<T> hb(nc, (T)0), ub((size_t)3 * nc, (T)0);
  for (int k = 0; k < nc; ++k) hb[s->lay.cell_new[E->cv0[i] + k] - c0] = (T)(h[k] - s->href);
  for (int l = 0; l < E->ne[i]; ++l) ub[s->lay.edge_new[E->ce0[i] + l] - 3 * c0] = (T)u[l];
  HIPTRY(hipMemcpyAsync((T *)s->h0 + c0, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice,
                        s->stream), "imx_ens_set_state");
  HIPTRY(hipMemcpyAsync((T *)s->u0 + (size_t)3 * c0, ub.data(), ub.size() * sizeof(T),
                        hipMemcpyHostToDevice, s->stream), "imx_ens_set_state");
  HIPTRY(hipStreamSynchronize(s->stream), "imx_ens_set_state");
  return IMX_OK;
}

template <typename T>
static int ens_download(imx_ens *E, int i, double *h, double *u) {
  imx_swm *s = E->s;
  const int nc = E->nv[i], c0 = E->c0[i];
  std::vector<T> hb(nc), ub((size_t)3 * nc);
  HIPTRY(hipMemcpyAsync(hb.data(), (const T *)s->h0 + c0, hb.size() * sizeof(T),
                        hipMemcpyDeviceToHost, s->stream), "imx_ens_get_state");
  HIPTRY(hipMemcpyAsync(ub.data(), (const T *)s->u0 + (size_t)3 * c0, ub.size() * sizeof(T),
                        hipMemcpyDeviceToHost, s->stream), "imx_ens_get_state");
  HIPTRY(hipStreamSynchronize(s->stream), "imx_ens_get_state");
  if (h)
    for (int k = 0; k < nc; ++k) h[k] = (double)hb[s->lay.cell_new[E->cv0[i] + k] - c0] + s->href;
  if (u)
    for (int l = 0; l < E->ne[i]; ++l) u[l] = (double)ub[s->lay.edge_new[E->ce0[i] + l] - 3 * c0];
  return IMX_OK;
}

}  // namespace imx

extern "C" {

int imx_ens_create(int32_t n, const imx_mesh_desc *descs, const double *dt, const imx_swm_opts *o,
                   imx_ens **out) {
  if (!out || !descs || !dt || !o || n < 1) return imx::set_error(IMX_EINVAL, "imx_ens_create: bad argument");
  *out = nullptr;
  if (o->precision == IMX_PREC_FP64_REF)
    return imx::set_error(IMX_ENOTSUP, "imx_ens_create: ensembles run the fused FP32/FP64 kernels");
  imx_ens *E = new (std::nothrow) imx_ens();
  if (!E) return imx::set_error(IMX_ENOMEM, "imx_ens_create: out of memory");
  int rc = IMX_OK;
  try {
    E->n = n;
    int64_t NV = 0, NEd = 0, NTr = 0;
    bool gp = true;
    for (int i = 0; i < n; ++i) {
      int r = imx::validate_desc(descs[i]);
      if (r) {
        delete E;
        return r;
      }
      E->nv.push_back(descs[i].nv);
      E->ne.push_back(descs[i].ne);
      E->nt.push_back(descs[i].nt);
      E->cv0.push_back((int32_t)NV);
      E->ce0.push_back((int32_t)NEd);
      NV += descs[i].nv;
      NEd += descs[i].ne;
      NTr += descs[i].nt;
      gp = gp && descs[i].tr_trhx_area && descs[i].hx_hxtr_area && descs[i].hx_trskw;
    }
    if (NV >= (1 << 29)) {
      delete E;
      return imx::set_error(IMX_EINVAL, "imx_ens_create: too many cells in total");
    }
    const int MXn = IMX_MAXNB;
    std::vector<int32_t> v_nnb(NV), v_nb(MXn * NV), v_ed(MXn * NV), v_tr(MXn * NV), ttg(MXn * NV),
        tnr(MXn * NV), ed_v(2 * NEd), ed_sh(2 * NEd), tr_v(3 * NTr), tr_ed(3 * NTr), tr_tg(3 * NTr),
        group(NV);
    std::vector<double> v_p(3 * NV), areag(NV), leng(NEd), dleng(NEd), ed_c(3 * NEd), edhx_c(3 * NEd),
        tr_area(NTr), tr_c(3 * NTr), tr_lat(NTr), dtc(NV), dte(NEd), trhx, hxtr, trskw;
    if (gp) {
      trhx.resize(3 * NTr);
      hxtr.resize(MXn * NV);
      trskw.resize((size_t)MXn * MXn * NV);
    }
    int64_t ov = 0, oe = 0, ot = 0;
    for (int i = 0; i < n; ++i) {
      const imx_mesh_desc &d = descs[i];
      for (int c = 0; c < d.nv; ++c) {
        v_nnb[ov + c] = d.v_nnb[c];
        group[ov + c] = i;
        dtc[ov + c] = dt[i];
        areag[ov + c] = d.hx_areag[c];
        for (int q = 0; q < 3; ++q) v_p[3 * (ov + c) + q] = d.v_p[3 * c + q];
        for (int j = 0; j < MXn; ++j) {
          const int64_t k = MXn * (ov + c) + j;
          const bool on = j < d.v_nnb[c];
          v_nb[k] = on ? d.v_nb[MXn * c + j] + (int32_t)ov : 0;
          v_ed[k] = on ? d.v_ed[MXn * c + j] + (int32_t)oe : 0;
          v_tr[k] = on ? d.v_tr[MXn * c + j] + (int32_t)ot : 0;
          ttg[k] = d.hx_ttgout[MXn * c + j];
          tnr[k] = d.hx_tnrccw[MXn * c + j];
        }
        if (gp) {
          for (int j = 0; j < MXn; ++j) hxtr[MXn * (ov + c) + j] = d.hx_hxtr_area[MXn * c + j];
          for (int j = 0; j < MXn * MXn; ++j) trskw[(size_t)MXn * MXn * (ov + c) + j] = d.hx_trskw[MXn * MXn * c + j];
        }
      }
      for (int l = 0; l < d.ne; ++l) {
        for (int k = 0; k < 2; ++k) {
          ed_v[2 * (oe + l) + k] = d.ed_v[2 * l + k] + (int32_t)ov;
          ed_sh[2 * (oe + l) + k] = d.ed_sh[2 * l + k] + (int32_t)ot;
        }
        leng[oe + l] = d.ed_leng[l];
        dleng[oe + l] = d.edhx_leng[l];
        dte[oe + l] = dt[i];
        for (int q = 0; q < 3; ++q) {
          ed_c[3 * (oe + l) + q] = d.ed_c[3 * l + q];
          edhx_c[3 * (oe + l) + q] = d.edhx_c[3 * l + q];
        }
      }
      for (int k = 0; k < d.nt; ++k) {
        for (int q = 0; q < 3; ++q) {
          tr_v[3 * (ot + k) + q] = d.tr_v[3 * k + q] + (int32_t)ov;
          tr_ed[3 * (ot + k) + q] = d.tr_ed[3 * k + q] + (int32_t)oe;
          tr_tg[3 * (ot + k) + q] = d.tr_tg[3 * k + q];
          tr_c[3 * (ot + k) + q] = d.tr_c[3 * k + q];
          if (gp) trhx[3 * (ot + k) + q] = d.tr_trhx_area[3 * k + q];
        }
        tr_area[ot + k] = d.tr_areag[k];
        tr_lat[ot + k] = d.tr_clat[k];
      }
      ov += d.nv;
      oe += d.ne;
      ot += d.nt;
    }
    imx_mesh_desc U;
    std::memset(&U, 0, sizeof(U));
    U.nv = (int32_t)NV;
    U.ne = (int32_t)NEd;
    U.nt = (int32_t)NTr;
    U.glevel = -1;
    U.v_nnb = v_nnb.data();
    U.v_nb = v_nb.data();
    U.v_ed = v_ed.data();
    U.v_tr = v_tr.data();
    U.hx_ttgout = ttg.data();
    U.hx_tnrccw = tnr.data();
    U.v_p = v_p.data();
    U.hx_areag = areag.data();
    U.ed_v = ed_v.data();
    U.ed_sh = ed_sh.data();
    U.ed_leng = leng.data();
    U.edhx_leng = dleng.data();
    U.ed_c = ed_c.data();
    U.edhx_c = edhx_c.data();
    U.tr_v = tr_v.data();
    U.tr_ed = tr_ed.data();
    U.tr_tg = tr_tg.data();
    U.tr_areag = tr_area.data();
    U.tr_c = tr_c.data();
    U.tr_clat = tr_lat.data();
    if (gp) {
      U.tr_trhx_area = trhx.data();
      U.hx_hxtr_area = hxtr.data();
      U.hx_trskw = trskw.data();
    }
    // one layout per distinct mesh (instances sharing a mesh share arrays),
    // concatenated instance-major with offsets
    imx::Layout UL;
    UL.cell_new.resize(NV);
    UL.edge_new.resize(NEd);
    UL.tri_new.resize(NTr);
    UL.NE = 3 * (int)NV;
    UL.NT = 2 * (int)NV;
    {
      std::vector<const double *> keys;
      std::vector<imx::Layout> lays;
      int64_t c0 = 0;
      for (int i = 0; i < n; ++i) {
        int li = -1;
        for (size_t q = 0; q < keys.size(); ++q)
          if (keys[q] == descs[i].v_p) li = (int)q;
        if (li < 0) {
          imx::Layout L;
          std::string err;
          if (!imx::build_layout(descs[i], o->use_sfc != 0, L, err)) {
            rc = imx::set_error(IMX_EINVAL, "layout: " + err);
            break;
          }
          keys.push_back(descs[i].v_p);
          lays.push_back(std::move(L));
          li = (int)lays.size() - 1;
        }
        const imx::Layout &L = lays[li];
        for (int k = 0; k < descs[i].nv; ++k) UL.cell_new[E->cv0[i] + k] = (int32_t)c0 + L.cell_new[k];
        for (int l = 0; l < descs[i].ne; ++l) UL.edge_new[E->ce0[i] + l] = 3 * (int32_t)c0 + L.edge_new[l];
        int64_t t0 = 0;
        for (int j = 0; j < i; ++j) t0 += descs[j].nt;
        for (int k = 0; k < descs[i].nt; ++k) UL.tri_new[t0 + k] = 2 * (int32_t)c0 + L.tri_new[k];
        c0 += descs[i].nv;
      }
    }
    if (!rc) {
      E->s = new (std::nothrow) imx_swm();
      if (!E->s) rc = imx::set_error(IMX_ENOMEM, "imx_ens_create: out of memory");
      else rc = imx::create(U, *o, E->s, group.data(), dtc.data(), dte.data(), &UL);
    }
    if (!rc) {
      int64_t acc = 0;
      for (int i = 0; i < n; ++i) {
        E->c0.push_back((int32_t)acc);
        acc += E->nv[i];
      }
      E->area = areag;
    }
  } catch (const std::bad_alloc &) {
    rc = imx::set_error(IMX_ENOMEM, "imx_ens_create: host out of memory");
  }
  if (rc) {
    imx_ens_destroy(E);
    return rc;
  }
  *out = E;
  return IMX_OK;
}

int imx_ens_set_state(imx_ens *E, int32_t i, const double *h, const double *u) {
  if (!E || !h || !u || i < 0 || i >= E->n) return imx::set_error(IMX_EINVAL, "imx_ens_set_state: bad argument");
  imx_swm *s = E->s;
  if (!s->href_set) {  // FP32 storage reference: area mean of the first state set
    double sa = 0.0, sh = 0.0;
    for (int k = 0; k < E->nv[i]; ++k) {
      sa += E->area[E->cv0[i] + k];
      sh += E->area[E->cv0[i] + k] * h[k];
    }
    s->href = sh / sa;
    s->href_set = true;
  }
  return s->single ? imx::ens_upload<float>(E, i, h, u) : imx::ens_upload<double>(E, i, h, u);
}

int imx_ens_get_state(imx_ens *E, int32_t i, double *h, double *u) {
  if (!E || i < 0 || i >= E->n) return imx::set_error(IMX_EINVAL, "imx_ens_get_state: bad argument");
  return E->s->single ? imx::ens_download<float>(E, i, h, u) : imx::ens_download<double>(E, i, h, u);
}

int imx_ens_step(imx_ens *E, int32_t nsteps) {
  if (!E) return imx::set_error(IMX_EINVAL, "imx_ens_step: NULL handle");
  return imx_swm_step(E->s, nsteps);
}

int imx_ens_step_timed(imx_ens *E, int32_t nsteps, double *ms) {
  if (!E) return imx::set_error(IMX_EINVAL, "imx_ens_step_timed: NULL handle");
  return imx_swm_step_timed(E->s, nsteps, ms);
}

int imx_ens_diagnostics(imx_ens *E, int32_t i, double out[IMX_NDIAG]) {
  if (!E || !out || i < 0 || i >= E->n) return imx::set_error(IMX_EINVAL, "imx_ens_diagnostics: bad argument");
  return imx::diag_range(E->s, E->c0[i], E->c0[i] + E->nv[i], out);
}

int imx_ens_tune(imx_ens *E, int32_t nsteps, int32_t out[6]) {
  if (!E) return imx::set_error(IMX_EINVAL, "imx_ens_tune: NULL handle");
  return imx::tune(E->s, nsteps, out);
}

int imx_ens_info(imx_ens *E, int64_t out[8]) {
  if (!E || !out) return imx::set_error(IMX_EINVAL, "imx_ens_info: NULL argument");
  int rc = imx_swm_info(E->s, out);
  out[7] = E->n;
  return rc;
}

void imx_ens_destroy(imx_ens *E) {
  if (!E) return;
  imx_swm_destroy(E->s);
  delete E;
}

}  // extern "C"
