/*
 * swm_oracle.c -- TEST INFRASTRUCTURE ONLY: a plain-C, FP64 restatement of the
 * reference shallow-water step, used by tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py as the CHECKER. The product (libimx) never
 * links, loads or calls it.
 *
 * Every loop follows the reference's operation order so that, compiled with
 * -ffp-contract=off, results are bit-for-bit those of the reference Fortran
 * built with -ffp-contract=off on the same inputs (pinned against golden
 * vectors written by oracle/_ref/ref_harness, see tests/golden/).
 *
 * Scope: HTC (the reference default) or HC staggering, geodesic areas, the
 * trsk10 wrapper (the default) or the pxt16 / gass18 wrappers, optionally
 * with APVM, diffusion and hyperdiffusion (oracle_scheme), i.e. the branches
 *   scalar_hx2ed        src/swm_operators.f90:88-98
 *   scalar_elem_product src/swm_operators.f90:1546-1577
 *   vort_tr             src/swm_operators.f90:551-611
 *   scalar_hx2tr        src/swm_operators.f90:229-284 (trsk :258-263)
 *   scalar_tr2ed        src/swm_operators.f90:124-176 (:152-153)
 *   scalar_elem_divide  src/swm_operators.f90:1579-1615
 *   vector_edhx2hx_perot src/swm_operators.f90:454-509 (HTC :483-489)
 *   div_hx              src/swm_operators.f90:613-661
 *   kinetic_energy_hx   src/swm_operators.f90:1132-1217 (trsk :1162-1171)
 *   grad_ed             src/swm_operators.f90:511-548 (HTC :535-541)
 *   coriolis_ed         src/swm_operators.f90:1219-1424 (trsk :1283-1309)
 *   rel_vort_tr         src/swm_operators.f90:780-827
 *   grad_tr2ed          src/swm_operators.f90:1427-1463     (APVM)
 *   vector_edtr2tr_perot src/swm_operators.f90:408-452       (APVM)
 *   apvm                src/swm_operators.f90:1465-1502
 *   diffusion_ed        src/swm_operators.f90:855-895
 *   hyperdiffusion_ed   src/swm_operators.f90:930-975
 *   grad_ed_tg          src/swm_operators.f90:738-777 (HTC :760-769)
 *   alignment_coef, diameter_coef  src/swm_operators.f90:977-1072
 * pxt16: scalar_hx2tr bary (:265-269, bar_coord_tr src/smeshpack.f90:5200-5245),
 *   scalar_tr2ed distance-weighted (:163-168), kinetic_energy_hx perhx
 *   (:1172-1173), coriolis_ed pered HTC (:1330-1368)
 * HC: the useStagHC branches of scalar_hx2ed (:100-114, edhx%c%b of
 *   initialize_gridprop src/swm.f90:1007-1018, insidetrmesh and robdet
 *   src/smeshpack.f90:5329-5420, :6597-6629), vort_tr, div_hx, grad_ed,
 *   grad_ed_tg, rel_vort_tr, vector_edhx2hx_perot and coriolis_ed; the signs
 *   follow from hx%nr = hx%ttgout sign(ed%tg . edhx%nr)
 *   (src/smeshpack.f90:1085-1116)
 * gass18: scalar_edtr2trcc (:286-346), scalar_tr2ed rhombi (:155-162),
 *   q_ed = eta_ed / h_rhb (src/swm.f90:2469-2473), coriolis_ed gass
 *   (:1379-1416, gethxedgeconnection src/smeshpack.f90:4127-4244)
 *   alignind            src/smeshpack.f90:6398-6490
 *   tendency            src/swm.f90:2416-2598
 *   ode_rk4             src/swm.f90:2600-2683
 *   sumf_areas, sumfsq_areas, calc_energies  src/swm.f90:3394-3490, :2686-2752
 * Potential enstrophy is not in the reference; oracle_diagnostics defines it
 * as sum_t 1/2 A_t a^2 h_t q_t^2 (Ringler et al. 2010).
 */
#include "swm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MX IMX_MAXNB

/* epsilon(1.) of the reference (constants.f90:45) */
static const double EPS_SINGLE = 1.1920928955078125e-07;

static double arclen(const double *p, const double *q);

/* cross_product (src/smeshpack.f90:6560-6578) */
static void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot(const double *a, const double *b) {
  double s = 0.0;
  for (int q = 0; q < 3; ++q) s = s + a[q] * b[q];
  return s;
}

/* bar_coord_tr (src/smeshpack.f90:5200-5245) of point p in triangle k */
static void bary_at(const imx_mesh_desc *d, const double *p, int k, double *b) {
  const double *p1 = d->v_p + 3 * (d->tr_v[3 * k] - 1), *p2 = d->v_p + 3 * (d->tr_v[3 * k + 1] - 1),
               *p3 = d->v_p + 3 * (d->tr_v[3 * k + 2] - 1);
  double c[3];
  cross(p, p2, c);
  b[0] = fabs(dot(c, p3));
  cross(p, p3, c);
  b[1] = fabs(dot(c, p1));
  cross(p, p1, c);
  b[2] = fabs(dot(c, p2));
  const double area = b[0] + b[1] + b[2];
  for (int i = 0; i < 3; ++i) b[i] = b[i] / area;
}
static void bary(const imx_mesh_desc *d, int k, double *b) { bary_at(d, d->tr_c + 3 * k, k, b); }

/* sphtriarea (src/smeshpack.f90:4710-4762) */
static double sphtriarea(const double *p1, const double *p2, const double *p3) {
  double t[3];
  for (int q = 0; q < 3; ++q) t[q] = p1[q] - p2[q];
  const double n12 = sqrt(dot(t, t));
  for (int q = 0; q < 3; ++q) t[q] = p2[q] - p3[q];
  const double n23 = sqrt(dot(t, t));
  for (int q = 0; q < 3; ++q) t[q] = p1[q] - p3[q];
  const double n13 = sqrt(dot(t, t));
  const double eps100 = 1.1920928955078125e-07 / 100;
  if (n12 < eps100 || n23 < eps100 || n13 < eps100) return 0.0;
  const double a = arclen(p1, p2), b = arclen(p2, p3), c = arclen(p3, p1);
  const double s = (a + b + c) / 2;
  const double tmp = tan(s / 2) * tan((s - a) / 2) * tan((s - b) / 2) * tan((s - c) / 2);
  if (tmp < 0) return 0.0;
  return 4.0 * atan(sqrt(tmp));
}

/* ed%nr = v1 x v2 / |v1 x v2| (src/smeshpack.f90:725-726) */
static void ed_nr(const imx_mesh_desc *d, int l, double *nr) {
  const double *a = d->v_p + 3 * (d->ed_v[2 * l] - 1), *b = d->v_p + 3 * (d->ed_v[2 * l + 1] - 1);
  cross(a, b, nr);
  const double n = sqrt(dot(nr, nr));
  for (int q = 0; q < 3; ++q) nr[q] = nr[q] / n;
}

/* edhx%nr = tr(sh1)%c x tr(sh2)%c normalised, edhx%tg = edhx%nr x edhx%c
 * normalised (src/smeshpack.f90:1030-1034) */
static void edhx_nr(const imx_mesh_desc *d, int l, double *nr) {
  cross(d->tr_c + 3 * (d->ed_sh[2 * l] - 1), d->tr_c + 3 * (d->ed_sh[2 * l + 1] - 1), nr);
  const double n = sqrt(dot(nr, nr));
  for (int q = 0; q < 3; ++q) nr[q] = nr[q] / n;
}
static void edhx_tg(const imx_mesh_desc *d, int l, double *tg) {
  double nr[3];
  edhx_nr(d, l, nr);
  cross(nr, d->edhx_c + 3 * l, tg);
  const double n = sqrt(dot(tg, tg));
  for (int q = 0; q < 3; ++q) tg[q] = tg[q] / n;
}
/* ed%tg = ed%nr x ed%c normalised (src/smeshpack.f90:727-729) */
static void ed_tg(const imx_mesh_desc *d, int l, double *tg) {
  double nr[3];
  ed_nr(d, l, nr);
  cross(nr, d->ed_c + 3 * l, tg);
  const double n = sqrt(dot(tg, tg));
  for (int q = 0; q < 3; ++q) tg[q] = tg[q] / n;
}

/* robdet (src/smeshpack.f90:6597-6629): the determinant in binary128 */
static double robdet(const double *p1, const double *p2, const double *p3) {
  const __float128 q1[3] = {p1[0], p1[1], p1[2]}, q2[3] = {p2[0], p2[1], p2[2]},
                   q3[3] = {p3[0], p3[1], p3[2]};
  __float128 a[6];
  a[0] = q1[0] * q2[1] * q3[2];
  a[1] = q2[0] * q3[1] * q1[2];
  a[2] = q3[0] * q1[1] * q2[2];
  a[3] = -q3[0] * q2[1] * q1[2];
  a[4] = -q1[0] * q3[1] * q2[2];
  a[5] = -q2[0] * q1[1] * q3[2];
  const __float128 t = a[0] + a[1] + a[2] + a[3] + a[4] + a[5];
  return (double)t;
}

/* insidetrmesh (src/smeshpack.f90:5329-5420), 0-based triangle */
static int insidetrmesh(const imx_mesh_desc *d, const double *p, int tr) {
  const double tol = 1.1920928955078125e-07 / 10000.0;
  const double tol2 = 2.220446049250313e-16 * 100.0; /* epsilon(pi), pi double */
  for (int l = 0; l < 3; ++l) {
    const double *p1 = d->v_p + 3 * (d->tr_v[3 * tr + l] - 1);
    const double *p2 = d->v_p + 3 * (d->tr_v[3 * tr + (l + 1) % 3] - 1);
    double t[3];
    for (int q = 0; q < 3; ++q) t[q] = p[q] - p1[q];
    if (sqrt(dot(t, t)) < tol) return 1;
    double vn[3];
    cross(p1, p2, vn);
    const double vnn = sqrt(dot(vn, vn));
    for (int q = 0; q < 3; ++q) vn[q] = vn[q] / vnn;
    const double determ = dot(vn, p);
    if (determ < -tol) return 0;
    if (determ > tol) continue;
    const double determ2 = robdet(p, p1, p2);
    if (fabs(determ2) < tol2 || determ * determ2 < 0) continue;
    if (determ2 < -tol2) return 0;
  }
  return 1;
}

/* gethxedgeconnection(ed1, ed2) (src/smeshpack.f90:4127-4244), 0-based ids:
 * -2 when ed1 == ed2 (the reference's 0), the third edge of a triangle
 * holding both, the edge linking them through a neighbouring triangle, or -1 */
static int hxedgeconn(const imx_mesh_desc *d, int ed1, int ed2) {
  if (ed1 == ed2) return -2;
  const int t1 = d->ed_sh[2 * ed1] - 1, t2 = d->ed_sh[2 * ed1 + 1] - 1;
  int tr = -1;
  for (int i = 0; i < 3; ++i) {
    if (ed2 == d->tr_ed[3 * t1 + i] - 1) { tr = t1; break; }
    if (ed2 == d->tr_ed[3 * t2 + i] - 1) { tr = t2; break; }
  }
  if (tr >= 0) {
    int e3 = -1;
    for (int i = 0; i < 3; ++i) {
      const int e = d->tr_ed[3 * tr + i] - 1;
      if (e != ed1 && e != ed2) e3 = e;
    }
    if (e3 >= 0) return e3;
  }
  for (int j = 0; j < 2; ++j) {
    const int trj = d->ed_sh[2 * ed2 + j] - 1;
    for (int i = 0; i < 3; ++i) {
      const int et = d->tr_ed[3 * trj + i] - 1;
      if (et == ed2) continue;
      const int tr2 = (d->ed_sh[2 * et] - 1 == trj) ? d->ed_sh[2 * et + 1] - 1 : d->ed_sh[2 * et] - 1;
      for (int l = 0; l < 3; ++l)
        if (ed1 == d->tr_ed[3 * tr2 + l] - 1) return et;
    }
  }
  return -1;
}

/* dsign(1, x) */
static double sgn1(double x) { return x < 0.0 || (x == 0.0 && signbit(x)) ? -1.0 : 1.0; }

/* sign of dot_product(ed%nr, tr(sh2)%c - tr(sh1)%c), ed%nr = v1 x v2 / |v1 x v2|
 * (src/smeshpack.f90:725-726): the orientation of the dual edge
 * sh1 -> sh2 relative to the primal edge's normal */
static double ed_nr_sign(const imx_mesh_desc *d, int l) {
  const double *a = d->v_p + 3 * (d->ed_v[2 * l] - 1), *b = d->v_p + 3 * (d->ed_v[2 * l + 1] - 1);
  double nr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double *t1 = d->tr_c + 3 * (d->ed_sh[2 * l] - 1), *t2 = d->tr_c + 3 * (d->ed_sh[2 * l + 1] - 1);
  double dp = 0.0;
  for (int q = 0; q < 3; ++q) dp = dp + nr[q] * (t2[q] - t1[q]);
  return sgn1(dp);
}

/* proj_vec_sphere (src/smeshpack.f90:4419-4433) in place */
static void proj_sphere(double *v, const double *pc) {
  double vp = 0.0, pn = 0.0;
  for (int q = 0; q < 3; ++q) vp = vp + v[q] * pc[q];
  for (int q = 0; q < 3; ++q) pn = pn + pc[q] * pc[q];
  pn = sqrt(pn);
  for (int q = 0; q < 3; ++q) v[q] = v[q] - vp * pc[q] / pn;
}

/* getedindexonhx (src/smeshpack.f90:4099-4125): local index of edge ed in cell */
static int edindex(const imx_mesh_desc *d, int ed1, int cell0) {
  for (int i = 0; i < d->v_nnb[cell0]; ++i)
    if (d->v_ed[MX * cell0 + i] == ed1) return i;
  return -1;
}

/* div_hx of an edge field (signs of the staggering in w->st_c) */
static void div_hx(const imx_mesh_desc *d, double erad, const oracle_fields *w, const double *f,
                   double *out) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < d->nv; ++i) {
    double s = 0.0;
    for (int j = 0; j < d->v_nnb[i]; ++j) {
      int l = d->v_ed[MX * i + j] - 1;
      double signcor = w->st_c[MX * i + j];
      s = s + signcor * f[l] * d->edhx_leng[l];
    }
    out[i] = s / d->hx_areag[i] / erad;
  }
}

/* rel_vort_tr of an edge field (signs in w->st_t) */
static void rel_vort_tr(const imx_mesh_desc *d, double erad, const oracle_fields *w, const double *f,
                        double *out) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < d->nt; ++k) {
    double z = 0.0;
    for (int l = 0; l < 3; ++l) {
      int ed = d->tr_ed[3 * k + l] - 1;
      double signcor = w->st_t[3 * k + l];
      z = z + f[ed] * d->ed_leng[ed] * signcor;
    }
    out[k] = z / d->tr_areag[k] / erad;
  }
}

/* diffusion_ed: grad_ed(K div) - grad_ed_tg(K vort) (signs in w->st_g, st_tt) */
static void diffusion(const imx_mesh_desc *d, double erad, const double *kh, const double *kt,
                      const double *div, const double *vort, oracle_fields *w, double *out) {
  for (int i = 0; i < d->nv; ++i) w->cdiv[i] = kh[i] * div[i];
  for (int k = 0; k < d->nt; ++k) w->cvort[k] = kt[k] * vort[k];
#pragma omp parallel for schedule(static)
  for (int l = 0; l < d->ne; ++l) {
    double g = w->st_g[l] * (w->cdiv[d->ed_v[2 * l + 1] - 1] - w->cdiv[d->ed_v[2 * l] - 1]);
    w->gdiv[l] = g / d->ed_leng[l] / erad;
    const double f1 = w->cvort[d->ed_sh[2 * l] - 1], f2 = w->cvort[d->ed_sh[2 * l + 1] - 1];
    double gt = w->st_tt[l] * (f2 - f1);
    w->gvort[l] = gt / (d->edhx_leng[l] * erad);
  }
  for (int l = 0; l < d->ne; ++l) out[l] = w->gdiv[l] - w->gvort[l];
}

int oracle_tendency(const imx_mesh_desc *d, const oracle_params *p, const double *h,
                    const double *u, const double *bt, double *masseq, double *momeq,
                    oracle_fields *w) {
  const int nv = d->nv, ne = d->ne, nt = d->nt;
  const double erad = p->erad;
  double *h_ed = w->h_ed, *uh = w->uh, *eta = w->eta, *h_tr = w->h_tr, *q_tr = w->q_tr;
  double *q_ed = w->q_ed, *divuh = w->divuh, *ke = w->ke_hx, *ghbk = w->ghbk;
  double *grad = w->grad_ghbk, *uhq = w->uhq_perp, *divu = w->divu, *zeta = w->zeta;
  double *eta_ed = w->eta_ed, *vhx = w->v_hx;
  int bad = 0;

  const int hc = p->scheme && p->scheme->stag == 1;
  /* signs and the Perot edge points of the staggering */
  for (int l = 0; l < ne; ++l) {
    double s_e = 1.0;
    if (hc) {
      double a[3], b[3];
      ed_tg(d, l, a);
      edhx_nr(d, l, b);
      s_e = sgn1(dot(a, b)); /* ttgout = hx%nr sign(ed%tg . edhx%nr) */
      double dv[3];
      const double *v1 = d->v_p + 3 * (d->ed_v[2 * l] - 1), *v2 = d->v_p + 3 * (d->ed_v[2 * l + 1] - 1);
      for (int q = 0; q < 3; ++q) dv[q] = v2[q] - v1[q];
      w->st_g[l] = sgn1(dot(b, dv));
      double tg[3], dt3[3];
      edhx_tg(d, l, tg);
      const double *t1 = d->tr_c + 3 * (d->ed_sh[2 * l] - 1), *t2 = d->tr_c + 3 * (d->ed_sh[2 * l + 1] - 1);
      for (int q = 0; q < 3; ++q) dt3[q] = t2[q] - t1[q];
      w->st_tt[l] = sgn1(dot(tg, dt3));
      for (int q = 0; q < 3; ++q) w->st_tg[3 * l + q] = tg[q];
    } else {
      w->st_g[l] = 1.0;
      w->st_tt[l] = ed_nr_sign(d, l);
    }
    w->h_rhb[l] = s_e; /* scratch: per-edge factor, consumed below */
  }
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < d->v_nnb[i]; ++j)
      w->st_c[MX * i + j] = (double)d->hx_ttgout[MX * i + j] * (hc ? w->h_rhb[d->v_ed[MX * i + j] - 1] : 1.0);
  for (int k = 0; k < nt; ++k)
    for (int l = 0; l < 3; ++l)
      w->st_t[3 * k + l] = (double)d->tr_tg[3 * k + l] * (hc ? w->h_rhb[d->tr_ed[3 * k + l] - 1] : 1.0);
  const double *epos = hc ? d->edhx_c : d->ed_c;

  /* zero_vector(momeq), zero_vector(masseq) */
  for (int l = 0; l < ne; ++l) momeq[l] = 0.0;
  for (int i = 0; i < nv; ++i) masseq[i] = 0.0;

  /* scalar_hx2ed + scalar_elem_product: the average, or (HC with pxt16)
     barycentric at the dual edge midpoint in its containing triangle */
  const int wrap0 = p->scheme ? p->scheme->wrap : 0;
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l) {
    if (hc && wrap0 == 1) {
      const double *pc = d->edhx_c + 3 * l;
      const int kt = insidetrmesh(d, pc, d->ed_sh[2 * l] - 1) ? d->ed_sh[2 * l] - 1 : d->ed_sh[2 * l + 1] - 1;
      double b[3], f = 0.0;
      bary_at(d, pc, kt, b);
      for (int i = 0; i < 3; ++i) f = f + b[i] * h[d->tr_v[3 * kt + i] - 1];
      h_ed[l] = f;
    } else {
      int c1 = d->ed_v[2 * l] - 1, c2 = d->ed_v[2 * l + 1] - 1;
      h_ed[l] = (h[c1] + h[c2]) * 0.5;
    }
  }
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l) uh[l] = u[l] * h_ed[l];

  /* vort_tr: absolute vorticity at triangles */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double e = 0.0;
    for (int l = 0; l < 3; ++l) {
      int ed = d->tr_ed[3 * k + l] - 1;
      double signcor = w->st_t[3 * k + l];
      e = e + u[ed] * d->ed_leng[ed] * signcor;
    }
    e = e / d->tr_areag[k] / erad;
    /* the Coriolis term by fsphere (src/swm_operators.f90:596-605) */
    if (p->fsphere == 1) eta[k] = p->fcte;
    else if (p->fsphere == 2) eta[k] = e + p->fcte;
    else eta[k] = e + 2.0 * p->omega * sin(d->tr_clat[k]);
  }

  const int wrap = p->scheme ? p->scheme->wrap : 0;
  if (wrap == 2) {
    /* scalar_edtr2trcc(h_ed) (gass18) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nt; ++k) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) {
        const int ed = d->tr_ed[3 * k + i] - 1;
        const double a = sphtriarea(d->tr_c + 3 * k, d->v_p + 3 * (d->ed_v[2 * ed] - 1),
                                    d->v_p + 3 * (d->ed_v[2 * ed + 1] - 1));
        s = s + a * h_ed[ed];
      }
      h_tr[k] = s / d->tr_areag[k];
    }
  } else if (wrap == 1) {
    /* scalar_hx2tr (bary, pxt16) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nt; ++k) {
      double b[3], s = 0.0;
      bary(d, k, b);
      for (int i = 0; i < 3; ++i) s = s + b[i] * h[d->tr_v[3 * k + i] - 1];
      h_tr[k] = s;
    }
  } else {
    /* scalar_hx2tr (trsk) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nt; ++k) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s = s + p->trhx_area[3 * k + i] * h[d->tr_v[3 * k + i] - 1];
      h_tr[k] = s / d->tr_areag[k];
    }
  }

  /* scalar_tr2ed: simple average (trsk10), distance-weighted (pxt16 on HTC),
     area-weighted rhombi (gass18) */
#define TR2ED(f, out)                                                                  \
  for (int l = 0; l < ne; ++l) {                                                       \
    const int k1 = d->ed_sh[2 * l] - 1, k2 = d->ed_sh[2 * l + 1] - 1;                  \
    if (wrap == 0 || hc) {                                                             \
      out[l] = 0.5 * (f[k1] + f[k2]);                                                  \
    } else if (wrap == 2) {                                                            \
      const double a1 = d->tr_areag[k1], a2 = d->tr_areag[k2];                          \
      out[l] = (a1 * f[k1] + a2 * f[k2]) / (a1 + a2);                                  \
    } else {                                                                           \
      const double d1 = arclen(d->ed_c + 3 * l, d->tr_c + 3 * k1);                     \
      const double d2 = arclen(d->ed_c + 3 * l, d->tr_c + 3 * k2);                     \
      out[l] = (d2 * f[k1] + d1 * f[k2]) / (d1 + d2);                                  \
    }                                                                                  \
  }
  TR2ED(eta, eta_ed)

  /* PV at triangles: scalar_elem_divide, with the reference's zero guard */
  if (p->testcase < 2) {
    for (int k = 0; k < nt; ++k) q_tr[k] = 0.0;
  } else {
    for (int k = 0; k < nt; ++k) {
      if (fabs(h_tr[k]) < EPS_SINGLE / 100000.0) {
        bad = 1;
        q_tr[k] = 0.0;
      } else {
        q_tr[k] = eta[k] / h_tr[k];
      }
    }
  }
  if (bad) return -4;

  if (wrap == 2) {
    /* PV on rhombi: eta_ed / h_rhb, h_rhb = scalar_tr2ed(h_tr) */
    TR2ED(h_tr, w->h_rhb)
    for (int l = 0; l < ne; ++l) {
      if (fabs(w->h_rhb[l]) < EPS_SINGLE / 100000.0) return -4;
      q_ed[l] = eta_ed[l] / w->h_rhb[l];
    }
  } else {
    TR2ED(q_tr, q_ed)
  }
#undef TR2ED

  const oracle_scheme *sc = p->scheme;
  const int apvm = sc && sc->apvm;

  /* grad_tr2ed(q_tr) and its Perot reconstruction at triangles (APVM) */
  if (apvm) {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < ne; ++l) {
      const int k1 = d->ed_sh[2 * l] - 1, k2 = d->ed_sh[2 * l + 1] - 1;
      w->q_grad_ed[l] = ed_nr_sign(d, l) * (q_tr[k2] - q_tr[k1]) / d->edhx_leng[l] / erad;
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nt; ++k) {
      double v[3] = {0.0, 0.0, 0.0};
      const double *tc = d->tr_c + 3 * k;
      for (int l = 0; l < 3; ++l) {
        const int ed = d->tr_ed[3 * k + l] - 1;
        const double signcor = sgn1(-(double)d->tr_tg[3 * k + l]); /* tr%nr = -tr%tg */
        for (int q = 0; q < 3; ++q)
          v[q] = v[q] + signcor * w->q_grad_ed[ed] * (d->ed_c[3 * ed + q] - tc[q]) * d->ed_leng[ed];
      }
      for (int q = 0; q < 3; ++q) v[q] = v[q] / d->tr_areag[k];
      proj_sphere(v, tc);
      for (int q = 0; q < 3; ++q) w->q_grad_tr[3 * k + q] = v[q];
    }
  }

  /* vector_edhx2hx_perot(u): used by APVM, dead for plain trsk10 (computed on
     request, as the reference does it every step) */
  if (w->perot || apvm || wrap == 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nv; ++i) {
      double v[3] = {0.0, 0.0, 0.0};
      const double *pc = d->v_p + 3 * i;
      for (int l = 0; l < d->v_nnb[i]; ++l) {
        int ed = d->v_ed[MX * i + l] - 1;
        double signcor = w->st_c[MX * i + l];
        for (int q = 0; q < 3; ++q)
          v[q] = v[q] + signcor * u[ed] * (epos[3 * ed + q] - pc[q]) * d->edhx_leng[ed];
      }
      for (int q = 0; q < 3; ++q) v[q] = v[q] / d->hx_areag[i];
      /* proj_vec_sphere (src/smeshpack.f90:4419-4433) */
      double vp = v[0] * pc[0];
      vp = vp + v[1] * pc[1];
      vp = vp + v[2] * pc[2];
      double pn = pc[0] * pc[0];
      pn = pn + pc[1] * pc[1];
      pn = pn + pc[2] * pc[2];
      pn = sqrt(pn);
      for (int q = 0; q < 3; ++q) vhx[3 * i + q] = v[q] - vp * pc[q] / pn;
    }
  }

  /* div_hx(uh) -> masseq = -divuh */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
    double s = 0.0;
    for (int j = 0; j < d->v_nnb[i]; ++j) {
      int l = d->v_ed[MX * i + j] - 1;
      double signcor = w->st_c[MX * i + j];
      s = s + signcor * uh[l] * d->edhx_leng[l];
    }
    divuh[i] = s / d->hx_areag[i] / erad;
  }
  for (int i = 0; i < nv; ++i) masseq[i] = -divuh[i];

  /* kinetic_energy_hx (trsk, or perhx for pxt16) and the Bernoulli potential g(h+b)+K */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
    if (wrap == 1) {
      ke[i] = dot(vhx + 3 * i, vhx + 3 * i) / 2.0;
      continue;
    }
    double k = 0.0, cell_area = 0.0;
    for (int l = 0; l < d->v_nnb[i]; ++l) {
      int ed = d->v_ed[MX * i + l] - 1;
      double ed_area = 0.25 * d->ed_leng[ed] * d->edhx_leng[ed];
      cell_area = cell_area + ed_area;
      k = k + ed_area * (u[ed] * u[ed]);
    }
    ke[i] = k / cell_area;
  }
  for (int i = 0; i < nv; ++i) w->hbt[i] = h[i] + (bt ? bt[i] : 0.0);
  for (int i = 0; i < nv; ++i) ghbk[i] = p->grav * w->hbt[i] + ke[i];

  /* grad_ed (HTC) -> momeq = -grad */
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l) {
    double g = w->st_g[l] * (ghbk[d->ed_v[2 * l + 1] - 1] - ghbk[d->ed_v[2 * l] - 1]);
    grad[l] = g / d->ed_leng[l] / erad;
  }
  for (int l = 0; l < ne; ++l) momeq[l] = -grad[l];

  /* apvm: q_ed - pvspar dt (grad q at the edge's triangles) . (v at its cells) */
  if (apvm) {
    const double *qg = w->q_grad_tr;
    for (int l = 0; l < ne; ++l) {
      const int k1 = d->ed_sh[2 * l] - 1, k2 = d->ed_sh[2 * l + 1] - 1;
      const int c1 = d->ed_v[2 * l] - 1, c2 = d->ed_v[2 * l + 1] - 1;
      double dp = 0.0;
      for (int q = 0; q < 3; ++q)
        dp = dp + (qg[3 * k1 + q] + qg[3 * k2 + q]) * (vhx[3 * c1 + q] + vhx[3 * c2 + q]);
      q_ed[l] = q_ed[l] - sc->pvspar * sc->dt * dp;
    }
  }

  /* coriolis_ed pered (pxt16, HTC): Perot reconstruction of uh q at the
     edge's two cells, averaged, projected, minus its ed%nr component */
  if (wrap == 1) {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < ne; ++l) {
      double vv[2][3];
      for (int ic = 0; ic < 2; ++ic) {
        const int node = d->ed_v[2 * l + ic] - 1;
        const double *pc = d->v_p + 3 * node;
        double v[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < d->v_nnb[node]; ++j) {
          const int l2 = d->v_ed[MX * node + j] - 1;
          const double qtmp = 0.5 * (q_ed[l] + q_ed[l2]);
          const double signcor = w->st_c[MX * node + j];
          for (int q = 0; q < 3; ++q)
            v[q] = v[q] + signcor * uh[l2] * qtmp * (epos[3 * l2 + q] - pc[q]) * d->edhx_leng[l2];
        }
        for (int q = 0; q < 3; ++q) v[q] = v[q] / d->hx_areag[node];
        proj_sphere(v, pc);
        for (int q = 0; q < 3; ++q) vv[ic][q] = v[q];
      }
      double v[3], nr[3];
      for (int q = 0; q < 3; ++q) v[q] = 0.5 * vv[0][q] + 0.5 * vv[1][q];
      if (hc) {
        proj_sphere(v, d->edhx_c + 3 * l);
        uhq[l] = dot(v, w->st_tg + 3 * l);
      } else {
        proj_sphere(v, d->ed_c + 3 * l);
        ed_nr(d, l, nr);
        uhq[l] = -dot(v, nr);
      }
    }
  } else {
  /* coriolis_ed (trsk, PV form; gass18: q of the rhombus linking the two
     edges where there is one) */
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l) {
    double s = 0.0;
    for (int ic = 0; ic < 2; ++ic) {
      int node = d->ed_v[2 * l + ic] - 1;
      int ei = edindex(d, l + 1, node);
      for (int j = 0; j < d->v_nnb[node]; ++j) {
        int k = d->v_ed[MX * node + j] - 1;
        double sg = hc ? w->st_c[MX * node + j] * w->st_c[MX * node + ei]
                       : 1.0 * (double)d->hx_ttgout[MX * node + j] * (double)d->hx_tnrccw[MX * node + ei];
        double signcor = sg < 0 ? -1.0 : 1.0;
        double qtmp;
        const int e3 = wrap == 2 ? hxedgeconn(d, l, k) : -1;
        if (e3 >= 0) qtmp = q_ed[e3];
        else qtmp = 0.5 * (q_ed[l] + q_ed[k]);
        s = s + uh[k] * qtmp * d->edhx_leng[k] * p->trskw[MX * MX * node + MX * ei + j] * signcor;
      }
    }
    uhq[l] = s / d->ed_leng[l];
  }
  }
  for (int l = 0; l < ne; ++l) momeq[l] = momeq[l] - uhq[l];

  /* diagnostics of the step: div_hx(u), rel_vort_tr(u) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
    double s = 0.0;
    for (int j = 0; j < d->v_nnb[i]; ++j) {
      int l = d->v_ed[MX * i + j] - 1;
      double signcor = w->st_c[MX * i + j];
      s = s + signcor * u[l] * d->edhx_leng[l];
    }
    divu[i] = s / d->hx_areag[i] / erad;
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double z = 0.0;
    for (int l = 0; l < 3; ++l) {
      int ed = d->tr_ed[3 * k + l] - 1;
      double signcor = w->st_t[3 * k + l];
      z = z + u[ed] * d->ed_leng[ed] * signcor;
    }
    zeta[k] = z / d->tr_areag[k] / erad;
  }

  /* diffusion_ed / hyperdiffusion_ed: momeq + lapu, momeq - lap_lapu */
  if (sc && (sc->diffusion || sc->hyperdiffusion)) {
    diffusion(d, erad, sc->dif_hx, sc->dif_tr, divu, zeta, w, w->lapu);
    if (sc->diffusion)
      for (int l = 0; l < ne; ++l) momeq[l] = momeq[l] + w->lapu[l];
    if (sc->hyperdiffusion) {
      div_hx(d, erad, w, w->lapu, w->div_lap);
      rel_vort_tr(d, erad, w, w->lapu, w->vort_lap);
      diffusion(d, erad, sc->dif_hx, sc->dif_tr, w->div_lap, w->vort_lap, w, w->lap_lapu);
      for (int l = 0; l < ne; ++l) momeq[l] = momeq[l] - w->lap_lapu[l];
    }
  }

  if (p->testcase <= 1)
    for (int l = 0; l < ne; ++l) momeq[l] = 0.0;
  return 0;
}

int oracle_fields_alloc(const imx_mesh_desc *d, oracle_fields *w, int with_perot) {
  memset(w, 0, sizeof(*w));
  const size_t ne = (size_t)d->ne, nv = (size_t)d->nv, nt = (size_t)d->nt;
  w->h_ed = calloc(ne, 8);
  w->uh = calloc(ne, 8);
  w->eta = calloc(nt, 8);
  w->h_tr = calloc(nt, 8);
  w->q_tr = calloc(nt, 8);
  w->q_ed = calloc(ne, 8);
  w->eta_ed = calloc(ne, 8);
  w->divuh = calloc(nv, 8);
  w->ke_hx = calloc(nv, 8);
  w->ghbk = calloc(nv, 8);
  w->hbt = calloc(nv, 8);
  w->grad_ghbk = calloc(ne, 8);
  w->uhq_perp = calloc(ne, 8);
  w->divu = calloc(nv, 8);
  w->zeta = calloc(nt, 8);
  w->v_hx = calloc(3 * nv, 8);
  w->perot = with_perot;
  w->q_grad_ed = calloc(ne, 8);
  w->q_grad_tr = calloc(3 * nt, 8);
  w->lapu = calloc(ne, 8);
  w->lap_lapu = calloc(ne, 8);
  w->cdiv = calloc(nv, 8);
  w->cvort = calloc(nt, 8);
  w->gdiv = calloc(ne, 8);
  w->gvort = calloc(ne, 8);
  w->div_lap = calloc(nv, 8);
  w->vort_lap = calloc(nt, 8);
  w->h_rhb = calloc(ne, 8);
  w->st_c = calloc((size_t)MX * nv, 8);
  w->st_t = calloc(3 * nt, 8);
  w->st_g = calloc(ne, 8);
  w->st_tt = calloc(ne, 8);
  w->st_tg = calloc(3 * ne, 8);
  w->massf = calloc(4 * nv, 8);
  w->momf = calloc(4 * ne, 8);
  w->h_new = calloc(nv, 8);
  w->u_new = calloc(ne, 8);
  if (!w->h_ed || !w->uh || !w->eta || !w->h_tr || !w->q_tr || !w->q_ed || !w->eta_ed ||
      !w->divuh || !w->ke_hx || !w->ghbk || !w->hbt || !w->grad_ghbk || !w->uhq_perp || !w->divu ||
      !w->zeta || !w->v_hx || !w->massf || !w->momf || !w->h_new || !w->u_new || !w->q_grad_ed ||
      !w->q_grad_tr || !w->lapu || !w->lap_lapu || !w->cdiv || !w->cvort || !w->gdiv || !w->gvort ||
      !w->div_lap || !w->vort_lap || !w->h_rhb || !w->st_c || !w->st_t || !w->st_g || !w->st_tt ||
      !w->st_tg)
    return -2;
  return 0;
}

void oracle_fields_free(oracle_fields *w) {
  double *ptrs[] = {w->h_ed, w->uh, w->eta, w->h_tr, w->q_tr, w->q_ed, w->eta_ed,
                    w->divuh, w->ke_hx, w->ghbk, w->hbt, w->grad_ghbk, w->uhq_perp, w->divu,
                    w->zeta, w->v_hx, w->massf, w->momf, w->h_new, w->u_new, w->q_grad_ed,
                    w->q_grad_tr, w->lapu, w->lap_lapu, w->cdiv, w->cvort, w->gdiv, w->gvort,
                    w->div_lap, w->vort_lap, w->h_rhb, w->st_c, w->st_t, w->st_g, w->st_tt,
                    w->st_tg};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  memset(w, 0, sizeof(*w));
}

/* ode_rk4 (src/swm.f90:2600-2683): (h, u) <- RK4 step of length dt. */
int oracle_rk4(const imx_mesh_desc *d, const oracle_params *p, double dt, double *h,
               double *u, const double *bt, oracle_fields *w) {
  const int nv = d->nv, ne = d->ne;
  double *m0 = w->massf, *m1 = m0 + nv, *m2 = m1 + nv, *m3 = m2 + nv;
  double *f0 = w->momf, *f1 = f0 + ne, *f2 = f1 + ne, *f3 = f2 + ne;
  double *hn = w->h_new, *un = w->u_new;
  int rc;
  if ((rc = oracle_tendency(d, p, h, u, bt, m0, f0, w))) return rc;
  for (int l = 0; l < ne; ++l) un[l] = u[l] + dt * f0[l] / 2.0;
  for (int i = 0; i < nv; ++i) hn[i] = h[i] + dt * m0[i] / 2.0;
  if ((rc = oracle_tendency(d, p, hn, un, bt, m1, f1, w))) return rc;
  for (int l = 0; l < ne; ++l) un[l] = u[l] + dt * f1[l] / 2.0;
  for (int i = 0; i < nv; ++i) hn[i] = h[i] + dt * m1[i] / 2.0;
  if ((rc = oracle_tendency(d, p, hn, un, bt, m2, f2, w))) return rc;
  for (int l = 0; l < ne; ++l) un[l] = u[l] + dt * f2[l];
  for (int i = 0; i < nv; ++i) hn[i] = h[i] + dt * m2[i];
  if ((rc = oracle_tendency(d, p, hn, un, bt, m3, f3, w))) return rc;
  for (int l = 0; l < ne; ++l)
    u[l] = u[l] + dt * (f0[l] + 2.0 * f1[l] + 2.0 * f2[l] + f3[l]) / 6.0;
  for (int i = 0; i < nv; ++i)
    h[i] = h[i] + dt * (m0[i] + 2.0 * m1[i] + 2.0 * m2[i] + m3[i]) / 6.0;
  return 0;
}

/* sumf_areas / calc_energies / sumfsq_areas (src/swm.f90:3394-3490, :2686-2752)
 * evaluated on (h, u, bt) with h_ed, divu, h_tr, q_tr taken from the fields of
 * the LAST oracle_tendency call, as the reference does (it reads the module
 * globals written by the last tendency evaluation). */
void oracle_diagnostics(const imx_mesh_desc *d, const oracle_params *p, const double *h,
                        const double *u, const double *bt, const oracle_fields *w,
                        double out[8]) {
  const int nv = d->nv, ne = d->ne, nt = d->nt;
  double sa = 0.0, sm = 0.0;
  for (int i = 0; i < nv; ++i) {
    sa = sa + d->hx_areag[i];
    sm = sm + d->hx_areag[i] * h[i];
  }
  double mass = sm / sa;
  double pe = 0.0;
  for (int i = 0; i < nv; ++i) {
    double b = bt ? bt[i] : 0.0;
    pe = pe + d->hx_areag[i] * p->grav * h[i] * (h[i] * 0.5 + b);
  }
  /* mean of h+b (sumf_areas(hbt)); hbt is h+bt of the last tendency call */
  double sb = 0.0;
  sa = 0.0;
  for (int i = 0; i < nv; ++i) {
    sa = sa + d->hx_areag[i];
    sb = sb + d->hx_areag[i] * w->hbt[i];
  }
  double mean_height = sb / sa;
  double upe = 0.0;
  for (int i = 0; i < nv; ++i) {
    double b = bt ? bt[i] : 0.0;
    double mhb = mean_height - b;
    upe = upe + d->hx_areag[i] * p->grav * mhb * (mhb * 0.5 + b);
  }
  double ke = 0.0;
  for (int l = 0; l < ne; ++l)
    ke = ke + 0.5 * d->ed_leng[l] * d->edhx_leng[l] * w->h_ed[l] * (u[l] * u[l]);
  double pz = 0.0;
  for (int k = 0; k < nt; ++k)
    pz = pz + 0.5 * d->tr_areag[k] * p->erad * p->erad * w->h_tr[k] * w->q_tr[k] * w->q_tr[k];
  double s2 = 0.0;
  sa = 0.0;
  for (int i = 0; i < nv; ++i) {
    sa = sa + d->hx_areag[i];
    s2 = s2 + d->hx_areag[i] * (w->divu[i] * w->divu[i]);
  }
  int badflag = 0;
  for (int i = 0; i < nv; ++i)
    if (!isfinite(h[i])) badflag = 1;
  for (int l = 0; l < ne; ++l)
    if (!isfinite(u[l])) badflag = 1;
  out[0] = mass;
  out[1] = pe;
  out[2] = ke;
  out[3] = pe + ke;
  out[4] = (pe + ke) - upe;
  out[5] = pz;
  out[6] = sqrt(s2 / sa);
  out[7] = badflag;
}

/* arclen (src/smeshpack.f90:4624-4657) */
static double arclen(const double *p, const double *q) {
  double d = 0.0;
  for (int i = 0; i < 3; ++i) d = d + (p[i] + q[i]) * (p[i] + q[i]);
  if (d == 0.0) return 3.1415927410125732; /* pi of constants.f90, single-rounded */
  if (d >= 4.0) return 0.0;
  return 2.0 * atan(sqrt((4.0 - d) / d));
}

/* alignind (src/smeshpack.f90:6398-6490), 0-based cell i */
static double alignind(const imx_mesh_desc *d, int i) {
  const int n = d->v_nnb[i];
  const int *tr = d->v_tr + MX * i;
  double tmp = 0.0, meanl = 0.0;
#define TC(j) (d->tr_c + 3 * (tr[(j)-1] - 1))
  if (n == 6 || n == 4) {
    const int np = n == 6 ? 3 : 2;
    for (int j = 1; j <= np; ++j) {
      int j0 = n == 6 ? j + 4 : j + 3;
      if (j0 > n) j0 -= n; /* modint */
      const int i1 = j, i2 = j + 1, i3 = n == 6 ? j + 3 : j + 2;
      double t1 = arclen(TC(i1), TC(i2)), t2 = arclen(TC(i3), TC(j0));
      meanl = t1 + t2;
      tmp = tmp + fabs(t1 - t2);
      t1 = arclen(TC(i1), TC(j0));
      t2 = arclen(TC(i2), TC(i3));
      tmp = tmp + fabs(t1 - t2);
    }
    tmp = tmp / 6;
    meanl = meanl / 6;
    tmp = tmp / meanl;
  } else {
    tmp = 1;
  }
#undef TC
  return tmp;
}

/* the coefficient at cells for one K_max and kind */
static void coef_hx(const imx_mesh_desc *d, double kmax, int kind, double *c) {
  const int nv = d->nv;
  if (kind == 0) {
    for (int i = 0; i < nv; ++i) c[i] = kmax;
    return;
  }
  double *a = malloc((size_t)nv * 8);
  if (kind == 1) { /* alignment_coef */
    for (int i = 0; i < nv; ++i) a[i] = alignind(d, i);
    for (int i = 0; i < nv; ++i) {
      double sum = a[i];
      for (int j = 0; j < d->v_nnb[i]; ++j) sum = sum + a[d->v_nb[MX * i + j] - 1];
      sum = sum / (d->v_nnb[i] + 1.0);
      c[i] = sum;
    }
    double mx = c[0];
    for (int i = 1; i < nv; ++i) mx = c[i] > mx ? c[i] : mx;
    for (int i = 0; i < nv; ++i) c[i] = c[i] / mx;
    for (int i = 0; i < nv; ++i) c[i] = kmax * c[i];
  } else { /* diameter_coef */
    for (int i = 0; i < nv; ++i) {
      double dm = 0.0;
      for (int j = 0; j < d->v_nnb[i]; ++j)
        for (int k = 0; k < d->v_nnb[i]; ++k) {
          const double *pj = d->tr_c + 3 * (d->v_tr[MX * i + j] - 1);
          const double *pk = d->tr_c + 3 * (d->v_tr[MX * i + k] - 1);
          double al = arclen(pj, pk);
          dm = dm > al ? dm : al;
        }
      c[i] = dm;
    }
    double mx = c[0];
    for (int i = 1; i < nv; ++i) mx = c[i] > mx ? c[i] : mx;
    for (int i = 0; i < nv; ++i) c[i] = c[i] / mx;
    const double pw = log(10.0) / log(2.0);
    for (int i = 0; i < nv; ++i) c[i] = kmax * pow(c[i], pw);
  }
  free(a);
}

/* scalar_hx2tr (src/swm_operators.f90:258-269): trsk (area-weighted), or
 * barycentric for pxt16 */
static void hx2tr(const imx_mesh_desc *d, const double *trhx_area, int wrap, const double *f, double *out) {
  for (int k = 0; k < d->nt; ++k) {
    double s = 0.0;
    if (wrap == 1) {
      double b[3];
      bary(d, k, b);
      for (int i = 0; i < 3; ++i) s = s + b[i] * f[d->tr_v[3 * k + i] - 1];
      out[k] = s;
    } else {
      for (int i = 0; i < 3; ++i) s = s + trhx_area[3 * k + i] * f[d->tr_v[3 * k + i] - 1];
      out[k] = s / d->tr_areag[k];
    }
  }
}

int oracle_diffusion_coef(const imx_mesh_desc *d, const double *trhx_area, int wrap, double k2,
                          int k2kind, double k4, int k4kind, double *dif_hx, double *dif_tr) {
  /* gass18: scalar_hx2tr stops ("Gassmann's scheme needs edge values") */
  if (wrap == 2 && (k2 > 0.0 || k4 > 0.0)) return -5;
  if (k2 > 0.0) {
    coef_hx(d, k2, k2kind, dif_hx);
    hx2tr(d, trhx_area, wrap, dif_hx, dif_tr);
  }
  if (k4 > 0.0) {
    coef_hx(d, k4, k4kind, dif_hx);
    hx2tr(d, trhx_area, wrap, dif_hx, dif_tr);
    for (int i = 0; i < d->nv; ++i) dif_hx[i] = sqrt(dif_hx[i]);
    for (int k = 0; k < d->nt; ++k) dif_tr[k] = sqrt(dif_tr[k]);
  }
  return 0;
}
