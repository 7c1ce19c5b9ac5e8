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
 * Scope: HTC staggering, trsk10 wrapper, geodesic areas, no diffusion, no
 * APVM (the reference defaults, par/swm.par), i.e. the branches
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

/* getedindexonhx (src/smeshpack.f90:4099-4125): local index of edge ed in cell */
static int edindex(const imx_mesh_desc *d, int ed1, int cell0) {
  for (int i = 0; i < d->v_nnb[cell0]; ++i)
    if (d->v_ed[MX * cell0 + i] == ed1) return i;
  return -1;
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

  /* zero_vector(momeq), zero_vector(masseq) */
  for (int l = 0; l < ne; ++l) momeq[l] = 0.0;
  for (int i = 0; i < nv; ++i) masseq[i] = 0.0;

  /* scalar_hx2ed + scalar_elem_product */
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l) {
    int c1 = d->ed_v[2 * l] - 1, c2 = d->ed_v[2 * l + 1] - 1;
    h_ed[l] = (h[c1] + h[c2]) * 0.5;
  }
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l) uh[l] = u[l] * h_ed[l];

  /* vort_tr: absolute vorticity at triangles */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double e = 0.0;
    for (int l = 0; l < 3; ++l) {
      int ed = d->tr_ed[3 * k + l] - 1;
      double signcor = (double)d->tr_tg[3 * k + l];
      e = e + u[ed] * d->ed_leng[ed] * signcor;
    }
    e = e / d->tr_areag[k] / erad;
    eta[k] = e + 2.0 * p->omega * sin(d->tr_clat[k]);
  }

  /* scalar_hx2tr (trsk) */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double s = 0.0;
    for (int i = 0; i < 3; ++i) s = s + p->trhx_area[3 * k + i] * h[d->tr_v[3 * k + i] - 1];
    h_tr[k] = s / d->tr_areag[k];
  }

  /* scalar_tr2ed(eta) */
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l)
    eta_ed[l] = 0.5 * (eta[d->ed_sh[2 * l] - 1] + eta[d->ed_sh[2 * l + 1] - 1]);

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

  /* scalar_tr2ed(q) */
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l)
    q_ed[l] = 0.5 * (q_tr[d->ed_sh[2 * l] - 1] + q_tr[d->ed_sh[2 * l + 1] - 1]);

  /* vector_edhx2hx_perot(u): dead for trsk10 but part of the reference step */
  if (vhx) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nv; ++i) {
      double v[3] = {0.0, 0.0, 0.0};
      const double *pc = d->v_p + 3 * i;
      for (int l = 0; l < d->v_nnb[i]; ++l) {
        int ed = d->v_ed[MX * i + l] - 1;
        double signcor = (double)d->hx_ttgout[MX * i + l];
        for (int q = 0; q < 3; ++q)
          v[q] = v[q] + signcor * u[ed] * (d->ed_c[3 * ed + q] - pc[q]) * d->edhx_leng[ed];
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
      double signcor = (double)d->hx_ttgout[MX * i + j];
      s = s + signcor * uh[l] * d->edhx_leng[l];
    }
    divuh[i] = s / d->hx_areag[i] / erad;
  }
  for (int i = 0; i < nv; ++i) masseq[i] = -divuh[i];

  /* kinetic_energy_hx (trsk) and the Bernoulli potential g(h+b)+K */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
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
    double g = ghbk[d->ed_v[2 * l + 1] - 1] - ghbk[d->ed_v[2 * l] - 1];
    grad[l] = g / d->ed_leng[l] / erad;
  }
  for (int l = 0; l < ne; ++l) momeq[l] = -grad[l];

  /* coriolis_ed (trsk, PV form) */
#pragma omp parallel for schedule(static)
  for (int l = 0; l < ne; ++l) {
    double s = 0.0;
    for (int ic = 0; ic < 2; ++ic) {
      int node = d->ed_v[2 * l + ic] - 1;
      int ei = edindex(d, l + 1, node);
      for (int j = 0; j < d->v_nnb[node]; ++j) {
        int k = d->v_ed[MX * node + j] - 1;
        double sg = 1.0 * (double)d->hx_ttgout[MX * node + j] * (double)d->hx_tnrccw[MX * node + ei];
        double signcor = sg < 0 ? -1.0 : 1.0;
        double qtmp = 0.5 * (q_ed[l] + q_ed[k]);
        s = s + uh[k] * qtmp * d->edhx_leng[k] * p->trskw[MX * MX * node + MX * ei + j] * signcor;
      }
    }
    uhq[l] = s / d->ed_leng[l];
  }
  for (int l = 0; l < ne; ++l) momeq[l] = momeq[l] - uhq[l];

  /* diagnostics of the step: div_hx(u), rel_vort_tr(u) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
    double s = 0.0;
    for (int j = 0; j < d->v_nnb[i]; ++j) {
      int l = d->v_ed[MX * i + j] - 1;
      double signcor = (double)d->hx_ttgout[MX * i + j];
      s = s + signcor * u[l] * d->edhx_leng[l];
    }
    divu[i] = s / d->hx_areag[i] / erad;
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double z = 0.0;
    for (int l = 0; l < 3; ++l) {
      int ed = d->tr_ed[3 * k + l] - 1;
      double signcor = (double)d->tr_tg[3 * k + l];
      z = z + u[ed] * d->ed_leng[ed] * signcor;
    }
    zeta[k] = z / d->tr_areag[k] / erad;
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
  w->v_hx = with_perot ? calloc(3 * nv, 8) : NULL;
  w->massf = calloc(4 * nv, 8);
  w->momf = calloc(4 * ne, 8);
  w->h_new = calloc(nv, 8);
  w->u_new = calloc(ne, 8);
  if (!w->h_ed || !w->uh || !w->eta || !w->h_tr || !w->q_tr || !w->q_ed || !w->eta_ed ||
      !w->divuh || !w->ke_hx || !w->ghbk || !w->hbt || !w->grad_ghbk || !w->uhq_perp || !w->divu ||
      !w->zeta || (with_perot && !w->v_hx) || !w->massf || !w->momf || !w->h_new || !w->u_new)
    return -2;
  return 0;
}

void oracle_fields_free(oracle_fields *w) {
  double *ptrs[] = {w->h_ed, w->uh, w->eta, w->h_tr, w->q_tr, w->q_ed, w->eta_ed,
                    w->divuh, w->ke_hx, w->ghbk, w->hbt, w->grad_ghbk, w->uhq_perp, w->divu,
                    w->zeta, w->v_hx, w->massf, w->momf, w->h_new, w->u_new};
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
