/*
 * swm_oracle.h -- TEST INFRASTRUCTURE ONLY. CPU (plain C, FP64) restatement of
 * the reference shallow-water step used as the checker of libimx; see
 * swm_oracle.c for the reference routines each function follows.
 * The mesh is passed with the same flat layout as the product's C ABI
 * (imx_mesh_desc, 1-based reference numbering); gridprop arrays are required.
 */
#ifndef SWM_ORACLE_H
#define SWM_ORACLE_H
#include "../include/imx_swm.h"

typedef struct oracle_params {
  double erad, grav, omega;
  int testcase;
  const double *trhx_area; /* [nt*3] */
  const double *trskw;     /* [nv*6*6] */
} oracle_params;

/* module-global fields of swm_data written by tendency */
typedef struct oracle_fields {
  double *h_ed, *uh, *eta, *h_tr, *q_tr, *q_ed, *eta_ed, *divuh, *ke_hx, *ghbk, *hbt;
  double *grad_ghbk, *uhq_perp, *divu, *zeta, *v_hx;
  double *massf, *momf, *h_new, *u_new; /* ode_rk4 scratch */
} oracle_fields;

int oracle_fields_alloc(const imx_mesh_desc *d, oracle_fields *w, int with_perot);
void oracle_fields_free(oracle_fields *w);
int oracle_tendency(const imx_mesh_desc *d, const oracle_params *p, const double *h,
                    const double *u, const double *bt, double *masseq, double *momeq,
                    oracle_fields *w);
int oracle_rk4(const imx_mesh_desc *d, const oracle_params *p, double dt, double *h,
               double *u, const double *bt, oracle_fields *w);
void oracle_diagnostics(const imx_mesh_desc *d, const oracle_params *p, const double *h,
                        const double *u, const double *bt, const oracle_fields *w,
                        double out[8]);
#endif
