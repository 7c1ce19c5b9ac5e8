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

/* scheme options beyond the trsk10/HTC defaults (par/swm.par) */
typedef struct oracle_scheme {
  int stag;              /* 0 HTC, 1 HC (staggering)                        */
  int wrap;              /* 0 trsk10, 1 pxt16, 2 gass18 (method wrapper)    */
  int apvm;              /* pv_stab = apvm                                  */
  double pvspar;         /* its parameter                                   */
  double dt;             /* the module time step the APVM term reads        */
  int diffusion;         /* K2_max /= 0: momeq += lapu                      */
  int hyperdiffusion;    /* K4_max /= 0: momeq -= lap_lapu                  */
  const double *dif_hx;  /* dif_coef_hx [nv] (oracle_diffusion_coef)        */
  const double *dif_tr;  /* dif_coef_tr [nt]                                */
} oracle_scheme;

typedef struct oracle_params {
  double erad, grav, omega;
  int testcase;
  const double *trhx_area; /* [nt*3] */
  const double *trskw;     /* [nv*6*6] */
  const oracle_scheme *scheme; /* NULL: the defaults */
  int fsphere;             /* 0: f = 2 Omega sin(lat); 1: eta = fcte; 2: f = fcte */
  double fcte;
} oracle_params;

/* module-global fields of swm_data written by tendency */
typedef struct oracle_fields {
  double *h_ed, *uh, *eta, *h_tr, *q_tr, *q_ed, *eta_ed, *divuh, *ke_hx, *ghbk, *hbt;
  double *grad_ghbk, *uhq_perp, *divu, *zeta, *v_hx;
  double *massf, *momf, *h_new, *u_new; /* ode_rk4 scratch */
  double *q_grad_ed, *q_grad_tr;        /* APVM: [ne], [nt*3]              */
  double *lapu, *lap_lapu;              /* (hyper)diffusion: [ne], [ne]     */
  double *cdiv, *cvort, *gdiv, *gvort;  /* diffusion_ed scratch: [nv], [nt], [ne], [ne] */
  double *div_lap, *vort_lap;           /* hyperdiffusion scratch: [nv], [nt] */
  double *h_rhb;                        /* gass18: h at the rhombi [ne]     */
  /* staggering-dependent signs (filled by each oracle_tendency call) */
  double *st_c;                         /* [nv*6] div/Perot sign per cell slot */
  double *st_t;                         /* [nt*3] vorticity sign per triangle slot */
  double *st_g, *st_tt;                 /* [ne] grad_ed and grad_ed_tg signs */
  double *st_tg;                        /* [ne*3] HC: edhx%tg */
  int perot;                            /* compute v_hx even when unused    */
} oracle_fields;

int oracle_fields_alloc(const imx_mesh_desc *d, oracle_fields *w, int with_perot);
void oracle_fields_free(oracle_fields *w);
int oracle_tendency(const imx_mesh_desc *d, const oracle_params *p, const double *h,
                    const double *u, const double *bt, double *masseq, double *momeq,
                    oracle_fields *w);
int oracle_rk4(const imx_mesh_desc *d, const oracle_params *p, double dt, double *h,
               double *u, const double *bt, oracle_fields *w);
/* dif_coef_hx / dif_coef_tr as swm_tests sets them up (src/swm.f90:170-210);
 * kind: 0 const, 1 align (alignment_coef), 2 diam (diameter_coef). With
 * k4 > 0 both hold the square roots of the hyperdiffusion coefficients.
 * wrap as oracle_scheme; -5 for gass18 with diffusion (the reference stops). */
int oracle_diffusion_coef(const imx_mesh_desc *d, const double *trhx_area, int wrap, double k2,
                          int k2kind, double k4, int k4kind, double *dif_hx, double *dif_tr);
void oracle_diagnostics(const imx_mesh_desc *d, const oracle_params *p, const double *h,
                        const double *u, const double *bt, const oracle_fields *w,
                        double out[8]);
#endif
