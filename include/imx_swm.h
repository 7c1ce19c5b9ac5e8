/*
 * imx_swm.h -- C ABI of libimx, the MI355X-native shallow-water time step
 * (TRiSK/HTC C-grid on icosahedral Voronoi/Delaunay meshes, classical RK4).
 *
 * Drop-in for the reference Fortran path
 *     swm_tests -> ode_rk4 -> tendency -> swm_operators
 * (src/swm.f90:102-380, :2416-2683, src/swm_operators.f90). The reference has
 * no handle and no return codes: all state is module-global in swm_data
 * (src/swm_data.f90:31-334) and every error is `print*; stop`. Here every
 * call returns 0 on success or a negative IMX_E* code, and imx_last_error()
 * gives the message; the Fortran wrapper turns a non-zero code back into
 * `print*, msg; stop`, which reproduces the reference's convention.
 *
 * Numbering: every index handed across this boundary is in the CALLER's
 * numbering and 1-based, exactly as in the reference grid_structure
 * (src/datastruct.f90:403-519). Host arrays are caller-owned and never
 * retained past the call. One host thread per handle; calls are ordered on
 * the handle's HIP stream; *_get_*, *_diagnostics and *_tendency block.
 */
#ifndef IMX_SWM_H
#define IMX_SWM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMX_OK        0
#define IMX_EINVAL   (-1)  /* bad argument / inconsistent mesh                */
#define IMX_ENOMEM   (-2)  /* host or device allocation failed                */
#define IMX_EDEVICE  (-3)  /* HIP runtime error                               */
#define IMX_ESTATE   (-4)  /* numerical guard tripped (h_tr ~ 0, NaN state)   */
#define IMX_EIO      (-5)  /* file I/O                                        */
#define IMX_ENOTSUP  (-6)  /* option not supported by this build              */

#define IMX_MAXNB 6        /* cells of icosahedral grids have 5 or 6 edges    */

/* Arithmetic of the step. */
#define IMX_PREC_FP64_REF  0  /* FP64, the reference's operation order, one
                                 kernel per reference operator; bit-for-bit
                                 with the reference built -ffp-contract=off   */
#define IMX_PREC_FP64      1  /* FP64, fused cell/edge/triangle passes        */
#define IMX_PREC_FP32      2  /* FP32 stencil; FP64 geometry, compensated FP64
                                 mass-flux divergence, FP64 reductions        */

/* Mesh positions of the icosahedron (src/smeshpack.f90:1234-1279). */
#define IMX_POS_EQS 0
#define IMX_POS_POL 1

/*
 * Flat view of the reference grid_structure: the arrays the trsk10/HTC
 * tendency, initialize_gridprop and the diagnostics read. Per-cell lists are
 * padded to IMX_MAXNB with 0. Vectors are (x,y,z) per entity, unit sphere.
 * Replaces passing `type(grid_structure)` (src/swm.f90:102-107, :152).
 */
typedef struct imx_mesh_desc {
  int32_t nv, ne, nt;          /* cells (Voronoi nodes), edges, triangles      */
  int32_t glevel;              /* icosahedral level, -1 if not icosahedral     */
  /* cells: v(i)%..., hx(i)%...                                               */
  const int32_t *v_nnb;        /* [nv]          v%nnb                          */
  const int32_t *v_nb;         /* [nv*6]        v%nb   (neighbour cells, CCW)  */
  const int32_t *v_ed;         /* [nv*6]        v%ed   (edges, CCW)            */
  const int32_t *v_tr;         /* [nv*6]        v%tr   (triangles, CCW)        */
  const int32_t *hx_ttgout;    /* [nv*6]  +-1   hx%ttgout                      */
  const int32_t *hx_tnrccw;    /* [nv*6]  +-1   hx%tnrccw                      */
  const double  *v_p;          /* [nv*3]        v%p                            */
  const double  *hx_areag;     /* [nv]          hx%areag (unit sphere)         */
  /* edges: ed(l)%..., edhx(l)%...  (edhx%sh == ed%v, edhx%v == ed%sh)        */
  const int32_t *ed_v;         /* [ne*2]        ed%v   (cells, v1 -> v2)       */
  const int32_t *ed_sh;        /* [ne*2]        ed%sh  (triangles)             */
  const double  *ed_leng;      /* [ne]          ed%leng   (primal, geodesic)   */
  const double  *edhx_leng;    /* [ne]          edhx%leng (dual, geodesic)     */
  const double  *ed_c;         /* [ne*3]        ed%c%p    (primal midpoint)    */
  const double  *edhx_c;       /* [ne*3]        edhx%c%p  (dual midpoint)      */
  /* triangles: tr(k)%...                                                     */
  const int32_t *tr_v;         /* [nt*3]        tr%v                           */
  const int32_t *tr_ed;        /* [nt*3]        tr%ed                          */
  const int32_t *tr_tg;        /* [nt*3]  +-1   tr%tg                          */
  const double  *tr_areag;     /* [nt]          tr%areag                       */
  const double  *tr_c;         /* [nt*3]        tr%c%p (circumcentre)          */
  const double  *tr_clat;      /* [nt]          tr%c%lat                       */
  /* Optional initialize_gridprop products (src/swm.f90:979-1041). NULL: the
     library computes them on the host in FP64 with the reference formulas. */
  const double  *tr_trhx_area; /* [nt*3]        tr%trhx_area                   */
  const double  *hx_hxtr_area; /* [nv*6]        hx%hxtr_area (/areag)          */
  const double  *hx_trskw;     /* [nv*6*6]      hx%trskw(i,j) at [c][i][j]     */
} imx_mesh_desc;

typedef struct imx_swm_opts {
  int32_t precision;   /* IMX_PREC_*                                          */
  int32_t testcase;    /* reference semantics: <2 -> q=0 (swm.f90:2464),
                          <=1 -> momeq=0 (swm.f90:2592)                       */
  int32_t use_sfc;     /* reorder cells/edges/triangles along a Hilbert curve */
  int32_t use_graph;   /* capture the RK4 step in a hipGraph                  */
  double  erad;        /* sphere radius [m]        (constants.f90:55)         */
  double  grav;        /* gravity [m s^-2]         (constants.f90:61)         */
  double  omega;       /* rotation rate [s^-1]     (constants.f90:67)         */
  double  dt;          /* time step [s]                                       */
  double  h_ref;       /* FP32 mode stores h - h_ref; 0 -> area mean of the
                          first state set                                     */
} imx_swm_opts;

typedef struct imx_mesh imx_mesh;  /* host mesh built by the library */
typedef struct imx_swm  imx_swm;   /* one SWM instance; owns device memory */

/* Diagnostics vector written by imx_swm_diagnostics (all FP64). */
#define IMX_DIAG_MASS      0  /* sumf_areas(h)               swm.f90:3394      */
#define IMX_DIAG_PENERGY   1  /* sum A g h (h/2 + b)         swm.f90:2704      */
#define IMX_DIAG_KENERGY   2  /* sum 1/2 l d h_e u^2         swm.f90:2739      */
#define IMX_DIAG_TENERGY   3  /* P + K                       swm.f90:2745      */
#define IMX_DIAG_AENERGY   4  /* T - unavailable P           swm.f90:2747      */
#define IMX_DIAG_PENSTROPHY 5 /* sum_t 1/2 A_t a^2 h_t q_t^2 (not in reference) */
#define IMX_DIAG_RMSDIV    6  /* sqrt(sumfsq_areas(div u))   swm.f90:352       */
#define IMX_DIAG_BADFLAG   7  /* 1 if h or u holds NaN/Inf, else 0             */
#define IMX_NDIAG          8

/* Intermediate fields of the last imx_swm_tendency call (names of the
   reference module globals in swm_data, written by tendency). */
#define IMX_FIELD_H_ED      0  /* [ne] */
#define IMX_FIELD_UH        1  /* [ne] */
#define IMX_FIELD_ETA       2  /* [nt] */
#define IMX_FIELD_H_TR      3  /* [nt] */
#define IMX_FIELD_Q_TR      4  /* [nt] */
#define IMX_FIELD_Q_ED      5  /* [ne] */
#define IMX_FIELD_DIVUH     6  /* [nv] */
#define IMX_FIELD_KE_HX     7  /* [nv] */
#define IMX_FIELD_GHBK      8  /* [nv] */
#define IMX_FIELD_GRAD_GHBK 9  /* [ne] */
#define IMX_FIELD_UHQ_PERP 10  /* [ne] */
#define IMX_FIELD_DIVU     11  /* [nv] */
#define IMX_FIELD_ZETA     12  /* [nt] */
#define IMX_FIELD_Q_GRAD_ED 13 /* [ne]  grad_tr2ed(q_tr), with APVM            */
#define IMX_FIELD_LAPU     14  /* [ne]  diffusion_ed of u, with K2 or K4       */
#define IMX_FIELD_LAP_LAPU 15  /* [ne]  hyperdiffusion_ed of u, with K4        */
/* Fields of the last imx_swm_lte evaluation (the tendency's analysis fields
   under test_lterror = 1, src/swm.f90:2500-2503, :2546-2551); vectors are
   Cartesian, x y z per entry. */
#define IMX_FIELD_Q_HX     16  /* [nv]    scalar_tr2hx(q_tr)                   */
#define IMX_FIELD_VHQ_TR   17  /* [nt*3]  vector_edhx2tr_perot(uh) q_tr        */
#define IMX_FIELD_VH_HX    18  /* [nv*3]  v_hx h                               */
#define IMX_FIELD_VHQ_HX   19  /* [nv*3]  v_hx h q_hx                          */
#define IMX_FIELD_Q_GRAD_TR 20 /* [nt*3]  vector_edtr2tr_perot(q_grad_ed), APVM */

/* Scheme options of par/swm.par beyond the trsk10/HTC defaults
   (swmpars, src/swm_data.f90:413-424, :430-551), for FP64_REF handles:
     stag    IMX_STAG_HTC (the default) or IMX_STAG_HC: u at the Voronoi
             edge midpoints instead of the primal/dual edge crossings
     wrap    method wrapper:
             IMX_WRAP_TRSK10  TRiSK (Ringler et al. 2010), the default
             IMX_WRAP_PXT16   Peixoto 2016: barycentric h at triangles,
                              distance-weighted q at edges, Perot kinetic
                              energy, Perot (pered) Coriolis
             IMX_WRAP_GASS18  Gassmann 2018: h at triangles from edges, PV
                              on rhombi, Gassmann Coriolis
     apvm    pv_stab = apvm: q_ed -= pvspar dt (grad q) . v    (:1465-1502)
     k2      K2_max, diffusion momeq += div(K grad) u           (:855-895)
     k4      K4_max, hyperdiffusion momeq -= (div(K grad))^2 u  (:930-975)
   with the coefficient function of each (const, align: smoothed alignment
   index, diam: cell diameter; src/swm_operators.f90:977-1072) set up as
   swm_tests does (src/swm.f90:170-210): with k4 > 0 both Laplacian terms
   use the square roots of the hyperdiffusion coefficients; gass18 with
   diffusion is refused (the reference stops: scalar_hx2tr has no Gassmann
   branch). The APVM term reads the handle's time step. All zero: the
   defaults. */
#define IMX_COEF_CONST 0
#define IMX_COEF_ALIGN 1
#define IMX_COEF_DIAM  2
#define IMX_STAG_HTC 0
#define IMX_STAG_HC  1
#define IMX_WRAP_TRSK10 0
#define IMX_WRAP_PXT16  1
#define IMX_WRAP_GASS18 2
#define IMX_WRAP_NONE   3  /* the individual method choices below            */
/* With wrap IMX_WRAP_NONE (par/swm.par wrapper line "none",
   src/swm_data.f90:438-551), FP64_REF handles:
     sinterp   scalar interpolation: trsk (kite areas, simple averages),
               bary (pxt16's barycentric h and distance-weighted q), gass
               (gass18's edge-based h and rhombus PV)
     coriolis  trsk, pered (pxt16's), gass (gass18's linking-edge PV) or
               dtred (dual-triangle reconstruction: Perot vector of uh q at
               the two triangles, swm_operators.f90:1310-1329)
     recon     kinetic energy: trsk, perhx (pxt16's), gass (trsk K blended
               with the triangle K of kinetic_energy_tr by gasscoef; dubos is
               gass with gasscoef 0) or melv (uh u / h, :1174-1198)
     nopv      level model (inoPV = 1): the TRiSK and Gassmann Coriolis
               terms take the vorticity at edges and u instead of q and uh
   The wrappers are the triples (trsk, trsk, trsk), (bary, pered, perhx)
   and (gass, gass, trsk). As in the reference: nopv is refused with pered
   and dtred, gass interpolation with diffusion. The reference's hyb
   Coriolis (which rewrites shared method flags inside its parallel edge
   loop), its exact Coriolis (test-case closed forms) and the CLUST PV
   correction (which stops in the reference) are not offered. */
#define IMX_SINT_TRSK 0
#define IMX_SINT_BARY 1
#define IMX_SINT_GASS 2
#define IMX_COR_TRSK  0
#define IMX_COR_PERED 1
#define IMX_COR_GASS  2
#define IMX_COR_DTRED 3
#define IMX_REC_TRSK  0
#define IMX_REC_PERHX 1
#define IMX_REC_GASS  2
#define IMX_REC_MELV  3
typedef struct imx_scheme {
  int32_t stag;         /* IMX_STAG_*                                         */
  int32_t wrap;         /* IMX_WRAP_*                                         */
  int32_t apvm;         /* 0/1                                                */
  double  pvspar;       /* APVM parameter                                     */
  double  k2;           /* K2_max, 0 = no diffusion                           */
  int32_t diffus;       /* IMX_COEF_*                                         */
  double  k4;           /* K4_max, 0 = no hyperdiffusion                      */
  int32_t hyperdiffus;  /* IMX_COEF_*                                         */
  int32_t sinterp;      /* IMX_SINT_*, with IMX_WRAP_NONE                      */
  int32_t coriolis;     /* IMX_COR_*,  with IMX_WRAP_NONE                      */
  int32_t recon;        /* IMX_REC_*,  with IMX_WRAP_NONE                      */
  double  gasscoef;     /* recon IMX_REC_GASS: weight of the trsk K            */
  int32_t nopv;         /* 0/1                                                */
} imx_scheme;

/* ---- errors -------------------------------------------------------------- */
const char *imx_last_error(void);

/* ---- mesh (replaces meshbuild for icos/nopt grids, src/smeshpack.f90:77-368,
        and initialize_gridprop, src/swm.f90:979-1041) ------------------------ */
int  imx_mesh_build_icos(int32_t glevel, int32_t pos, imx_mesh **out);
/* meshbuild for mesh%kind = "read" (meshread, src/smeshpack.f90:7549-7700,
   then meshstruct): nodes from path (.xyz: count then x y z per line; .gmt:
   lon lat in degrees per line) and the neighbour rings from the .ngb file of
   the same name (6 indices per node, 0 for none). Without the .ngb file the
   nodes are triangulated as the reference's STRIPACK trmesh does
   (src/stripackm.f90:6843-7314); at most IMX_MAXNB neighbours per node. */
int  imx_mesh_build_read(const char *path, imx_mesh **out);
int  imx_mesh_desc_of(const imx_mesh *m, imx_mesh_desc *out);
int  imx_mesh_gridprop(imx_mesh *m);             /* fills the optional arrays */
int  imx_mesh_write(const imx_mesh *m, const char *path);  /* IMXBIN records  */
/* Reads what imx_mesh_write writes (the records of grid_structure that
   imx_mesh_desc and the closed-form test cases read): no counterpart in the
   reference, which rebuilds or meshloads its grids. */
int  imx_mesh_read(const char *path, imx_mesh **out);
/* Grid files in the reference's binary layout (meshload / meshwrite,
   src/smeshpack.f90:7702-8221): <dir>/<name>_header.dat and the ten data
   files (Fortran unformatted sequential records). Loading fills every array
   of imx_mesh_desc (call imx_mesh_gridprop after it); writing emits all 11
   files so the reference's meshload can read a library mesh. */
int  imx_mesh_load_ref(const char *dir, const char *name, imx_mesh **out);
int  imx_mesh_write_ref(const imx_mesh *m, const char *dir, const char *name);
/* plot_scalarfield's lat-lon sampling (src/interpack.f90:7376-7514):
   field f at position pos (0 cells, 1 triangles, 2/6 primal edges, 3 dual
   edges; the reference's nearest rules) on nlat x nlon cell-centred
   points; out[3*nlat*nlon] = (lon, lat, value) float32, longitude fastest,
   the stream the reference writes to data/<name>_<grid>.dat. */
int  imx_remap_ll(const imx_mesh_desc *d, int32_t pos, const double *f, int32_t nlon,
                  int32_t nlat, float *out);
/* mkdir -p for the hosts' output folders (grid/, data/): no counterpart in
   the reference, whose runs expect the folders to exist. */
int  imx_make_dir(const char *path);
void imx_mesh_destroy(imx_mesh *m);
/* initialize_gridprop on caller arrays: trhx_intersec_areas
   (src/smeshpack.f90:5080-5154), hxtr_intersec_areas (:4991-5078) and
   calc_trisk_weights (src/interpack.f90:3978-4098). */
int  imx_gridprop(const imx_mesh_desc *d, double *tr_trhx_area,
                  double *hx_hxtr_area, double *hx_trskw);

/* ---- initial states -------------------------------------------------------- */
/* Williamson TC2 as initialize_fields case(2) with HTC (src/swm.f90:1205-1247)
   on a library-built mesh (needs ed%c%lat, ed%tg and v%lat). */
int  imx_ic_tc2(const imx_mesh *m, double erad, double *h, double *u);
/* the same with HC staggering: u along edhx%nr at the Voronoi edge
   midpoints (src/swm.f90:1262-1270) */
int  imx_ic_tc2_hc(const imx_mesh *m, double erad, double *h, double *u);
/* TC2's exact potential vorticity q_tr_exact at the triangle circumcentres
   (src/swm.f90:1293-1296), for the error columns of write_evol_error. */
int  imx_ic_tc2_pv(const imx_mesh *m, double erad, double *q_tr);
/* initialize_fields (src/swm.f90:1044-1700) for test case tc, staggering
   stag (IMX_STAG_*): h [nv], u [ne] and the bottom topography bt [nv]
   (zero where the case has none), bit-identical to the reference on
   library-built meshes. Cases 0 (cos lon cos^2 lat at rest), 1, 2, 5
   (conical mountain), 51 (Gaussian mountain, fluid depth hollgw minus the
   bump: the Hollingsworth line of par/swm.par), 6 (Rossby-Haurwitz), 21,
   22, 23, 52, 53 (Galewsky jets with and without the perturbation).
   IMX_ENOTSUP for other cases. */
int  imx_ic_testcase(const imx_mesh *m, int32_t tc, int32_t stag, double erad, double hollgw,
                     double *h, double *u, double *bt);
/* The Coriolis parameter a test case runs with (the module globals fsphere
   and fcte that initialize_fields sets, src/swm_data.f90:42-43):
   IMX_F_VARIABLE f = 2 Omega sin(lat) (the default), IMX_F_LINEAR the
   absolute vorticity is the constant fcte alone (11: the linearised
   equations), IMX_F_CONSTANT f = fcte (33, 35, 40, 41 with fcte = 0; 42 with
   1.4584e-4). Apply it with imx_swm_set_coriolis. */
#define IMX_F_VARIABLE 0
#define IMX_F_LINEAR   1
#define IMX_F_CONSTANT 2
int  imx_ic_testcase_coriolis(int32_t tc, int32_t *fsphere, double *fcte);
/* Seeded spectral state: real spherical harmonics l=1..lmax, coefficients
   N(0,1)*l^-2 from SplitMix64 + Box-Muller; h = h0 + h', h' area-rms =
   hfrac*h0; u = -(psi(t2)-psi(t1))/(l_e a) * sign, scaled to area-rms
   speed urms. Not in the reference (new test case). */
int  imx_ic_spectral(const imx_mesh_desc *d, uint64_t seed, int32_t lmax,
                     double h0, double hfrac, double urms, double erad,
                     double *h, double *u);

/* ---- the time step (module swm) ------------------------------------------- */
void imx_swm_default_opts(imx_swm_opts *o);
int  imx_swm_create(const imx_mesh_desc *d, const imx_swm_opts *o, imx_swm **out);
/* Scheme options (imx_scheme); d is the mesh the handle was created on
   (the coefficients and the APVM reconstruction read its geometry).
   FP64_REF handles take every option. Fused (FP32/FP64) handles take, HTC
   or HC, with APVM, diffusion and hyperdiffusion, every method triple of
   the scalar interpolation trsk or bary, the Coriolis term trsk, pered or
   dtred and any kinetic energy; gass18's interpolation, Coriolis term or
   both (trsk's or bary's for the other) with any kinetic energy and without diffusion; and nopv with
   trsk10's methods (with or without diffusion and hyperdiffusion); they
   return IMX_ENOTSUP for other method triples and nopv with other
   options. */
int  imx_swm_set_scheme(imx_swm *s, const imx_mesh_desc *d, const imx_scheme *sc);
int  imx_swm_set_state(imx_swm *s, const double *h, const double *u);
int  imx_swm_get_state(imx_swm *s, double *h, double *u);
/* Bottom topography bt%f [nv] in caller order (module global bt of swm_data,
   set by initialize_fields, e.g. the mountain of test case 5,
   src/swm.f90:1327-1377): the Bernoulli potential becomes g (h + b) + K
   (src/swm.f90:2531-2536) and the potential and available energies take b
   (calc_energies, :2686-2752). NULL restores b = 0. h stays the fluid
   depth. */
int  imx_swm_set_topography(imx_swm *s, const double *bt);
/* The Coriolis term of vort_tr (src/swm_operators.f90:596-605): fsphere
   IMX_F_VARIABLE (f = 2 Omega sin(lat) at the circumcentres, the default),
   IMX_F_LINEAR (eta = fcte, the relative vorticity left out) or
   IMX_F_CONSTANT (eta = zeta + fcte). Every path; the potential enstrophy
   takes the same eta. */
int  imx_swm_set_coriolis(imx_swm *s, int32_t fsphere, double fcte);
/* subroutine tendency(h, u, masseq, momeq)       src/swm.f90:2416-2598 */
int  imx_swm_tendency(imx_swm *s, const double *h, const double *u,
                      double *masseq, double *momeq);
/* An intermediate of the last tendency evaluation (after a step: its 4th
   RK stage, the module globals the reference leaves behind). FP64_REF keeps
   every IMX_FIELD_*; fused handles keep only IMX_FIELD_Q_TR. */
int  imx_swm_get_field(imx_swm *s, int32_t which, double *out);
/* Linearised operator for normal-mode analysis (swm_normal_mode_analysis,
   src/swm.f90:738-976): M(r, i) = (T(x + pert e_i) - T(x)) / pert over the
   nv + ne unknowns (h then u; rows masseq then momeq), as 1-based COO
   triplets sorted by row then column, exact zeros dropped. Columns with
   disjoint stencils share one tendency evaluation (column colouring), which
   leaves every entry bit-identical to the one-column-at-a-time difference.
   With cap smaller than the entry count nothing is written to row, col
   and val, and *nnz (and *ncolors) report the count: a size query. */
int  imx_swm_jacobian(imx_swm *s, const imx_mesh_desc *d, const double *h, const double *u,
                      double pert, int64_t cap, int32_t *row, int32_t *col, double *val,
                      int64_t *nnz, int32_t *ncolors);
/* Local truncation error analysis, swm_horiz_loc_trunc_er
   (src/swm.f90:382-734), on an FP64_REF handle created with testcase 2 or
   one of the Hollingsworth cases 32-35 on mesh m, with the handle's scheme
   options. Sets the test case's state, bottom and Coriolis term
   (initialize_fields, src/swm.f90:1205-1326, :1918-2068) as the resident
   state, evaluates the tendency with the analysis fields (IMX_FIELD_Q_HX..),
   and compares 18 fields with the case's closed forms. imx_swm_lte_hol takes
   the layer depth hollgw of cases 32-35 (the Hollingsworth line of
   par/swm.par); imx_swm_lte passes 0, which those cases refuse.
   err[3 r + 0..2] = error_calc's (max relative, relative L2, max) of row r,
   rows in the order the reference prints them (IMX_LTE_*). Fields the
   tendency does not compute under the handle's options are zero, as in the
   reference: q_grad_tr without APVM, lapu without K2 or K4.
   mim[IMX_LTE_NMIM]: the mimetic sums it prints (src/swm.f90:631-718),
   see IMX_MIM_*. */
#define IMX_LTE_H_ED            0
#define IMX_LTE_DIVUH           1
#define IMX_LTE_KENERGY         2
#define IMX_LTE_ABS_VORT        3
#define IMX_LTE_H_TR            4
#define IMX_LTE_Q_TR            5
#define IMX_LTE_Q_HX            6
#define IMX_LTE_Q_ED            7
#define IMX_LTE_GRAD_GHBK       8
#define IMX_LTE_GRAD_GHBK_EXACT 9  /* grad_ed of the exact g h + K           */
#define IMX_LTE_VHQ_TR         10
#define IMX_LTE_Q_GRAD_TR      11
#define IMX_LTE_VH_HX          12
#define IMX_LTE_VHQ_HX         13
#define IMX_LTE_UHQ_PERP       14
#define IMX_LTE_LAPU           15
#define IMX_LTE_MASSEQ         16
#define IMX_LTE_MOMEQ          17
#define IMX_LTE_NROW           18
#define IMX_MIM_HDIVU          0  /* sum of h div u                          */
#define IMX_MIM_PRODRULE       1  /* sum of h div u + u grad h               */
#define IMX_MIM_DIVUH          2  /* sum of div(hu)                          */
#define IMX_MIM_PRODDIFF       3  /* |[2] - [1]|                             */
#define IMX_MIM_CROSSRULE      4  /* sum of k . rot(grad h)                  */
#define IMX_MIM_CORIOLIS       5  /* sum of Coriolis energy                  */
#define IMX_MIM_CORIOLIS_REL   6  /* [5] / total energy                      */
#define IMX_LTE_NMIM           7
int  imx_swm_lte(imx_swm *s, const imx_mesh *m, double *err, double *mim);
int  imx_swm_lte_hol(imx_swm *s, const imx_mesh *m, double hollgw, double *err, double *mim);
/* nsteps x ode_rk4 on the resident state            src/swm.f90:2600-2683 */
int  imx_swm_step(imx_swm *s, int32_t nsteps);
/* as imx_swm_step, returns device time (HIP events on the handle's stream)  */
int  imx_swm_step_timed(imx_swm *s, int32_t nsteps, double *ms);
/* per-kernel average device time [ms] over nsteps, one entry per kernel of
   the step (names via imx_swm_kernel_name); returns the count in *nk        */
int  imx_swm_profile(imx_swm *s, int32_t nsteps, double *ms, int32_t *nk);
const char *imx_swm_kernel_name(imx_swm *s, int32_t k);
int  imx_swm_diagnostics(imx_swm *s, double out[IMX_NDIAG]);
/* No counterpart in the reference: for callers that cannot allocate (code
   under HIP graph capture, tight serving loops). The diagnostics of
   imx_swm_diagnostics written into the caller's device buffer
   out_dev[IMX_NDIAG], enqueued on the handle's stream without waiting and
   without allocating: the reduction partials live in a static device array
   of the library (4 MiB per GPU, held until the process ends), enough for
   meshes up to about 2.79 million cells (larger ones return IMX_EINVAL).
   The array is shared by every call on a GPU: calls that overlap race. */
int  imx_swm_diagnostics_dev(imx_swm *s, double *out_dev);
int  imx_swm_info(imx_swm *s, int64_t out[8]);   /* nv ne nt bytes_dev ... */
void imx_swm_destroy(imx_swm *s);

/* Autotuner (no counterpart in the reference): times every block size
   (64/128/256/512) x XCD-remap choice of each fused kernel over nsteps steps
   and keeps the fastest; out = {bs, remap} for tri, cell, edge. The state is
   restored. FAST precisions only. */
int  imx_swm_tune(imx_swm *s, int32_t nsteps, int32_t out[6]);

/* ---- ensembles: many independent instances in one set of launches --------
   (no counterpart in the reference; the batched configuration). Instance i
   has its own mesh descs[i] and time step dt[i]; all instances advance
   together, one launch per fused kernel per RK stage, one hipGraph per step.
   FAST precisions only. States and diagnostics are per instance, in that
   instance's own (caller, 1-based) numbering. */
typedef struct imx_ens imx_ens;
int  imx_ens_create(int32_t n, const imx_mesh_desc *descs, const double *dt,
                    const imx_swm_opts *o, imx_ens **out);
int  imx_ens_set_state(imx_ens *e, int32_t i, const double *h, const double *u);
int  imx_ens_get_state(imx_ens *e, int32_t i, double *h, double *u);
int  imx_ens_step(imx_ens *e, int32_t nsteps);
int  imx_ens_step_timed(imx_ens *e, int32_t nsteps, double *ms);
int  imx_ens_diagnostics(imx_ens *e, int32_t i, double out[IMX_NDIAG]);
int  imx_ens_tune(imx_ens *e, int32_t nsteps, int32_t out[6]);
int  imx_ens_info(imx_ens *e, int64_t out[8]);  /* as imx_swm_info, out[7] = n */
void imx_ens_destroy(imx_ens *e);

/* ---- learned correction (no counterpart in the reference) -----------------
   Per-edge MLP 48 -> 128 -> 128 -> 128 -> 2, activation x*clamp(0.5+0.4255x, 0, 1) (hard-sigmoid GELU), BF16 weights
   and activations on MFMA with FP32 accumulation; FP32 master parameters
   updated by AdamW. Parameters are flat FP32: W1[48][128] b1[128]
   W2[128][128] b2 W3[128][128] b3 W4[128][2] b4[2] (W stored [in][out]). */
typedef struct imx_mlp imx_mlp;
int  imx_mlp_create(uint64_t seed, imx_mlp **out);
int  imx_mlp_nparams(void);
int  imx_mlp_get_params(imx_mlp *m, float *p);
/* replaces the parameters; the AdamW moments are kept (a checkpoint
   restore continues the same optimisation) */
int  imx_mlp_set_params(imx_mlp *m, const float *p);
/* draws the parameters imx_mlp_create(seed) starts from and clears the
   AdamW state: a new fit without a new handle */
int  imx_mlp_init(imx_mlp *m, uint64_t seed);
/* host rows X[n][48] -> Y[n][2] */
int  imx_mlp_forward(imx_mlp *m, int64_t n, const float *X, float *Y);
/* L = sum_e w[e] |Y[e] - T[e]|^2: gradient into grad[nparams] (may be NULL),
   loss into *loss; one AdamW step (lr, decoupled weight decay wd on the
   matrices) when apply != 0. Output o with bit o of gmask set takes T[e][o]
   as dL/dY[e][o] itself (a loss defined outside the network) and adds
   nothing to *loss. */
int  imx_mlp_train(imx_mlp *m, int64_t n, const float *X, const float *T,
                   const float *w, float lr, float wd, int32_t apply,
                   double *loss, float *grad, int32_t gmask);
int  imx_mlp_forward_timed(imx_mlp *m, int64_t n, int32_t reps, double *ms);
void imx_mlp_destroy(imx_mlp *m);

/* Learned correction bound to a coarse FP32/FP64 fused handle (d: the mesh
   the handle was created from). No counterpart in the reference: the
   coarse model with a learned per-step correction is a new capability.
   imx_learn_sample saves the state, takes one plain step and builds the 48
   raw features per edge (copied to raw[ne][48] in caller edge order when
   raw != NULL). imx_learn_set_scaling gives the feature mean/std and the
   target scales (scaled target = tscale * physical). imx_learn_train fits
   the MLP to the correction du[ne], dh[nv] of that step (`epochs` AdamW
   steps); the network's second output is an edge mass flux whose
   divergence is the h correction, so corrections conserve total mass.
   loss[2] receives the scaled u and h losses. imx_learn_step runs corrected
   steps (one graph launch each): plain step + MLP correction; a
   correction whose RMS (scaled units) exceeds `bound` is dropped and the
   plain step kept (IMX_LIMIT_FALLBACK, the default), or scaled down to RMS
   `bound` (IMX_LIMIT_SCALE, imx_learn_set_limiter); a non-finite one, or
   bound 0, keeps the plain step exactly in both modes; *nlimited counts
   the limited steps. */
#define IMX_LIMIT_FALLBACK 0
#define IMX_LIMIT_SCALE    1
typedef struct imx_learn imx_learn;
int  imx_learn_create(imx_swm *coarse, imx_mlp *m, const imx_mesh_desc *d, imx_learn **out);
int  imx_learn_sample(imx_learn *l, float *raw);
int  imx_learn_set_scaling(imx_learn *l, const float mu[48], const float sig[48],
                           float tscale_u, float tscale_h);
int  imx_learn_train(imx_learn *l, const double *du, const double *dh, int32_t epochs,
                     float lr, float wd, double loss[2]);
/* Device sample store for shuffled training: imx_learn_reserve allocates
   room for `cap` samples (normalised features + scaled targets, about
   110 bytes per edge); imx_learn_store keeps the last sample with its
   targets (ring buffer); imx_learn_fit runs nsteps AdamW steps, each on one
   stored sample drawn uniformly from `seed`, lr linear from lr0 to lr1,
   loss[2] = mean scaled (u, h) losses of the last min(nsteps,16) steps. */
int  imx_learn_reserve(imx_learn *l, int32_t cap);
int  imx_learn_store(imx_learn *l, const double *du, const double *dh, int32_t *slot);
int  imx_learn_fit(imx_learn *l, int32_t nsteps, float lr0, float lr1, float wd, uint64_t seed,
                   double loss[2]);
int  imx_learn_step(imx_learn *l, int32_t nsteps, double bound, int32_t *nlimited);
int  imx_learn_set_limiter(imx_learn *l, int32_t mode);
int  imx_learn_last_rms(imx_learn *l, double *rms);
void imx_learn_destroy(imx_learn *l);

/* ---- measurement utility (no counterpart in the reference) ----------------
   Streams `bytes` from one device buffer to another `reps` times and returns
   the achieved read+write bandwidth [GB/s]; also the known-size kernel used
   to scale profiler memory counters. Allocates and frees its own buffers. */
int  imx_copy_probe(int64_t bytes, int32_t reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif /* IMX_SWM_H */
