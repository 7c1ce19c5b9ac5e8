// Local truncation error analysis of the tendency on TC2 (host side of
// imx_swm_lte): the closed forms of initialize_fields case(2) under
// test_lterror = 1 (src/swm.f90:1205-1326), error_calc / error_calc_vec
// (src/swm.f90:3492-3611) and the mimetic sums of swm_horiz_loc_trunc_er
// (src/swm.f90:626-718). Every field of the tendency comes from the GPU; this
// file holds the analysis the reference runs once after it, in its
// expression and summation order, so the numbers match it bit for bit.
#include <cmath>
#include <limits>
#include <vector>

#include "imx_internal.h"

namespace imx {

namespace {

// sumfsq_areas (src/swm.f90:3441-3490): area-weighted mean of f^2 with the
// weights of the field's position (cells, triangles or edges).
enum Pos { CELL = 0, TRI = 1, EDGE = 2 };

double sumfsq_areas(const HostMesh &m, Pos pos, const double *f) {
  double s = 0.0, a = 0.0;
  if (pos == CELL) {
    for (int i = 0; i < m.nv; ++i) {
      a = a + m.hx_areag[i];
      s = s + m.hx_areag[i] * (f[i] * f[i]);
    }
  } else if (pos == TRI) {
    for (int k = 0; k < m.nt; ++k) {
      a = a + m.tr_areag[k];
      s = s + m.tr_areag[k] * (f[k] * f[k]);
    }
  } else {
    for (int l = 0; l < m.ne; ++l) {
      const double w = m.ed_leng[l] * m.edhx_leng[l] * 0.5;
      a = a + w;
      s = s + w * (f[l] * f[l]);
    }
  }
  return s / a;
}

// error_calc (src/swm.f90:3492-3543): out = (max relative, relative L2, max)
void error_calc(const HostMesh &m, Pos pos, int n, const double *f, const double *fref, double *out) {
  std::vector<double> dif(n);
  for (int i = 0; i < n; ++i) dif[i] = f[i] - fref[i];
  const double refsum = sumfsq_areas(m, pos, fref);
  const double e2 = (refsum > 0) ? std::sqrt(sumfsq_areas(m, pos, dif.data())) / std::sqrt(refsum)
                                 : std::sqrt(sumfsq_areas(m, pos, dif.data()));
  double refmax = -std::numeric_limits<double>::max(), emax = -std::numeric_limits<double>::max();
  for (int i = 0; i < n; ++i) {
    refmax = std::max(refmax, std::fabs(fref[i]));
    emax = std::max(emax, std::fabs(dif[i]));
  }
  out[0] = (refmax > 0) ? emax / refmax : emax;
  out[1] = e2;
  out[2] = emax;
}

// error_calc_vec (src/swm.f90:3545-3611) on x y z vectors per entry
void error_calc_vec(const HostMesh &m, Pos pos, int n, const double *v, const double *vref, double *out) {
  std::vector<double> fd(n), fr(n);
  for (int i = 0; i < n; ++i) {
    const double d[3] = {v[3 * i] - vref[3 * i], v[3 * i + 1] - vref[3 * i + 1], v[3 * i + 2] - vref[3 * i + 2]};
    fd[i] = norm3(d);
    fr[i] = norm3(vref + 3 * i);
  }
  const double refsum = sumfsq_areas(m, pos, fr.data());
  const double e2 = (refsum > 0) ? std::sqrt(sumfsq_areas(m, pos, fd.data())) / std::sqrt(refsum)
                                 : std::sqrt(sumfsq_areas(m, pos, fd.data()));
  double refmax = -std::numeric_limits<double>::max(), emax = -std::numeric_limits<double>::max();
  for (int i = 0; i < n; ++i) {
    refmax = std::max(refmax, std::fabs(fr[i]));
    emax = std::max(emax, std::fabs(fd[i]));
  }
  // with refmax <= 0 the reference leaves errormax unset; TC2's exact
  // vectors are never all zero
  out[0] = (refmax > 0) ? emax / refmax : emax;
  out[1] = e2;
  out[2] = emax;
}

// cos(x) as a call of its own. The compilers fuse sin(x) and cos(x) of one
// argument into one sincos call, whose cos differs from cos at some
// arguments (glibc); the reference's closed forms at edges take the
// separate cos, its initial states (imx_ic.cpp) the fused one.
double cos_alone(double x) {
  volatile double v = x;
  return std::cos(v);
}

}  // namespace

// wachspress_coords_hxv(v(i)%p, mesh, i) (src/interpack.f90:2310-2415): the
// Wachspress coordinates of cell node i in the polygon of its ring
// circumcentres, in the order of v%tr.
void wachspress_node(const imx_mesh_desc &d, int i, double *w) {
  const int n = d.v_nnb[i];
  const double *p = d.v_p + 3 * i;
  const double ar = d.hx_areag[i];
  double A[IMX_MAXNB], B[IMX_MAXNB];
  auto tc = [&](int j) { return d.tr_c + 3 * (d.v_tr[IMX_MAXNB * i + j] - 1); };
  for (int j = 0; j < n; ++j) {
    const double *p1 = tc((j + n - 1) % n), *p2 = tc(j), *p3 = tc((j + 1) % n);
    B[j] = sphtriarea(p1, p2, p3) / ar;
    A[j] = sphtriarea(p2, p3, p) / ar;
  }
  double sw = 0.0;
  for (int j = 0; j < n; ++j) {
    double x = B[j];
    const int jm = (j + n - 1) % n;
    for (int k = 0; k < n; ++k)
      if (k != j && k != jm) x = x * A[k];
    w[j] = x;
    sw = sw + x;
  }
  for (int j = 0; j < n; ++j) w[j] = w[j] / sw;
}

int lte_analysis(const HostMesh &m, int tc, double hollgw, bool hc, bool perhx, double erad, double grav,
                 const LteFields &f, double *err, double *mim) {
  const int nv = m.nv, ne = m.ne, nt = m.nt;
  if (hc && ((int)m.edhx_nr.size() < 3 * ne || (int)m.edhx_tg.size() < 3 * ne))
    return set_error(IMX_EINVAL, "imx_swm_lte: the mesh has no Voronoi edge vectors");
  if ((int)m.ed_tg.size() < 3 * ne || (int)m.ed_nr.size() < 3 * ne || (int)m.hx_nr.size() < IMX_MAXNB * nv)
    return set_error(IMX_EINVAL, "imx_swm_lte: the mesh has no edge vectors");
  // constants of TC2 (src/swm.f90:1206-1209) and of the Hollingsworth
  // cases (:1918-1952): the same zonal flow, over a bottom in balance with
  // it under a layer of constant depth hollgw; 33 and 35 without rotation
  const double pi = 3.1415927410125732;  // constants.f90:31 (single-rounded)
  const double pi2 = 2.0 * pi, day2sec = 86400.0, gravi = 1.0 / 9.80616, omega = 7.292e-5;
  const bool hol = tc >= 32 && tc <= 35, fs = tc == 33 || tc == 35;
  const double u0 = pi2 * erad / (12.0 * day2sec);
  const double h0 = 2.94e4 * gravi;
  const double h_ct = fs ? ((u0 * u0) / 2.0) * gravi : (erad * omega * u0 + (u0 * u0) / 2.0) * gravi;
  const double eta_ct = fs ? (2. * u0 / erad) : (2. * u0 / erad + 2. * omega);
  const double grad_ct = u0 * eta_ct;

  // exact fields at cells
  std::vector<double> ke_x(nv), q_hx_x(nv), vh_x((size_t)3 * nv), vhq_x((size_t)3 * nv), zero(std::max({nv, ne, nt}), 0.0);
  for (int i = 0; i < nv; ++i) {
    const double lat = m.v_lat[i];
    const double c = u0 * std::cos(lat);
    ke_x[i] = c * c / 2.0;
    double v[3];
    vec_sph2cart(c, 0.0, &m.v_p[3 * i], v);
    const double q = eta_ct * std::sin(lat) / f.h[i];
    q_hx_x[i] = q;
    for (int k = 0; k < 3; ++k) {
      vh_x[3 * i + k] = v[k] * f.h[i];
      vhq_x[3 * i + k] = vh_x[3 * i + k] * q;
    }
  }
  // exact fields at edges (primal midpoints with HTC, Voronoi midpoints with HC)
  std::vector<double> h_ed_x(ne), grad_x(ne), q_ed_x(ne), uhq_x(ne), lapu_x(ne);
  for (int l = 0; l < ne; ++l) {
    const double *p = hc ? &m.edhx_c[3 * l] : &m.ed_c[3 * l];
    double lat = m.ed_clat[l];
    if (hc) {
      double lon;
      cart2sph(p[0], p[1], p[2], &lon, &lat);
    }
    const double s = std::sin(lat), c = cos_alone(lat);
    const double *dir = hc ? &m.edhx_nr[3 * l] : &m.ed_tg[3 * l];   // u's direction
    const double *prp = hc ? &m.edhx_tg[3 * l] : &m.ed_nr[3 * l];   // uhq_perp's
    double ved[3], g[3], lp[3];
    vec_sph2cart(u0 * c, 0.0, p, ved);
    h_ed_x[l] = hol ? hollgw : h0 - h_ct * (s * s);
    vec_sph2cart(0.0, -grad_ct * s * c, p, g);
    grad_x[l] = dot3(g, dir);
    q_ed_x[l] = eta_ct * s / h_ed_x[l];
    uhq_x[l] = hc ? +eta_ct * s * dot3(ved, prp) : -eta_ct * s * dot3(ved, prp);
    vec_sph2cart(u0 * (1. - 2. * (c * c)) / ((erad * erad) * c), 0.0, p, lp);
    lapu_x[l] = dot3(lp, dir);
  }
  // exact fields at triangle circumcentres
  std::vector<double> eta_x(nt), h_tr_x(nt), q_tr_x(nt), qg_x((size_t)3 * nt), vhq_tr_x((size_t)3 * nt);
  for (int k = 0; k < nt; ++k) {
    const double lat = m.tr_clat[k];
    const double s = std::sin(lat), c = std::cos(lat);
    if (!hol) {  // TC2's grad q at the circumcentre (the others leave it zero)
      const double dd = h0 - h_ct * (s * s);
      const double vt = eta_ct * c * (h0 + h_ct * (s * s)) / (erad * (dd * dd));
      vec_sph2cart(0., vt, &m.tr_c[3 * k], &qg_x[3 * k]);
    }
    eta_x[k] = eta_ct * s;
    h_tr_x[k] = hol ? hollgw : h0 - h_ct * (s * s);
    q_tr_x[k] = eta_x[k] / h_tr_x[k];
    double v[3];
    vec_sph2cart(u0 * c, 0.0, &m.tr_c[3 * k], v);
    for (int q = 0; q < 3; ++q) vhq_tr_x[3 * k + q] = v[q] * eta_x[k];
  }
  // grad_ed of the exact potential g (h + b) + K, grad_edge_Cgrid
  // (src/diffoperpack.f90:226-330, trsk) over erad
  std::vector<double> ghbk_x(nv), grad_gx(ne);
  for (int i = 0; i < nv; ++i) ghbk_x[i] = grav * (f.h[i] + f.bt[i]) + ke_x[i];
  auto grad_cells = [&](const double *hf, int l) {
    const int v1 = m.ed_v[2 * l] - 1, v2 = m.ed_v[2 * l + 1] - 1;
    double g = hf[v2] - hf[v1];
    if (hc) {
      const double dv[3] = {m.v_p[3 * v2] - m.v_p[3 * v1], m.v_p[3 * v2 + 1] - m.v_p[3 * v1 + 1],
                            m.v_p[3 * v2 + 2] - m.v_p[3 * v1 + 2]};
      const double sg = std::signbit(dot3(&m.edhx_nr[3 * l], dv)) ? -1.0 : 1.0;
      g = sg * g;
    }
    return g / m.ed_leng[l];
  };
  for (int l = 0; l < ne; ++l) grad_gx[l] = grad_cells(ghbk_x.data(), l) / erad;

  // the table (src/swm.f90:465-623)
  const double *zf = zero.data();
  error_calc(m, EDGE, ne, f.h_ed, h_ed_x.data(), err + 3 * IMX_LTE_H_ED);
  error_calc(m, CELL, nv, f.divuh, zf, err + 3 * IMX_LTE_DIVUH);
  error_calc(m, CELL, nv, f.ke_hx, ke_x.data(), err + 3 * IMX_LTE_KENERGY);
  error_calc(m, TRI, nt, f.eta, eta_x.data(), err + 3 * IMX_LTE_ABS_VORT);
  error_calc(m, TRI, nt, f.h_tr, h_tr_x.data(), err + 3 * IMX_LTE_H_TR);
  error_calc(m, TRI, nt, f.q_tr, q_tr_x.data(), err + 3 * IMX_LTE_Q_TR);
  error_calc(m, CELL, nv, f.q_hx, q_hx_x.data(), err + 3 * IMX_LTE_Q_HX);
  error_calc(m, EDGE, ne, f.q_ed, q_ed_x.data(), err + 3 * IMX_LTE_Q_ED);
  error_calc(m, EDGE, ne, f.grad_ghbk, grad_x.data(), err + 3 * IMX_LTE_GRAD_GHBK);
  error_calc(m, EDGE, ne, grad_gx.data(), grad_x.data(), err + 3 * IMX_LTE_GRAD_GHBK_EXACT);
  error_calc_vec(m, TRI, nt, f.vhq_tr, vhq_tr_x.data(), err + 3 * IMX_LTE_VHQ_TR);
  std::vector<double> z3((size_t)3 * std::max(nv, nt), 0.0);
  error_calc_vec(m, TRI, nt, f.q_grad_tr ? f.q_grad_tr : z3.data(), qg_x.data(), err + 3 * IMX_LTE_Q_GRAD_TR);
  error_calc_vec(m, CELL, nv, f.vh_hx, vh_x.data(), err + 3 * IMX_LTE_VH_HX);
  error_calc_vec(m, CELL, nv, f.vhq_hx, vhq_x.data(), err + 3 * IMX_LTE_VHQ_HX);
  error_calc(m, EDGE, ne, f.uhq_perp, uhq_x.data(), err + 3 * IMX_LTE_UHQ_PERP);
  error_calc(m, EDGE, ne, f.lapu ? f.lapu : zf, lapu_x.data(), err + 3 * IMX_LTE_LAPU);
  error_calc(m, CELL, nv, f.masseq, zf, err + 3 * IMX_LTE_MASSEQ);
  error_calc(m, EDGE, ne, f.momeq, zf, err + 3 * IMX_LTE_MOMEQ);

  // mimetic properties, normalised by maxh = h0 and maxvel = u0
  // (src/swm.f90:636-718)
  const double maxh = h0, maxvel = u0;
  double prod = 0.0;
  for (int i = 0; i < nv; ++i) {
    // div_cell_Cgrid (src/diffoperpack.f90:122-224): pos 6 with HTC
    // (hx%ttgout), pos 3 with HC (hx%nr)
    double dv = 0.0;
    for (int j = 0; j < m.v_nnb[i]; ++j) {
      const int k = m.v_ed[IMX_MAXNB * i + j] - 1;
      const double sg = hc ? (double)m.hx_nr[IMX_MAXNB * i + j] : (double)m.hx_ttgout[IMX_MAXNB * i + j];
      const double nrv = sg * f.u[k];
      dv = dv + nrv * m.edhx_leng[k];
    }
    dv = dv / m.hx_areag[i];
    const double hdivu = f.h[i] * dv / maxh / maxvel;
    prod = prod + m.hx_areag[i] * hdivu;
  }
  mim[IMX_MIM_HDIVU] = prod;
  for (int l = 0; l < ne; ++l) {
    const double ugradh = f.u[l] * grad_cells(f.h, l) / maxh / maxvel;
    prod = prod + ugradh * m.ed_leng[l] * m.edhx_leng[l];
  }
  mim[IMX_MIM_PRODRULE] = prod;
  double sdiv = 0.0;
  for (int i = 0; i < nv; ++i) sdiv = sdiv + m.hx_areag[i] * f.divuh[i] / maxh / maxvel;
  mim[IMX_MIM_DIVUH] = sdiv;
  mim[IMX_MIM_PRODDIFF] = std::fabs(sdiv - prod);
  // cross rule: grad_edge_Cgrid of h_ed, which the reference reads at the
  // edge's cell indices (h_ed%f(ed%v(1:2)), src/swm.f90:667-669); then the
  // circulation of that gradient around each triangle
  std::vector<double> gh(ne);
  for (int l = 0; l < ne; ++l) gh[l] = grad_cells(f.h_ed, l);
  double cross = 0.0;
  for (int k = 0; k < nt; ++k) {
    double c = 0.0;
    for (int l = 0; l < 3; ++l) {
      const int ed = m.tr_ed[3 * k + l] - 1;
      double sg = (double)m.tr_tg[3 * k + l];
      if (hc) {
        const double *t = &m.ed_tg[3 * ed];
        sg = std::signbit((double)m.tr_tg[3 * k + l] * dot3(&m.edhx_nr[3 * ed], t)) ? -1.0 : 1.0;
      }
      c = c + gh[ed] * m.ed_leng[ed] * sg;
    }
    c = c / m.tr_areag[k] / maxh / maxh;
    cross = cross + m.tr_areag[k] * c;
  }
  mim[IMX_MIM_CROSSRULE] = cross;
  double cor = 0.0;
  for (int l = 0; l < ne; ++l) cor = cor + (f.uh[l] * f.uhq_perp[l]) * m.edhx_leng[l] * m.ed_leng[l] * 0.5;
  mim[IMX_MIM_CORIOLIS] = cor;
  // total energy as calc_energies forms it after the tendency
  // (src/swm.f90:2686-2752): the kinetic part from the tendency's h_ed, or
  // with Perot's cell reconstruction (pxt16) the sum of ke_hx over the areas
  double pe = 0.0, ke = 0.0;
  for (int i = 0; i < nv; ++i) pe = pe + m.hx_areag[i] * grav * f.h[i] * (f.h[i] * 0.5 + f.bt[i]);
  if (perhx) {
    for (int i = 0; i < nv; ++i) ke = ke + f.ke_hx[i] * m.hx_areag[i];
  } else {
    for (int l = 0; l < ne; ++l) ke = ke + 0.5 * m.ed_leng[l] * m.edhx_leng[l] * f.h_ed[l] * (f.u[l] * f.u[l]);
  }
  mim[IMX_MIM_CORIOLIS_REL] = cor / (pe + ke);
  return IMX_OK;
}

}  // namespace imx
