// Initial states for libimx: the reference's Williamson TC2 and the seeded
// spectral state used by the throughput and accuracy runs (host, FP64).
#include <cmath>
#include <cstring>
#include <vector>

#include "imx_internal.h"

namespace imx {

// convert_vec_sph2cart (src/smeshpack.f90:4379-4417)
void vec_sph2cart(double vlon, double vlat, const double *p, double *v) {
  double r2 = p[0] * p[0];
  r2 = r2 + p[1] * p[1];
  double rho = std::sqrt(r2);
  if (rho == 0.0) {
    v[0] = vlat;
    v[1] = vlon;
    v[2] = 0.0;
    return;
  }
  v[0] = -vlon * (p[1] / rho) - (vlat * p[0] * p[2]) / rho;
  v[1] = vlon * (p[0] / rho) - (vlat * p[1] * p[2]) / rho;
  v[2] = vlat * rho;
}

// SplitMix64 (Steele, Lea, Flood 2014) and Box-Muller normals.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }  // (0,1)
  double normal() {
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

// Real orthonormal spherical-harmonic expansion
//   f = sum_{l=1..L} sum_{m=0..l} Pbar_lm(sin lat) (a_lm cos m lon + b_lm sin m lon)
// with Pbar the orthonormal associated Legendre functions (standard three-term
// recurrence in l at fixed m).
struct SphExpansion {
  int L;
  std::vector<double> a, b;  // index l*(L+1)+m
  SphExpansion(int L_, SplitMix64 &rng) : L(L_), a((L_ + 1) * (L_ + 1), 0.0), b((L_ + 1) * (L_ + 1), 0.0) {
    for (int l = 1; l <= L; ++l)
      for (int m = 0; m <= l; ++m) {
        double amp = 1.0 / ((double)l * (double)l);
        a[l * (L + 1) + m] = amp * rng.normal();
        b[l * (L + 1) + m] = (m > 0) ? amp * rng.normal() : 0.0;
      }
  }
  double eval(const double *p, std::vector<double> &P) const {
    const double x = p[2];  // sin(lat) on the unit sphere
    const double rho = std::sqrt(std::max(0.0, 1.0 - x * x));
    const double lon = std::atan2(p[1], p[0]);
    const int W = L + 1;
    P.assign((size_t)W * W, 0.0);
    // Pbar_00 = 1/sqrt(4 pi); Pbar_mm = -sqrt((2m+1)/(2m)) rho Pbar_{m-1,m-1}
    P[0] = 1.0 / std::sqrt(4.0 * M_PI);
    for (int m = 1; m <= L; ++m)
      P[m * W + m] = -std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * rho * P[(m - 1) * W + (m - 1)];
    for (int m = 0; m < L; ++m) P[(m + 1) * W + m] = x * std::sqrt(2.0 * m + 3.0) * P[m * W + m];
    for (int m = 0; m <= L; ++m)
      for (int l = m + 2; l <= L; ++l) {
        double al = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        double bl = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        P[l * W + m] = al * (x * P[(l - 1) * W + m] - bl * P[(l - 2) * W + m]);
      }
    double f = 0.0;
    for (int l = 1; l <= L; ++l)
      for (int m = 0; m <= l; ++m) {
        double nrm = (m == 0) ? 1.0 : std::sqrt(2.0);
        f += nrm * P[l * W + m] * (a[l * W + m] * std::cos(m * lon) + b[l * W + m] * std::sin(m * lon));
      }
    return f;
  }
};

}  // namespace imx

extern "C" {

// initialize_fields, case(2) with HTC staggering (src/swm.f90:1205-1247).
int imx_ic_tc2(const imx_mesh *m, double erad, double *h, double *u) {
  if (!m || !h || !u) return imx::set_error(IMX_EINVAL, "imx_ic_tc2: NULL argument");
  const double pi = 3.1415927410125732;  // constants.f90:31 (single-rounded)
  const double pi2 = 2.0 * pi;
  const double day2sec = 86400.0;
  const double grav = 9.80616, gravi = 1.0 / 9.80616, omega = 7.292e-5;
  (void)grav;
  const double u0 = pi2 * erad / (12.0 * day2sec);
  const double h0 = 2.94e4 * gravi;
  const double h_ct = (erad * omega * u0 + (u0 * u0) / 2.0) * gravi;
  for (int i = 0; i < m->nv; ++i) {
    double s = std::sin(m->v_lat[i]);
    h[i] = h0 - h_ct * (s * s);
  }
  for (int l = 0; l < m->ne; ++l) {
    double v[3];
    imx::vec_sph2cart(u0 * std::cos(m->ed_clat[l]), 0.0, &m->ed_c[3 * l], v);
    u[l] = imx::dot3(v, &m->ed_tg[3 * l]);
  }
  return IMX_OK;
}

// initialize_fields, case(2) with HC staggering (src/swm.f90:1262-1270): u
// is the component of the zonal flow along edhx%nr at the Voronoi edge
// midpoint, whose latitude cart2sph gives (midpoint_edhx, smeshpack.f90:6311-6334).
int imx_ic_tc2_hc(const imx_mesh *m, double erad, double *h, double *u) {
  if (!m || !h || !u) return imx::set_error(IMX_EINVAL, "imx_ic_tc2_hc: NULL argument");
  if (m->edhx_c.size() < (size_t)3 * m->ne || m->edhx_nr.size() < (size_t)3 * m->ne)
    return imx::set_error(IMX_EINVAL, "imx_ic_tc2_hc: the mesh has no Voronoi edge vectors");
  const double pi = 3.1415927410125732;  // constants.f90:31 (single-rounded)
  const double pi2 = 2.0 * pi;
  const double day2sec = 86400.0;
  const double gravi = 1.0 / 9.80616, omega = 7.292e-5;
  const double u0 = pi2 * erad / (12.0 * day2sec);
  const double h0 = 2.94e4 * gravi;
  const double h_ct = (erad * omega * u0 + (u0 * u0) / 2.0) * gravi;
  for (int i = 0; i < m->nv; ++i) {
    double s = std::sin(m->v_lat[i]);
    h[i] = h0 - h_ct * (s * s);
  }
  for (int l = 0; l < m->ne; ++l) {
    const double *c = &m->edhx_c[3 * l];
    double lon, lat, v[3];
    imx::cart2sph(c[0], c[1], c[2], &lon, &lat);
    imx::vec_sph2cart(u0 * std::cos(lat), 0.0, c, v);
    u[l] = imx::dot3(v, &m->edhx_nr[3 * l]);
  }
  return IMX_OK;
}

// TC2's exact potential vorticity at triangle circumcentres
// (src/swm.f90:1293-1296): eta_ct sin(lat) / (h0 - h_ct sin(lat)^2).
int imx_ic_tc2_pv(const imx_mesh *m, double erad, double *q) {
  if (!m || !q) return imx::set_error(IMX_EINVAL, "imx_ic_tc2_pv: NULL argument");
  const double pi = 3.1415927410125732;  // constants.f90:31 (single-rounded)
  const double pi2 = 2.0 * pi;
  const double day2sec = 86400.0;
  const double gravi = 1.0 / 9.80616, omega = 7.292e-5;
  const double u0 = pi2 * erad / (12.0 * day2sec);
  const double h0 = 2.94e4 * gravi;
  const double h_ct = (erad * omega * u0 + (u0 * u0) / 2.0) * gravi;
  const double eta_ct = 2.0 * u0 / erad + 2.0 * omega;
  for (int k = 0; k < m->nt; ++k) {
    const double s = std::sin(m->tr_clat[k]);
    q[k] = eta_ct * s / (h0 - h_ct * (s * s));
  }
  return IMX_OK;
}

int imx_ic_spectral(const imx_mesh_desc *d, uint64_t seed, int32_t lmax, double h0,
                    double hfrac, double urms, double erad, double *h, double *u) {
  if (!d || !h || !u) return imx::set_error(IMX_EINVAL, "imx_ic_spectral: NULL argument");
  if (lmax < 1 || lmax > 64) return imx::set_error(IMX_EINVAL, "imx_ic_spectral: lmax must be in 1..64");
  imx::SplitMix64 rng(seed);
  imx::SphExpansion hs(lmax, rng);
  imx::SphExpansion ps(lmax, rng);
  // h' at cell nodes, scaled to area-rms hfrac*h0
  imx::parallel_ranges(d->nv, [&](int i0, int i1) {
    std::vector<double> P;
    for (int i = i0; i < i1; ++i) h[i] = hs.eval(d->v_p + 3 * i, P);
  });
  double sa = 0.0, s2 = 0.0;
  for (int i = 0; i < d->nv; ++i) {
    sa += d->hx_areag[i];
    s2 += d->hx_areag[i] * h[i] * h[i];
  }
  double hs_rms = std::sqrt(s2 / sa);
  double hscale = (hs_rms > 0) ? hfrac * h0 / hs_rms : 0.0;
  for (int i = 0; i < d->nv; ++i) h[i] = h0 + hscale * h[i];
  // streamfunction at circumcentres; u = -d psi / dn along ed%nr
  std::vector<double> psi(d->nt);
  imx::parallel_ranges(d->nt, [&](int k0, int k1) {
    std::vector<double> P;
    for (int k = k0; k < k1; ++k) psi[k] = ps.eval(d->tr_c + 3 * k, P);
  });
  for (int l = 0; l < d->ne; ++l) {
    const double *p1 = d->v_p + 3 * (d->ed_v[2 * l] - 1);
    const double *p2 = d->v_p + 3 * (d->ed_v[2 * l + 1] - 1);
    int t1 = d->ed_sh[2 * l] - 1, t2 = d->ed_sh[2 * l + 1] - 1;
    double n[3], dc[3];
    imx::cross3(p1, p2, n);
    for (int q = 0; q < 3; ++q) dc[q] = d->tr_c[3 * t2 + q] - d->tr_c[3 * t1 + q];
    double s = imx::dot3(n, dc) >= 0 ? 1.0 : -1.0;
    u[l] = -s * (psi[t2] - psi[t1]) / (d->edhx_leng[l] * erad);
  }
  // scale to rms speed sqrt(2 <K>) = urms, K as kinetic_energy_hx (trsk)
  double ka = 0.0;
  sa = 0.0;
  for (int i = 0; i < d->nv; ++i) {
    double ke = 0.0, ca = 0.0;
    for (int j = 0; j < d->v_nnb[i]; ++j) {
      int e = d->v_ed[IMX_MAXNB * i + j] - 1;
      double w = 0.25 * d->ed_leng[e] * d->edhx_leng[e];
      ca += w;
      ke += w * u[e] * u[e];
    }
    ka += d->hx_areag[i] * ke / ca;
    sa += d->hx_areag[i];
  }
  double urms0 = std::sqrt(2.0 * ka / sa);
  double uscale = (urms0 > 0) ? urms / urms0 : 0.0;
  for (int l = 0; l < d->ne; ++l) u[l] *= uscale;
  return IMX_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------
// The test cases of initialize_fields (src/swm.f90:1043-2414) that define a
// state analytically or from a tabulated balance: 0, 1, 2, 5, 51, 6, 11, 12,
// 21, 22, 23, 52, 53, 32-35 and 40-42. Every expression keeps the reference's operation order,
// its single-precision literals (0.4, 2., 7.0, epsilon(1.)) and the
// single-rounded pi, so the states equal the reference's bit for bit.
namespace imx {
namespace {

constexpr double kPi = 3.1415927410125732;  // constants.f90:31 (single-rounded)
constexpr double kPio2 = kPi / 2.0;
constexpr double kPiby2 = kPi * 0.5;
constexpr double kDeg2rad = kPi / 180.0;
constexpr double kEpsSingle = 1.1920928955078125e-07;  // epsilon(1.)
constexpr double kGrav = 9.80616, kGravi = 1.0 / 9.80616, kOmega = 7.292e-5, kDay2sec = 86400.0;

// real(8)**integer as the Fortran runtime evaluates it: binary powering
// from 1 (a product by 1 is exact, so x**2 = x*x and x**3 = x*(x*x))
double ipow(double b, int n) {
  double r = 1.0;
  for (;;) {
    if (n & 1) r *= b;
    n >>= 1;
    if (n == 0) break;
    b *= b;
  }
  return r;
}

// position, latitude, longitude and the direction u is measured along, at
// edge l: ed%c and ed%tg (HTC) or edhx%c and edhx%nr (HC)
struct EdgePoint {
  const double *p, *dir;
  double lat, lon;
};

EdgePoint edge_point(const imx_mesh &m, int l, bool hc) {
  EdgePoint e;
  if (hc) {
    e.p = &m.edhx_c[3 * l];
    e.dir = &m.edhx_nr[3 * l];
    cart2sph(e.p[0], e.p[1], e.p[2], &e.lon, &e.lat);
  } else {
    e.p = &m.ed_c[3 * l];
    e.dir = &m.ed_tg[3 * l];
    e.lat = m.ed_clat[l];
    e.lon = m.ed_clon[l];
  }
  return e;
}

double along(const EdgePoint &e, double vlon, double vlat) {
  double v[3];
  vec_sph2cart(vlon, vlat, e.p, v);
  return dot3(v, e.dir);
}

// case(1): cosine bell of 1000 m radius erad/3 at (0, 0) in a solid-body
// wind of one revolution per 12 days
void tc1(const imx_mesh &m, bool hc, double erad, double *h, double *u) {
  const double u0 = 2.0 * kPi * erad / (12.0 * kDay2sec);
  const double h0 = 1000.0, rmax = erad / 3.0;
  const double p0[3] = {std::cos(0.0) * std::cos(0.0), std::cos(0.0) * std::sin(0.0), std::sin(0.0)};
  for (int i = 0; i < m.nv; ++i) {
    const double r = erad * arclen(&m.v_p[3 * i], p0);
    h[i] = (r < rmax) ? (h0 * 0.5) * (1.0 + std::cos(kPi * r / rmax)) : 0.0;
  }
  for (int l = 0; l < m.ne; ++l) {
    const EdgePoint e = edge_point(m, l, hc);
    u[l] = along(e, u0 * std::cos(e.lat), 0.0);
  }
}

// case(2): steady zonal geostrophic flow
void tc2(const imx_mesh &m, bool hc, double erad, double *h, double *u) {
  const double u0 = 2.0 * kPi * erad / (12.0 * kDay2sec);
  const double h0 = 2.94e4 * kGravi;
  const double h_ct = (erad * kOmega * u0 + (u0 * u0) / 2.0) * kGravi;
  for (int i = 0; i < m.nv; ++i) {
    const double s = std::sin(m.v_lat[i]);
    h[i] = h0 - h_ct * (s * s);
  }
  for (int l = 0; l < m.ne; ++l) {
    const EdgePoint e = edge_point(m, l, hc);
    u[l] = along(e, u0 * std::cos(e.lat), 0.0);
  }
}

// case(5) / case(51): zonal flow over a conical (5) or Gaussian (51) mountain
void tc5(const imx_mesh &m, bool hc, double erad, double hollgw, bool gauss, double *h, double *u, double *bt) {
  const double u0 = 20.0, h0 = 5960.0;
  const double h_ct = (erad * kOmega * u0 + (u0 * u0) / 2.0) * kGravi;
  const double lon0 = -kPi * 0.5, lat0 = kPi / 6.0, rmax = kPi / 9.0;
  if (!gauss) {
    for (int i = 0; i < m.nv; ++i) {
      const double lon = m.v_lon[i], lat = m.v_lat[i];
      const double r = std::sqrt((lon - lon0) * (lon - lon0) + (lat - lat0) * (lat - lat0));
      const double s = std::sin(lat);
      const double hh = h0 - h_ct * (s * s);
      bt[i] = (r < rmax) ? 2000.0 * (1.0 - r / rmax) : 0.0;
      h[i] = hh - bt[i];
    }
  } else {
    const double cl = std::cos(lat0);
    const double p0[3] = {cl * std::cos(lon0), cl * std::sin(lon0), std::sin(lat0)};
    const double c04 = (double)0.4f;  // default-real literal
    for (int i = 0; i < m.nv; ++i) {
      const double s = std::sin(m.v_lat[i]);
      double b = h0 - h_ct * (s * s);
      double d[3];
      for (int q = 0; q < 3; ++q) d[q] = m.v_p[3 * i + q] - p0[q];
      const double r = std::sqrt(dot3(d, d));
      const double rmaxc = 2.0 * (1.0 - std::cos(rmax));
      const double g = std::exp(-(c04 * (r * r) / (rmaxc * rmaxc)));
      bt[i] = b + 20.0 * g;
      h[i] = hollgw - 20.0 * g;
    }
  }
  for (int l = 0; l < m.ne; ++l) {
    const EdgePoint e = edge_point(m, l, hc);
    u[l] = along(e, u0 * std::cos(e.lat), 0.0);
  }
}

// case(6): Rossby-Haurwitz wave of wavenumber 4
void tc6(const imx_mesh &m, bool hc, double erad, double *h, double *u) {
  const double K_RH = 7.848e-6, w = 7.848e-6, h0 = 8.0e3;
  const int M = 4;
  auto h_tc6 = [&](double lon, double lat) {
    const double c = std::cos(lat);
    double atmp;
    if (std::fabs(lat) > kPio2 - kEpsSingle) {
      atmp = 0.0;
    } else {
      atmp = 0.5 * w * (2.0 * kOmega + w) * (c * c) +
             0.25 * (K_RH * K_RH) * ipow(c, M * 2) *
                 ((double)(M + 1) * (c * c) + (double)(2.0f * (float)(M * M) - (float)M - 2.0f) -
                  (double)(2.0f * (float)(M * M)) / (c * c));
    }
    const double btmp = 2.0 * ((kOmega + w) * K_RH / (double)((M + 1) * (M + 2))) * ipow(c, M) *
                        ((double)((float)(M * M) + 2.0f * (float)M + 2.0f) - (double)((M + 1) * (M + 1)) * (c * c));
    const double ctmp = 0.25 * (K_RH * K_RH) * ipow(c, 2 * M) * ((double)(M + 1) * (c * c) - (double)(M + 2));
    return h0 + kGravi * (erad * erad) *
                    (atmp + btmp * std::cos((double)M * lon) + ctmp * std::cos((double)(2.0f * (float)M) * lon));
  };
  for (int l = 0; l < m.ne; ++l) {
    const EdgePoint e = edge_point(m, l, hc);
    const double c = std::cos(e.lat), s = std::sin(e.lat);
    const double ut = erad * w * c + erad * K_RH * ipow(c, M - 1) * ((double)M * (s * s) - (c * c)) *
                                         std::cos((double)M * e.lon);
    const double vt = -(erad * K_RH * (double)M * ipow(c, M - 1) * s * std::sin((double)M * e.lon));
    u[l] = along(e, ut, vt);
  }
  for (int i = 0; i < m.nv; ++i) h[i] = h_tc6(m.v_lon[i], m.v_lat[i]);
}

// case(21, 22, 23, 52, 53): Galewsky et al. barotropic jet, h in gradient
// balance tabulated by trapezoidal integration in latitude and interpolated
// (cubic, linear next to the poles); 21, 23 and 53 add the bump
bool galewsky(const imx_mesh &m, bool hc, int tc, double erad, double *h, double *u, std::string &err) {
  double u00, lat0, lat1;
  if (tc == 23) {
    u00 = 200.0;
    lat0 = kPi / 7.0;
    lat1 = kPi / 2.0 - lat0;
  } else if (tc == 52 || tc == 53) {
    u00 = 80.0;
    lat0 = -(5.0 * kDeg2rad);
    lat1 = -(45.0 * kDeg2rad);
  } else {
    u00 = 80.0;
    lat0 = kPi / 7.0;
    lat1 = kPi / 2.0 - lat0;
  }
  const double en = std::exp(-4.0 / ((lat1 - lat0) * (lat1 - lat0)));
  const double umen = u00 / en;
  const double rearth = erad, rotatn = kOmega;
  const int nygg = 4 * (int)std::floor(std::sqrt((float)m.nv));
  std::vector<double> hgg((size_t)nygg + 2, 0.0);  // 1-based, hgg[1..nygg+1]
  const double dygg = kPi / (double)nygg;
  double totvol = 0.0, totarea = 0.0;
  hgg[1] = 0.0;
  auto jet = [&](double l) {
    const double den = (l - lat0) * (l - lat1);
    return (den < 0.0) ? umen * std::exp(1.0 / den) : 0.0;
  };
  for (int j = 2; j <= nygg; ++j) {
    const double l1 = (double)(j - 2) * dygg - kPiby2;
    const double l2 = (double)(j - 1) * dygg - kPiby2;
    double u1 = jet(l1), u2 = jet(l2);
    u1 = u1 * (2.0 * rotatn * std::sin(l1) + std::tan(l1) * u1 / rearth);
    u2 = u2 * (2.0 * rotatn * std::sin(l2) + std::tan(l2) * u2 / rearth);
    hgg[j] = hgg[j - 1] - rearth * 0.5 * (u1 + u2) * dygg;
    totarea = totarea + std::cos(l2) * dygg;
    totvol = totvol + hgg[j] * std::cos(l2) * dygg;
  }
  hgg[nygg + 1] = hgg[nygg];
  totvol = totvol / (totarea * kGrav);
  for (int j = 1; j <= nygg + 1; ++j) hgg[j] = hgg[j] + (1.0e4 - totvol) * kGrav;
  for (int j = 1; j <= nygg + 1; ++j) hgg[j] = hgg[j] * kGravi;
  for (int i = 0; i < m.nv; ++i) {
    const double l1 = m.v_lat[i] + kPiby2;
    const int jy = (int)std::floor(l1 / dygg) + 1;
    const double beta = (l1 - (double)(jy - 1) * dygg) / dygg;
    if (jy < 1 || jy > nygg) {
      // a node exactly at the north pole: the reference's cubic stencil
      // reads past its table there
      err = "Galewsky state: a node lies on the pole (latitude table index " + std::to_string(jy) + ")";
      return false;
    }
    if (jy == 1 || jy == nygg) {
      const double cc2 = 1.0 - beta, cc3 = beta;
      h[i] = cc2 * hgg[jy] + cc3 * hgg[jy + 1];
    } else {
      const double cc1 = -(beta * (beta - 1.0) * (beta - 2.0) / 6.0);
      const double cc2 = 0.5 * (beta + 1.0) * (beta - 1.0) * (beta - 2.0);
      const double cc3 = -(0.5 * (beta + 1.0) * beta * (beta - 2.0));
      const double cc4 = (beta + 1.0) * beta * (beta - 1.0) / 6.0;
      h[i] = cc1 * hgg[jy - 1] + cc2 * hgg[jy] + cc3 * hgg[jy + 1] + cc4 * hgg[jy + 2];
    }
  }
  for (int l = 0; l < m.ne; ++l) {
    const EdgePoint e = edge_point(m, l, hc);
    u[l] = along(e, jet(e.lat), 0.0);
  }
  if (tc == 21 || tc == 23 || tc == 53) {
    const double hpert = 120.0, alpha = 1.0 / 3.0, beta = 1.0 / 15.0;
    const double lat2 = (tc == 53) ? -(25.0 * kDeg2rad) : 0.5 * kPiby2;
    for (int i = 0; i < m.nv; ++i) {
      const double l2 = m.v_lat[i];
      double l1 = m.v_lon[i];
      const double clat = std::cos(l2);
      if (tc == 53) l1 = l1 + 120.0 * kDeg2rad;
      const double a = l1 / alpha, b = (lat2 - l2) / beta;
      const double e1 = std::exp(-(a * a));
      const double e2 = std::exp(-(b * b));
      h[i] = h[i] + hpert * clat * e1 * e2;
    }
  }
  return true;
}

// case(11, 12): the linearised equations at rest, for normal modes
void tc11(const imx_mesh &m, double *h, double *u) {
  for (int i = 0; i < m.nv; ++i) h[i] = 1e5 * kGravi;
  std::fill(u, u + m.ne, 0.0);
}

// case(32, 33, 34, 35): Hollingsworth's test, a layer of depth hollgw over
// a bottom in balance with the zonal flow of case 2; 33 and 35 on a
// non-rotating sphere (f-sphere with f = 0, the balance without Coriolis)
void hollingsworth(const imx_mesh &m, bool hc, int tc, double erad, double hollgw, double *h, double *u,
                   double *bt) {
  const double u0 = 2.0 * kPi * erad / (12.0 * kDay2sec);
  const double h0 = 2.94e4 * kGravi;
  const bool fs = (tc == 33 || tc == 35);
  const double fcte = 0.0;
  const double h_ct = fs ? ((u0 * u0) / 2.0) * kGravi : (erad * kOmega * u0 + (u0 * u0) / 2.0) * kGravi;
  for (int i = 0; i < m.nv; ++i) {
    h[i] = hollgw;
    const double s = std::sin(m.v_lat[i]);
    double b = h0 - h_ct * (s * s);
    if (fs) b = b - (erad / kGrav) * fcte * s;
    bt[i] = b;
  }
  for (int l = 0; l < m.ne; ++l) {
    const EdgePoint e = edge_point(m, l, hc);
    u[l] = along(e, u0 * std::cos(e.lat), 0.0);
  }
}

// the rotation of cases 41 and 42 that moves (lon0, lat0) to the pole:
// r = R p, summed in column order as matmul does
struct Rotation {
  double R[3][3];
  Rotation(double lon0, double lat0) {
    R[0][0] = std::sin(lat0) * std::cos(lon0);
    R[0][1] = std::sin(lat0) * std::sin(lon0);
    R[0][2] = -std::cos(lat0);
    R[1][0] = -std::sin(lon0);
    R[1][1] = std::cos(lon0);
    R[1][2] = 0.0;
    R[2][0] = std::cos(lat0) * std::cos(lon0);
    R[2][1] = std::cos(lat0) * std::sin(lon0);
    R[2][2] = std::sin(lat0);
  }
  void apply(const double *p, double *r) const {
    for (int i = 0; i < 3; ++i) {
      double a = 0.0;
      for (int j = 0; j < 3; ++j) a = a + R[i][j] * p[j];
      r[i] = a;
    }
  }
  void apply_t(const double *p, double *r) const {
    for (int i = 0; i < 3; ++i) {
      double a = 0.0;
      for (int j = 0; j < 3; ++j) a = a + R[j][i] * p[j];
      r[i] = a;
    }
  }
};

// case(40, 41, 42): steady localised flows, h0 (2 - sin(lat)^n) about the
// pole (40) or about (1 deg, 3 deg) (41, 42, through the rotation). 40 and
// 41: u = u0 sin^k cos with f = 0; 42: on the f-sphere f = fcte, u from the
// gradient-wind balance, at rest south of the rotated equator
void localised(const imx_mesh &m, bool hc, int tc, double erad, double *h, double *u, double *bt) {
  const double h0 = (tc == 40 ? 10000.0 : 100000.0) * kGravi;
  const int k = (tc == 42) ? 160 : 120;
  const int n = k * 2 + 2;
  const double u0 = std::sqrt((double)n * h0 * kGrav);
  const double fcte = (tc == 42) ? 1.4584e-4 : 0.0;
  const Rotation rot(1.0 * kDeg2rad, 3.0 * kDeg2rad);
  for (int i = 0; i < m.nv; ++i) {
    double lat = m.v_lat[i];
    if (tc != 40) {
      double p[3], lon;
      rot.apply(&m.v_p[3 * i], p);
      cart2sph(p[0], p[1], p[2], &lon, &lat);
    }
    if (tc == 42 && !(lat > 0.0)) h[i] = 2.0 * h0;
    else h[i] = h0 * (2.0 - ipow(std::sin(lat), n));
    bt[i] = 0.0;
  }
  for (int l = 0; l < m.ne; ++l) {
    const EdgePoint e = edge_point(m, l, hc);
    if (tc == 40) {
      u[l] = along(e, u0 * ipow(std::sin(e.lat), k) * std::cos(e.lat), 0.0);
      continue;
    }
    double p[3], lon, lat;
    rot.apply(e.p, p);
    cart2sph(p[0], p[1], p[2], &lon, &lat);
    double ut;
    if (tc == 41) {
      ut = u0 * ipow(std::sin(lat), k) * std::cos(lat);
    } else {
      // the HC branch tests lat > eps, the HTC one |lat| > eps
      const bool away = hc ? lat > kEpsSingle : std::fabs(lat) > kEpsSingle;
      const double F = away ? erad * fcte * std::cos(lat) / std::sin(lat) : 0.0;
      const double c = std::cos(lat);
      const double C = -(kGrav * h0 * (double)n * ipow(std::sin(lat), n - 2) * (c * c));
      ut = (lat > 0.0) ? (-F + std::sqrt(F * F - 4.0 * C)) / 2.0 : 0.0;
    }
    double v[3], w[3];
    vec_sph2cart(ut, 0.0, p, v);
    rot.apply_t(v, w);
    u[l] = dot3(w, e.dir);
  }
}

}  // namespace
}  // namespace imx

extern "C" {

int imx_ic_testcase_coriolis(int32_t tc, int32_t *fsphere, double *fcte) {
  if (!fsphere || !fcte) return imx::set_error(IMX_EINVAL, "imx_ic_testcase_coriolis: NULL argument");
  *fsphere = IMX_F_VARIABLE;
  *fcte = 0.0;
  switch (tc) {
    case 11: *fsphere = IMX_F_LINEAR; *fcte = 1.4584e-4; break;
    case 33: case 35: case 40: case 41: *fsphere = IMX_F_CONSTANT; break;
    case 42: *fsphere = IMX_F_CONSTANT; *fcte = 1.4584e-4; break;
    default: break;
  }
  return IMX_OK;
}

int imx_ic_testcase(const imx_mesh *m, int32_t tc, int32_t stag, double erad, double hollgw, double *h, double *u,
                    double *bt) {
  if (!m || !h || !u || !bt) return imx::set_error(IMX_EINVAL, "imx_ic_testcase: NULL argument");
  if (stag != IMX_STAG_HTC && stag != IMX_STAG_HC) return imx::set_error(IMX_EINVAL, "imx_ic_testcase: bad stag");
  const bool hc = stag == IMX_STAG_HC;
  if (hc && (m->edhx_c.size() < (size_t)3 * m->ne || m->edhx_nr.size() < (size_t)3 * m->ne))
    return imx::set_error(IMX_EINVAL, "imx_ic_testcase: the mesh has no Voronoi edge vectors");
  if (m->v_lat.size() < (size_t)m->nv || m->ed_clat.size() < (size_t)m->ne)
    return imx::set_error(IMX_EINVAL, "imx_ic_testcase: the mesh has no latitudes");
  std::fill(bt, bt + m->nv, 0.0);
  std::string err;
  switch (tc) {
    case 0:  // time discretisation check: h = cos(lon) cos(lat)^2, u = 0
      for (int i = 0; i < m->nv; ++i) {
        const double c = std::cos(m->v_lat[i]);
        h[i] = std::cos(m->v_lon[i]) * (c * c);
      }
      std::fill(u, u + m->ne, 0.0);
      break;
    case 1: imx::tc1(*m, hc, erad, h, u); break;
    case 2: imx::tc2(*m, hc, erad, h, u); break;
    case 5: imx::tc5(*m, hc, erad, hollgw, false, h, u, bt); break;
    case 51: imx::tc5(*m, hc, erad, hollgw, true, h, u, bt); break;
    case 6: imx::tc6(*m, hc, erad, h, u); break;
    case 11:
    case 12: imx::tc11(*m, h, u); break;
    case 32:
    case 33:
    case 34:
    case 35:
      if (hollgw < imx::kEpsSingle)
        return imx::set_error(IMX_EINVAL, "imx_ic_testcase: the Hollingsworth test cases need a positive layer "
                                          "thickness (hollgw)");
      imx::hollingsworth(*m, hc, tc, erad, hollgw, h, u, bt);
      break;
    case 40:
    case 41:
    case 42: imx::localised(*m, hc, tc, erad, h, u, bt); break;
    case 21:
    case 22:
    case 23:
    case 52:
    case 53:
      if (!imx::galewsky(*m, hc, tc, erad, h, u, err)) return imx::set_error(IMX_EINVAL, "imx_ic_testcase: " + err);
      break;
    default:
      return imx::set_error(IMX_ENOTSUP, "imx_ic_testcase: test case " + std::to_string(tc) +
                                             " is not part of this build (0, 1, 2, 5, 6, 11, 12, 21, 22, 23, 32, 33, 34, 35, 40, 41, "
                                             "42, 51, 52, 53 are)");
  }
  return IMX_OK;
}

}  // extern "C"
