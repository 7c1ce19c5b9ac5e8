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
