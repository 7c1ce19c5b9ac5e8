// Lat-lon remap for plotting, as plot_scalarfield / scalar_remap_geo2ll
// (src/interpack.f90:7376-7514, :1026-1084): a field on the grid is sampled
// on nlat x nlon cell-centred lat-lon points with the reference's nearest
// rules (src/interpack.f90:205-300):
//   cells (pos 0)            value of the nearest vertex of the triangle
//                            containing the point (getnearnode :3636)
//   triangles (pos 1)        value of the containing triangle (gettr)
//   Voronoi edges (pos 3)    nearest dual-edge midpoint among the
//                            containing triangle's 3 edges, or among the
//                            nearest node's edges away from the
//                            circumcentre (getnearhxedge :3685)
//   primal edges (pos 2, 6)  the same with primal-edge midpoints
//                            (getneartredge :3750)
// and written as float32 triples (lon, lat, value), longitude fastest, the
// stream plot_scalarfield writes. The containing triangle is found by a
// visibility walk on the triangulation from the previous point's triangle
// (the reference searches a lat-lon bucket grid; both give the unique
// containing triangle away from edges). Host code: this is output, not
// the time step.
#include <cmath>
#include <vector>

#include "imx_internal.h"

namespace imx {

namespace {

// p . (a x b): > 0 when p lies left of the great circle a -> b
inline double orient(const double *a, const double *b, const double *p) {
  double c[3];
  cross3(a, b, c);
  return dot3(c, p);
}

struct Walker {
  const imx_mesh_desc &d;
  explicit Walker(const imx_mesh_desc &dd) : d(dd) {}
  const double *vp(int v) const { return d.v_p + 3 * v; }
  // containing triangle of p (0-based), starting from triangle t
  int locate(const double *p, int t) const {
    for (int it = 0; it < 4 * d.nt + 16; ++it) {
      int out = -1;
      double worst = 0.0;
      for (int q = 0; q < 3; ++q) {
        const int a = d.tr_v[3 * t + q] - 1, b = d.tr_v[3 * t + (q + 1) % 3] - 1;
        const double s = orient(vp(a), vp(b), p);
        if (s < worst) {
          worst = s;
          out = q;
        }
      }
      if (out < 0) return t;
      // cross the edge (v_q, v_q+1): the other triangle of that edge
      const int a = d.tr_v[3 * t + out], b = d.tr_v[3 * t + (out + 1) % 3];
      int e = -1;
      for (int q = 0; q < 3; ++q) {
        const int l = d.tr_ed[3 * t + q] - 1;
        const int x = d.ed_v[2 * l], y = d.ed_v[2 * l + 1];
        if ((x == a && y == b) || (x == b && y == a)) e = l;
      }
      if (e < 0) return t;
      t = (d.ed_sh[2 * e] - 1 == t) ? d.ed_sh[2 * e + 1] - 1 : d.ed_sh[2 * e] - 1;
    }
    return t;
  }
};

inline double dist(const double *a, const double *b) {
  double c[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  return norm3(c);
}

}  // namespace

bool remap_ll(const imx_mesh_desc &d, int pos, const double *f, int nlon, int nlat, float *out,
              std::string &err) {
  if (pos != 0 && pos != 1 && pos != 2 && pos != 3 && pos != 6) {
    err = "field position " + std::to_string(pos) + " is not plottable (0, 1, 2, 3, 6)";
    return false;
  }
  const double pi = 4.0 * (double)0.78539816339744830961566f;  // constants.f90:31
  const double deg2rad = pi / 180.0;
  const double dlat = 180.0 / (double)nlat, dlon = 360.0 / (double)nlon;
  double mincdist = 1e300;
  for (int l = 0; l < d.ne; ++l) mincdist = std::min(mincdist, d.edhx_leng[l]);
  const int MX = IMX_MAXNB;
  Walker W(d);
  parallel_ranges(nlat, [&](int j0, int j1) {
    int t = 0;
    for (int j = j0; j < j1; ++j) {
      double tlat = -90 + dlat / 2.0;
      for (int q = 0; q < j; ++q) tlat = tlat + dlat;  // the reference's accumulation
      double tlon = -180 + dlon / 2.0;
      for (int i = 0; i < nlon; ++i) {
        const double lon = tlon * deg2rad, lat = tlat * deg2rad;
        const double coslat = std::cos(lat);
        const double p[3] = {coslat * std::cos(lon), coslat * std::sin(lon), std::sin(lat)};
        t = W.locate(p, t);
        double v = 0.0;
        auto nearnode = [&]() {
          int best = -1;
          double dmin = 10000000.0;
          for (int q = 0; q < 3; ++q) {
            const int vv = d.tr_v[3 * t + q] - 1;
            const double dd = dist(d.v_p + 3 * vv, p);
            if (dd < dmin) {
              best = vv;
              dmin = dd;
            }
          }
          return best;
        };
        if (pos == 0) {
          v = f[nearnode()];
        } else if (pos == 1) {
          v = f[t];
        } else {
          const double *ctr = (pos == 3) ? d.edhx_c : d.ed_c;
          int best = -1;
          double dmin = 100000.0;
          if (arclen(d.tr_c + 3 * t, p) < mincdist / 2.0) {
            for (int q = 0; q < 3; ++q) {
              const int l = d.tr_ed[3 * t + q] - 1;
              const double dd = arclen(ctr + 3 * l, p);
              if (dd < dmin) {
                best = l;
                dmin = dd;
              }
            }
          } else {
            const int nd = nearnode();
            for (int q = 0; q < d.v_nnb[nd]; ++q) {
              const int l = d.v_ed[MX * nd + q] - 1;
              const double dd = arclen(ctr + 3 * l, p);
              if (dd < dmin) {
                best = l;
                dmin = dd;
              }
            }
          }
          v = f[best];
        }
        float *o = out + 3 * ((size_t)j * nlon + i);
        o[0] = (float)tlon;
        o[1] = (float)tlat;
        o[2] = (float)v;
        tlon = tlon + dlon;
      }
    }
  });
  return true;
}

}  // namespace imx

extern "C" {

int imx_remap_ll(const imx_mesh_desc *d, int32_t pos, const double *f, int32_t nlon, int32_t nlat, float *out) {
  if (!d || !f || !out || nlon < 1 || nlat < 1)
    return imx::set_error(IMX_EINVAL, "imx_remap_ll: bad argument");
  if (!d->tr_v || !d->tr_ed || !d->ed_v || !d->ed_sh || !d->tr_c || !d->ed_c || !d->edhx_c)
    return imx::set_error(IMX_EINVAL, "imx_remap_ll: the mesh descriptor lacks triangle or edge tables");
  std::string err;
  if (!imx::remap_ll(*d, pos, f, nlon, nlat, out, err)) return imx::set_error(IMX_EINVAL, "imx_remap_ll: " + err);
  return IMX_OK;
}

}  // extern "C"
