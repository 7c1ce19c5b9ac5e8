// Host-side icosahedral mesh builder, reference-convention structuring and
// grid-property precomputation for libimx. Compiled with -ffp-contract=off so
// that every FP64 expression rounds exactly as written.
//
// What is kept alike with the reference, and why:
//  * the 12 base nodes (eqs/pol) and the refinement rule "normalised mean of
//    the two end points" (src/smeshpack.f90:1234-1279, :1659-1665): they fix
//    the node coordinates, which are part of the mesh definition;
//  * the numbering RULES of meshstruct (src/smeshpack.f90:706-886): edge l
//    runs from its lower to its higher node, triangles are owned by their
//    lowest node, ed%sh is filled in triangle order; the sign tables
//    tr%tg/nr, hx%tg/nr, hx%ttgout, hx%tnrccw (:840-860, :1071-1104). The C-grid
//    operators are written against these conventions;
//  * the geometry formulas (arclen, sphtriarea, circumcentre, sphhxarea,
//    midpoints) and initialize_gridprop (kite areas, TRiSK weights of
//    Thuburn et al. 2009 eq. 33): they define the discretisation.
// The triangulation is computed as the reference computes it, with STRIPACK's
// incremental Delaunay insertion restated in imx_trmesh.cpp, for the
// refinement levels (midpoint order) and for the final node set (neighbour
// rings), so icosahedral grids carry the reference's numbering.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <atomic>

#include "imx_internal.h"

namespace imx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

// ---------------------------------------------------------------- geometry --
// Fortran dot_product over 3 elements: ((a1 b1 + a2 b2) + a3 b3).
double dot3(const double *a, const double *b) {
  double s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

double norm3(const double *a) { return std::sqrt(dot3(a, a)); }

// arclen (src/smeshpack.f90:4624-4657): geodesic distance on the unit sphere.
double arclen(const double *p, const double *q) {
  double s[3] = {p[0] + q[0], p[1] + q[1], p[2] + q[2]};
  double d = dot3(s, s);
  if (d == 0.0) return 3.1415927410125732;  // the reference's single-rounded pi
  if (d >= 4.0) return 0.0;
  return 2.0 * std::atan(std::sqrt((4.0 - d) / d));
}

// epsilon(1.) of the reference (constants.f90:45): single-precision epsilon.
static const double kEpsSingle = 1.1920928955078125e-07;

static inline void sub3(const double *a, const double *b, double *c) {
  c[0] = a[0] - b[0];
  c[1] = a[1] - b[1];
  c[2] = a[2] - b[2];
}

// sphtriarea (src/smeshpack.f90:4710-4762): L'Huilier's formula.
double sphtriarea(const double *p1, const double *p2, const double *p3) {
  double t[3];
  sub3(p1, p2, t);
  double n12 = norm3(t);
  sub3(p2, p3, t);
  double n23 = norm3(t);
  sub3(p1, p3, t);
  double n13 = norm3(t);
  if (n12 < kEpsSingle / 100.0 || n23 < kEpsSingle / 100.0 || n13 < kEpsSingle / 100.0)
    return 0.0;
  double a = arclen(p1, p2);
  double b = arclen(p2, p3);
  double c = arclen(p3, p1);
  double s = (a + b + c) / 2.0;
  double tmp = std::tan(s / 2.0) * std::tan((s - a) / 2.0);
  tmp = tmp * std::tan((s - b) / 2.0);
  tmp = tmp * std::tan((s - c) / 2.0);
  if (tmp < 0.0) return 0.0;
  return 4.0 * std::atan(std::sqrt(tmp));
}

// det (src/smeshpack.f90:6580): (p1 x p2) . p3
double det3(const double *p1, const double *p2, const double *p3) {
  double c[3];
  cross3(p1, p2, c);
  return dot3(c, p3);
}

// cart2sph (src/smeshpack.f90:4279-4314)
void cart2sph(double x, double y, double z, double *lon, double *lat) {
  double pnrm = x * x;
  pnrm = pnrm + y * y;
  pnrm = pnrm + z * z;
  pnrm = std::sqrt(pnrm);
  *lon = (x != 0.0 || y != 0.0) ? std::atan2(y, x) : 0.0;
  *lat = (pnrm != 0.0) ? std::asin(z / pnrm) : 0.0;
}

static inline double fsign1(double x) { return std::signbit(x) ? -1.0 : 1.0; }

// p <- p / norm(p), element by element (as `p=p/norm(p)`).
static inline void normalize3(double *p) {
  double n = norm3(p);
  p[0] = p[0] / n;
  p[1] = p[1] / n;
  p[2] = p[2] / n;
}

// ------------------------------------------------------------- gridprop ----
// initialize_gridprop for HTC/trsk10 (src/swm.f90:979-1041):
//   trhx_intersec_areas (src/smeshpack.f90:5080-5154),
//   hxtr_intersec_areas (src/smeshpack.f90:4991-5078) normalised by areag,
//   calc_trisk_weights  (src/interpack.f90:3978-4098).
void gridprop(const imx_mesh_desc &d, double *trhx, double *hxtr, double *trskw) {
  const int MX = IMX_MAXNB;
  parallel_ranges(d.nt, [&](int k0, int k1) {
  for (int k = k0; k < k1; ++k) {
    const double *pc = d.tr_c + 3 * k;
    for (int j = 0; j < 3; ++j) {
      int jm = (j + 2) % 3;  // modint(j-1,3) in 1-based terms
      int ed1 = d.tr_ed[3 * k + jm] - 1;
      int ed2 = d.tr_ed[3 * k + j] - 1;
      const double *ped1 = d.ed_c + 3 * ed1;
      const double *ped2 = d.ed_c + 3 * ed2;
      const double *pv = d.v_p + 3 * (d.tr_v[3 * k + j] - 1);
      double area1 = sphtriarea(pc, ped1, pv);
      if (det3(pc, ped1, pv) < 0.0) area1 = -area1;
      double area2 = sphtriarea(pc, pv, ped2);
      if (det3(pc, pv, ped2) < 0.0) area2 = -area2;
      trhx[3 * k + j] = area1 + area2;
    }
  }
  });
  parallel_ranges(d.nv, [&](int i0, int i1) {
  for (int i = i0; i < i1; ++i) {
    int n = d.v_nnb[i];
    const double *pc = d.v_p + 3 * i;
    double *R = hxtr + MX * i;
    for (int j = 0; j < MX; ++j) R[j] = 0.0;
    for (int j = 0; j < n; ++j) {
      int ed1 = d.v_ed[MX * i + j] - 1;
      int ed2 = d.v_ed[MX * i + (j + 1) % n] - 1;
      const double *ped1 = d.edhx_c + 3 * ed1;
      const double *ped2 = d.edhx_c + 3 * ed2;
      const double *pv = d.tr_c + 3 * (d.v_tr[MX * i + j] - 1);
      double area1 = sphtriarea(pc, ped1, pv);
      if (det3(pc, ped1, pv) < 0.0) area1 = -area1;
      double area2 = sphtriarea(pc, pv, ped2);
      if (det3(pc, pv, ped2) < 0.0) area2 = -area2;
      R[j] = area1 + area2;
    }
    for (int j = 0; j < n; ++j) R[j] = R[j] / d.hx_areag[i];
    double *W = trskw + MX * MX * i;
    for (int j = 0; j < MX * MX; ++j) W[j] = 0.0;
    // w(a,b) = sum of R over kites a..b-1 (cyclic) - 1/2, w(a,a) = 0
    for (int a = 0; a < n; ++a) {
      for (int b = 0; b < n; ++b) {
        if (a == b) {
          W[MX * a + b] = 0.0;
          continue;
        }
        double rs = 0.0;
        if (b < a) {
          for (int k = a; k < n; ++k) rs = rs + R[k];
          for (int k = 0; k < b; ++k) rs = rs + R[k];
        } else {
          for (int k = a; k < b; ++k) rs = rs + R[k];
        }
        W[MX * a + b] = rs - 0.5;
      }
    }
  }
  });
}

void desc_of(const HostMesh &m, imx_mesh_desc &d) {
  std::memset(&d, 0, sizeof(d));
  d.nv = m.nv;
  d.ne = m.ne;
  d.nt = m.nt;
  d.glevel = m.glevel;
  d.v_nnb = m.v_nnb.data();
  d.v_nb = m.v_nb.data();
  d.v_ed = m.v_ed.data();
  d.v_tr = m.v_tr.data();
  d.hx_ttgout = m.hx_ttgout.data();
  d.hx_tnrccw = m.hx_tnrccw.data();
  d.v_p = m.v_p.data();
  d.hx_areag = m.hx_areag.data();
  d.ed_v = m.ed_v.data();
  d.ed_sh = m.ed_sh.data();
  d.ed_leng = m.ed_leng.data();
  d.edhx_leng = m.edhx_leng.data();
  d.ed_c = m.ed_c.data();
  d.edhx_c = m.edhx_c.data();
  d.tr_v = m.tr_v.data();
  d.tr_ed = m.tr_ed.data();
  d.tr_tg = m.tr_tg.data();
  d.tr_areag = m.tr_areag.data();
  d.tr_c = m.tr_c.data();
  d.tr_clat = m.tr_clat.data();
  d.tr_trhx_area = m.tr_trhx_area.empty() ? nullptr : m.tr_trhx_area.data();
  d.hx_hxtr_area = m.hx_hxtr_area.empty() ? nullptr : m.hx_hxtr_area.data();
  d.hx_trskw = m.hx_trskw.empty() ? nullptr : m.hx_trskw.data();
}

// ----------------------------------------------------------------- IMXBIN --
BinWriter::BinWriter(const char *path) {
  f = std::fopen(path, "wb");
  if (f) std::fwrite("IMXBIN01", 1, 8, f);
}
BinWriter::~BinWriter() {
  if (f) std::fclose(f);
}
static void put_hdr(FILE *f, const char *name, const char *dt, int64_t n) {
  char nm[24];
  std::memset(nm, ' ', sizeof(nm));
  std::memcpy(nm, name, std::min<size_t>(std::strlen(name), sizeof(nm)));
  std::fwrite(nm, 1, 24, f);
  std::fwrite(dt, 1, 4, f);
  std::fwrite(&n, 8, 1, f);
}
void BinWriter::put_i4(const char *name, const int32_t *a, int64_t n) {
  put_hdr(f, name, "i4  ", n);
  std::fwrite(a, 4, (size_t)n, f);
}
void BinWriter::put_f8(const char *name, const double *a, int64_t n) {
  put_hdr(f, name, "f8  ", n);
  std::fwrite(a, 8, (size_t)n, f);
}
void BinWriter::put(const char *name, const std::vector<int32_t> &a) {
  put_i4(name, a.data(), (int64_t)a.size());
}
void BinWriter::put(const char *name, const std::vector<double> &a) {
  put_f8(name, a.data(), (int64_t)a.size());
}

// ------------------------------------------------------- icosahedral build --
// Base icosahedron (src/smeshpack.f90:1197-1279).
static void icos_base(int pos, double x[12], double y[12], double z[12]) {
  if (pos == IMX_POS_EQS) {
    const double radius = 1.0;
    const double sqrt5 = std::sqrt(5.0);
    const double phi = (1.0 + sqrt5) * 0.5;
    const double ratio = std::sqrt(10.0 + (2.0 * sqrt5)) / (4.0 * phi);
    const double a = (radius / ratio) * 0.5;
    const double b = (radius / ratio) / (2.0 * phi);
    const double tab[12][3] = {{0, b, -a},  {0, -b, -a}, {a, 0, -b},  {-a, 0, -b},
                               {b, a, 0},   {-b, a, 0},  {-a, 0, b},  {a, 0, b},
                               {b, -a, 0},  {-b, -a, 0}, {0, b, a},   {0, -b, a}};
    for (int i = 0; i < 12; ++i) {
      x[i] = tab[i][0];
      y[i] = tab[i][1];
      z[i] = tab[i][2];
    }
  } else {
    // Nodes on the poles. pi is the reference's single-rounded value and the
    // longitude step 2/5.*pi takes 2/5 in single precision (constants.f90:31,
    // smeshpack.f90:1261-1278).
    const double pi = 3.1415927410125732;
    const double dlon = (double)(2.0f / 5.0f) * pi;
    double lat = std::atan(0.5);
    x[0] = 0; y[0] = 0; z[0] = 1;
    double lon = -pi;
    for (int i = 1; i <= 5; ++i) {
      double cl = std::cos(lat);
      x[i] = cl * std::cos(lon);
      y[i] = cl * std::sin(lon);
      z[i] = std::sin(lat);
      lon = lon + dlon;
    }
    x[6] = 0; y[6] = 0; z[6] = -1;
    lat = -lat;
    lon = -4.0 * pi / 5.0;
    for (int i = 7; i <= 11; ++i) {
      double cl = std::cos(lat);
      x[i] = cl * std::cos(lon);
      y[i] = cl * std::sin(lon);
      z[i] = std::sin(lat);
      lon = lon + dlon;
    }
  }
}

static bool structure(HostMesh &m, int nv, const std::vector<double> &p, const std::vector<int32_t> &nnb,
                      const std::vector<int32_t> &nb0, std::string &err);

// meshbuild for icos/nopt grids (src/smeshpack.f90:77-368 with icosanodes
// :1297-1392): the 12 base nodes, then glevel refinements that each
// triangulate the current nodes (trmesh) and append the normalised midpoint
// of every arc i-j with j > i, node by node in ring order (meshrefin
// :1587-1675); finally the mesh is structured from the triangulation of all
// nodes (meshstruct). The hierarchical construction (hrchy > 0) repeats the
// same refinements level by level and yields the same mesh.
static bool build_icos(int glevel, int pos, HostMesh &m, std::string &err) {
  const int MX = IMX_MAXNB;
  const size_t nfin = (size_t)10 * ((size_t)1 << (2 * glevel)) + 2;
  std::vector<double> x(nfin), y(nfin), z(nfin);
  icos_base(pos, x.data(), y.data(), z.data());
  int nv = 12;
  std::vector<int32_t> nnb, nb;
  int64_t restarts = 0;
  for (int lev = 1; lev <= glevel; ++lev) {
    int64_t r = 0;
    if (!trmesh_rings(nv, x.data(), y.data(), z.data(), nnb, nb, MX, &r, err)) return false;
    restarts += r;
    int k = nv;
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nnb[i]; ++j) {
        const int o = nb[(size_t)i * MX + j];
        if (o <= i) continue;
        double qx = 0.5 * (x[i] + x[o]), qy = 0.5 * (y[i] + y[o]), qz = 0.5 * (z[i] + z[o]);
        double r2 = qx * qx;
        r2 = r2 + qy * qy;
        r2 = r2 + qz * qz;
        const double rr = std::sqrt(r2);
        x[k] = qx / rr;
        y[k] = qy / rr;
        z[k] = qz / rr;
        ++k;
      }
    if ((size_t)k != (size_t)10 * ((size_t)1 << (2 * lev)) + 2) {
      err = "refinement to level " + std::to_string(lev) + " gave " + std::to_string(k) + " nodes";
      return false;
    }
    nv = k;
  }
  int64_t r = 0;
  if (!trmesh_rings(nv, x.data(), y.data(), z.data(), nnb, nb, MX, &r, err)) return false;
  restarts += r;
  std::vector<double> p((size_t)3 * nv);
  for (int i = 0; i < nv; ++i) {
    p[3 * i] = x[i];
    p[3 * i + 1] = y[i];
    p[3 * i + 2] = z[i];
  }
  m.glevel = glevel;
  m.pos = (pos == IMX_POS_POL) ? "pol" : "eqs";
  m.trmesh_restarts = restarts;
  return structure(m, nv, p, nnb, nb, err);
}

// ---- meshstruct-equivalent structuring (src/smeshpack.f90:370-1195) ----
// From nodes and their CCW neighbour rings (0-based, IMX_MAXNB slots per
// node, in the order meshstruct takes them).
static bool structure(HostMesh &m, int nv, const std::vector<double> &p, const std::vector<int32_t> &nnb,
                      const std::vector<int32_t> &nb0, std::string &err) {
  const int MX = IMX_MAXNB;
  m.nv = nv;
  m.nt = 2 * nv - 4;
  m.ne = 3 * nv - 6;
  const int ne = m.ne, nt = m.nt;
  m.v_p = p;
  m.v_p.resize((size_t)3 * nv);
  m.v_lat.resize(nv);
  m.v_lon.resize(nv);
  for (int i = 0; i < nv; ++i)
    cart2sph(m.v_p[3 * i], m.v_p[3 * i + 1], m.v_p[3 * i + 2], &m.v_lon[i], &m.v_lat[i]);
  m.v_nnb = nnb;
  m.v_nb.assign((size_t)nv * MX, 0);
  for (size_t q = 0; q < (size_t)nv * MX; ++q) m.v_nb[q] = (q % MX) < (size_t)nnb[q / MX] ? nb0[q] + 1 : 0;

  // Edges: edge l runs v1 -> v2 with v1 < v2, numbered in (node, ring) order.
  m.v_ed.assign((size_t)nv * MX, 0);
  m.ed_v.assign((size_t)2 * ne, 0);
  m.ed_sh.assign((size_t)2 * ne, 0);
  m.ed_c.assign((size_t)3 * ne, 0);
  m.ed_nr.assign((size_t)3 * ne, 0);
  m.ed_tg.assign((size_t)3 * ne, 0);
  m.ed_leng.assign(ne, 0);
  m.ed_clat.assign(ne, 0);
  m.ed_clon.assign(ne, 0);
  int l = 0;
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < nnb[i]; ++j) {
      int k = nb0[(size_t)i * MX + j];
      if (k > i) {
        m.v_ed[(size_t)i * MX + j] = l + 1;
        m.ed_v[2 * l] = i + 1;
        m.ed_v[2 * l + 1] = k + 1;
        const double *pi = &m.v_p[3 * i], *pk = &m.v_p[3 * k];
        double c[3] = {(pi[0] + pk[0]) / 2.0, (pi[1] + pk[1]) / 2.0, (pi[2] + pk[2]) / 2.0};
        normalize3(c);
        std::memcpy(&m.ed_c[3 * l], c, sizeof(c));
        cart2sph(c[0], c[1], c[2], &m.ed_clon[l], &m.ed_clat[l]);
        double nr[3], tg[3];
        cross3(pi, pk, nr);
        normalize3(nr);
        cross3(nr, c, tg);
        normalize3(tg);
        std::memcpy(&m.ed_nr[3 * l], nr, sizeof(nr));
        std::memcpy(&m.ed_tg[3 * l], tg, sizeof(tg));
        m.ed_leng[l] = arclen(pi, pk);
        ++l;
      }
    }
  if (l != ne) {
    err = "edge count mismatch";
    return false;
  }
  // slot of node b in the ring of node a (-1 if absent)
  auto slot_of = [&](int a, int b) -> int {
    for (int j = 0; j < nnb[a]; ++j)
      if (nb0[(size_t)a * MX + j] == b) return j;
    return -1;
  };
  // edge id of (a, b), numbered at its lower node
  auto edge_of = [&](int a, int b) -> int {
    const int lo = std::min(a, b), j = slot_of(lo, std::max(a, b));
    return j < 0 ? -1 : m.v_ed[(size_t)lo * MX + j] - 1;
  };
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < nnb[i]; ++j) {
      int k = nb0[(size_t)i * MX + j];
      if (k < i) {
        const int e = edge_of(i, k);
        if (e < 0) {
          err = "edge lookup failed";
          return false;
        }
        m.v_ed[(size_t)i * MX + j] = e + 1;
      }
    }

  // Triangles: (i, nb(j), nb(j+1)) is created by its lowest node i.
  m.v_tr.assign((size_t)nv * MX, 0);
  m.tr_v.assign((size_t)3 * nt, 0);
  // A triangle met again at a higher node i is found in the ring of its
  // lowest node lo, where it was numbered: rotating the CCW triple (i, b, c)
  // to start at lo gives the two ring slots (x at j', y at j'+1).
  int kt = 0;
  for (int i = 0; i < nv; ++i) {
    int n = nnb[i];
    for (int j = 0; j < n; ++j) {
      int b = nb0[(size_t)i * MX + j], c = nb0[(size_t)i * MX + (j + 1) % n];
      if (std::min(b, c) > i) {
        m.v_tr[(size_t)i * MX + j] = kt + 1;
        m.tr_v[3 * kt] = i + 1;
        m.tr_v[3 * kt + 1] = b + 1;
        m.tr_v[3 * kt + 2] = c + 1;
        ++kt;
      } else {
        const int lo = std::min(b, c), x = (lo == b) ? c : i, y = (lo == b) ? i : b;
        const int nl = nnb[lo];
        int found = -1;
        for (int q = 0; q < nl && found < 0; ++q)
          if (nb0[(size_t)lo * MX + q] == x && nb0[(size_t)lo * MX + (q + 1) % nl] == y) found = q;
        if (found < 0 || m.v_tr[(size_t)lo * MX + found] == 0) {
          err = "triangle lookup failed";
          return false;
        }
        m.v_tr[(size_t)i * MX + j] = m.v_tr[(size_t)lo * MX + found];
      }
    }
  }
  if (kt != nt) {
    err = "triangle count mismatch";
    return false;
  }
  m.tr_ed.assign((size_t)3 * nt, 0);
  m.tr_tg.assign((size_t)3 * nt, 0);
  m.tr_nr.assign((size_t)3 * nt, 0);
  for (int k = 0; k < nt; ++k)
    for (int e = 0; e < 3; ++e) {
      int a = m.tr_v[3 * k + e] - 1, b = m.tr_v[3 * k + (e + 1) % 3] - 1;
      int ed = edge_of(a, b);
      if (ed < 0) {
        err = "triangle edge lookup failed";
        return false;
      }
      m.tr_ed[3 * k + e] = ed + 1;
      if (m.ed_sh[2 * ed] == 0) m.ed_sh[2 * ed] = k + 1;
      else m.ed_sh[2 * ed + 1] = k + 1;
      double nr[3];
      cross3(&m.v_p[3 * a], &m.v_p[3 * b], nr);
      nr[0] = -nr[0];
      nr[1] = -nr[1];
      nr[2] = -nr[2];
      normalize3(nr);
      double dp = dot3(nr, &m.ed_nr[3 * ed]);
      if (std::fabs(std::fabs(dp) - 1.0) > kEpsSingle * 100) {
        err = "triangle and edge normals not colinear";
        return false;
      }
      m.tr_nr[3 * k + e] = dp > 0 ? 1 : -1;
      m.tr_tg[3 * k + e] = dp > 0 ? -1 : 1;
    }

  // Circumcentres and areas (trcircumc :4659, sphtriarea :4710).
  m.tr_c.assign((size_t)3 * nt, 0);
  m.tr_clat.assign(nt, 0);
  m.tr_clon.assign(nt, 0);
  m.tr_areag.assign(nt, 0);
  parallel_ranges(nt, [&](int k0, int k1) {
  for (int k = k0; k < k1; ++k) {
    const double *p1 = &m.v_p[3 * (m.tr_v[3 * k] - 1)];
    const double *p2 = &m.v_p[3 * (m.tr_v[3 * k + 1] - 1)];
    const double *p3 = &m.v_p[3 * (m.tr_v[3 * k + 2] - 1)];
    double e1[3], e2[3], c[3];
    sub3(p2, p1, e1);
    sub3(p3, p1, e2);
    cross3(e1, e2, c);
    normalize3(c);
    std::memcpy(&m.tr_c[3 * k], c, sizeof(c));
    cart2sph(c[0], c[1], c[2], &m.tr_clon[k], &m.tr_clat[k]);
    m.tr_areag[k] = sphtriarea(p1, p2, p3);
  }
  });

  // Voronoi edges (src/smeshpack.f90:1016-1049).
  m.edhx_c.assign((size_t)3 * ne, 0);
  m.edhx_nr.assign((size_t)3 * ne, 0);
  m.edhx_tg.assign((size_t)3 * ne, 0);
  m.edhx_leng.assign(ne, 0);
  parallel_ranges(ne, [&](int e0, int e1) {
  for (int e = e0; e < e1; ++e) {
    const double *c1 = &m.tr_c[3 * (m.ed_sh[2 * e] - 1)];
    const double *c2 = &m.tr_c[3 * (m.ed_sh[2 * e + 1] - 1)];
    double c[3] = {(c1[0] + c2[0]) / 2.0, (c1[1] + c2[1]) / 2.0, (c1[2] + c2[2]) / 2.0};
    normalize3(c);
    std::memcpy(&m.edhx_c[3 * e], c, sizeof(c));
    double nr[3], tg[3];
    cross3(c1, c2, nr);
    normalize3(nr);
    cross3(nr, c, tg);
    normalize3(tg);
    std::memcpy(&m.edhx_nr[3 * e], nr, sizeof(nr));
    std::memcpy(&m.edhx_tg[3 * e], tg, sizeof(tg));
    m.edhx_leng[e] = arclen(c1, c2);
  }
  });

  // Voronoi cells: sign tables and areas (src/smeshpack.f90:1071-1151, sphhxarea :4765).
  m.hx_tg.assign((size_t)nv * MX, 0);
  m.hx_nr.assign((size_t)nv * MX, 0);
  m.hx_ttgout.assign((size_t)nv * MX, 0);
  m.hx_tnrccw.assign((size_t)nv * MX, 0);
  m.hx_areag.assign(nv, 0);
  std::atomic<int> bad{0};
  parallel_ranges(nv, [&](int i0, int i1) {
  for (int i = i0; i < i1; ++i) {
    int n = nnb[i];
    double area = 0.0;
    for (int j = 0; j < n; ++j) {
      int j1 = m.v_tr[(size_t)i * MX + (j + n - 1) % n] - 1;
      int j2 = m.v_tr[(size_t)i * MX + j] - 1;
      int ed = m.v_ed[(size_t)i * MX + j] - 1;
      double nr[3];
      cross3(&m.tr_c[3 * j1], &m.tr_c[3 * j2], nr);
      nr[0] = -nr[0];
      nr[1] = -nr[1];
      nr[2] = -nr[2];
      normalize3(nr);
      double dp = dot3(nr, &m.edhx_nr[3 * ed]);
      if (std::fabs(std::fabs(dp) - 1.0) > kEpsSingle * 100) bad = 1;
      int hnr = dp > 0 ? 1 : -1, htg = dp > 0 ? -1 : 1;
      m.hx_nr[(size_t)i * MX + j] = hnr;
      m.hx_tg[(size_t)i * MX + j] = htg;
      double w[3];
      for (int q = 0; q < 3; ++q) w[q] = m.edhx_nr[3 * ed + q] * (double)hnr;
      m.hx_ttgout[(size_t)i * MX + j] = (int)fsign1(dot3(&m.ed_tg[3 * ed], w));
      for (int q = 0; q < 3; ++q) w[q] = m.edhx_tg[3 * ed + q] * (double)htg;
      m.hx_tnrccw[(size_t)i * MX + j] = (int)fsign1(dot3(&m.ed_nr[3 * ed], w));
      area = area + sphtriarea(&m.v_p[3 * i], &m.tr_c[3 * j1], &m.tr_c[3 * j2]);
    }
    m.hx_areag[i] = area;
  }
  });
  if (bad) {
    err = "hexagon edge normal not colinear";
    return false;
  }
  return true;
}

bool build_from_rings(HostMesh &m, int nv, const std::vector<double> &p, const std::vector<int32_t> &nnb,
                      const std::vector<int32_t> &nb0, std::string &err) {
  return structure(m, nv, p, nnb, nb0, err);
}

}  // namespace imx

// ------------------------------------------------------------------ C ABI --
extern "C" {

const char *imx_last_error(void) { return imx::g_last_error.c_str(); }

int imx_mesh_build_icos(int32_t glevel, int32_t pos, imx_mesh **out) {
  if (!out) return imx::set_error(IMX_EINVAL, "imx_mesh_build_icos: out is NULL");
  *out = nullptr;
  if (glevel < 0 || glevel > 11)
    return imx::set_error(IMX_EINVAL, "imx_mesh_build_icos: glevel must be in 0..11");
  if (pos != IMX_POS_EQS && pos != IMX_POS_POL)
    return imx::set_error(IMX_EINVAL, "imx_mesh_build_icos: pos must be IMX_POS_EQS or IMX_POS_POL");
  imx_mesh *m = new (std::nothrow) imx_mesh();
  if (!m) return imx::set_error(IMX_ENOMEM, "imx_mesh_build_icos: out of memory");
  std::string err;
  try {
    if (!imx::build_icos(glevel, pos, *m, err)) {
      delete m;
      return imx::set_error(IMX_EINVAL, "imx_mesh_build_icos: " + err);
    }
  } catch (const std::bad_alloc &) {
    delete m;
    return imx::set_error(IMX_ENOMEM, "imx_mesh_build_icos: out of memory");
  }
  *out = m;
  return IMX_OK;
}

int imx_mesh_build_read(const char *path, imx_mesh **out) {
  if (!path || !out) return imx::set_error(IMX_EINVAL, "imx_mesh_build_read: NULL argument");
  *out = nullptr;
  const std::string fn(path);
  const std::string ext = fn.size() > 4 ? fn.substr(fn.size() - 3) : "";
  if (ext != "xyz" && ext != "gmt")
    return imx::set_error(IMX_EINVAL, "imx_mesh_build_read: unknown kind of file format. Extension: " + ext);
  FILE *f = std::fopen(path, "r");
  if (!f) return imx::set_error(IMX_EIO, "imx_mesh_build_read: " + fn + " not found");
  std::vector<double> p;
  int n = 0;
  bool ok = true;
  if (ext == "xyz") {
    // meshread (src/smeshpack.f90:7549-7700): the node count, then x y z
    ok = std::fscanf(f, "%d", &n) == 1 && n >= 4;
    p.resize((size_t)3 * std::max(n, 0));
    for (int i = 0; ok && i < n; ++i)
      ok = std::fscanf(f, "%lf%*[ ,\t]%lf%*[ ,\t]%lf", &p[3 * i], &p[3 * i + 1], &p[3 * i + 2]) == 3;
  } else {
    // lon lat in degrees per line, to Cartesian with sph2cart (:4251-4277)
    const double deg2rad = 3.1415927410125732 / 180.0;  // constants.f90:38 (single-rounded pi)
    double lon, lat;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
      if (std::sscanf(line, "%lf%*[ ,\t]%lf", &lon, &lat) != 2) continue;
      lon = lon * deg2rad;
      lat = lat * deg2rad;
      const double c = std::cos(lat);
      p.push_back(c * std::cos(lon));
      p.push_back(c * std::sin(lon));
      p.push_back(std::sin(lat));
    }
    n = (int)(p.size() / 3);
    ok = n >= 4;
  }
  std::fclose(f);
  if (!ok) return imx::set_error(IMX_EIO, "imx_mesh_build_read: cannot read the nodes of " + fn);
  // the Delaunay structure: one line of 6 neighbour indices per node, 0 for
  // none, in ring order; without the file meshstruct triangulates the nodes
  // (deltriread = 0, src/smeshpack.f90:497-520)
  const std::string ngb = fn.substr(0, fn.size() - 4) + ".ngb";
  const int MX = IMX_MAXNB;
  std::vector<int32_t> nnb(n, 0), nb((size_t)MX * n, 0);
  f = std::fopen(ngb.c_str(), "r");
  if (f) {
    for (int i = 0; ok && i < n; ++i) {
      int v[MX];
      for (int j = 0; ok && j < MX; ++j) ok = std::fscanf(f, "%d", &v[j]) == 1;
      for (int j = 0; ok && j < MX; ++j)
        if (v[j] > 0) {
          if (v[j] > n) ok = false;
          else nb[(size_t)MX * i + nnb[i]++] = v[j] - 1;
        }
      ok = ok && nnb[i] >= 3;
    }
    std::fclose(f);
    if (!ok) return imx::set_error(IMX_EIO, "imx_mesh_build_read: bad neighbour list in " + ngb);
  } else {
    std::vector<double> x(n), y(n), z(n);
    for (int i = 0; i < n; ++i) {
      x[i] = p[3 * i];
      y[i] = p[3 * i + 1];
      z[i] = p[3 * i + 2];
    }
    std::string terr;
    int64_t restarts = 0;
    try {
      if (!imx::trmesh_rings(n, x.data(), y.data(), z.data(), nnb, nb, MX, &restarts, terr))
        return imx::set_error(IMX_EINVAL, "imx_mesh_build_read: " + terr);
    } catch (const std::bad_alloc &) {
      return imx::set_error(IMX_ENOMEM, "imx_mesh_build_read: out of memory");
    }
  }
  auto m = std::make_unique<imx_mesh>();
  m->kind = "read";
  m->glevel = -1;
  std::string err;
  try {
    if (!imx::build_from_rings(*m, n, p, nnb, nb, err))
      return imx::set_error(IMX_EINVAL, "imx_mesh_build_read: " + err);
  } catch (const std::bad_alloc &) {
    return imx::set_error(IMX_ENOMEM, "imx_mesh_build_read: out of memory");
  }
  *out = m.release();
  return IMX_OK;
}

int imx_mesh_desc_of(const imx_mesh *m, imx_mesh_desc *out) {
  if (!m || !out) return imx::set_error(IMX_EINVAL, "imx_mesh_desc_of: NULL argument");
  imx::desc_of(*m, *out);
  return IMX_OK;
}

int imx_mesh_gridprop(imx_mesh *m) {
  if (!m) return imx::set_error(IMX_EINVAL, "imx_mesh_gridprop: NULL mesh");
  m->tr_trhx_area.assign((size_t)3 * m->nt, 0.0);
  m->hx_hxtr_area.assign((size_t)IMX_MAXNB * m->nv, 0.0);
  m->hx_trskw.assign((size_t)IMX_MAXNB * IMX_MAXNB * m->nv, 0.0);
  imx_mesh_desc d;
  imx::desc_of(*m, d);
  imx::gridprop(d, m->tr_trhx_area.data(), m->hx_hxtr_area.data(), m->hx_trskw.data());
  return IMX_OK;
}

int imx_gridprop(const imx_mesh_desc *d, double *trhx, double *hxtr, double *trskw) {
  if (!d || !trhx || !hxtr || !trskw) return imx::set_error(IMX_EINVAL, "imx_gridprop: NULL argument");
  imx::gridprop(*d, trhx, hxtr, trskw);
  return IMX_OK;
}

int imx_mesh_write(const imx_mesh *m, const char *path) {
  if (!m || !path) return imx::set_error(IMX_EINVAL, "imx_mesh_write: NULL argument");
  imx::BinWriter w(path);
  if (!w.ok()) return imx::set_error(IMX_EIO, std::string("imx_mesh_write: cannot open ") + path);
  std::vector<int32_t> dims = {m->nv, m->ne, m->nt, m->glevel};
  w.put("dims", dims);
  w.put("v_p", m->v_p);
  w.put("v_lat", m->v_lat);
  w.put("v_lon", m->v_lon);
  w.put("v_nnb", m->v_nnb);
  w.put("v_nb", m->v_nb);
  w.put("v_ed", m->v_ed);
  w.put("v_tr", m->v_tr);
  w.put("hx_ttgout", m->hx_ttgout);
  w.put("hx_tnrccw", m->hx_tnrccw);
  w.put("hx_tg", m->hx_tg);
  w.put("hx_nr", m->hx_nr);
  w.put("hx_areag", m->hx_areag);
  w.put("ed_v", m->ed_v);
  w.put("ed_sh", m->ed_sh);
  w.put("ed_leng", m->ed_leng);
  w.put("edhx_leng", m->edhx_leng);
  w.put("ed_c", m->ed_c);
  w.put("ed_clat", m->ed_clat);
  w.put("ed_clon", m->ed_clon);
  w.put("ed_tg", m->ed_tg);
  w.put("ed_nr", m->ed_nr);
  w.put("edhx_c", m->edhx_c);
  w.put("edhx_nr", m->edhx_nr);
  w.put("edhx_tg", m->edhx_tg);
  w.put("tr_v", m->tr_v);
  w.put("tr_ed", m->tr_ed);
  w.put("tr_tg", m->tr_tg);
  w.put("tr_nr", m->tr_nr);
  w.put("tr_areag", m->tr_areag);
  w.put("tr_c", m->tr_c);
  w.put("tr_clat", m->tr_clat);
  w.put("tr_clon", m->tr_clon);
  if (!m->tr_trhx_area.empty()) {
    w.put("tr_trhx_area", m->tr_trhx_area);
    w.put("hx_hxtr_area", m->hx_hxtr_area);
    w.put("hx_trskw", m->hx_trskw);
  }
  return IMX_OK;
}

int imx_mesh_read(const char *path, imx_mesh **out) {
  if (!path || !out) return imx::set_error(IMX_EINVAL, "imx_mesh_read: NULL argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) return imx::set_error(IMX_EIO, std::string("imx_mesh_read: cannot open ") + path);
  char magic[8];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "IMXBIN01", 8) != 0) {
    std::fclose(f);
    return imx::set_error(IMX_EIO, std::string("imx_mesh_read: not an IMXBIN file: ") + path);
  }
  auto m = std::make_unique<imx_mesh>();
  std::vector<int32_t> dims;
  std::map<std::string, std::vector<int32_t> *> iv = {
      {"v_nnb", &m->v_nnb}, {"v_nb", &m->v_nb}, {"v_ed", &m->v_ed}, {"v_tr", &m->v_tr},
      {"hx_ttgout", &m->hx_ttgout}, {"hx_tnrccw", &m->hx_tnrccw}, {"hx_tg", &m->hx_tg},
      {"hx_nr", &m->hx_nr}, {"ed_v", &m->ed_v}, {"ed_sh", &m->ed_sh}, {"tr_v", &m->tr_v},
      {"tr_ed", &m->tr_ed}, {"tr_tg", &m->tr_tg}, {"tr_nr", &m->tr_nr}, {"dims", &dims}};
  std::map<std::string, std::vector<double> *> rv = {
      {"v_p", &m->v_p}, {"v_lat", &m->v_lat}, {"v_lon", &m->v_lon}, {"hx_areag", &m->hx_areag},
      {"ed_leng", &m->ed_leng}, {"edhx_leng", &m->edhx_leng}, {"ed_c", &m->ed_c},
      {"ed_clat", &m->ed_clat}, {"ed_clon", &m->ed_clon}, {"ed_tg", &m->ed_tg}, {"ed_nr", &m->ed_nr},
      {"edhx_c", &m->edhx_c}, {"edhx_nr", &m->edhx_nr}, {"edhx_tg", &m->edhx_tg},
      {"tr_areag", &m->tr_areag}, {"tr_c", &m->tr_c}, {"tr_clat", &m->tr_clat},
      {"tr_clon", &m->tr_clon}, {"tr_trhx_area", &m->tr_trhx_area},
      {"hx_hxtr_area", &m->hx_hxtr_area}, {"hx_trskw", &m->hx_trskw}};
  std::string bad;
  for (;;) {
    char name[25] = {0}, dt[4];
    int64_t n = 0;
    if (std::fread(name, 1, 24, f) != 24) break;
    if (std::fread(dt, 1, 4, f) != 4 || std::fread(&n, 8, 1, f) != 1 || n < 0 || n > (int64_t)1 << 31) {
      bad = "truncated record header";
      break;
    }
    std::string nm(name);
    nm.erase(nm.find_last_not_of(' ') + 1);
    const bool isf = std::memcmp(dt, "f8  ", 4) == 0, isi = std::memcmp(dt, "i4  ", 4) == 0;
    if (!isf && !isi) {
      bad = "unknown record type of " + nm;
      break;
    }
    const size_t bytes = (size_t)n * (isf ? 8 : 4);
    std::vector<char> buf(bytes);
    if (bytes && std::fread(buf.data(), 1, bytes, f) != bytes) {
      bad = "truncated record " + nm;
      break;
    }
    if (isi && iv.count(nm)) {
      iv[nm]->resize(n);
      std::memcpy(iv[nm]->data(), buf.data(), bytes);
    } else if (isf && rv.count(nm)) {
      rv[nm]->resize(n);
      std::memcpy(rv[nm]->data(), buf.data(), bytes);
    }
  }
  std::fclose(f);
  if (!bad.empty()) return imx::set_error(IMX_EIO, "imx_mesh_read: " + bad);
  if (dims.size() < 3) return imx::set_error(IMX_EINVAL, "imx_mesh_read: no dims record");
  m->nv = dims[0];
  m->ne = dims[1];
  m->nt = dims[2];
  m->glevel = dims.size() > 3 ? dims[3] : -1;
  const size_t nv = m->nv, ne = m->ne, nt = m->nt, X = IMX_MAXNB;
  const bool ok = m->v_p.size() == 3 * nv && m->v_lat.size() == nv && m->hx_areag.size() == nv &&
                  m->v_nnb.size() == nv && m->v_nb.size() == X * nv && m->v_ed.size() == X * nv &&
                  m->v_tr.size() == X * nv && m->hx_ttgout.size() == X * nv &&
                  m->hx_tnrccw.size() == X * nv && m->ed_v.size() == 2 * ne && m->ed_sh.size() == 2 * ne &&
                  m->ed_leng.size() == ne && m->edhx_leng.size() == ne && m->ed_c.size() == 3 * ne &&
                  m->edhx_c.size() == 3 * ne && m->tr_v.size() == 3 * nt && m->tr_ed.size() == 3 * nt &&
                  m->tr_tg.size() == 3 * nt && m->tr_areag.size() == nt && m->tr_c.size() == 3 * nt &&
                  m->tr_clat.size() == nt;
  if (!ok) return imx::set_error(IMX_EINVAL, "imx_mesh_read: a required record is missing or has the wrong size");
  imx_mesh_desc d;
  imx::desc_of(*m, d);
  if (!m->tr_trhx_area.empty() && (m->tr_trhx_area.size() != 3 * nt || m->hx_hxtr_area.size() != X * nv ||
                                   m->hx_trskw.size() != X * X * nv))
    return imx::set_error(IMX_EINVAL, "imx_mesh_read: gridprop records of the wrong size");
  *out = m.release();
  return IMX_OK;
}

void imx_mesh_destroy(imx_mesh *m) { delete m; }

}  // extern "C"

namespace imx {

// alignind (src/smeshpack.f90:6398-6490): regularity of a hexagonal (or
// square) cell from its dual edges; 1 for other cells. As in the reference,
// the mean length keeps only the last pair's sum.
static double alignind(const imx_mesh_desc &d, int i) {
  const int n = d.v_nnb[i];
  auto tc = [&](int j) { return d.tr_c + 3 * (d.v_tr[IMX_MAXNB * i + j - 1] - 1); };  // j 1-based
  if (n != 6 && n != 4) return 1.0;
  double tmp = 0.0, meanl = 0.0;
  const int pairs = n == 6 ? 3 : 2;
  for (int j = 1; j <= pairs; ++j) {
    const int i1 = j, i2 = j + 1, i3 = n == 6 ? j + 3 : j + 2;
    int i0 = n == 6 ? j + 4 : j + 3;
    if (i0 > n) i0 -= n;  // modint
    double t1 = arclen(tc(i1), tc(i2));
    double t2 = arclen(tc(i3), tc(i0));
    meanl = t1 + t2;
    tmp = tmp + std::fabs(t1 - t2);
    t1 = arclen(tc(i1), tc(i0));
    t2 = arclen(tc(i2), tc(i3));
    tmp = tmp + std::fabs(t1 - t2);
  }
  tmp = tmp / 6;
  meanl = meanl / 6;
  return tmp / meanl;
}

// alignment_coef / diameter_coef (src/swm_operators.f90:977-1072), or const
static void coef_at_cells(const imx_mesh_desc &d, double kmax, int kind, std::vector<double> &c) {
  const int nv = d.nv;
  c.assign(nv, kmax);
  if (kind == IMX_COEF_CONST) return;
  if (kind == IMX_COEF_ALIGN) {
    std::vector<double> a(nv);
    for (int i = 0; i < nv; ++i) a[i] = alignind(d, i);
    for (int i = 0; i < nv; ++i) {  // smoothing over the cell and its neighbours
      double sum = a[i];
      for (int j = 0; j < d.v_nnb[i]; ++j) sum = sum + a[d.v_nb[IMX_MAXNB * i + j] - 1];
      c[i] = sum / (d.v_nnb[i] + 1.0);
    }
  } else {  // the largest circumcentre distance within the cell
    for (int i = 0; i < nv; ++i) {
      double dm = 0.0;
      for (int j = 0; j < d.v_nnb[i]; ++j)
        for (int k = 0; k < d.v_nnb[i]; ++k)
          dm = std::max(dm, arclen(d.tr_c + 3 * (d.v_tr[IMX_MAXNB * i + j] - 1),
                                   d.tr_c + 3 * (d.v_tr[IMX_MAXNB * i + k] - 1)));
      c[i] = dm;
    }
  }
  double mx = c[0];
  for (int i = 1; i < nv; ++i) mx = std::max(mx, c[i]);
  for (int i = 0; i < nv; ++i) c[i] = c[i] / mx;
  if (kind == IMX_COEF_ALIGN) {
    for (int i = 0; i < nv; ++i) c[i] = kmax * c[i];
  } else {
    const double p = std::log(10.0) / std::log(2.0);
    for (int i = 0; i < nv; ++i) c[i] = kmax * std::pow(c[i], p);
  }
}

void bary_circumcentre(const imx_mesh_desc &d, int k, double b[3]) { bary_at(d, d.tr_c + 3 * k, k, b); }

void bary_at(const imx_mesh_desc &d, const double *p, int k, double b[3]) {
  const double *p1 = d.v_p + 3 * (d.tr_v[3 * k] - 1), *p2 = d.v_p + 3 * (d.tr_v[3 * k + 1] - 1),
               *p3 = d.v_p + 3 * (d.tr_v[3 * k + 2] - 1);
  b[0] = std::fabs(det3(p, p2, p3));
  b[1] = std::fabs(det3(p, p3, p1));
  b[2] = std::fabs(det3(p, p1, p2));
  const double area = b[0] + b[1] + b[2];
  for (int i = 0; i < 3; ++i) b[i] = b[i] / area;
}

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

bool insidetrmesh(const imx_mesh_desc &d, const double *p, int tr) {
  const double tol = kEpsSingle / 10000.0;
  const double tol2 = 2.220446049250313e-16 * 100.0;  // epsilon(pi) of the FP64 pi
  for (int l = 0; l < 3; ++l) {
    const double *p1 = d.v_p + 3 * (d.tr_v[3 * tr + l] - 1);
    const double *p2 = d.v_p + 3 * (d.tr_v[3 * tr + (l + 1) % 3] - 1);
    double t[3];
    sub3(p, p1, t);
    if (norm3(t) < tol) return true;
    double vn[3];
    cross3(p1, p2, vn);
    const double vnn = norm3(vn);
    for (int q = 0; q < 3; ++q) vn[q] = vn[q] / vnn;
    const double determ = dot3(vn, p);
    if (determ < -tol) return false;
    if (determ > tol) continue;
    const double determ2 = robdet(p, p1, p2);
    if (std::fabs(determ2) < tol2 || determ * determ2 < 0) continue;
    if (determ2 < -tol2) return false;
  }
  return true;
}

int hxedgeconn(const imx_mesh_desc &d, int ed1, int ed2) {
  if (ed1 == ed2) return -2;
  // getedgeintriangle: the third edge of a triangle holding both
  const int t1 = d.ed_sh[2 * ed1] - 1, t2 = d.ed_sh[2 * ed1 + 1] - 1;
  int tr = -1;
  for (int i = 0; i < 3 && tr < 0; ++i) {
    if (ed2 == d.tr_ed[3 * t1 + i] - 1) tr = t1;
    else if (ed2 == d.tr_ed[3 * t2 + i] - 1) tr = t2;
  }
  if (tr >= 0) {
    int e3 = -1;
    for (int i = 0; i < 3; ++i) {
      const int e = d.tr_ed[3 * tr + i] - 1;
      if (e != ed1 && e != ed2) e3 = e;
    }
    if (e3 >= 0) return e3;
  }
  // an edge of ed2's triangles whose other triangle holds ed1
  for (int j = 0; j < 2; ++j) {
    const int trj = d.ed_sh[2 * ed2 + j] - 1;
    for (int i = 0; i < 3; ++i) {
      const int et = d.tr_ed[3 * trj + i] - 1;
      if (et == ed2) continue;
      const int tr2 = (d.ed_sh[2 * et] - 1 == trj) ? d.ed_sh[2 * et + 1] - 1 : d.ed_sh[2 * et] - 1;
      for (int l = 0; l < 3; ++l)
        if (ed1 == d.tr_ed[3 * tr2 + l] - 1) return et;
    }
  }
  return -1;
}

// scalar_hx2tr (src/swm_operators.f90:258-269): trsk area weights, or the
// barycentric coordinates of the circumcentre (bary interpolation)
static void hx2tr(const imx_mesh_desc &d, const double *trhx, bool bary, const std::vector<double> &f,
                  std::vector<double> &out) {
  out.assign(d.nt, 0.0);
  for (int k = 0; k < d.nt; ++k) {
    double s = 0.0;
    if (bary) {
      double b[3];
      bary_circumcentre(d, k, b);
      for (int i = 0; i < 3; ++i) s = s + b[i] * f[d.tr_v[3 * k + i] - 1];
      out[k] = s;
    } else {
      for (int i = 0; i < 3; ++i) s = s + trhx[3 * k + i] * f[d.tr_v[3 * k + i] - 1];
      out[k] = s / d.tr_areag[k];
    }
  }
}

void diffusion_coef(const imx_mesh_desc &d, const double *trhx, bool bary, double k2, int k2kind,
                    double k4, int k4kind, std::vector<double> &hx, std::vector<double> &tr) {
  hx.assign(d.nv, 0.0);
  tr.assign(d.nt, 0.0);
  if (k2 > 0.0) {
    coef_at_cells(d, k2, k2kind, hx);
    hx2tr(d, trhx, bary, hx, tr);
  }
  if (k4 > 0.0) {
    coef_at_cells(d, k4, k4kind, hx);
    hx2tr(d, trhx, bary, hx, tr);
    for (double &x : hx) x = std::sqrt(x);
    for (double &x : tr) x = std::sqrt(x);
  }
}

}  // namespace imx
