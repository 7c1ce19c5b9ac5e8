// Grid files in the reference's binary layout: the 11 Fortran unformatted
// sequential files written by meshwrite and read by meshload
// (src/smeshpack.f90:7702-8221), <dir>/<name>_header.dat, _vert_coord.dat,
// _vert_nb.dat, _ed_cc.dat, _ed_nb.dat, _edhx_cc.dat, _edhx_nb.dat,
// _tr_cc.dat, _tr_nb.dat, _hx.dat and _quad.dat.
//
// Record framing: each record is a 4-byte little-endian length, the data,
// and the length again. A record longer than 2^31-9 bytes is split into
// subrecords whose leading marker is negative while more follow (the
// framing gfortran and flang write). The record contents follow the
// reference's write lists field by field (i4 = int32, r8 = float64,
// character(len=16|128) blank-padded), so the layout has to match theirs.
//
// Loading fills every array the shallow-water path uses. Writing also emits
// the fields the path does not use, computed from their definitions in
// meshstruct (src/smeshpack.f90:645-1151): nbd = -p_i.p_j, nbdg = arclen,
// chord lengths, circumradii, planar triangle areas, triangle angles, and the
// Voronoi cell centroid and planar area from its (node, circumcentre,
// circumcentre) sub-triangles. The alignment index is written as 0 and the
// regular lat-lon search grid as empty (nlat = 0); the shallow-water path
// uses neither.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "imx_internal.h"

namespace imx {

namespace {

const int MX = IMX_MAXNB;
const int64_t kMaxSub = 2147483639;  // longest subrecord payload

struct RecReader {
  FILE *f = nullptr;
  std::string path;
  explicit RecReader(const std::string &p) : path(p) { f = std::fopen(p.c_str(), "rb"); }
  ~RecReader() {
    if (f) std::fclose(f);
  }
  bool ok() const { return f != nullptr; }
  // next logical record into buf; false at end of file or on a framing error
  bool next(std::vector<char> &buf, std::string &err) {
    buf.clear();
    for (;;) {
      int32_t lead = 0, trail = 0;
      if (std::fread(&lead, 4, 1, f) != 1) {
        if (buf.empty()) err = path + ": unexpected end of file";
        else err = path + ": truncated record";
        return false;
      }
      const int64_t n = lead < 0 ? -(int64_t)lead : lead;
      const size_t o = buf.size();
      buf.resize(o + (size_t)n);
      if (n && std::fread(buf.data() + o, 1, (size_t)n, f) != (size_t)n) {
        err = path + ": truncated record";
        return false;
      }
      if (std::fread(&trail, 4, 1, f) != 1 || (trail < 0 ? -(int64_t)trail : trail) != n) {
        err = path + ": bad record marker";
        return false;
      }
      if (lead >= 0) return true;
    }
  }
};

struct Cursor {
  const std::vector<char> &b;
  size_t o = 0;
  bool bad = false;
  explicit Cursor(const std::vector<char> &buf) : b(buf) {}
  int32_t i4() {
    int32_t v = 0;
    if (o + 4 > b.size()) {
      bad = true;
      return 0;
    }
    std::memcpy(&v, b.data() + o, 4);
    o += 4;
    return v;
  }
  double r8() {
    double v = 0;
    if (o + 8 > b.size()) {
      bad = true;
      return 0;
    }
    std::memcpy(&v, b.data() + o, 8);
    o += 8;
    return v;
  }
  std::string str(size_t n) {
    if (o + n > b.size()) {
      bad = true;
      return "";
    }
    std::string s(b.data() + o, n);
    o += n;
    const size_t e = s.find_last_not_of(' ');
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
  }
  bool done() const { return !bad && o == b.size(); }
};

struct RecWriter {
  FILE *f = nullptr;
  std::vector<char> buf;
  explicit RecWriter(const std::string &p) { f = std::fopen(p.c_str(), "wb"); }
  ~RecWriter() {
    if (f) std::fclose(f);
  }
  bool ok() const { return f != nullptr; }
  void i4(int32_t v) { put(&v, 4); }
  void r8(double v) { put(&v, 8); }
  void str(const std::string &s, size_t n) {
    std::string t = s.substr(0, n);
    t.resize(n, ' ');
    put(t.data(), n);
  }
  void put(const void *p, size_t n) {
    const size_t o = buf.size();
    buf.resize(o + n);
    std::memcpy(buf.data() + o, p, n);
  }
  bool end() {  // flush buf as one record (subrecords when long)
    int64_t left = (int64_t)buf.size(), o = 0;
    bool good = true;
    do {
      const int64_t n = left > kMaxSub ? kMaxSub : left;
      const bool more = left > n;
      const int32_t lead = more ? -(int32_t)n : (int32_t)n;
      const int32_t trail = (o > 0) ? -(int32_t)n : (int32_t)n;
      good = good && std::fwrite(&lead, 4, 1, f) == 1;
      good = good && (n == 0 || std::fwrite(buf.data() + o, 1, (size_t)n, f) == (size_t)n);
      good = good && std::fwrite(&trail, 4, 1, f) == 1;
      o += n;
      left -= n;
    } while (left > 0);
    buf.clear();
    return good;
  }
};

std::string file_of(const std::string &dir, const std::string &name, const char *suffix) {
  std::string d = dir;
  if (!d.empty() && d.back() != '/') d += '/';
  return d + name + suffix;
}

// spherical angles of triangle (a, b, c) at its three vertices
void tri_angles(const double *a, const double *b, const double *c, double *ang) {
  const double *P[3] = {a, b, c};
  for (int k = 0; k < 3; ++k) {
    const double *p = P[k], *q = P[(k + 1) % 3], *r = P[(k + 2) % 3];
    double tq[3], tr[3], x[3];
    const double dq = dot3(p, q), dr = dot3(p, r);
    for (int i = 0; i < 3; ++i) {
      tq[i] = q[i] - dq * p[i];
      tr[i] = r[i] - dr * p[i];
    }
    cross3(tq, tr, x);
    ang[k] = std::atan2(norm3(x), dot3(tq, tr));
  }
}

}  // namespace

bool load_ref(const std::string &dir, const std::string &name, HostMesh &m, std::string &err) {
  std::vector<char> b;
  // header
  {
    RecReader r(file_of(dir, name, "_header.dat"));
    if (!r.ok()) {
      err = "cannot open " + r.path;
      return false;
    }
    std::vector<std::string> s(3);
    for (int k = 0; k < 3; ++k) {
      if (!r.next(b, err)) return false;
      Cursor c(b);
      s[k] = c.str(16);
    }
    m.kind = s[0];
    m.pos = s[1];
    m.optm = s[2];
    if (!r.next(b, err)) return false;
    {
      Cursor c(b);
      m.glevel = c.i4();
    }
    if (!r.next(b, err)) return false;
    Cursor c(b);
    m.nv = c.i4();
    m.nt = c.i4();
    m.ne = c.i4();
    if (c.bad || m.nv < 4 || m.nt != 2 * m.nv - 4 || m.ne != 3 * m.nv - 6) {
      err = r.path + ": counts are not those of a closed spherical triangulation";
      return false;
    }
  }
  const int nv = m.nv, ne = m.ne, nt = m.nt;
  // vertices
  {
    RecReader r(file_of(dir, name, "_vert_coord.dat"));
    if (!r.ok()) {
      err = "cannot open " + r.path;
      return false;
    }
    if (!r.next(b, err)) return false;
    if (Cursor(b).i4() != nv) {
      err = r.path + ": vertex count differs from the header";
      return false;
    }
    if (!r.next(b, err)) return false;
    Cursor c(b);
    m.v_p.assign((size_t)3 * nv, 0);
    m.v_lon.assign(nv, 0);
    m.v_lat.assign(nv, 0);
    m.v_nnb.assign(nv, 0);
    for (int i = 0; i < nv; ++i) {
      c.i4();
      for (int q = 0; q < 3; ++q) m.v_p[3 * i + q] = c.r8();
      m.v_lon[i] = c.r8();
      m.v_lat[i] = c.r8();
      m.v_nnb[i] = c.i4();
      if (m.v_nnb[i] < 3 || m.v_nnb[i] > MX) {
        err = r.path + ": a node has " + std::to_string(m.v_nnb[i]) + " neighbours (3..6 supported)";
        return false;
      }
    }
    if (!c.done()) {
      err = r.path + ": record length does not match the vertex list";
      return false;
    }
  }
  {
    RecReader r(file_of(dir, name, "_vert_nb.dat"));
    if (!r.ok()) {
      err = "cannot open " + r.path;
      return false;
    }
    m.v_nb.assign((size_t)MX * nv, 0);
    m.v_ed.assign((size_t)MX * nv, 0);
    m.v_tr.assign((size_t)MX * nv, 0);
    for (int i = 0; i < nv; ++i) {
      if (!r.next(b, err)) return false;
      Cursor c(b);
      for (int j = 0; j < m.v_nnb[i]; ++j) {
        c.i4();
        c.i4();
        const int n = c.i4();
        m.v_nb[(size_t)MX * i + j] = c.i4();
        c.r8();
        c.r8();
        m.v_ed[(size_t)MX * i + j] = c.i4();
        m.v_tr[(size_t)MX * i + j] = c.i4();
        if (n != m.v_nnb[i]) c.bad = true;
      }
      if (!c.done()) {
        err = r.path + ": neighbour record of node " + std::to_string(i + 1) + " is malformed";
        return false;
      }
    }
  }
  // primal and dual edges
  auto load_edges = [&](const char *cc, const char *nb, std::vector<double> &ctr, std::vector<double> *clat,
                        std::vector<double> *clon, std::vector<int32_t> *ev, std::vector<int32_t> *esh,
                        std::vector<double> &tg, std::vector<double> &nr, std::vector<double> &leng) -> bool {
    RecReader r1(file_of(dir, name, cc));
    if (!r1.ok()) {
      err = "cannot open " + r1.path;
      return false;
    }
    if (!r1.next(b, err)) return false;
    if (Cursor(b).i4() != ne) {
      err = r1.path + ": edge count differs from the header";
      return false;
    }
    if (!r1.next(b, err)) return false;
    Cursor c(b);
    ctr.assign((size_t)3 * ne, 0);
    if (clat) clat->assign(ne, 0);
    if (clon) clon->assign(ne, 0);
    for (int l = 0; l < ne; ++l) {
      c.i4();
      for (int q = 0; q < 3; ++q) ctr[3 * l + q] = c.r8();
      const double lon = c.r8(), lat = c.r8();
      if (clon) (*clon)[l] = lon;
      if (clat) (*clat)[l] = lat;
    }
    if (!c.done()) {
      err = r1.path + ": record length does not match the edge list";
      return false;
    }
    RecReader r2(file_of(dir, name, nb));
    if (!r2.ok()) {
      err = "cannot open " + r2.path;
      return false;
    }
    if (!r2.next(b, err)) return false;
    Cursor d(b);
    if (ev) ev->assign((size_t)2 * ne, 0);
    if (esh) esh->assign((size_t)2 * ne, 0);
    tg.assign((size_t)3 * ne, 0);
    nr.assign((size_t)3 * ne, 0);
    leng.assign(ne, 0);
    for (int l = 0; l < ne; ++l) {
      d.i4();
      const int v1 = d.i4(), v2 = d.i4(), s1 = d.i4(), s2 = d.i4();
      if (ev) {
        (*ev)[2 * l] = v1;
        (*ev)[2 * l + 1] = v2;
      }
      if (esh) {
        (*esh)[2 * l] = s1;
        (*esh)[2 * l + 1] = s2;
      }
      for (int q = 0; q < 3; ++q) tg[3 * l + q] = d.r8();
      for (int q = 0; q < 3; ++q) nr[3 * l + q] = d.r8();
      d.r8();  // planar length
      leng[l] = d.r8();
    }
    if (!d.done()) {
      err = r2.path + ": record length does not match the edge list";
      return false;
    }
    return true;
  };
  if (!load_edges("_ed_cc.dat", "_ed_nb.dat", m.ed_c, &m.ed_clat, &m.ed_clon, &m.ed_v, &m.ed_sh, m.ed_tg,
                  m.ed_nr, m.ed_leng))
    return false;
  if (!load_edges("_edhx_cc.dat", "_edhx_nb.dat", m.edhx_c, nullptr, nullptr, nullptr, nullptr, m.edhx_tg,
                  m.edhx_nr, m.edhx_leng))
    return false;
  // triangles
  {
    RecReader r(file_of(dir, name, "_tr_cc.dat"));
    if (!r.ok()) {
      err = "cannot open " + r.path;
      return false;
    }
    if (!r.next(b, err)) return false;
    if (Cursor(b).i4() != nt) {
      err = r.path + ": triangle count differs from the header";
      return false;
    }
    if (!r.next(b, err)) return false;
    Cursor c(b);
    m.tr_c.assign((size_t)3 * nt, 0);
    m.tr_clon.assign(nt, 0);
    m.tr_clat.assign(nt, 0);
    for (int k = 0; k < nt; ++k) {
      c.i4();
      for (int q = 0; q < 3; ++q) m.tr_c[3 * k + q] = c.r8();
      m.tr_clon[k] = c.r8();
      m.tr_clat[k] = c.r8();
    }
    if (!c.done()) {
      err = r.path + ": record length does not match the triangle list";
      return false;
    }
  }
  {
    RecReader r(file_of(dir, name, "_tr_nb.dat"));
    if (!r.ok()) {
      err = "cannot open " + r.path;
      return false;
    }
    if (!r.next(b, err)) return false;
    Cursor c(b);
    m.tr_v.assign((size_t)3 * nt, 0);
    m.tr_ed.assign((size_t)3 * nt, 0);
    m.tr_tg.assign((size_t)3 * nt, 0);
    m.tr_nr.assign((size_t)3 * nt, 0);
    m.tr_areag.assign(nt, 0);
    for (int k = 0; k < nt; ++k) {
      c.i4();
      for (int q = 0; q < 3; ++q) m.tr_v[3 * k + q] = c.i4();
      for (int q = 0; q < 3; ++q) m.tr_ed[3 * k + q] = c.i4();
      for (int q = 0; q < 3; ++q) c.i4();  // neighbour triangles
      for (int q = 0; q < 3; ++q) m.tr_tg[3 * k + q] = c.i4();
      for (int q = 0; q < 3; ++q) m.tr_nr[3 * k + q] = c.i4();
      c.r8();  // radp
      c.r8();  // radg
      c.r8();  // areap
      m.tr_areag[k] = c.r8();
      for (int q = 0; q < 3; ++q) c.r8();  // angles
    }
    if (!c.done()) {
      err = r.path + ": record length does not match the triangle list";
      return false;
    }
  }
  // Voronoi cells
  {
    RecReader r(file_of(dir, name, "_hx.dat"));
    if (!r.ok()) {
      err = "cannot open " + r.path;
      return false;
    }
    if (!r.next(b, err)) return false;
    if (Cursor(b).i4() != nv) {
      err = r.path + ": cell count differs from the header";
      return false;
    }
    m.hx_tg.assign((size_t)MX * nv, 0);
    m.hx_nr.assign((size_t)MX * nv, 0);
    m.hx_ttgout.assign((size_t)MX * nv, 0);
    m.hx_tnrccw.assign((size_t)MX * nv, 0);
    m.hx_areag.assign(nv, 0);
    for (int i = 0; i < nv; ++i) {
      if (!r.next(b, err)) return false;  // barycentre
      if (!r.next(b, err)) return false;
      Cursor c(b);
      for (int j = 0; j < m.v_nnb[i]; ++j) {
        m.hx_tg[(size_t)MX * i + j] = c.i4();
        m.hx_nr[(size_t)MX * i + j] = c.i4();
        m.hx_ttgout[(size_t)MX * i + j] = c.i4();
        m.hx_tnrccw[(size_t)MX * i + j] = c.i4();
      }
      if (!c.done()) {
        err = r.path + ": sign record of cell " + std::to_string(i + 1) + " is malformed";
        return false;
      }
      if (!r.next(b, err)) return false;
      Cursor a(b);
      a.r8();
      m.hx_areag[i] = a.r8();
      a.r8();
      if (!a.done()) {
        err = r.path + ": area record of cell " + std::to_string(i + 1) + " is malformed";
        return false;
      }
    }
  }
  // index sanity: everything the kernels gather through must be in range
  auto in = [](int32_t v, int n) { return v >= 1 && v <= n; };
  for (int i = 0; i < nv; ++i)
    for (int j = 0; j < m.v_nnb[i]; ++j)
      if (!in(m.v_nb[(size_t)MX * i + j], nv) || !in(m.v_ed[(size_t)MX * i + j], ne) ||
          !in(m.v_tr[(size_t)MX * i + j], nt)) {
        err = "node " + std::to_string(i + 1) + " has an out-of-range neighbour, edge or triangle";
        return false;
      }
  for (int l = 0; l < ne; ++l)
    if (!in(m.ed_v[2 * l], nv) || !in(m.ed_v[2 * l + 1], nv) || !in(m.ed_sh[2 * l], nt) ||
        !in(m.ed_sh[2 * l + 1], nt)) {
      err = "edge " + std::to_string(l + 1) + " has an out-of-range node or triangle";
      return false;
    }
  for (int k = 0; k < nt; ++k)
    for (int q = 0; q < 3; ++q)
      if (!in(m.tr_v[3 * k + q], nv) || !in(m.tr_ed[3 * k + q], ne)) {
        err = "triangle " + std::to_string(k + 1) + " has an out-of-range node or edge";
        return false;
      }
  return true;
}

bool write_ref(const HostMesh &m, const std::string &dir, const std::string &name, std::string &err) {
  const int nv = m.nv, ne = m.ne, nt = m.nt;
  // derived fields not kept in HostMesh
  double vmin = 1e300, vmax = 0, vsum = 0, cmin = 1e300, cmax = 0, csum = 0;
  double tamin = 1e300, tamax = 0, tasum = 0, hamin = 1e300, hamax = 0, hasum = 0;
  double angmin = 1e300, angmax = 0, angsum = 0;
  int maxvnb = 0;
  for (int i = 0; i < nv; ++i) {
    maxvnb = std::max(maxvnb, m.v_nnb[i]);
    for (int j = 0; j < m.v_nnb[i]; ++j) {
      const double d = arclen(&m.v_p[3 * i], &m.v_p[3 * (m.v_nb[(size_t)MX * i + j] - 1)]);
      vmin = std::min(vmin, d);
      vmax = std::max(vmax, d);
      vsum += d;
    }
    hamin = std::min(hamin, m.hx_areag[i]);
    hamax = std::max(hamax, m.hx_areag[i]);
    hasum += m.hx_areag[i];
  }
  for (int l = 0; l < ne; ++l) {
    cmin = std::min(cmin, m.edhx_leng[l]);
    cmax = std::max(cmax, m.edhx_leng[l]);
    csum += m.edhx_leng[l];
  }
  std::vector<double> ang((size_t)3 * nt);
  for (int k = 0; k < nt; ++k) {
    tri_angles(&m.v_p[3 * (m.tr_v[3 * k] - 1)], &m.v_p[3 * (m.tr_v[3 * k + 1] - 1)],
               &m.v_p[3 * (m.tr_v[3 * k + 2] - 1)], &ang[3 * k]);
    for (int q = 0; q < 3; ++q) {
      angmin = std::min(angmin, ang[3 * k + q]);
      angmax = std::max(angmax, ang[3 * k + q]);
      angsum += ang[3 * k + q];
    }
    tamin = std::min(tamin, m.tr_areag[k]);
    tamax = std::max(tamax, m.tr_areag[k]);
    tasum += m.tr_areag[k];
  }
  auto open = [&](const char *suffix, RecWriter *&w) -> bool {
    w = new RecWriter(file_of(dir, name, suffix));
    if (!w->ok()) {
      err = "cannot create " + file_of(dir, name, suffix);
      delete w;
      w = nullptr;
      return false;
    }
    return true;
  };
  RecWriter *w = nullptr;
  bool good = true;
  auto close = [&]() {
    good = good && std::ferror(w->f) == 0;
    delete w;
    w = nullptr;
  };
  // header
  if (!open("_header.dat", w)) return false;
  w->str(m.kind.empty() ? "icos" : m.kind, 16);
  good = good && w->end();
  w->str(m.pos.empty() ? "eqs" : m.pos, 16);
  good = good && w->end();
  w->str(m.optm.empty() ? "nopt" : m.optm, 16);
  good = good && w->end();
  w->i4(m.glevel);
  good = good && w->end();
  w->i4(nv);
  w->i4(nt);
  w->i4(ne);
  good = good && w->end();
  w->i4(0);  // nlat: no regular search grid
  w->r8(0.0);
  good = good && w->end();
  w->i4(maxvnb);
  w->i4(0);
  w->r8(vmin);
  w->r8(vmax);
  w->r8(vsum / std::max(1, 2 * ne));
  w->r8(cmin);
  w->r8(cmax);
  w->r8(csum / std::max(1, ne));
  good = good && w->end();
  w->r8(tamin);
  w->r8(tamax);
  w->r8(tasum / std::max(1, nt));
  w->r8(hamin);
  w->r8(hamax);
  w->r8(hasum / std::max(1, nv));
  good = good && w->end();
  w->r8(angmin);
  w->r8(angmax);
  w->r8(angsum / std::max(1, 3 * nt));
  good = good && w->end();
  w->str(name, 128);
  good = good && w->end();
  close();
  // vertices
  if (!open("_vert_coord.dat", w)) return false;
  w->i4(nv);
  good = good && w->end();
  for (int i = 0; i < nv; ++i) {
    w->i4(i + 1);
    for (int q = 0; q < 3; ++q) w->r8(m.v_p[3 * i + q]);
    w->r8(m.v_lon[i]);
    w->r8(m.v_lat[i]);
    w->i4(m.v_nnb[i]);
  }
  good = good && w->end();
  close();
  if (!open("_vert_nb.dat", w)) return false;
  for (int i = 0; i < nv; ++i) {
    for (int j = 0; j < m.v_nnb[i]; ++j) {
      const int k = m.v_nb[(size_t)MX * i + j];
      w->i4(i + 1);
      w->i4(j + 1);
      w->i4(m.v_nnb[i]);
      w->i4(k);
      w->r8(-dot3(&m.v_p[3 * i], &m.v_p[3 * (k - 1)]));
      w->r8(arclen(&m.v_p[3 * i], &m.v_p[3 * (k - 1)]));
      w->i4(m.v_ed[(size_t)MX * i + j]);
      w->i4(m.v_tr[(size_t)MX * i + j]);
    }
    good = good && w->end();
  }
  close();
  // primal edges
  if (!open("_ed_cc.dat", w)) return false;
  w->i4(ne);
  good = good && w->end();
  for (int l = 0; l < ne; ++l) {
    w->i4(l + 1);
    for (int q = 0; q < 3; ++q) w->r8(m.ed_c[3 * l + q]);
    w->r8(m.ed_clon[l]);
    w->r8(m.ed_clat[l]);
  }
  good = good && w->end();
  close();
  if (!open("_ed_nb.dat", w)) return false;
  for (int l = 0; l < ne; ++l) {
    const double *p1 = &m.v_p[3 * (m.ed_v[2 * l] - 1)], *p2 = &m.v_p[3 * (m.ed_v[2 * l + 1] - 1)];
    double d[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
    w->i4(l + 1);
    w->i4(m.ed_v[2 * l]);
    w->i4(m.ed_v[2 * l + 1]);
    w->i4(m.ed_sh[2 * l]);
    w->i4(m.ed_sh[2 * l + 1]);
    for (int q = 0; q < 3; ++q) w->r8(m.ed_tg[3 * l + q]);
    for (int q = 0; q < 3; ++q) w->r8(m.ed_nr[3 * l + q]);
    w->r8(norm3(d));
    w->r8(m.ed_leng[l]);
  }
  good = good && w->end();
  close();
  // dual edges: ends are the circumcentres of the edge's two triangles
  if (!open("_edhx_cc.dat", w)) return false;
  w->i4(ne);
  good = good && w->end();
  for (int l = 0; l < ne; ++l) {
    const double *c = &m.edhx_c[3 * l];
    double lon, lat;
    cart2sph(c[0], c[1], c[2], &lon, &lat);
    w->i4(l + 1);
    for (int q = 0; q < 3; ++q) w->r8(c[q]);
    w->r8(lon);
    w->r8(lat);
  }
  good = good && w->end();
  close();
  if (!open("_edhx_nb.dat", w)) return false;
  for (int l = 0; l < ne; ++l) {
    const int t1 = m.ed_sh[2 * l], t2 = m.ed_sh[2 * l + 1];
    const double *c1 = &m.tr_c[3 * (t1 - 1)], *c2 = &m.tr_c[3 * (t2 - 1)];
    double d[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
    w->i4(l + 1);
    w->i4(t1);
    w->i4(t2);
    w->i4(m.ed_v[2 * l]);
    w->i4(m.ed_v[2 * l + 1]);
    for (int q = 0; q < 3; ++q) w->r8(m.edhx_tg[3 * l + q]);
    for (int q = 0; q < 3; ++q) w->r8(m.edhx_nr[3 * l + q]);
    w->r8(norm3(d));
    w->r8(m.edhx_leng[l]);
  }
  good = good && w->end();
  close();
  // triangles
  if (!open("_tr_cc.dat", w)) return false;
  w->i4(nt);
  good = good && w->end();
  for (int k = 0; k < nt; ++k) {
    w->i4(k + 1);
    for (int q = 0; q < 3; ++q) w->r8(m.tr_c[3 * k + q]);
    w->r8(m.tr_clon[k]);
    w->r8(m.tr_clat[k]);
  }
  good = good && w->end();
  for (int k = 0; k < nt; ++k) {  // barycentres: normalised vertex mean
    double c[3] = {0, 0, 0}, lon, lat;
    for (int v = 0; v < 3; ++v)
      for (int q = 0; q < 3; ++q) c[q] += m.v_p[3 * (m.tr_v[3 * k + v] - 1) + q] / 3.0;
    const double n = norm3(c);
    for (int q = 0; q < 3; ++q) c[q] /= n;
    cart2sph(c[0], c[1], c[2], &lon, &lat);
    w->i4(k + 1);
    for (int q = 0; q < 3; ++q) w->r8(c[q]);
    w->r8(lon);
    w->r8(lat);
  }
  good = good && w->end();
  close();
  if (!open("_tr_nb.dat", w)) return false;
  for (int k = 0; k < nt; ++k) {
    const double *p1 = &m.v_p[3 * (m.tr_v[3 * k] - 1)], *p2 = &m.v_p[3 * (m.tr_v[3 * k + 1] - 1)],
                 *p3 = &m.v_p[3 * (m.tr_v[3 * k + 2] - 1)];
    w->i4(k + 1);
    for (int q = 0; q < 3; ++q) w->i4(m.tr_v[3 * k + q]);
    for (int q = 0; q < 3; ++q) w->i4(m.tr_ed[3 * k + q]);
    for (int q = 0; q < 3; ++q) {  // neighbour across edge q: the other triangle of that edge
      const int l = m.tr_ed[3 * k + q] - 1;
      w->i4(m.ed_sh[2 * l] == k + 1 ? m.ed_sh[2 * l + 1] : m.ed_sh[2 * l]);
    }
    for (int q = 0; q < 3; ++q) w->i4(m.tr_tg[3 * k + q]);
    for (int q = 0; q < 3; ++q) w->i4(m.tr_nr[3 * k + q]);
    const double *c = &m.tr_c[3 * k];
    double d[3] = {c[0] - p1[0], c[1] - p1[1], c[2] - p1[2]};
    double e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double e2[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    double x[3];
    cross3(e1, e2, x);
    w->r8(norm3(d));
    w->r8(arclen(c, p1));
    w->r8(norm3(x) / 2);
    w->r8(m.tr_areag[k]);
    for (int q = 0; q < 3; ++q) w->r8(ang[3 * k + q]);
  }
  good = good && w->end();
  close();
  // Voronoi cells
  if (!open("_hx.dat", w)) return false;
  w->i4(nv);
  good = good && w->end();
  for (int i = 0; i < nv; ++i) {
    const int n = m.v_nnb[i];
    const double *p = &m.v_p[3 * i];
    double cen[3] = {0, 0, 0}, ap = 0;
    for (int j = 0; j < n; ++j) {
      const double *a = &m.tr_c[3 * (m.v_tr[(size_t)MX * i + (j + n - 1) % n] - 1)];
      const double *bb = &m.tr_c[3 * (m.v_tr[(size_t)MX * i + j] - 1)];
      const double A = sphtriarea(p, a, bb);
      for (int q = 0; q < 3; ++q) cen[q] += A * (p[q] + a[q] + bb[q]) / 3.0;
      double e1[3] = {a[0] - p[0], a[1] - p[1], a[2] - p[2]}, e2[3] = {bb[0] - p[0], bb[1] - p[1], bb[2] - p[2]};
      double x[3];
      cross3(e1, e2, x);
      ap += norm3(x) / 2;
    }
    const double nn = norm3(cen);
    for (int q = 0; q < 3; ++q) cen[q] /= nn;
    double lon, lat;
    cart2sph(cen[0], cen[1], cen[2], &lon, &lat);
    w->i4(i + 1);
    for (int q = 0; q < 3; ++q) w->r8(cen[q]);
    w->r8(lon);
    w->r8(lat);
    good = good && w->end();
    for (int j = 0; j < n; ++j) {
      w->i4(m.hx_tg[(size_t)MX * i + j]);
      w->i4(m.hx_nr[(size_t)MX * i + j]);
      w->i4(m.hx_ttgout[(size_t)MX * i + j]);
      w->i4(m.hx_tnrccw[(size_t)MX * i + j]);
    }
    good = good && w->end();
    w->r8(ap);
    w->r8(m.hx_areag[i]);
    w->r8(0.0);  // alignment index: not used by the shallow-water path
    good = good && w->end();
  }
  close();
  // regular search grid: empty
  if (!open("_quad.dat", w)) return false;
  w->i4(0);
  good = good && w->end();
  close();
  if (!good) err = "write error in " + file_of(dir, name, "_*.dat");
  return good;
}

}  // namespace imx

extern "C" {

int imx_mesh_load_ref(const char *dir, const char *name, imx_mesh **out) {
  if (!dir || !name || !out) return imx::set_error(IMX_EINVAL, "imx_mesh_load_ref: NULL argument");
  *out = nullptr;
  imx_mesh *m = new (std::nothrow) imx_mesh();
  if (!m) return imx::set_error(IMX_ENOMEM, "imx_mesh_load_ref: out of memory");
  std::string err;
  bool ok = false;
  try {
    ok = imx::load_ref(dir, name, *m, err);
  } catch (const std::bad_alloc &) {
    err = "out of memory";
  }
  if (!ok) {
    delete m;
    return imx::set_error(IMX_EIO, "imx_mesh_load_ref: " + err);
  }
  *out = m;
  return IMX_OK;
}

int imx_make_dir(const char *path) {
  if (!path || !*path) return imx::set_error(IMX_EINVAL, "imx_make_dir: empty path");
  std::string p(path), cur;
  size_t i = 0;
  while (i <= p.size()) {  // every prefix ending at '/' and the whole path
    const size_t j = p.find('/', i);
    cur = p.substr(0, j == std::string::npos ? p.size() : j);
    if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
      return imx::set_error(IMX_EIO, "imx_make_dir: cannot create " + cur + ": " + std::strerror(errno));
    if (j == std::string::npos) break;
    i = j + 1;
  }
  return IMX_OK;
}

int imx_mesh_write_ref(const imx_mesh *m, const char *dir, const char *name) {
  if (!m || !dir || !name) return imx::set_error(IMX_EINVAL, "imx_mesh_write_ref: NULL argument");
  int rc = imx_make_dir(dir);
  if (rc) return rc;
  std::string err;
  if (!imx::write_ref(*m, dir, name, err)) return imx::set_error(IMX_EIO, "imx_mesh_write_ref: " + err);
  return IMX_OK;
}

}  // extern "C"
