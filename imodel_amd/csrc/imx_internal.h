// Internal declarations shared by the host-side translation units of libimx.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/imx_swm.h"

namespace imx {

// fn(i0, i1) over [0, n) split across up to 16 host threads; each index is
// handled by exactly one thread, so per-element results do not depend on
// the split (reductions stay serial in the callers).
template <typename F>
void parallel_ranges(int n, F fn) {
  const int hw = (int)std::thread::hardware_concurrency();
  const int nt = std::max(1, std::min({16, hw > 0 ? hw : 1, n / 4096 + 1}));
  if (nt == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const int i0 = (int)((int64_t)n * t / nt), i1 = (int)((int64_t)n * (t + 1) / nt);
    th.emplace_back([=] { fn(i0, i1); });
  }
  for (auto &x : th) x.join();
}

// Thread-local last error; every C entry point returns a code and leaves the
// message here.
int set_error(int code, const std::string &msg);

// Host mesh in the reference's conventions (1-based indices, per-cell lists
// padded to IMX_MAXNB). Field names follow grid_structure.
struct HostMesh {
  int32_t nv = 0, ne = 0, nt = 0, glevel = -1;
  std::string kind = "icos", pos = "eqs", optm = "nopt";  // grid-file header fields
  std::vector<double> v_p, v_lat, v_lon, hx_areag;
  std::vector<int32_t> v_nnb, v_nb, v_ed, v_tr, hx_ttgout, hx_tnrccw, hx_tg, hx_nr;
  std::vector<int32_t> ed_v, ed_sh;
  std::vector<double> ed_leng, edhx_leng, ed_c, ed_clat, ed_clon, ed_tg, ed_nr;
  std::vector<double> edhx_c, edhx_nr, edhx_tg;
  std::vector<int32_t> tr_v, tr_ed, tr_tg, tr_nr;
  std::vector<double> tr_areag, tr_c, tr_clat, tr_clon;
  std::vector<double> tr_trhx_area, hx_hxtr_area, hx_trskw;  // gridprop
  int64_t trmesh_restarts = 0;  // located-walk restarts of the triangulation
};

// Delaunay triangulation of n unit vectors in STRIPACK TRMESH's insertion
// order (imx_trmesh.cpp); returns each node's CCW neighbour ring (0-based,
// maxnb slots per node) starting where meshstruct starts reading it, and
// the number of walk restarts (non-zero: the numbering may differ from the
// reference's, whose restarts are clock-seeded).
bool trmesh_rings(int n, const double *x, const double *y, const double *z, std::vector<int32_t> &nnb,
                  std::vector<int32_t> &nb, int maxnb, int64_t *restarts, std::string &err);

// Spherical geometry with the reference's expression order
// (src/smeshpack.f90: arclen :4624, sphtriarea :4710, cross_product :6560,
// det :6580, norm :7061, cart2sph :4279). Evaluated without FMA contraction.
double dot3(const double *a, const double *b);
void cross3(const double *a, const double *b, double *c);
double norm3(const double *a);
double arclen(const double *p, const double *q);
double sphtriarea(const double *p1, const double *p2, const double *p3);
double det3(const double *p1, const double *p2, const double *p3);
void cart2sph(double x, double y, double z, double *lon, double *lat);

void gridprop(const imx_mesh_desc &d, double *trhx, double *hxtr, double *trskw);
// convert_vec_sph2cart (src/smeshpack.f90:4379-4417)
void vec_sph2cart(double vlon, double vlat, const double *p, double *v);
// Wachspress coordinates of cell node i (0-based) in its ring of
// circumcentres, in v%tr order (wachspress_coords_hxv, interpack.f90:2310-2415)
void wachspress_node(const imx_mesh_desc &d, int i, double *w);
// Radius, in cell rings around a row's cells, of a stencil superset of the
// handle's tendency: 1 for TRiSK, one more ring each for APVM, for
// hyperdiffusion and for the gass18 Coriolis.
int swm_stencil_rings(const imx_swm *s);
// dif_coef_hx [nv] / dif_coef_tr [nt] as swm_tests sets them up
// (src/swm.f90:170-210), caller numbering; kind IMX_COEF_*. With k4 > 0
// both hold the square roots of the hyperdiffusion coefficients.
void diffusion_coef(const imx_mesh_desc &d, const double *trhx, bool bary, double k2, int k2kind,
                    double k4, int k4kind, std::vector<double> &hx, std::vector<double> &tr);
// bar_coord_tr (src/smeshpack.f90:5200-5245) of point p in triangle k, and
// of triangle k's circumcentre
void bary_at(const imx_mesh_desc &d, const double *p, int k, double b[3]);
void bary_circumcentre(const imx_mesh_desc &d, int k, double b[3]);
// insidetrmesh (src/smeshpack.f90:5329-5420), with robdet's binary128
// determinant near the edges (:6597-6629); 0-based triangle
bool insidetrmesh(const imx_mesh_desc &d, const double *p, int tr);
// gethxedgeconnection(ed1, ed2) (src/smeshpack.f90:4127-4244), 0-based:
// -2 if ed1 == ed2, the linking edge, or -1 if the edges are far apart
int hxedgeconn(const imx_mesh_desc &d, int ed1, int ed2);
// Internal numbering: cells along a Hilbert curve (or caller order); edge of
// owning cell c -> 3c+k, triangle of owning cell c -> 2c+k. Maps are caller
// (0-based) -> internal; NE = 3 nv and NT = 2 nv slots (some are dummies).
struct Layout {
  std::vector<int32_t> cell_new, edge_new, tri_new;
  int NE = 0, NT = 0;
};
// group (optional, per caller cell): cells are ordered by (group, Hilbert key),
// so each group's cells -- and the edges and triangles they own -- are
// contiguous internal ranges.
bool build_layout(const imx_mesh_desc &d, bool use_sfc, Layout &L, std::string &err,
                  const int32_t *group = nullptr);
void desc_of(const HostMesh &m, imx_mesh_desc &d);
// meshstruct on nodes p [3 nv] and their CCW neighbour rings (0-based,
// IMX_MAXNB slots per node): edges, triangles, geometry and sign tables.
bool build_from_rings(HostMesh &m, int nv, const std::vector<double> &p, const std::vector<int32_t> &nnb,
                      const std::vector<int32_t> &nb0, std::string &err);
// Grid files in the reference's meshwrite/meshload layout (imx_gridio.cpp).
bool load_ref(const std::string &dir, const std::string &name, HostMesh &m, std::string &err);
bool write_ref(const HostMesh &m, const std::string &dir, const std::string &name, std::string &err);

// IMXBIN container: "IMXBIN01", then records (name[24], dtype[4], int64 n, data).
struct BinWriter {
  FILE *f = nullptr;
  explicit BinWriter(const char *path);
  ~BinWriter();
  bool ok() const { return f != nullptr; }
  void put(const char *name, const std::vector<int32_t> &a);
  void put(const char *name, const std::vector<double> &a);
  void put_i4(const char *name, const int32_t *a, int64_t n);
  void put_f8(const char *name, const double *a, int64_t n);
};

// Fields of one LTE evaluation in caller numbering (vectors x y z per entry;
// q_grad_tr and lapu NULL when the options leave them uncomputed).
struct LteFields {
  const double *h, *u, *h_ed, *uh, *divuh, *ke_hx, *eta, *h_tr, *q_tr, *q_hx, *q_ed, *grad_ghbk,
      *vhq_tr, *q_grad_tr, *vh_hx, *vhq_hx, *uhq_perp, *lapu, *masseq, *momeq, *bt;
};
// Closed forms of test case tc (2, or the Hollingsworth cases 32-35 with
// layer depth hollgw), error_calc of each row and the mimetic sums of
// swm_horiz_loc_trunc_er (imx_lte.cpp); perhx: Perot's cell reconstruction
// of the kinetic energy, which changes calc_energies' kinetic energy.
int lte_analysis(const HostMesh &m, int tc, double hollgw, bool hc, bool perhx, double erad, double grav,
                 const LteFields &f, double *err, double *mim);

}  // namespace imx

// The opaque handle of the C ABI is the host mesh itself.
struct imx_mesh : public imx::HostMesh {};
