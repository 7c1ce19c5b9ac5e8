// Space-filling-curve ordering of cells, edges and triangles for coalesced,
// cache-friendly gathers on the GPU.
//
// Cells are sorted by the 3-D Hilbert index of their unit position vector
// (Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004:
// transpose form, 21 bits per axis). Edges and triangles take their
// "first-touch" order: walking cells in Hilbert order, each cell numbers the
// not-yet-numbered edges (triangles) of its ring in ring order, so the ~3 edges
// and ~2 triangles a cell introduces sit next to it in memory.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "imx_internal.h"

namespace imx {

static uint64_t hilbert3(double x, double y, double z) {
  const int B = 21;
  uint32_t X[3];
  const double v[3] = {x, y, z};
  for (int i = 0; i < 3; ++i) {
    double t = (v[i] + 1.0) * 0.5 * (double)(1u << B);
    if (t < 0) t = 0;
    if (t > (double)((1u << B) - 1)) t = (double)((1u << B) - 1);
    X[i] = (uint32_t)t;
  }
  // inverse undo of the reflections/exchanges (Skilling's AxestoTranspose)
  const uint32_t M = 1u << (B - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];  // Gray encode
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t key = 0;
  for (int b = B - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

// new_of_old maps caller (0-based) index -> internal index.
void sfc_order(const imx_mesh_desc &d, bool use_sfc, std::vector<int32_t> &cell_new,
               std::vector<int32_t> &edge_new, std::vector<int32_t> &tri_new) {
  const int nv = d.nv, ne = d.ne, nt = d.nt, MX = IMX_MAXNB;
  std::vector<int32_t> cell_old(nv);
  std::iota(cell_old.begin(), cell_old.end(), 0);
  if (use_sfc) {
    std::vector<uint64_t> key(nv);
    for (int i = 0; i < nv; ++i) key[i] = hilbert3(d.v_p[3 * i], d.v_p[3 * i + 1], d.v_p[3 * i + 2]);
    std::stable_sort(cell_old.begin(), cell_old.end(),
                     [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  }
  cell_new.assign(nv, -1);
  for (int i = 0; i < nv; ++i) cell_new[cell_old[i]] = i;
  edge_new.assign(ne, -1);
  tri_new.assign(nt, -1);
  if (!use_sfc) {
    std::iota(edge_new.begin(), edge_new.end(), 0);
    std::iota(tri_new.begin(), tri_new.end(), 0);
    return;
  }
  int32_t ecount = 0, tcount = 0;
  for (int ci = 0; ci < nv; ++ci) {
    int c = cell_old[ci];
    for (int j = 0; j < d.v_nnb[c]; ++j) {
      int e = d.v_ed[MX * c + j] - 1;
      if (edge_new[e] < 0) edge_new[e] = ecount++;
      int t = d.v_tr[MX * c + j] - 1;
      if (tri_new[t] < 0) tri_new[t] = tcount++;
    }
  }
  // entities not reached through a ring (never on a closed sphere mesh)
  for (int e = 0; e < ne; ++e)
    if (edge_new[e] < 0) edge_new[e] = ecount++;
  for (int t = 0; t < nt; ++t)
    if (tri_new[t] < 0) tri_new[t] = tcount++;
}

}  // namespace imx
