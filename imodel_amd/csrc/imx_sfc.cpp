// Internal numbering of libimx: cells along a space-filling curve, edges and
// triangles numbered by an owning cell.
//
// Cells are sorted by the 3-D Hilbert index of their unit position vector
// (Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004:
// transpose form, 21 bits per axis).
//
// Every edge is owned by one of its two cells and every triangle by one of
// its three vertices, with at most 3 edges and 2 triangles per cell. Edge e of
// owner c gets internal id 3c+k and triangle t gets 2c+k (k = rank of the
// entity in the owner's CCW ring), so a cell's own edges and triangles sit
// next to it in memory and no cell->edge or cell->triangle table is needed:
// a cell finds the id of ring entity j from the neighbour id it already
// holds plus a few bits. Ids of slots no cell fills (pentagons own fewer) are
// dummies that every kernel skips.
//  * Edges: an Euler-circuit orientation (Hierholzer) of the cell graph with
//    the 12 odd-degree pentagons paired by virtual edges gives every cell
//    out-degree deg/2 (hexagons 3, pentagons 2 or 3); a cell owns its
//    out-edges.
//  * Triangles: a b-matching triangle -> vertex with capacity 2, greedy in
//    Hilbert order, completed by augmenting paths.
#include <algorithm>
#include <cstdint>
#include <deque>
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
  const uint32_t M = 1u << (B - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {  // inverse undo
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

bool build_layout(const imx_mesh_desc &d, bool use_sfc, Layout &L, std::string &err,
                  const int32_t *group) {
  const int nv = d.nv, ne = d.ne, nt = d.nt, MX = IMX_MAXNB;
  // ---- cells
  std::vector<int32_t> cell_old(nv);
  std::iota(cell_old.begin(), cell_old.end(), 0);
  if (use_sfc || group) {
    std::vector<uint64_t> key(nv, 0);
    if (use_sfc)
      for (int i = 0; i < nv; ++i) key[i] = hilbert3(d.v_p[3 * i], d.v_p[3 * i + 1], d.v_p[3 * i + 2]);
    std::stable_sort(cell_old.begin(), cell_old.end(), [&](int32_t a, int32_t b) {
      if (group && group[a] != group[b]) return group[a] < group[b];
      return key[a] < key[b];
    });
  }
  L.cell_new.assign(nv, -1);
  for (int i = 0; i < nv; ++i) L.cell_new[cell_old[i]] = i;

  // ---- edge owners: closed trails (Hierholzer) over cells in internal order
  std::vector<int32_t> odd;
  for (int ci = 0; ci < nv; ++ci)
    if (d.v_nnb[cell_old[ci]] & 1) odd.push_back(cell_old[ci]);
  if (odd.size() & 1) {
    err = "odd number of odd-degree cells";
    return false;
  }
  const int nvirt = (int)odd.size() / 2;
  std::vector<int32_t> virt_of(nv, -1);  // virtual edge id (ne + k) at an odd cell
  std::vector<int32_t> virt_a(nvirt), virt_b(nvirt);
  for (int k = 0; k < nvirt; ++k) {
    virt_a[k] = odd[2 * k];
    virt_b[k] = odd[2 * k + 1];
    virt_of[odd[2 * k]] = ne + k;
    virt_of[odd[2 * k + 1]] = ne + k;
  }
  std::vector<char> used((size_t)ne + nvirt, 0);
  std::vector<int32_t> tail((size_t)ne + nvirt, -1);
  std::vector<int32_t> ptr(nv, 0);  // next ring slot to try (slot nnb = virtual edge)
  auto other = [&](int32_t e, int32_t v) -> int32_t {
    if (e >= ne) return virt_a[e - ne] == v ? virt_b[e - ne] : virt_a[e - ne];
    return d.ed_v[2 * e] - 1 == v ? d.ed_v[2 * e + 1] - 1 : d.ed_v[2 * e] - 1;
  };
  auto next_edge = [&](int32_t v) -> int32_t {
    const int n = d.v_nnb[v];
    while (ptr[v] <= n) {
      const int s = ptr[v]++;
      const int32_t e = (s < n) ? d.v_ed[MX * v + s] - 1 : virt_of[v];
      if (e >= 0 && !used[e]) return e;
    }
    return -1;
  };
  for (int ci = 0; ci < nv; ++ci) {
    const int32_t s = cell_old[ci];
    for (;;) {
      int32_t v = s;
      int32_t e = next_edge(v);
      if (e < 0) break;
      while (e >= 0) {
        used[e] = 1;
        tail[e] = v;
        v = other(e, v);
        e = next_edge(v);
      }
      if (v != s) {
        err = "edge orientation trail did not close";
        return false;
      }
    }
  }
  // ranks of owned edges in the owner's ring
  L.edge_new.assign(ne, -1);
  for (int ci = 0; ci < nv; ++ci) {
    const int32_t c = cell_old[ci];
    int k = 0;
    for (int j = 0; j < d.v_nnb[c]; ++j) {
      const int32_t e = d.v_ed[MX * c + j] - 1;
      if (tail[e] == c) {
        if (k > 2) {
          err = "a cell owns more than 3 edges";
          return false;
        }
        L.edge_new[e] = 3 * ci + k++;
      }
    }
  }
  for (int e = 0; e < ne; ++e)
    if (L.edge_new[e] < 0) {
      err = "edge without owner";
      return false;
    }

  // ---- triangle owners: capacity-2 b-matching
  std::vector<int32_t> own(nt, -1), load(nv, 0);
  for (int ci = 0; ci < nv; ++ci) {  // greedy in Hilbert order
    const int32_t c = cell_old[ci];
    for (int j = 0; j < d.v_nnb[c] && load[c] < 2; ++j) {
      const int32_t t = d.v_tr[MX * c + j] - 1;
      if (own[t] < 0) {
        own[t] = c;
        ++load[c];
      }
    }
  }
  std::vector<int32_t> par_t(nt, -1), par_v(nt, -1), seen(nt, -1);
  for (int t0 = 0; t0 < nt; ++t0) {
    if (own[t0] >= 0) continue;
    std::deque<int32_t> q;
    q.push_back(t0);
    seen[t0] = t0;
    int32_t end_t = -1, end_v = -1;
    while (!q.empty() && end_t < 0) {
      const int32_t x = q.front();
      q.pop_front();
      for (int l = 0; l < 3 && end_t < 0; ++l) {
        const int32_t v = d.tr_v[3 * x + l] - 1;
        if (v == own[x]) continue;
        if (load[v] < 2) {
          end_t = x;
          end_v = v;
          break;
        }
        for (int j = 0; j < d.v_nnb[v]; ++j) {
          const int32_t z = d.v_tr[MX * v + j] - 1;
          if (own[z] == v && seen[z] != t0) {
            seen[z] = t0;
            par_t[z] = x;
            par_v[z] = v;
            q.push_back(z);
          }
        }
      }
    }
    if (end_t < 0) {
      err = "triangle ownership has no augmenting path";
      return false;
    }
    // shift ownership back along the path: end_t takes the free vertex, each
    // parent takes the vertex its child held
    int32_t cur = end_t;
    own[cur] = end_v;
    ++load[end_v];
    while (cur != t0) {
      own[par_t[cur]] = par_v[cur];
      cur = par_t[cur];
    }
  }
  L.tri_new.assign(nt, -1);
  for (int ci = 0; ci < nv; ++ci) {
    const int32_t c = cell_old[ci];
    int k = 0;
    for (int j = 0; j < d.v_nnb[c]; ++j) {
      const int32_t t = d.v_tr[MX * c + j] - 1;
      if (own[t] == c && L.tri_new[t] < 0) {
        if (k > 1) {
          err = "a cell owns more than 2 triangles";
          return false;
        }
        L.tri_new[t] = 2 * ci + k++;
      }
    }
  }
  for (int t = 0; t < nt; ++t)
    if (L.tri_new[t] < 0) {
      err = "triangle without owner";
      return false;
    }
  L.NE = 3 * nv;
  L.NT = 2 * nv;
  return true;
}

}  // namespace imx
