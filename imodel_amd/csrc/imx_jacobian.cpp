// Linearised shallow-water operator for normal-mode analysis
// (swm_normal_mode_analysis, src/swm.f90:738-976): the one-sided finite
// difference Jacobian M(:, i) = (T(x + pert e_i) - T(x)) / pert of the
// tendency T = (masseq, momeq) over all nv + ne unknowns (h, then u).
//
// The reference evaluates one tendency per unknown. Here columns whose
// rows cannot overlap share one evaluation (Curtis-Powell-Reid column
// colouring): row r of the tendency reads only unknowns in a stencil S(r),
// so when every row's stencil holds at most one column of a colour, the
// tendency of the state perturbed at all columns of that colour gives, in
// row r, exactly (bit for bit) the value of the single-column perturbation
// of the column of that colour in S(r). S(r) is taken as a superset of
// what the TRiSK tendency reads:
//   masseq_c: h at c and its neighbours, u on edges among them;
//   momeq_e (cells c1, c2): h at c1, c2 and all their neighbours, u on
//   every edge whose two cells are in that set (the cells' own edges for
//   K and the flux, and the edges between consecutive neighbours that
//   close the triangles of the vorticity in q).
// APVM (q at an edge reads grad q at its triangles and v at its cells) and
// hyperdiffusion (the Laplacian of the Laplacian) each reach one cell ring
// further; diffusion stays within the TRiSK set. The sets are therefore
// the cells within R rings of the row's cells, R = 1, 2 or 3
// (swm_stencil_rings), with the edges among them.
// A few tens of colours replace nv + ne evaluations.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "imx_internal.h"

namespace imx {

namespace {

// stencil superset of every row (0..nv-1 masseq, nv..nv+ne-1 momeq), as
// column ids (0..nv-1 h, nv.. u)
void stencils(const imx_mesh_desc &d, int rings, std::vector<std::vector<int32_t>> &S) {
  const int nv = d.nv, ne = d.ne, MX = IMX_MAXNB;
  S.assign((size_t)nv + ne, {});
  std::vector<int32_t> mark(nv, -1);
  std::vector<int32_t> cells;
  auto close_over = [&](int row, const std::vector<int32_t> &cs) {
    std::vector<int32_t> &s = S[row];
    for (int c : cs) s.push_back(c);
    // edges with both ends in the set: scan the edges of each member
    for (int c : cs)
      for (int j = 0; j < d.v_nnb[c]; ++j) {
        const int e = d.v_ed[MX * c + j] - 1;
        const int a = d.ed_v[2 * e] - 1, b = d.ed_v[2 * e + 1] - 1;
        if (mark[a] == row && mark[b] == row) s.push_back(nv + e);
      }
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
  };
  auto add_cell = [&](int row, int c) {
    if (mark[c] != row) {
      mark[c] = row;
      cells.push_back(c);
    }
  };
  // the row's cells, then `rings` rings of neighbours around them
  auto grow = [&](int row) {
    size_t a = 0;
    for (int r = 0; r < rings; ++r) {
      const size_t b = cells.size();
      for (size_t i = a; i < b; ++i) {
        const int c = cells[i];
        for (int j = 0; j < d.v_nnb[c]; ++j) add_cell(row, d.v_nb[MX * c + j] - 1);
      }
      a = b;
    }
  };
  for (int c = 0; c < nv; ++c) {
    cells.clear();
    add_cell(c, c);
    grow(c);
    close_over(c, cells);
  }
  for (int e = 0; e < ne; ++e) {
    const int row = nv + e;
    cells.clear();
    for (int q = 0; q < 2; ++q) add_cell(row, d.ed_v[2 * e + q] - 1);
    grow(row);
    // mark is indexed by cell with the row id as stamp; rows >= nv reuse it
    close_over(row, cells);
  }
}

}  // namespace

}  // namespace imx

extern "C" {

int imx_swm_jacobian(imx_swm *s, const imx_mesh_desc *d, const double *h, const double *u, double pert,
                     int64_t cap, int32_t *row, int32_t *col, double *val, int64_t *nnz, int32_t *ncolors) {
  if (!s || !d || !h || !u || !nnz || !(pert != 0.0))
    return imx::set_error(IMX_EINVAL, "imx_swm_jacobian: bad argument");
  const int nv = d->nv, ne = d->ne, n = nv + ne;
  std::vector<std::vector<int32_t>> S;
  imx::stencils(*d, imx::swm_stencil_rings(s), S);
  // column -> rows that may read it
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  for (int r = 0; r < n; ++r)
    for (int c : S[r]) ++cnt[c + 1];
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> rows_of(cnt[n]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < n; ++r)
      for (int c : S[r]) rows_of[pos[c]++] = r;
  }
  // greedy colouring of the column intersection graph
  std::vector<int32_t> color(n, -1), seen;
  int ncol = 0;
  for (int i = 0; i < n; ++i) {
    for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k)
      for (int j : S[rows_of[k]])
        if (color[j] >= 0) {
          if ((int)seen.size() <= color[j]) seen.resize(color[j] + 1, -1);
          seen[color[j]] = i;
        }
    int c = 0;
    while (c < (int)seen.size() && seen[c] == i) ++c;
    color[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  std::vector<std::vector<int32_t>> members(ncol);
  for (int i = 0; i < n; ++i) members[color[i]].push_back(i);
  // base tendency and one evaluation per colour
  std::vector<double> m0(nv), f0(ne), m1(nv), f1(ne), hp(h, h + nv), up(u, u + ne);
  int rc = imx_swm_tendency(s, h, u, m0.data(), f0.data());
  if (rc) return rc;
  struct Ent {
    int32_t r, c;
    double v;
  };
  std::vector<Ent> out;
  for (int c = 0; c < ncol; ++c) {
    for (int i : members[c]) {
      if (i < nv) hp[i] = h[i] + pert;
      else up[i - nv] = u[i - nv] + pert;
    }
    rc = imx_swm_tendency(s, hp.data(), up.data(), m1.data(), f1.data());
    if (rc) return rc;
    for (int i : members[c]) {
      if (i < nv) hp[i] = h[i];
      else up[i - nv] = u[i - nv];
      for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) {
        const int r = rows_of[k];
        const double v = r < nv ? (m1[r] - m0[r]) / pert : (f1[r - nv] - f0[r - nv]) / pert;
        if (v != 0.0) out.push_back({r, i, v});
      }
    }
  }
  std::sort(out.begin(), out.end(), [](const Ent &a, const Ent &b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  *nnz = (int64_t)out.size();
  if (ncolors) *ncolors = ncol;
  if (cap < (int64_t)out.size()) return IMX_OK;  // size query: nothing written
  for (size_t k = 0; k < out.size(); ++k) {
    if (row) row[k] = out[k].r + 1;
    if (col) col[k] = out[k].c + 1;
    if (val) val[k] = out[k].v;
  }
  return IMX_OK;
}

}  // extern "C"
