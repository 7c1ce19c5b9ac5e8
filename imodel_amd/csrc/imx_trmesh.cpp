// Delaunay triangulation of nodes on the unit sphere, in the insertion and
// edge-swap order of the reference's mesh generator, so that every index
// table built from it (neighbour rings, edges, triangles) carries the
// reference's numbering.
//
// The reference triangulates with STRIPACK's TRMESH (Renka, ACM TOMS 772,
// vendored as src/stripackm.f90): nodes 1..3 form the first triangle, node k
// is located from the already-inserted node nearest to it (the NEAR/NEXT/DIST
// lists), added inside its triangle, outside the hull or as the node that
// closes the sphere, and then the circumcircle test swaps the arcs opposite
// to it. Which neighbour a node's ring starts at (LEND) depends on that whole
// history, and the icosahedral midpoints lie exactly on the arcs they bisect,
// so the orientation predicates run on rounding noise. Matching the
// reference therefore needs the same steps, the same predicate expressions
// and the same tie rules, evaluated without FMA contraction (this file is
// compiled with -ffp-contract=off). Steps are kept alike for that reason;
// the data layout (one adjacency arena per call, 1-based as the linked lists
// store signed node ids) is this file's own.
//
// Followed: trmesh src/stripackm.f90:6843-7314, addnod :38-302, trfind
// :5507-6060, intadd :3822-3923, bdyadd :497-637, covsph :865-976, insert
// :3402-3459, swap :5129-5245, swptst :5248-5345, left :4175-4235, lstptr
// :4238-4317.
//
// The reference's TRFIND restarts a failed walk from a node drawn by NRAND
// (:4102-4172), seeded from the system clock, so a run that needs it is not
// reproducible even by the reference. Here the restart node comes from a
// fixed sequence and every restart is counted; the reference's numbering is
// only claimed for triangulations with no restart (the icosahedral grids
// tested have none).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "imx_internal.h"

namespace imx {

namespace {

struct Delaunay {
  const double *x, *y, *z;          // node coordinates, 1-based views
  std::vector<int32_t> list, lptr;  // arc arena: signed node id, successor
  std::vector<int32_t> lend;        // per node: arena slot of its last arc
  int32_t lnew = 1;                 // next free arena slot
  int64_t restarts = 0;
  uint32_t restart_state = 12345u;

  // orientation of p against the plane through the origin, a and b
  // (the reference's LEFT and TRFIND's DET share this expression)
  double det(int a, int b, double xp, double yp, double zp) const {
    const double x1 = x[a], y1 = y[a], z1 = z[a], x2 = x[b], y2 = y[b], z2 = z[b];
    return xp * (y1 * z2 - y2 * z1) - yp * (x1 * z2 - x2 * z1) + zp * (x1 * y2 - x2 * y1);
  }

  // slot of nb in the ring whose last slot is lpl (the last slot when absent)
  int32_t find_arc(int32_t lpl, int32_t nb) const {
    int32_t lp = lptr[lpl];
    for (;;) {
      if (list[lp] == nb) break;
      lp = lptr[lp];
      if (lp == lpl) break;
    }
    return lp;
  }

  void insert_after(int32_t k, int32_t lp) {
    const int32_t s = lptr[lp];
    lptr[lp] = lnew;
    list[lnew] = k;
    lptr[lnew] = s;
    ++lnew;
  }

  // circumcircle test: does n4 lie inside the circle through n1 n2 n3
  // (the arc n1-n2 is to be replaced by n3-n4)
  bool swap_test(int n1, int n2, int n3, int n4) const {
    const double x4 = x[n4], y4 = y[n4], z4 = z[n4];
    const double dx1 = x[n1] - x4, dx2 = x[n2] - x4, dx3 = x[n3] - x4;
    const double dy1 = y[n1] - y4, dy2 = y[n2] - y4, dy3 = y[n3] - y4;
    const double dz1 = z[n1] - z4, dz2 = z[n2] - z4, dz3 = z[n3] - z4;
    return dx3 * (dy2 * dz1 - dy1 * dz2) - dy3 * (dx2 * dz1 - dx1 * dz2) + dz3 * (dx2 * dy1 - dx1 * dy2) > 0.0;
  }

  // replace arc io1-io2 by in1-in2; returns the slot of in1 in in2's ring,
  // or 0 when in1-in2 is already an arc (nothing changed)
  int32_t swap_arc(int in1, int in2, int io1, int io2) {
    int32_t lp = find_arc(lend[in1], in2);
    if (std::abs(list[lp]) == in2) return 0;
    lp = find_arc(lend[io1], in2);
    int32_t lph = lptr[lp];
    lptr[lp] = lptr[lph];
    if (lend[io1] == lph) lend[io1] = lp;
    lp = find_arc(lend[in1], io1);
    int32_t lpsav = lptr[lp];
    lptr[lp] = lph;
    list[lph] = in2;
    lptr[lph] = lpsav;
    lp = find_arc(lend[io2], in1);
    lph = lptr[lp];
    lptr[lp] = lptr[lph];
    if (lend[io2] == lph) lend[io2] = lp;
    lp = find_arc(lend[in2], io2);
    lpsav = lptr[lp];
    lptr[lp] = lph;
    list[lph] = in1;
    lptr[lph] = lpsav;
    return lph;
  }

  // k strictly inside triangle (i1, i2, i3)
  void add_interior(int k, int i1, int i2, int i3) {
    insert_after(k, find_arc(lend[i1], i2));
    insert_after(k, find_arc(lend[i2], i3));
    insert_after(k, find_arc(lend[i3], i1));
    list[lnew] = i1;
    list[lnew + 1] = i2;
    list[lnew + 2] = i3;
    lptr[lnew] = lnew + 1;
    lptr[lnew + 1] = lnew + 2;
    lptr[lnew + 2] = lnew;
    lend[k] = lnew + 2;
    lnew += 3;
  }

  // k outside the hull, seeing the boundary nodes i1 .. i2 (CCW)
  void add_boundary(int k, int i1, int i2) {
    const int n1 = i1, n2 = i2;
    int32_t lp = lend[n1];
    int32_t lsav = lptr[lp];
    lptr[lp] = lnew;
    list[lnew] = -k;
    lptr[lnew] = lsav;
    lend[n1] = lnew;
    ++lnew;
    int nxt = -list[lp];
    list[lp] = nxt;
    const int nsav = nxt;
    for (;;) {
      lp = lend[nxt];
      insert_after(k, lp);
      if (nxt == n2) break;
      nxt = -list[lp];
      list[lp] = nxt;
    }
    lsav = lnew;
    list[lnew] = n1;
    lptr[lnew] = lnew + 1;
    ++lnew;
    nxt = nsav;
    while (nxt != n2) {
      list[lnew] = nxt;
      lptr[lnew] = lnew + 1;
      ++lnew;
      lp = lend[nxt];
      nxt = list[lp];
    }
    list[lnew] = -n2;
    lptr[lnew] = lsav;
    lend[k] = lnew;
    ++lnew;
  }

  // k sees every boundary node: the triangulation closes over the sphere
  void add_cover(int k, int n0) {
    const int nst = n0;
    int nxt = nst;
    for (;;) {
      const int32_t lp = lend[nxt];
      insert_after(k, lp);
      nxt = -list[lp];
      list[lp] = nxt;
      if (nxt == nst) break;
    }
    const int32_t lsav = lnew;
    for (;;) {
      const int32_t lp = lend[nxt];
      list[lnew] = nxt;
      lptr[lnew] = lnew + 1;
      ++lnew;
      nxt = list[lp];
      if (nxt == nst) break;
    }
    lptr[lnew - 1] = lsav;
    lend[k] = lnew - 1;
  }

  int restart_node(int n) {
    ++restarts;
    restart_state = restart_state * 1664525u + 1013904223u;
    return 1 + (int)(restart_state % (uint32_t)n);
  }

  // Locate p among nodes 1..n starting at nst: returns (i1, i2, i3) of the
  // triangle containing p, or i3 = 0 with (i1, i2) the first and last
  // boundary node visible from p, i1 = i2 when all are, i1 = 0 when the
  // nodes are coplanar.
  void locate(int nst, const double *p, int n, int &i1, int &i2, int &i3) {
    const double eps = 2.220446049250313e-16;  // epsilon(1.0d0)
    const double tol = 100.0 * eps;
    const double xp = p[0], yp = p[1], zp = p[2];
    int n0 = nst, n1 = 0, n2 = 0, n3 = 0, n4 = 0, n1s = 0, n2s = 0, nf = 0, nl = 0, nxt = 0;
    int32_t lp = 0;
    double b1 = 0, b2 = 0, b3 = 0, s12 = 0, ptn1 = 0, ptn2 = 0, q[3];
    if (n0 < 1 || n < n0) n0 = restart_node(n);

  L2:
    lp = lend[n0];
    nl = list[lp];
    lp = lptr[lp];
    nf = list[lp];
    n1 = nf;
    if (0 < nl) {
    L3:
      if (det(n0, n1, xp, yp, zp) < 0.0) {
        lp = lptr[lp];
        n1 = list[lp];
        if (n1 == nl) goto L6;
        goto L3;
      }
    } else {
      nl = -nl;
      if (det(n0, nf, xp, yp, zp) < 0.0) {
        n1 = n0;
        n2 = nf;
        goto L9;
      }
      if (det(nl, n0, xp, yp, zp) < 0.0) {
        n1 = nl;
        n2 = n0;
        goto L9;
      }
    }
  L4:
    lp = lptr[lp];
    n2 = std::abs(list[lp]);
    if (det(n0, n2, xp, yp, zp) < 0.0) goto L7;
    n1 = n2;
    if (n1 != nl) goto L4;
    if (det(n0, nf, xp, yp, zp) < 0.0) goto L6;
    {
      volatile double c = std::fabs(x[n0] * xp + y[n0] * yp + z[n0] * zp);  // STORE: no extra precision
      if (c < 1.0 - 4.0 * eps) {
        for (;;) {
          if (det(n1, n0, xp, yp, zp) < 0.0) break;
          lp = lptr[lp];
          n1 = std::abs(list[lp]);
          if (n1 == nl) {
            i1 = i2 = i3 = 0;
            return;
          }
        }
      }
    }
    n0 = n1;
    goto L2;
  L6:
    n2 = nf;
  L7:
    n3 = n0;
    n1s = n1;
    n2s = n2;
  L8:
    b3 = det(n1, n2, xp, yp, zp);
    if (b3 < 0.0) {
      lp = find_arc(lend[n2], n1);
      if (list[lp] < 0) goto L9;
      lp = lptr[lp];
      n4 = std::abs(list[lp]);
      if (det(n0, n4, xp, yp, zp) < 0.0) {
        n3 = n2;
        n2 = n4;
        n1s = n1;
        if (n2 != n2s && n2 != n0) goto L8;
      } else {
        n3 = n1;
        n1 = n4;
        n2s = n2;
        if (n1 != n1s && n1 != n0) goto L8;
      }
      n0 = restart_node(n);
      goto L2;
    }
    if (b3 >= eps) {
      b1 = det(n2, n3, xp, yp, zp);
      b2 = det(n3, n1, xp, yp, zp);
      if (b1 < -tol || b2 < -tol) {
        n0 = restart_node(n);
        goto L2;
      }
    } else {
      s12 = x[n1] * x[n2] + y[n1] * y[n2] + z[n1] * z[n2];
      ptn1 = xp * x[n1] + yp * y[n1] + zp * z[n1];
      ptn2 = xp * x[n2] + yp * y[n2] + zp * z[n2];
      b1 = ptn1 - s12 * ptn2;
      b2 = ptn2 - s12 * ptn1;
      if (b1 < -tol || b2 < -tol) {
        n0 = restart_node(n);
        goto L2;
      }
    }
    i1 = n1;
    i2 = n2;
    i3 = n3;
    return;
  L9:
    n1s = n1;
    n2s = n2;
    nl = 0;
  L10:
    lp = lend[n2];
    lp = lptr[lp];
    nxt = list[lp];
    if (det(n2, nxt, xp, yp, zp) >= 0.0) {
      s12 = x[n1] * x[n2] + y[n1] * y[n2] + z[n1] * z[n2];
      q[0] = x[n1] - s12 * x[n2];
      q[1] = y[n1] - s12 * y[n2];
      q[2] = z[n1] - s12 * z[n2];
      if (xp * q[0] + yp * q[1] + zp * q[2] >= 0.0) goto L11;
      if (x[nxt] * q[0] + y[nxt] * q[1] + z[nxt] * q[2] >= 0.0) goto L11;
      nl = n2;
    }
    n1 = n2;
    n2 = nxt;
    if (n2 != n1s) goto L10;
    i1 = n1s;
    i2 = n1s;
    i3 = 0;
    return;
  L11:
    nf = n2;
    if (nl == 0) {
      n2 = n2s;
      n1 = n1s;
    L12:
      lp = lend[n1];
      nxt = -list[lp];
      if (0.0 <= det(nxt, n1, xp, yp, zp)) {
        s12 = x[n1] * x[n2] + y[n1] * y[n2] + z[n1] * z[n2];
        q[0] = x[n2] - s12 * x[n1];
        q[1] = y[n2] - s12 * y[n1];
        q[2] = z[n2] - s12 * z[n1];
        if (xp * q[0] + yp * q[1] + zp * q[2] >= 0.0) goto L13;
        if (x[nxt] * q[0] + y[nxt] * q[1] + z[nxt] * q[2] >= 0.0) goto L13;
        nf = n1;
      }
      n2 = n1;
      n1 = nxt;
      if (n1 != n1s) goto L12;
      i1 = n1;
      i2 = n1;
      i3 = 0;
      return;
    L13:
      nl = n1;
    }
    i1 = nf;
    i2 = nl;
    i3 = 0;
  }

  // insert node k (nodes 1..k-1 triangulated), search from nst
  bool add_node(int nst, int k, std::string &err) {
    const int km1 = k - 1;
    const int ist = nst < 1 ? km1 : nst;
    const double p[3] = {x[k], y[k], z[k]};
    int i1, i2, i3;
    locate(ist, p, km1, i1, i2, i3);
    if (i1 == 0) {
      err = "trmesh: the nodes are coplanar";
      return false;
    }
    if (i3 != 0) {
      for (int l : {i1, i2, i3})
        if (p[0] == x[l] && p[1] == y[l] && p[2] == z[l]) {
          err = "trmesh: node " + std::to_string(l) + " equals node " + std::to_string(k);
          return false;
        }
      add_interior(k, i1, i2, i3);
    } else if (i1 != i2) {
      add_boundary(k, i1, i2);
    } else {
      add_cover(k, i1);
    }
    // restore the Delaunay property around k by swapping the arcs opposite
    // to it, walking k's ring
    const int32_t lpf = lptr[lend[k]];
    int io2 = list[lpf];
    int32_t lpo1 = lptr[lpf];
    int io1 = std::abs(list[lpo1]);
    for (;;) {
      int32_t lp = find_arc(lend[io1], io2);
      if (0 <= list[lp]) {
        lp = lptr[lp];
        const int in1 = std::abs(list[lp]);
        const int32_t lpo1s = lpo1;
        if (!swap_test(in1, k, io1, io2)) {
          if (lpo1 == lpf || list[lpo1] < 0) break;
          io2 = io1;
          lpo1 = lptr[lpo1];
          io1 = std::abs(list[lpo1]);
          continue;
        }
        lpo1 = swap_arc(in1, k, io1, io2);
        if (lpo1 != 0) {
          io1 = in1;
          continue;
        }
        lpo1 = lpo1s;
      }
      if (lpo1 == lpf || list[lpo1] < 0) break;
      io2 = io1;
      lpo1 = lptr[lpo1];
      io1 = std::abs(list[lpo1]);
    }
    return true;
  }
};

inline double neg_dot(const double *x, const double *y, const double *z, int a, int b) {
  return -(x[a] * x[b] + y[a] * y[b] + z[a] * z[b]);
}

}  // namespace

bool trmesh_rings(int n, const double *px, const double *py, const double *pz, std::vector<int32_t> &nnb,
                  std::vector<int32_t> &nb, int maxnb, int64_t *restarts, std::string &err) {
  if (n < 4) {
    err = "trmesh: fewer than 4 nodes";
    return false;
  }
  Delaunay T;
  T.x = px - 1;
  T.y = py - 1;
  T.z = pz - 1;
  const double *x = T.x, *y = T.y, *z = T.z;
  const size_t arena = 6 * (size_t)(n - 2) + 1;
  T.list.assign(arena, 0);
  T.lptr.assign(arena, 0);
  T.lend.assign((size_t)n + 1, 0);
  // node v's two-arc ring in slots s, s+1: a, then boundary arc -b
  auto ring2 = [&](int v, int32_t s, int a, int b) {
    T.list[s] = a;
    T.lptr[s] = s + 1;
    T.list[s + 1] = -b;
    T.lptr[s + 1] = s;
    T.lend[v] = s + 1;
  };
  if (!(T.det(1, 2, x[3], y[3], z[3]) >= 0.0)) {  // 3 right of 1->2
    ring2(1, 1, 3, 2);
    ring2(2, 3, 1, 3);
    ring2(3, 5, 2, 1);
  } else if (!(T.det(2, 1, x[3], y[3], z[3]) >= 0.0)) {
    ring2(1, 1, 2, 3);
    ring2(2, 3, 3, 1);
    ring2(3, 5, 1, 2);
  } else {
    err = "trmesh: the first three nodes are collinear";
    return false;
  }
  T.lnew = 7;
  // near[k]: for an uninserted k the inserted node nearest to it; for an
  // inserted node the head of the list (linked by next) of uninserted nodes
  // it is nearest to. dist = -cos(angle).
  std::vector<int32_t> near((size_t)n + 1, 0), next((size_t)n + 1, 0);
  std::vector<double> dist((size_t)n + 1, 0.0);
  for (int k = n; k >= 4; --k) {
    const double d1 = neg_dot(x, y, z, k, 1), d2 = neg_dot(x, y, z, k, 2), d3 = neg_dot(x, y, z, k, 3);
    int c;
    double d;
    if (d1 <= d2 && d1 <= d3) {
      c = 1;
      d = d1;
    } else if (d2 <= d1 && d2 <= d3) {
      c = 2;
      d = d2;
    } else {
      c = 3;
      d = d3;
    }
    near[k] = c;
    dist[k] = d;
    next[k] = near[c];
    near[c] = k;
  }
  int i0 = 1;
  for (int k = 4; k <= n; ++k) {
    if (!T.add_node(near[k], k, err)) return false;
    // unlink k from the list of its nearest node
    int i = near[k];
    if (near[i] == k) {
      near[i] = next[k];
    } else {
      i = near[i];
      for (;;) {
        i0 = i;
        i = next[i0];
        if (i == k) break;
      }
      next[i0] = next[k];
    }
    near[k] = 0;
    // uninserted nodes now nearer to k than to k's neighbours move to k
    const int32_t lpl = T.lend[k];
    int32_t lp = lpl;
    do {
      lp = T.lptr[lp];
      const int j = std::abs(T.list[lp]);
      i = near[j];
      while (i != 0) {
        const int nexti = next[i];
        const double d = neg_dot(x, y, z, i, k);
        if (d < dist[i]) {
          near[i] = k;
          dist[i] = d;
          if (i == near[j]) near[j] = nexti;
          else next[i0] = nexti;
          next[i] = near[k];
          near[k] = i;
        } else {
          i0 = i;
        }
        i = nexti;
      }
    } while (lp != lpl);
  }
  // neighbour rings in the order the mesh structuring reads them: from the
  // arc after the last one, CCW
  nnb.assign(n, 0);
  nb.assign((size_t)n * maxnb, 0);
  for (int v = 1; v <= n; ++v) {
    const int32_t lpl = T.lend[v];
    int32_t lp = T.lptr[lpl];
    int c = 0;
    for (;;) {
      if (c >= maxnb) {
        err = "trmesh: node " + std::to_string(v) + " has more than " + std::to_string(maxnb) + " neighbours";
        return false;
      }
      nb[(size_t)(v - 1) * maxnb + c++] = std::abs(T.list[lp]) - 1;
      if (lp == lpl) break;
      lp = T.lptr[lp];
    }
    nnb[v - 1] = c;
  }
  if (restarts) *restarts = T.restarts;
  return true;
}

}  // namespace imx
