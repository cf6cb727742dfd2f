// Element order for locality, host side.
//
// The kernels process elements in blocks of 16 consecutive indices and gather
// neighbours' face traces through the connectivity, so they run fastest when
// consecutive elements are close in space.  A caller's numbering need not be
// (a structured mesh numbered row by row, or the order the adaptive driver's
// pointer_mesh leaves after remeshing, src/cmal_s.f:164-208), so the library
// renumbers elements internally along a space-filling curve of their
// centroids (Hilbert or Z-order), separately inside each ensemble member.
//
// The centroids come from the mesh operands alone: vect holds each element's
// half edge vectors from its first node, (ax, ay) to the second node and
// (bx, by) to the third, so an element's nodes, edge midpoints and centroid
// are known relative to its first node; two elements sharing face f meet at
// the same edge midpoint, which fixes their relative position.  A breadth-
// first walk over the faces places every element of a connected component.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/spoc_abi.h"
#include "error.h"

namespace spoc {

namespace {

// midpoint of local edge k (0-based; edge k joins local nodes k and k+1 mod
// 3) minus the centroid, both relative to the first node
void mid_minus_centroid(const double *v, int k, double &dx, double &dy) {
  const double ax = v[0], bx = v[1], ay = v[2], by = v[3];
  const double cx = 2.0 * (ax + bx) / 3.0, cy = 2.0 * (ay + by) / 3.0;
  double mx, my;
  if (k == 0) {
    mx = ax;
    my = ay;
  } else if (k == 1) {
    mx = ax + bx;
    my = ay + by;
  } else {
    mx = bx;
    my = by;
  }
  dx = mx - cx;
  dy = my - cy;
}

constexpr int kBits = 20;  // the curves' grid: 2^20 cells a side

// distance along the Hilbert curve of cell (x, y) of the 2^kBits grid
uint64_t hilbert(uint32_t x, uint32_t y) {
  const uint32_t n = 1u << kBits;
  uint64_t d = 0;
  for (uint32_t s = n >> 1; s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {  // rotate the quadrant
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

uint64_t spread(uint32_t v) {  // bits of v to the even bit positions
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

}  // namespace

// perm[new] = old (0-based) for elements [mem_e0[m], mem_e0[m+1]) of each
// member m, along the Hilbert curve (hilbert) or the Z-order curve; the
// connectivity is the validated 1-based arrays of mesh_set.
void locality_order(int ne, int nf, const int *itm, const int *tseg, const int *ssit,
                    const double *vect, const std::vector<int> &mem_e0, bool use_hilbert,
                    std::vector<int> &perm) {
  std::vector<double> cx(ne, 0.0), cy(ne, 0.0);
  std::vector<char> seen(ne, 0);
  std::vector<int> queue;
  queue.reserve(ne);
  double off = 0.0;  // components of one member side by side
  for (int root = 0; root < ne; root++) {
    if (seen[root]) continue;
    seen[root] = 1;
    cx[root] = off;
    cy[root] = 0.0;
    size_t head = queue.size();
    queue.push_back(root);
    double xmax = off;
    while (head < queue.size()) {
      const int a = queue[head++];
      for (int k = 0; k < 3; k++) {
        const int f = itm[3 * a + k] - 1;
        const int s1 = tseg[2 * f] - 1, s2 = tseg[2 * f + 1] - 1;
        if (s2 < 0) continue;  // wall
        const int b = s1 == a ? s2 : s1;
        if (b < 0 || b >= ne || seen[b]) continue;
        const int kb = (s1 == a ? ssit[2 * f + 1] : ssit[2 * f]) - 1;
        if (kb < 0 || kb > 2) continue;
        double da_x, da_y, db_x, db_y;
        mid_minus_centroid(vect + 5 * (size_t)a, k, da_x, da_y);
        mid_minus_centroid(vect + 5 * (size_t)b, kb, db_x, db_y);
        cx[b] = cx[a] + da_x - db_x;
        cy[b] = cy[a] + da_y - db_y;
        xmax = std::max(xmax, cx[b]);
        seen[b] = 1;
        queue.push_back(b);
      }
    }
    off = xmax + 1.0;
  }
  (void)nf;
  perm.resize(ne);
  std::vector<uint64_t> code(ne);
  for (size_t m = 0; m + 1 < mem_e0.size(); m++) {
    const int e0 = mem_e0[m], e1 = mem_e0[m + 1];
    if (e1 <= e0) continue;
    double x0 = cx[e0], x1 = cx[e0], y0 = cy[e0], y1 = cy[e0];
    for (int e = e0; e < e1; e++) {
      x0 = std::min(x0, cx[e]);
      x1 = std::max(x1, cx[e]);
      y0 = std::min(y0, cy[e]);
      y1 = std::max(y1, cy[e]);
    }
    // one square grid of 2^kBits cells a side over the member's extent
    const double span = std::max(x1 - x0, y1 - y0);
    const double scale = span > 0.0 ? (double)((1u << kBits) - 1) / span : 0.0;
    for (int e = e0; e < e1; e++) {
      const uint32_t ix = (uint32_t)std::lround((cx[e] - x0) * scale);
      const uint32_t iy = (uint32_t)std::lround((cy[e] - y0) * scale);
      code[e] = use_hilbert ? hilbert(ix, iy) : (spread(ix) | (spread(iy) << 1));
      perm[e] = e;
    }
    std::stable_sort(perm.begin() + e0, perm.begin() + e1,
                     [&](int a, int b) { return code[a] < code[b]; });
  }
}

// Share of the interior faces whose two elements lie in the same block of
// 16 consecutive internal elements (the kernels' unit of work); inv[e] is
// caller element e's internal position, or null for the caller's numbering.
double block_share(int nf, const int *tseg, const int *inv) {
  long inner = 0, same = 0;
  for (int f = 0; f < nf; f++) {
    const int a = tseg[2 * f] - 1, b = tseg[2 * f + 1] - 1;
    if (b < 0) continue;
    const int ia = inv ? inv[a] : a, ib = inv ? inv[b] : b;
    inner++;
    same += (ia >> 4) == (ib >> 4);
  }
  return inner ? (double)same / (double)inner : 1.0;
}

// Face order following the element order: faces numbered in order of first
// appearance over the renumbered elements' local edges.  fperm[new] = old.
void face_order(int ne, int nf, const int *itm, const std::vector<int> &perm,
                std::vector<int> &fperm) {
  fperm.clear();
  fperm.reserve(nf);
  std::vector<char> seen(nf, 0);
  for (int en = 0; en < ne; en++) {
    const int e = perm[en];
    for (int k = 0; k < 3; k++) {
      const int f = itm[3 * e + k] - 1;
      if (!seen[f]) {
        seen[f] = 1;
        fperm.push_back(f);
      }
    }
  }
  for (int f = 0; f < nf; f++)  // faces no element names (not expected)
    if (!seen[f]) fperm.push_back(f);
}

}  // namespace spoc

extern "C" int spoc_mesh_order(int mode, int nmem, const int *ne_m, const int *nf_m,
                               const int *itm, const int *tseg, const int *ssit,
                               const double *vect, int *perm) {
  return spoc::guarded([&]() -> int {
    if (nmem < 1 || !ne_m || !nf_m || !itm || !tseg || !ssit || !vect || !perm ||
        (mode != SPOC_ORDER_HILBERT && mode != SPOC_ORDER_ZCURVE))
      return spoc::report(SPOC_EINVAL, "spoc_mesh_order: bad arguments");
    long ne = 0, nf = 0;
    std::vector<int> e0(nmem + 1, 0);
    for (int m = 0; m < nmem; m++) {
      if (ne_m[m] <= 0 || nf_m[m] <= 0)
        return spoc::report(SPOC_EINVAL, "spoc_mesh_order: member %d: ne=%d nf=%d", m, ne_m[m],
                            nf_m[m]);
      ne += ne_m[m];
      nf += nf_m[m];
      e0[m + 1] = (int)ne;
      if (ne > (1L << 30) || nf > (1L << 30))
        return spoc::report(SPOC_EINVAL, "spoc_mesh_order: mesh too large");
    }
    // the members' own 1-based numbering, shifted to the union's
    std::vector<int> uitm(3 * ne), utseg(2 * nf);
    long eo = 0, fo = 0;
    for (int m = 0; m < nmem; m++) {
      for (long i = 3 * eo; i < 3 * (eo + ne_m[m]); i++) {
        if (itm[i] < 1 || itm[i] > nf_m[m])
          return spoc::report(SPOC_EINVAL, "spoc_mesh_order: itm out of range in member %d", m);
        uitm[i] = itm[i] + (int)fo;
      }
      for (long i = 2 * fo; i < 2 * (fo + nf_m[m]); i++) {
        if (tseg[i] < 0 || tseg[i] > ne_m[m] || (i % 2 == 0 && tseg[i] == 0))
          return spoc::report(SPOC_EINVAL, "spoc_mesh_order: tseg out of range in member %d", m);
        utseg[i] = tseg[i] == 0 ? 0 : tseg[i] + (int)eo;
      }
      eo += ne_m[m];
      fo += nf_m[m];
    }
    std::vector<int> p;
    spoc::locality_order((int)ne, (int)nf, uitm.data(), utseg.data(), ssit, vect, e0,
                         mode == SPOC_ORDER_HILBERT, p);
    for (long i = 0; i < ne; i++) perm[i] = p[i] + 1;
    return SPOC_OK;
  });
}
