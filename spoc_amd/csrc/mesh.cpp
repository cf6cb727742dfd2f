// Host-side mesh operands of cal_var, built as the reference driver builds
// them, so that every integer array (face numbering, orientation, local
// edges, wall flags) equals the reference's and every geometry word is the
// same IEEE expression.  Arrays are 1-based int32 and column-major like the
// Fortran ones: in(3, ne), itm(3, ne), tseg(2, nf), ...
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <unordered_map>
#include <vector>

#include "../../include/spoc_abi.h"
#include "error.h"

namespace {

// Faces in order of first encounter over (element, local edge)
// (src/cmal_s.f:164-208).  Local edge k of element i runs from node
// in(k, i) to in(k+1, i) (k = 3 wraps to 1).  An edge a -> b joins the face
// of the lowest-numbered earlier element that holds b -> a; the reference
// finds it by scanning b's incident elements in increasing order, stopping
// at element i itself (b is one of its nodes), which is the first element
// holding b -> a.  Otherwise the edge opens face nf + 1 with first element
// i and node pair (a, b).
int pointer_mesh(int ne, const int *in, int *itm, int *tseg, int *pseg) {
  std::unordered_map<uint64_t, int> first;  // directed edge -> face of its first element
  first.reserve(3 * (size_t)ne * 2);
  auto key = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
  int nf = 0;
  for (int i = 0; i < ne; i++)
    for (int k = 0; k < 3; k++) {
      const int a = in[3 * i + k], b = in[3 * i + (k + 1) % 3];
      auto it = first.find(key(b, a));
      if (it != first.end()) {
        const int f = it->second;
        tseg[2 * (f - 1) + 1] = i + 1;
        itm[3 * i + k] = f;
      } else {
        nf++;
        itm[3 * i + k] = nf;
        tseg[2 * (nf - 1)] = i + 1;
        tseg[2 * (nf - 1) + 1] = 0;
        pseg[2 * (nf - 1)] = a;
        pseg[2 * (nf - 1) + 1] = b;
      }
      first.emplace(key(a, b), itm[3 * i + k]);
    }
  return nf;
}

// Orientation signs and local edge numbers of every face (src/mat1.f:28-64).
int cal_its(int ne, int nf, const int *itm, const int *tseg, int *its, int *ssit) {
  for (size_t t = 0; t < 3 * (size_t)ne; t++) its[t] = 0;
  for (int f = 1; f <= nf; f++) {
    int tk = tseg[2 * (f - 1)], j = 1;
    while (j < 4 && itm[3 * (tk - 1) + j - 1] != f) j++;
    if (j == 4) return f;
    its[3 * (tk - 1) + j - 1] = 1;
    ssit[2 * (f - 1)] = j;
    ssit[2 * (f - 1) + 1] = 0;
    tk = tseg[2 * (f - 1) + 1];
    if (tk > 0) {
      j = 1;
      while (j < 4 && itm[3 * (tk - 1) + j - 1] != f) j++;
      if (j == 4) return f;
      its[3 * (tk - 1) + j - 1] = -1;
      ssit[2 * (f - 1) + 1] = j;
    }
  }
  return 0;
}

int cit(int k, int np) {  // src/chmal.f:146-154
  int k1 = k + np;
  if (k1 > 3) k1 -= 3;
  if (k1 < 1) k1 += 3;
  return k1;
}

}  // namespace

extern "C" {

int spoc_mesh_build(int nn, int ne, const int *in, const double *x, const double *y, int nc,
                    double f0, double beta, int ld, int check, int *nf_out, int *itm, int *its,
                    int *tseg, int *pseg, int *ssit, double *norm, double *vect, double *normt,
                    int *bdl, double *fc, double *deta) {
  return spoc::guarded([&]() -> int {
    if (nn <= 0 || ne <= 0 || !in || !x || !y || !nf_out || !itm || !its || !tseg || !pseg ||
        !ssit || !norm || !vect || !normt || !bdl || nc < 0 || (fc && ld < (nc + 1) * (nc + 2) / 2))
      return spoc::report(SPOC_EINVAL, "spoc_mesh_build: bad arguments");
    for (size_t t = 0; t < 3 * (size_t)ne; t++)
      if (in[t] < 1 || in[t] > nn)
        return spoc::report(SPOC_EINVAL, "spoc_mesh_build: in(%d,%d)=%d not a node", (int)(t % 3) + 1,
                            (int)(t / 3) + 1, in[t]);
    const int nf = pointer_mesh(ne, in, itm, tseg, pseg);
    *nf_out = nf;
    if (check) {  // src/cmal_s.f:222-277: at most one wall face, and it is edge 1
      std::vector<int> cnt(ne + 1, 0);
      for (int f = 1; f <= nf; f++) {
        if (tseg[2 * (f - 1) + 1] != 0) continue;
        const int tk = tseg[2 * (f - 1)];
        if (++cnt[tk] > 1)
          return spoc::report(SPOC_EMESH, "triangle %d has more than one face on the wall", tk);
        if (itm[3 * (tk - 1)] != f)
          return spoc::report(SPOC_EMESH, "triangle %d: its wall face is not its edge 1", tk);
      }
    }
    if (const int bad = cal_its(ne, nf, itm, tseg, its, ssit))
      return spoc::report(SPOC_EMESH, "face %d is not an edge of its elements", bad);
    for (int f = 0; f < nf; f++) {  // src/mat1.f:7-23
      const int p1 = pseg[2 * f] - 1, p2 = pseg[2 * f + 1] - 1;
      norm[2 * f] = y[p2] - y[p1];
      norm[2 * f + 1] = -(x[p2] - x[p1]);
    }
    for (int e = 0; e < ne; e++) {  // src/shwwnl.f:384-397, src/mat1.f:69-83
      const int i1 = in[3 * e] - 1, i2 = in[3 * e + 1] - 1, i3 = in[3 * e + 2] - 1;
      const double ax = (x[i2] - x[i1]) * 0.5, bx = (x[i3] - x[i1]) * 0.5;
      const double ay = (y[i2] - y[i1]) * 0.5, by = (y[i3] - y[i1]) * 0.5;
      const double d = ax * by - ay * bx;
      vect[5 * e] = ax;
      vect[5 * e + 1] = bx;
      vect[5 * e + 2] = ay;
      vect[5 * e + 3] = by;
      vect[5 * e + 4] = 1.0 / d;
      if (deta) deta[e] = d;
      const int s1 = itm[3 * e] - 1;
      const double dnx = norm[2 * s1], dny = norm[2 * s1 + 1];
      const double dis = std::sqrt(dnx * dnx + dny * dny);
      normt[2 * e] = dnx / dis;
      normt[2 * e + 1] = dny / dis;
      bdl[e] = 0;
      if (fc) {
        double *c = fc + (size_t)ld * e;
        for (int m = 0; m < ld; m++) c[m] = 0.0;
        c[0] = f0 + 0.5 * beta * (y[i3] + y[i2]);
        c[1] = 0.5 * beta * (y[i3] - y[i1]);
        c[nc + 1] = 0.5 * beta * (y[i2] - y[i1]);
      }
    }
    for (int f = 0; f < nf; f++)
      if (tseg[2 * f + 1] == 0) bdl[tseg[2 * f] - 1] = 1;
    return SPOC_OK;
  });
}

int spoc_mesh_changemaille(int nn, int ne, int *in) {
  return spoc::guarded([&]() -> int {
    if (nn <= 0 || ne <= 0 || !in)
      return spoc::report(SPOC_EINVAL, "spoc_mesh_changemaille: bad arguments");
    for (size_t t = 0; t < 3 * (size_t)ne; t++)
      if (in[t] < 1 || in[t] > nn)
        return spoc::report(SPOC_EINVAL, "spoc_mesh_changemaille: in(%d,%d)=%d not a node",
                            (int)(t % 3) + 1, (int)(t / 3) + 1, in[t]);
    std::vector<int> itm(3 * (size_t)ne), tseg(6 * (size_t)ne), pseg(6 * (size_t)ne);
    const int nf = pointer_mesh(ne, in, itm.data(), tseg.data(), pseg.data());
    auto I = [&](int k, int e) -> int & { return in[3 * (e - 1) + k - 1]; };
    auto M = [&](int k, int e) -> int & { return itm[3 * (e - 1) + k - 1]; };
    auto T = [&](int s, int f) -> int & { return tseg[2 * (f - 1) + s - 1]; };
    std::vector<int> sbd, pebd(ne + 1, 0);
    for (int f = 1; f <= nf; f++)
      if (T(2, f) == 0) {
        sbd.push_back(f);
        pebd[T(1, f)]++;
      }
    // corner triangles: flip the diagonal shared with their interior
    // neighbour (src/chmal.f:70-118)
    for (int i = 1; i <= ne; i++) {
      if (pebd[i] <= 1) continue;
      int j = 1, ns1 = M(j, i);
      while (T(2, ns1) == 0) {
        if (++j > 3)  // every face on the wall
          return spoc::report(SPOC_EMESH, "triangle %d has all its faces on the wall", i);
        ns1 = M(j, i);
      }
      const int tk1 = T(1, ns1), tk2 = T(2, ns1);
      int j1 = 1, j2 = 1;
      while (M(j1, tk1) != ns1) j1++;
      while (M(j2, tk2) != ns1) j2++;
      const int k1 = cit(j1, 1), k2 = cit(j2, -1), k3 = cit(j2, 1), k4 = cit(j1, -1);
      const int i2 = I(k2, tk2), i4 = I(k4, tk1);
      I(k1, tk1) = i2;
      I(k3, tk2) = i4;
      const int mk3 = M(k3, tk2), mk1 = M(k1, tk1);
      M(j1, tk1) = mk3;
      M(j2, tk2) = mk1;
      M(k1, tk1) = ns1;
      M(k3, tk2) = ns1;
      pseg[2 * (ns1 - 1)] = i2;
      pseg[2 * (ns1 - 1) + 1] = i4;
      if (i == tk1) T(1, M(j2, tk2)) = tk2;
      else T(1, M(j1, tk1)) = tk1;
    }
    // rotate each wall triangle so that its wall face is its edge 1
    // (src/chmal.f:120-141)
    for (int ns1 : sbd) {
      const int tk1 = T(1, ns1);
      int j = 1;
      while (M(j, tk1) != ns1) j++;
      int pp[3], pf[3];
      for (int k = 1; k <= 3; k++) {
        pp[k - 1] = I(cit(k, j - 1), tk1);
        pf[k - 1] = M(cit(k, j - 1), tk1);
      }
      for (int k = 1; k <= 3; k++) {
        I(k, tk1) = pp[k - 1];
        M(k, tk1) = pf[k - 1];
      }
    }
    return SPOC_OK;
  });
}

}  // extern "C"
