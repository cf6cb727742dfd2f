// Mesh refinement and coarsening of the adaptive driver, host side: the
// refinement hierarchy of src/adapt.f (COMMON /adap_*/), built by init_adap
// (src/adapt.f:380-449) and changed once per adaptation by the sequence of
// global_refinment (src/adapt.f:89-119): the per-element flags mapped onto
// the hierarchy, deraffiner (1393-1583), adapmaille (532-788) with
// check_level (1350-1392), protridouble (789-998) and elimpptri (999-1121);
// plus inter_only_new (1233-1258).
//
// The hierarchy keeps every red triangle ever made (refined ones stay, with
// their four children in tenfan); the triangles of the mesh are its leaves
// (tutili) plus the green halves that protridouble cuts from a leaf next to
// a finer neighbour, which are made again at every adaptation.  Elements,
// nodes and neighbours are numbered from 1 as in the reference, with 0 for
// "none" (a wall face, no children, no parent), and every array keeps an
// unused slot 0 so the code indexes the way the routines do.  The node
// coordinates of new midpoints are 0.5 (x_a + x_b) with the reference's
// operand order, so the meshes are equal to the reference's bit for bit.
//
// Where the reference reads outside its arrays the restatement fixes a
// value (these reads only matter when triangles are removed):
//  - deraffiner compares tlevel and fenlev of tmt(j, i) = 0 (a wall face):
//    no neighbour, nothing to re-point;
//  - global_refinment hands deraffiner the per-element coarsening flags
//    fderaff (indexed there by hierarchy number, src/adapt.f:104,117) of
//    which only the first ne are set: the others count as false;
//  - deraffiner renumbers tabsol entries of green halves past its table:
//    they are made 0 (elimpptri rewrites tabsol before any use).
#include <algorithm>
#include <cmath>
#include <set>
#include <memory>
#include <unordered_map>
#include <vector>

#include "../../include/spoc_abi.h"
#include "error.h"

struct spoc_amr {
  int nn_ab = 0, ne_ab = 0, ndoub = 0;
  int ne = 0;  // triangles of the current mesh (entries of tabsol)
  // per hierarchy triangle (slot 0 unused)
  std::vector<int> in_ab, tmt, tenfan;  // 3, 3, 4 per triangle
  std::vector<int> tlevel, torigi;
  std::vector<char> tutili, tdoubl;
  std::vector<int> tabsol;  // mesh triangle -> hierarchy triangle
  std::vector<double> x, y;  // hierarchy nodes (slot 0 unused)

  int &IN(int j, int e) { return in_ab[3 * e + j - 1]; }
  int &TMT(int j, int e) { return tmt[3 * e + j - 1]; }
  int &TEN(int j, int e) { return tenfan[4 * e + j - 1]; }
  // room for triangles 1..n and nodes 1..m
  void reserve_tri(int n) {
    if ((int)tlevel.size() > n) return;
    const size_t s = (size_t)n + 1;
    in_ab.resize(3 * s, 0);
    tmt.resize(3 * s, 0);
    tenfan.resize(4 * s, 0);
    tlevel.resize(s, 0);
    torigi.resize(s, 0);
    tutili.resize(s, 0);
    tdoubl.resize(s, 0);
  }
  void reserve_node(int m) {
    if ((int)x.size() > m) return;
    x.resize((size_t)m + 1, 0.0);
    y.resize((size_t)m + 1, 0.0);
  }
};

namespace {

// the edge after edge k of a triangle: cit(3, k, 1, k1) (src/adapt.f:1336-1345)
inline int next3(int k) { return k == 3 ? 1 : k + 1; }

// check_level: a flagged triangle flags its coarser neighbours, transitively
// (src/adapt.f:1350-1392).
void check_level(spoc_amr &A, int n, std::vector<char> &fconv) {
  std::vector<int> pe;
  for (int i = 1; i <= n; i++)
    if (fconv[i]) pe.push_back(i);
  for (size_t ind = 0; ind < pe.size(); ind++) {
    const int i = pe[ind];
    for (int j = 1; j <= 3; j++) {
      const int tk = A.TMT(j, i);
      if (tk > 0 && !fconv[tk] && A.tlevel[tk] < A.tlevel[i]) {
        fconv[tk] = 1;
        pe.push_back(tk);
      }
    }
  }
}

// The reference's exchange sort of the refinement list by level
// (src/adapt.f:581-589: for i < j, swap a(i) and a(j) when level(a(i)) >
// level(a(j))), whose order among equal levels fixes the new numbering.
// Pass i swaps at each later position that holds a level below the one
// then at i; those positions are found from per-level sets of positions,
// so the passes cost O(levels log n) each instead of O(n), with the same
// result.
void exchange_sort(std::vector<int> &a, const std::vector<int> &level) {
  const int n = (int)a.size();
  int lmax = 0;
  for (int e : a) lmax = std::max(lmax, level[e]);
  std::vector<std::set<int>> at((size_t)lmax + 1);
  for (int i = 0; i < n; i++) at[level[a[i]]].insert(i);
  for (int i = 0; i < n - 1; i++) {
    int cur = level[a[i]], p = i;
    for (;;) {
      int q = n;
      for (int l = 0; l < cur; l++) {
        auto it = at[l].upper_bound(p);
        if (it != at[l].end() && *it < q) q = *it;
      }
      if (q == n) break;
      const int lq = level[a[q]];
      at[cur].erase(i);
      at[lq].erase(q);
      std::swap(a[i], a[q]);
      at[lq].insert(i);
      at[cur].insert(q);
      cur = lq;
      p = q;
    }
    at[cur].erase(i);
  }
}

// the face of triangle t that looks at triangle s (the reference's search
// `do while (tmt(j, t) .ne. s)`); 0 if none
int face_towards(spoc_amr &A, int t, int s) {
  for (int j = 1; j <= 3; j++)
    if (A.TMT(j, t) == s) return j;
  return 0;
}

int add_mid(spoc_amr &A, int &nnnew, int a, int b) {
  nnnew++;
  A.reserve_node(nnnew);
  A.x[nnnew] = 0.5 * (A.x[a] + A.x[b]);
  A.y[nnnew] = 0.5 * (A.y[a] + A.y[b]);
  return nnnew;
}

// deraffiner (src/adapt.f:1393-1583): the four children of a triangle go
// when all four are flagged; same-level neighbours of a removed child point
// to its parent; the hierarchy and its nodes are renumbered without the
// removed ones.
void deraffiner(spoc_amr &A, std::vector<char> &fconv_ab, std::vector<char> &fder) {
  const int ne_ab = A.ne_ab;
  for (int i = 1; i <= ne_ab; i++)
    if (fder[i] && fconv_ab[i]) fder[i] = 0;
  std::vector<char> fenlev((size_t)ne_ab + 1, 0);
  for (int i = 1; i <= ne_ab; i++) {
    if (A.TEN(1, i) <= 0) continue;
    int k = 0;
    for (int j = 1; j <= 4; j++)
      if (fder[A.TEN(j, i)]) k++;
    if (k == 4) {
      for (int j = 1; j <= 4; j++) fenlev[A.TEN(j, i)] = 1;
      A.tutili[i] = 1;
      for (int j = 1; j <= 4; j++) A.TEN(j, i) = 0;
    }
  }
  for (int i = 1; i <= ne_ab; i++) {
    if (fenlev[i]) continue;
    for (int j = 1; j <= 3; j++) {
      const int tk = A.TMT(j, i);
      if (tk > 0 && A.tlevel[tk] == A.tlevel[i] && fenlev[tk]) A.TMT(j, i) = A.torigi[tk];
    }
  }
  std::vector<int> peremp((size_t)ne_ab + 1, 0);
  int tk = 0;
  for (int i = 1; i <= ne_ab; i++) peremp[i] = fenlev[i] ? 0 : ++tk;
  const int nemax = tk;
  auto map = [&](int t) { return (t >= 0 && t <= ne_ab) ? peremp[t] : 0; };
  tk = 0;
  for (int i = 1; i <= ne_ab; i++) {
    if (fenlev[i]) continue;
    tk++;
    for (int j = 1; j <= 3; j++) A.IN(j, tk) = A.IN(j, i);
    for (int j = 1; j <= 3; j++) A.TMT(j, tk) = map(A.TMT(j, i));
    for (int j = 1; j <= 4; j++) A.TEN(j, tk) = map(A.TEN(j, i));
    A.torigi[tk] = map(A.torigi[i]);
    A.tutili[tk] = A.tutili[i];
    A.tlevel[tk] = A.tlevel[i];
    fconv_ab[tk] = fconv_ab[i];
  }
  for (int i = 1; i <= A.ne; i++) A.tabsol[i] = map(A.tabsol[i]);
  A.ne_ab = nemax;
  // nodes no remaining triangle uses
  std::vector<char> used((size_t)A.nn_ab + 1, 0);
  for (int i = 1; i <= A.ne_ab; i++)
    for (int j = 1; j <= 3; j++) used[A.IN(j, i)] = 1;
  std::vector<int> ppremp((size_t)A.nn_ab + 1, 0);
  int k = 0;
  for (int i = 1; i <= A.nn_ab; i++)
    if (used[i]) {
      ppremp[i] = ++k;
      A.x[k] = A.x[i];
      A.y[k] = A.y[i];
    }
  A.nn_ab = k;
  for (int i = 1; i <= A.ne_ab; i++)
    for (int j = 1; j <= 3; j++) A.IN(j, i) = ppremp[A.IN(j, i)];
}

// adapmaille (src/adapt.f:532-788): red refinement of the flagged
// triangles, coarsest first, each into four (corner children 1-3 on the
// parent's edges, child 4 inside), the children of two refined neighbours
// linked across their common edge; then every leaf with two refined
// neighbours is flagged, check_level, and again until nothing is flagged.
int adapmaille(spoc_amr &A, std::vector<char> &fconv_ab, int &nnnew, int &nemaxnew) {
  nnnew = A.nn_ab;
  nemaxnew = A.ne_ab;
  check_level(A, nemaxnew, fconv_ab);
  std::vector<int> peraff;
  for (int i = 1; i <= A.ne_ab; i++)
    if (fconv_ab[i]) peraff.push_back(i);
  for (int pass = 0; !peraff.empty(); pass++) {
    // a consistent hierarchy differs by at most one level across a face, so
    // the closure ends within as many passes as there are levels
    if (pass > 64)
      return spoc::report(SPOC_EMESH, "spoc_amr_adapt: the refinement closure does not end");
    const int n = (int)peraff.size();
    exchange_sort(peraff, A.tlevel);
    for (int ind = 0; ind < n; ind++) {
      const int i = peraff[ind];
      A.tutili[i] = 0;
      const int i1 = A.IN(1, i), i2 = A.IN(2, i), i3 = A.IN(3, i);
      const int i4 = add_mid(A, nnnew, i1, i2);
      const int i5 = add_mid(A, nnnew, i2, i3);
      const int i6 = add_mid(A, nnnew, i3, i1);
      A.reserve_tri(nemaxnew + 4);
      for (int j = 1; j <= 4; j++) A.TEN(j, i) = nemaxnew + j;
      for (int c = 1; c <= 4; c++)
        for (int j = 1; j <= 4; j++) A.TEN(j, nemaxnew + c) = 0;
      const int c1 = nemaxnew + 1, c2 = c1 + 1, c3 = c1 + 2, c4 = c1 + 3;
      const int lev = A.tlevel[i] + 1;
      const int corner[4][3] = {{i1, i4, i6}, {i4, i2, i5}, {i6, i5, i3}, {i5, i6, i4}};
      const int nbr[4][3] = {{A.TMT(1, i), c4, A.TMT(3, i)},
                             {A.TMT(1, i), A.TMT(2, i), c4},
                             {c4, A.TMT(2, i), A.TMT(3, i)},
                             {c3, c1, c2}};
      for (int c = 0; c < 4; c++) {
        const int t = c1 + c;
        A.tlevel[t] = lev;
        A.torigi[t] = i;
        A.tutili[t] = 1;
        for (int j = 1; j <= 3; j++) {
          A.IN(j, t) = corner[c][j - 1];
          A.TMT(j, t) = nbr[c][j - 1];
        }
      }
      nemaxnew += 4;
      // a refined neighbour: link the two children pairs along the edge
      const int tk2 = i;
      for (int j = 1; j <= 3; j++) {
        const int tk1 = A.TMT(j, i), j2 = j;
        if (tk1 > 0 && !A.tutili[tk1]) {
          const int j1 = face_towards(A, tk1, tk2);
          if (j1 == 0)  // (the reference's search would run off the face list)
            return spoc::report(SPOC_EMESH, "spoc_amr_adapt: triangle %d does not face %d", tk1, tk2);
          const int k2 = next3(j2), k1 = next3(j1);
          A.TMT(j1, A.TEN(j1, tk1)) = A.TEN(k2, tk2);
          A.TMT(j1, A.TEN(k1, tk1)) = A.TEN(j2, tk2);
          A.TMT(j2, A.TEN(j2, tk2)) = A.TEN(k1, tk1);
          A.TMT(j2, A.TEN(k2, tk2)) = A.TEN(j1, tk1);
        }
      }
    }
    fconv_ab.assign((size_t)nemaxnew + 1, 0);
    for (int i = 1; i <= nemaxnew; i++) {
      if (!A.tutili[i]) continue;
      int k = 0;
      for (int j = 1; j <= 3; j++) {
        const int tk = A.TMT(j, i);
        if (tk > 0 && !A.tutili[tk]) k++;
      }
      if (k >= 2) fconv_ab[i] = 1;
    }
    check_level(A, nemaxnew, fconv_ab);
    peraff.clear();
    for (int i = 1; i <= nemaxnew; i++)
      if (fconv_ab[i]) peraff.push_back(i);
  }
  return 0;
}

// protridouble (src/adapt.f:789-998): a leaf next to a finer leaf is cut in
// two at the midpoint of that edge; each of the two finer children along
// the edge makes one half (the midpoint is made twice and merged by
// elimpptri).
int protridouble(spoc_amr &A, int &nnnew, int nemaxnew) {
  for (int i = 1; i <= nemaxnew; i++) A.tdoubl[i] = 0;
  A.ndoub = 0;
  for (int i = 1; i <= nemaxnew; i++) {
    if (!A.tutili[i]) continue;
    for (int j = 1; j <= 3; j++) {
      const int tk1 = A.TMT(j, i);
      if (!(tk1 > 0 && A.tlevel[i] > A.tlevel[tk1])) continue;
      const int i1 = A.IN(1, tk1), i2 = A.IN(2, tk1), i3 = A.IN(3, tk1);
      const int tk2 = A.torigi[i];
      const int j1 = face_towards(A, tk1, tk2), j2 = face_towards(A, tk2, tk1);
      if (tk2 == 0 || j1 == 0 || j2 == 0)  // (the reference's searches would run off)
        return spoc::report(SPOC_EMESH, "spoc_amr_adapt: triangle %d and the parent %d of %d do "
                            "not face each other", tk1, tk2, i);
      const int k2 = next3(j2);
      const bool first = A.TEN(j2, tk2) == i;
      if (!first && A.TEN(k2, tk2) != i) continue;
      int a, b, tri[3];
      if (j1 == 1) a = i1, b = i2;
      else if (j1 == 2) a = i3, b = i2;
      else if (j1 == 3) a = i1, b = i3;
      else continue;
      const int i4 = add_mid(A, nnnew, a, b);
      if (first) {
        if (j1 == 1) tri[0] = i4, tri[1] = i2, tri[2] = i3;
        else if (j1 == 2) tri[0] = i4, tri[1] = i3, tri[2] = i1;
        else tri[0] = i1, tri[1] = i2, tri[2] = i4;
      } else {
        if (j1 == 1) tri[0] = i1, tri[1] = i4, tri[2] = i3;
        else if (j1 == 2) tri[0] = i1, tri[1] = i2, tri[2] = i4;
        else tri[0] = i4, tri[1] = i2, tri[2] = i3;
      }
      A.ndoub++;
      const int k = nemaxnew + A.ndoub;
      A.reserve_tri(k);
      A.tdoubl[k] = 1;
      A.torigi[k] = tk1;
      for (int q = 1; q <= 3; q++) A.IN(q, k) = tri[q - 1];
      A.tutili[tk1] = 0;
      A.tutili[k] = 1;
    }
  }
  return 0;
}

// elimpptri (src/adapt.f:999-1121): new nodes within 1e-3 (squared distance
// below 1e-6) of an earlier node are merged into it; nodes are renumbered;
// the mesh is the leaves and green halves in hierarchy order (tabsol).
// The reference compares each node with every later new node; here the
// candidates come from a grid of 1e-3 cells (a pair within 1e-3 lies at most
// two cells apart, allowing for the rounding of x / 1e-3), and the nodes are
// taken in the reference's order, which gives the same merges.
void elimpptri(spoc_amr &A, int &nnnew, int nemaxnew) {
  const int nn_ab = A.nn_ab;
  std::vector<char> ppelim((size_t)nnnew + 1, 0);
  std::vector<int> ppremp((size_t)nnnew + 1, 0);
  // new nodes (> nn_ab) by cell, each cell's list in increasing number
  const double h = 1e-3;
  struct Key {
    long long cx, cy;
    bool operator==(const Key &o) const { return cx == o.cx && cy == o.cy; }
  };
  struct KeyHash {
    size_t operator()(const Key &k) const {
      return std::hash<long long>()(k.cx * 0x9E3779B97F4A7C15LL ^ k.cy);
    }
  };
  auto cell = [&](int i) {
    return Key{(long long)std::floor(A.x[i] / h), (long long)std::floor(A.y[i] / h)};
  };
  std::unordered_map<Key, std::vector<int>, KeyHash> grid;
  for (int jn = nn_ab + 1; jn <= nnnew; jn++) grid[cell(jn)].push_back(jn);
  std::vector<int> cand;
  for (int i = 1; i <= nnnew - 1; i++) {
    if (ppelim[i]) continue;
    const int i2 = i > nn_ab ? i : nn_ab;
    const Key c = cell(i);
    cand.clear();
    for (long long dx = -2; dx <= 2; dx++)
      for (long long dy = -2; dy <= 2; dy++) {
        auto it = grid.find(Key{c.cx + dx, c.cy + dy});
        if (it == grid.end()) continue;
        for (int jn : it->second)
          if (jn > i2) cand.push_back(jn);
      }
    for (int jn : cand) {
      const double ddx = A.x[i] - A.x[jn], ddy = A.y[i] - A.y[jn];
      const double dd = ddx * ddx + ddy * ddy;
      if (dd < 1.0e-6) {
        ppelim[jn] = 1;
        ppremp[jn] = i;
      }
    }
  }
  int j = 0;
  for (int i = 1; i <= nnnew; i++) {
    if (!ppelim[i]) {
      ppremp[i] = ++j;
      A.x[j] = A.x[i];
      A.y[j] = A.y[i];
    } else {
      ppremp[i] = ppremp[ppremp[i]];
    }
  }
  nnnew = j;
  A.nn_ab = nnnew;
  for (int i = A.ne_ab + 1; i <= nemaxnew + A.ndoub; i++)
    for (int q = 1; q <= 3; q++) A.IN(q, i) = ppremp[A.IN(q, i)];
  A.ne_ab = nemaxnew;
  int nenew = 0;
  A.tabsol.assign(1, 0);
  for (int i = 1; i <= A.ne_ab + A.ndoub; i++)
    if (A.tutili[i]) {
      nenew++;
      A.tabsol.push_back(i);
    }
  A.ne = nenew;
}

}  // namespace

extern "C" {

int spoc_amr_create(spoc_amr **out, int nn, int ne, int nf, const double *xgr, const double *ygr,
                    const int *in, const int *tseg, const int *itm) {
  return spoc::guarded([&]() -> int {
    if (!out || nn < 3 || ne < 1 || nf < 1 || !xgr || !ygr || !in || !tseg || !itm)
      return spoc::report(SPOC_EINVAL, "spoc_amr_create: bad arguments");
    *out = nullptr;
    for (int i = 0; i < 3 * ne; i++)
      if (in[i] < 1 || in[i] > nn || itm[i] < 1 || itm[i] > nf)
        return spoc::report(SPOC_EINVAL, "spoc_amr_create: in or itm out of range at %d", i);
    for (int i = 0; i < 2 * nf; i++)
      if (tseg[i] < 0 || tseg[i] > ne || (i % 2 == 0 && tseg[i] == 0))
        return spoc::report(SPOC_EINVAL, "spoc_amr_create: tseg out of range at %d", i);
    std::unique_ptr<spoc_amr> A(new spoc_amr());  // (freed on every early return)
    A->reserve_tri(ne);
    A->reserve_node(nn);
    // init_adap (src/adapt.f:380-449)
    for (int i = 1; i <= ne; i++) {
      A->tlevel[i] = 0;
      A->torigi[i] = 0;
      A->tutili[i] = 1;
      A->tdoubl[i] = 0;
      for (int j = 1; j <= 3; j++) A->IN(j, i) = in[3 * (i - 1) + j - 1];
    }
    A->tabsol.assign((size_t)ne + 1, 0);
    for (int i = 1; i <= ne; i++) A->tabsol[i] = i;
    A->ne_ab = A->ne = ne;
    A->nn_ab = nn;
    for (int i = 1; i <= nn; i++) {
      A->x[i] = xgr[i - 1];
      A->y[i] = ygr[i - 1];
    }
    A->ndoub = 0;
    for (int i = 1; i <= ne; i++)
      for (int j = 1; j <= 3; j++) {
        const int f = itm[3 * (i - 1) + j - 1], s1 = tseg[2 * (f - 1)], s2 = tseg[2 * (f - 1) + 1];
        int t = -1;
        if (s1 == i) t = s2;
        if (s2 == i) t = s1;
        if (t < 0)
          return spoc::report(SPOC_EMESH, "spoc_amr_create: face %d of element %d does not list it",
                              f, i);
        A->TMT(j, i) = t;
      }
    *out = A.release();
    return 0;
  });
}

int spoc_amr_adapt(spoc_amr *A, int ne, const int *fconv, const int *fderaff, int *nn_new,
                   int *ne_new) {
  return spoc::guarded([&]() -> int {
    if (!A || ne != A->ne || !fconv || !fderaff || !nn_new || !ne_new)
      return spoc::report(SPOC_EINVAL, "spoc_amr_adapt: bad arguments (ne must be the mesh's %d)",
                          A ? A->ne : 0);
    // the flags onto the hierarchy (src/adapt.f:89-101); a green half's
    // parent is a leaf again
    std::vector<char> fconv_ab((size_t)A->ne_ab + 1, 0);
    for (int i = 1; i <= ne; i++) {
      const int tk = A->tabsol[i];
      int tk1 = tk;
      if (tk > A->ne_ab) {
        tk1 = A->torigi[tk];
        A->tutili[tk1] = 1;
      }
      if (fconv[i - 1]) fconv_ab[tk1] = 1;
    }
    // deraffiner reads the per-element flags by hierarchy number
    std::vector<char> fder((size_t)std::max(A->ne_ab, ne) + 1, 0);
    for (int i = 1; i <= ne; i++) fder[i] = fderaff[i - 1] != 0;
    // A family goes when its four children are flagged (and not flagged for
    // refinement).  If one of them has children of its own, the reference
    // removes it all the same and its children are left without a parent,
    // overlapping the parent that comes back: refused here, with the
    // hierarchy as it was.
    for (int i = 1; i <= A->ne_ab; i++) {
      if (A->TEN(1, i) <= 0) continue;
      int k = 0, inner = 0;
      for (int j = 1; j <= 4; j++) {
        const int c = A->TEN(j, i);
        if (fder[c] && !fconv_ab[c]) k++;
        if (A->TEN(1, c) > 0) inner++;
      }
      if (k == 4 && inner) {
        // (a green half's parent was made a leaf again above: undo that)
        for (int e = 1; e <= ne; e++)
          if (A->tabsol[e] > A->ne_ab) A->tutili[A->torigi[A->tabsol[e]]] = 0;
        return spoc::report(SPOC_EINVAL, "spoc_amr_adapt: the coarsening flags remove the children "
                            "of %d, of which one is refined", i);
      }
    }
    deraffiner(*A, fconv_ab, fder);
    int nnnew = 0, nemaxnew = 0, rc;
    // (after an error here the hierarchy is half-changed: destroy it)
    if ((rc = adapmaille(*A, fconv_ab, nnnew, nemaxnew))) return rc;
    if ((rc = protridouble(*A, nnnew, nemaxnew))) return rc;
    elimpptri(*A, nnnew, nemaxnew);
    *nn_new = A->nn_ab;
    *ne_new = A->ne;
    return 0;
  });
}

int spoc_amr_mesh(const spoc_amr *A, double *xgr, double *ygr, int *in, int *tabsol) {
  return spoc::guarded([&]() -> int {
    if (!A || !xgr || !ygr || !in) return spoc::report(SPOC_EINVAL, "spoc_amr_mesh: bad arguments");
    spoc_amr &B = const_cast<spoc_amr &>(*A);
    for (int i = 1; i <= A->nn_ab; i++) {
      xgr[i - 1] = A->x[i];
      ygr[i - 1] = A->y[i];
    }
    for (int i = 1; i <= A->ne; i++) {
      const int t = A->tabsol[i];
      for (int j = 1; j <= 3; j++) in[3 * (i - 1) + j - 1] = B.IN(j, t);
      if (tabsol) tabsol[i - 1] = t;
    }
    return 0;
  });
}

int spoc_amr_levels(const spoc_amr *A, int *level) {
  return spoc::guarded([&]() -> int {
    if (!A || !level) return spoc::report(SPOC_EINVAL, "spoc_amr_levels: bad arguments");
    for (int i = 1; i <= A->ne; i++) {
      const int t = A->tabsol[i];
      // a green half is at its parent's level
      level[i - 1] = A->tlevel[t > A->ne_ab ? A->torigi[t] : t];
    }
    return 0;
  });
}

void spoc_amr_destroy(spoc_amr *A) { delete A; }

int spoc_amr_inter_only_new(int ne, const int *in, int nenew, const int *innew, int *tinter) {
  return spoc::guarded([&]() -> int {
    if (ne < 0 || nenew < 0 || (ne && !in) || (nenew && (!innew || !tinter)))
      return spoc::report(SPOC_EINVAL, "spoc_amr_inter_only_new: bad arguments");
    // src/adapt.f:1233-1258: the old triangles, in order, that reappear with
    // the same nodes in the same order as the next unmatched new triangle
    for (int i = 0; i < nenew; i++) tinter[i] = 0;
    int tk = 1;
    for (int i = 1; i <= ne && tk <= nenew; i++) {
      int k = 0;
      for (int j = 0; j < 3; j++)
        if (in[3 * (i - 1) + j] == innew[3 * (tk - 1) + j]) k++;
      if (k == 3) {
        tinter[tk - 1] = i;
        tk++;
      }
    }
    return 0;
  });
}

}  // extern "C"
