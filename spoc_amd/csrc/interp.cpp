// Fields of the adaptive driver carried onto a new mesh, host side:
// interpol + interlin (src/adapt.f:1662-1824) for the modal state and calh2
// (src/adapt.f:303-373) for fields given as functions of (x, y).  Both
// L2-project onto each new element: values at the triangle rule's points,
// the projection integrals (inttri2, src/nonlin.f:91-107) and the two
// triangular solves with the mass matrix's Cholesky factor (dtrsm3,
// src/dtrsm1.f:1-46), in the reference's operation order, so the results
// equal the reference's bit for bit.
//
// interpol locates each quadrature point in the old mesh by a walk that
// starts from the element where the previous point was found (the first
// from element 1); the element a point on an edge is assigned to depends on
// that path, so the walk runs in the reference's order.  The walk does not
// depend on the field, so all fields are carried in one pass (the
// reference calls interpol once per field, each time starting from element
// 1, along the same path).
#include <cmath>
#include <vector>

#include "../../include/spoc_abi.h"
#include "error.h"

namespace {

double po(int n, double x) {  // src/legendre2.f:5-30
  double p0 = 1.0, p1 = x, p = p0;
  if (n == 0) return p0;
  if (n == 1) return p1;
  for (int i = 2; i <= n; i++) {
    const double a0 = (double)i, a2 = (double)(2 * i - 1), a3 = (double)(i - 1);
    p = (a2 * x * p1 - a3 * p0) / a0;
    p0 = p1;
    p1 = p;
  }
  return p;
}

struct Rule {
  int nc, nm, ng;
  const double *wi, *xi, *yi, *pi2, *amm;
  int ld_g, ld_m;
};

// b <- (A^T A)^-1 b with A the upper factor (src/dtrsm1.f:9-38).
void dtrsm3(const Rule &R, double *b) {
  const int m = R.nm;
  double diag[64];
  for (int i = 0; i < m; i++) diag[i] = 1.0 / R.amm[i + (size_t)R.ld_m * i];
  for (int i = 0; i < m; i++) {
    double t = b[i];
    for (int k = 0; k < i; k++) t = t - R.amm[k + (size_t)R.ld_m * i] * b[k];
    b[i] = t * diag[i];
  }
  for (int k = m - 1; k >= 0; k--) {
    b[k] = b[k] * diag[k];
    for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * R.amm[i + (size_t)R.ld_m * k];
  }
}

// Projection of the point values mff onto the modes, then the solve.
void project(const Rule &R, const double *mff, double *b) {
  for (int m = 0; m < R.nm; m++) {  // src/nonlin.f:35-57, 91-107
    double ans = 0.0;
    for (int q = 0; q < R.ng; q++) ans = ans + R.wi[q] * mff[q] * R.pi2[q + (size_t)R.ld_g * m];
    b[m] = ans;
  }
  dtrsm3(R, b);
}

// Modal expansion of element e at local coordinates (x, y)
// (fsol, src/mat1.f:495-511).
double fsol(const Rule &R, double x, double y, const double *a) {
  double s = 0.0;
  int j = 0;
  for (int k1 = 0; k1 <= R.nc; k1++)
    for (int k2 = 0; k2 <= R.nc - k1; k2++) s = s + a[j++] * po(k1, x) * po(k2, y);
  return s;
}

// Every entry of the n 1-based indices a lies in [lo, hi]; else the
// message names the array and the first entry outside.
bool in_range(const int *a, size_t n, int lo, int hi, const char *fn, const char *name) {
  for (size_t i = 0; i < n; i++)
    if (a[i] < lo || a[i] > hi) {
      spoc::report(SPOC_EINVAL, "%s: %s[%zu]=%d outside [%d, %d]", fn, name, i, a[i], lo, hi);
      return false;
    }
  return true;
}

}  // namespace

extern "C" {

int spoc_interpol(int nc, int ng, const double *wi, const double *xi, const double *yi,
                  const double *pi2, int ld_g, const double *amm, int ld_m, int nn, int ne,
                  int nf, const double *xgr, const double *ygr, const int *in, const int *itm,
                  const int *tseg, const int *pseg, int nnnew, int nenew, const double *xgrnew,
                  const double *ygrnew, const int *innew, const int *tinter, int nfield, int ld,
                  const double *const *u, double *const *unew) {
  return spoc::guarded([&]() -> int {
    const int nm = (nc + 1) * (nc + 2) / 2;
    if (nc < 0 || nm > 64 || ng <= 0 || ng > 64 || ld_g < ng || ld_m < nm || ld < nm || nn <= 0 ||
        ne <= 0 || nf <= 0 || nnnew <= 0 || nenew <= 0 || nfield <= 0 || !wi || !xi || !yi ||
        !pi2 || !amm || !xgr || !ygr || !in || !itm || !tseg || !pseg || !xgrnew || !ygrnew ||
        !innew || !tinter || !u || !unew)
      return spoc::report(SPOC_EINVAL, "spoc_interpol: bad arguments");
    // the walk indexes nodes, faces and elements with these: check them first
    const char *fn = "spoc_interpol";
    if (!in_range(in, 3 * (size_t)ne, 1, nn, fn, "in") ||
        !in_range(itm, 3 * (size_t)ne, 1, nf, fn, "itm") ||
        !in_range(pseg, 2 * (size_t)nf, 1, nn, fn, "pseg") ||
        !in_range(innew, 3 * (size_t)nenew, 1, nnnew, fn, "innew") ||
        !in_range(tinter, (size_t)nenew, 0, ne, fn, "tinter"))
      return SPOC_EINVAL;
    for (int f = 0; f < nf; f++)
      if (tseg[2 * f] < 1 || tseg[2 * f] > ne || tseg[2 * f + 1] < 0 || tseg[2 * f + 1] > ne)
        return spoc::report(SPOC_EINVAL, "spoc_interpol: tseg(:,%d)=(%d,%d) out of range", f + 1,
                            tseg[2 * f], tseg[2 * f + 1]);
    const Rule R{nc, nm, ng, wi, xi, yi, pi2, amm, ld_g, ld_m};
    auto X = [&](int n) { return xgr[n - 1]; };  // 1-based node access
    auto Y = [&](int n) { return ygr[n - 1]; };
    std::vector<double> mff((size_t)nfield * ng), b(nm);
    const long max_steps = 4L * ne + 64;  // a walk crosses each element at most a few times
    int tk = 1;
    for (int i = 0; i < nenew; i++) {
      if (tinter[i] != 0) {  // an element of the old mesh, kept: copy it
        const int t = tinter[i];
        for (int f = 0; f < nfield; f++)
          for (int m = 0; m < nm; m++)
            unew[f][(size_t)ld * i + m] = u[f][(size_t)ld * (t - 1) + m];
        continue;
      }
      const int i1 = innew[3 * i] - 1, i2 = innew[3 * i + 1] - 1, i3 = innew[3 * i + 2] - 1;
      const double agx = (xgrnew[i2] - xgrnew[i1]) * 0.5, bgx = (xgrnew[i3] - xgrnew[i1]) * 0.5;
      const double agy = (ygrnew[i2] - ygrnew[i1]) * 0.5, bgy = (ygrnew[i3] - ygrnew[i1]) * 0.5;
      const double xof = xgrnew[i1], yof = ygrnew[i1];
      for (int j = 0; j < ng; j++) {
        const double x = xi[j], y = yi[j];
        double xdp = agx * (x + 1.0) + bgx * (y + 1.0) + xof;
        double ydp = agy * (x + 1.0) + bgy * (y + 1.0) + yof;
        // interlin (src/adapt.f:1747-1824)
        double lab[3];
        long steps = 0;
        for (;;) {
          if (++steps > max_steps)
            return spoc::report(SPOC_EMESH, "spoc_interpol: the walk for point %d of new element "
                                "%d does not end", j + 1, i + 1);
          const int *n = in + 3 * (tk - 1);
          const double x1 = X(n[0]), y1 = Y(n[0]), x2 = X(n[1]), y2 = Y(n[1]);
          const double x3 = X(n[2]), y3 = Y(n[2]);
          const double aire = (x2 - x3) * (y2 - y1) - (x2 - x1) * (y2 - y3);
          lab[0] = (xdp - x2) * (ydp - y3) - (xdp - x3) * (ydp - y2);
          lab[1] = (xdp - x3) * (ydp - y1) - (xdp - x1) * (ydp - y3);
          lab[2] = (xdp - x1) * (ydp - y2) - (xdp - x2) * (ydp - y1);
          for (int k = 0; k < 3; k++) {
            lab[k] = lab[k] / aire;
            if (std::fabs(lab[k]) < 1.0e-9) lab[k] = 0.0;
          }
          if (lab[0] >= 0 && lab[1] >= 0 && lab[2] >= 0) break;
          const int *fm = itm + 3 * (tk - 1);
          const int ns1 = lab[0] < 0.0 ? fm[1] : (lab[1] < 0.0 ? fm[2] : fm[0]);
          int tk1 = tseg[2 * (ns1 - 1)];
          if (tk1 == tk) tk1 = tseg[2 * (ns1 - 1) + 1];
          if (tk1 > 0) {
            tk = tk1;
          } else {  // outside the mesh: onto the wall edge, and again
            const int n1 = pseg[2 * (ns1 - 1)], n2 = pseg[2 * (ns1 - 1) + 1];
            const double xa = X(n1), ya = Y(n1), xb = X(n2), yb = Y(n2);
            const double d = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb);
            double pr = (xa - xdp) * (xa - xb) + (ya - ydp) * (ya - yb);
            pr = pr / d;
            xdp = xa * (1.0 - pr) + xb * pr;
            ydp = ya * (1.0 - pr) + yb * pr;
          }
        }
        const double lx = 2.0 * lab[1] - 1.0, ly = 2.0 * lab[2] - 1.0;
        for (int f = 0; f < nfield; f++)
          mff[(size_t)f * ng + j] = fsol(R, lx, ly, u[f] + (size_t)ld * (tk - 1));
      }
      for (int f = 0; f < nfield; f++) {
        project(R, mff.data() + (size_t)f * ng, b.data());
        for (int m = 0; m < nm; m++) unew[f][(size_t)ld * i + m] = b[m];
      }
    }
    return SPOC_OK;
  });
}

int spoc_calh2(int nc, int ng, const double *wi, const double *xi, const double *yi,
               const double *pi2, int ld_g, const double *amm, int ld_m, double (*ff)(double x,
               double y, void *user), void *user, int nn, int ne, const double *xgr,
               const double *ygr, const int *in, const int *tinter, int ne0, int ld,
               const double *h0, double *h1) {
  return spoc::guarded([&]() -> int {
    const int nm = (nc + 1) * (nc + 2) / 2;
    if (nc < 0 || nm > 64 || ng <= 0 || ng > 64 || ld_g < ng || ld_m < nm || ld < nm || nn <= 0 ||
        ne <= 0 || ne0 < 0 || !ff || !wi || !xi || !yi || !pi2 || !amm || !xgr || !ygr || !in ||
        !tinter || !h1)
      return spoc::report(SPOC_EINVAL, "spoc_calh2: bad arguments");
    if (!in_range(in, 3 * (size_t)ne, 1, nn, "spoc_calh2", "in") ||
        !in_range(tinter, (size_t)ne, 0, h0 ? ne0 : 0, "spoc_calh2", "tinter"))
      return SPOC_EINVAL;
    const Rule R{nc, nm, ng, wi, xi, yi, pi2, amm, ld_g, ld_m};
    std::vector<double> mff(ng), b(nm);
    for (int i = 0; i < ne; i++) {
      if (tinter[i] != 0) {
        for (int m = 0; m < nm; m++) h1[(size_t)ld * i + m] = h0[(size_t)ld * (tinter[i] - 1) + m];
        continue;
      }
      const int i1 = in[3 * i] - 1, i2 = in[3 * i + 1] - 1, i3 = in[3 * i + 2] - 1;
      const double agx = (xgr[i2] - xgr[i1]) * 0.5, agy = (ygr[i2] - ygr[i1]) * 0.5;
      const double bgx = (xgr[i3] - xgr[i1]) * 0.5, bgy = (ygr[i3] - ygr[i1]) * 0.5;
      const double xof = xgr[i1], yof = ygr[i1];
      for (int j = 0; j < ng; j++) {
        const double x = xi[j], y = yi[j];
        mff[j] = ff(agx * (x + 1.0) + bgx * (y + 1.0) + xof, agy * (x + 1.0) + bgy * (y + 1.0) + yof,
                    user);
      }
      project(R, mff.data(), b.data());
      for (int m = 0; m < nm; m++) h1[(size_t)ld * i + m] = b[m];
    }
    return SPOC_OK;
  });
}

}  // extern "C"
