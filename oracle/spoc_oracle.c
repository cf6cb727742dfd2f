/*
 * CPU FP64 restatement of SPOC's RK4 right-hand side -- TEST INFRASTRUCTURE
 * (the oracle).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it; the product never does.
 *
 * Pinned by: (1) bit-exact comparison against golden vectors written by the
 * reference Fortran compiled from its own sources (oracle/_ref;
 * tests/golden/make_goldens.py writes them, tests/test_oracle_golden.py
 * compares); (2) closed forms and
 * invariants (Legendre values, quadrature exactness, Cholesky vs numpy,
 * reduced basis vanishing on edge 1, rot_vec involution, mass conservation),
 * tests/test_oracle_math.py.
 *
 * Compile with -ffp-contract=off: the reference, built with flang -O2 and no
 * -march, contains no fused multiply-adds.
 */
#include "spoc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LDM ORC_LDM
#define LDF ORC_LDF
#define LDG ORC_LDG
#define M2(a, i, j) (a)[(i) + LDM * (j)]             /* 0-based (i, j) of a (45,45) */
#define PICO(t, j, m, k) (t)->pico[(j) + LDF * ((m) + LDM * (k))]
#define PI2(t, q, m) (t)->pi2[(q) + LDG * (m)]

/* Legendre P_n by three-term recurrence: src/legendre2.f:5-30. */
double orc_po(int n, double x) {
  double p0 = 1.0, p1 = x, po = p0;
  if (n == 0) return p0;
  if (n == 1) return p1;
  for (int i = 2; i <= n; i++) {
    double a0 = (double)i, a2 = (double)(2 * i - 1), a3 = (double)(i - 1);
    po = (a2 * x * p1 - a3 * p0) / a0;
    p0 = p1;
    p1 = po;
  }
  return po;
}

/* Legendre P'_n: src/legendre2.f:32-62. */
double orc_dpo(int n, double x) {
  double p0 = 1.0, p1 = x, dp0 = 0.0, dp1 = 1.0, dpo = dp0;
  if (n == 0) return dp0;
  if (n == 1) return dp1;
  for (int i = 2; i <= n; i++) {
    double a0 = (double)i, a2 = (double)(2 * i - 1), a3 = (double)(i - 1);
    dpo = (a2 * x * dp1 + a2 * p1 - a3 * dp0) / a0;
    double p2 = (a2 * x * p1 - a3 * p0) / a0;
    p0 = p1;
    p1 = p2;
    dp0 = dp1;
    dp1 = dpo;
  }
  return dpo;
}

/* ip(nc, i) = sum_{j<i} (nc+1-j): src/mat1.f:372-381. */
static int orc_ip(int nc, int i) {
  int ip = 0;
  for (int j = 0; j < i; j++) ip += nc + 1 - j;
  return ip;
}

/* Operator integrand kinds for cal_mat (src/mat1.f:513-550). */
enum { F_MASS, F_DX, F_DY };

/* cal_mat with inttri: src/mat1.f:90-142, 266-279. */
static void orc_cal_mat(orc_tables *t, double *mat, int kind) {
  int nc = t->nc;
  memset(mat, 0, sizeof(double) * LDM * LDM);
  int i1 = 0;
  for (int k1 = 0; k1 <= nc; k1++)
    for (int k2 = 0; k2 <= nc - k1; k2++) {
      int i2 = 0;
      for (int k3 = 0; k3 <= nc; k3++)
        for (int k4 = 0; k4 <= nc - k3; k4++) {
          double ans = 0.0;
          for (int j = 0; j < t->ng; j++) {
            double x = t->xi[j], y = t->yi[j], f;
            if (kind == F_MASS)
              f = orc_po(k1, x) * orc_po(k2, y) * orc_po(k3, x) * orc_po(k4, y);
            else if (kind == F_DX)
              f = orc_po(k1, x) * orc_po(k2, y) * orc_dpo(k3, x) * orc_po(k4, y);
            else
              f = orc_po(k1, x) * orc_po(k2, y) * orc_po(k3, x) * orc_dpo(k4, y);
            ans = ans + t->wi[j] * f;
          }
          M2(mat, i1, i2) = ans;
          i2++;
        }
      i1++;
    }
}

/* cal_matl: reduced-basis mass matrix C^T amm C, src/mat1.f:285-368. */
static void orc_cal_matl(orc_tables *t, const double *amm, double *amml) {
  int nc = t->nc;
  memset(amml, 0, sizeof(double) * LDM * LDM);
  /* The four sign patterns of the reference, in its loop order. */
  static const int jstart[4] = {0, 1, 0, 1}, lstart[4] = {0, 0, 1, 1};
  for (int pass = 0; pass < 4; pass++)
    for (int i = 0; i <= nc - 1; i++)
      for (int j = jstart[pass]; j <= nc - 1 - i; j += 2)
        for (int k = 0; k <= nc - 1; k++)
          for (int l = lstart[pass]; l <= nc - 1 - k; l += 2) {
            int i1 = orc_ip(nc - 1, i) + j + 1, i2 = orc_ip(nc, i) + j + 2;
            int j1 = orc_ip(nc - 1, k) + l + 1, j2 = orc_ip(nc, k) + l + 2;
            int i0 = orc_ip(nc, i) + 1, j0 = orc_ip(nc, k) + 1;
            double a22 = M2(amm, i2 - 1, j2 - 1), a02 = M2(amm, i0 - 1, j2 - 1);
            double a20 = M2(amm, i2 - 1, j0 - 1), a00 = M2(amm, i0 - 1, j0 - 1);
            double v;
            if (pass == 0) v = a22 + a02 + a20 + a00;
            else if (pass == 1) v = a22 - a02 + a20 - a00;
            else if (pass == 2) v = a22 + a02 - a20 - a00;
            else v = a22 - a02 - a20 + a00;
            M2(amml, i1 - 1, j1 - 1) = v;
          }
}

/* Unblocked upper Cholesky as taken by the vendored DPOTRF for n <= 64
 * (src/solve.f:108-113 -> DPOTF2 src/solve.f:383-403, with DDOT
 * src/solve.f:1378, DGEMV('T') src/solve2.f:541-551, DSCAL src/solve2.f:575).
 * DDOT's unrolling by 5 adds left to right, i.e. sequentially. */
static int orc_dpotf2(int n, double *a) {
  for (int j = 0; j < n; j++) {
    double dot = 0.0;
    for (int i = 0; i < j; i++) dot = dot + M2(a, i, j) * M2(a, i, j);
    double ajj = M2(a, j, j) - dot;
    if (ajj <= 0.0) {
      M2(a, j, j) = ajj;
      return j + 1;
    }
    ajj = sqrt(ajj);
    M2(a, j, j) = ajj;
    if (j < n - 1) {
      if (j > 0)
        for (int c = j + 1; c < n; c++) {
          double temp = 0.0;
          for (int i = 0; i < j; i++) temp = temp + M2(a, i, c) * M2(a, i, j);
          M2(a, j, c) = M2(a, j, c) + (-1.0) * temp;
        }
      double r = 1.0 / ajj;
      for (int c = j + 1; c < n; c++) M2(a, j, c) = r * M2(a, j, c);
    }
  }
  return 0;
}

/* Table set-up in the driver's order: src/shwwnl.f:127-156 (cal_p2
 * src/nonlin.f:10-30, cal_pico src/mat1.f:214-259, calip src/mat1.f:383-410). */
int orc_setup(orc_tables *t, int nc, int ng, int ngf, const double *wi, const double *xi,
              const double *yi, const double *wif, const double *tif) {
  memset(t, 0, sizeof(*t));
  if (nc < 1 || nc > 8 || ng < 1 || ng > LDG || ngf < 1 || ngf > LDF) return -1;
  t->nc = nc;
  t->ng = ng;
  t->ngf = ngf;
  t->nm = (nc + 1) * (nc + 2) / 2;
  for (int i = 0; i < ng; i++) {
    t->wi[i] = wi[i];
    t->xi[i] = xi[i];
    t->yi[i] = yi[i];
  }
  for (int i = 0; i < ngf; i++) {
    t->wif[i] = wif[i];
    t->tif[i] = tif[i];
  }
  int i1 = 0;
  for (int k1 = 0; k1 <= nc; k1++)
    for (int k2 = 0; k2 <= nc - k1; k2++) {
      for (int q = 0; q < ng; q++) PI2(t, q, i1) = orc_po(k1, xi[q]) * orc_po(k2, yi[q]);
      for (int j = 0; j < ngf; j++) {
        double tj = tif[j];
        PICO(t, j, i1, 0) = orc_po(k1, tj) * orc_po(k2, -1.0);
        PICO(t, j, i1, 1) = orc_po(k1, -tj) * orc_po(k2, tj);
        PICO(t, j, i1, 2) = orc_po(k1, -1.0) * orc_po(k2, -tj);
      }
      i1++;
    }
  for (int i = 0; i <= nc; i++) t->cip[0 + 2 * i] = orc_ip(nc, i);
  for (int i = 0; i <= nc - 1; i++) t->cip[1 + 2 * i] = orc_ip(nc - 1, i);
  double amm[LDM * LDM];
  orc_cal_mat(t, amm, F_MASS);
  orc_cal_mat(t, t->agx, F_DX);
  orc_cal_mat(t, t->agy, F_DY);
  orc_cal_matl(t, amm, t->amml);
  memcpy(t->ammo, amm, sizeof(amm));
  memcpy(t->amm, amm, sizeof(amm));
  t->info1 = orc_dpotf2(t->nm, t->amm);
  t->info2 = orc_dpotf2(t->nm - nc - 1, t->amml);
  return (t->info1 || t->info2) ? 1 : 0;
}

/* Face traces: src/calflu.f:6-71.  a0 has leading dimension ld. */
void orc_cal_flu(const orc_tables *t, int nf, const int *tseg, double *fbdy, const double *a0,
                 int ld, const int *ssit) {
  int nm = t->nm, ngf = t->ngf;
  for (int i = 0; i < nf; i++) {
    double *f = fbdy + LDF * i;
    for (int j = 0; j < ngf; j++) f[j] = 0.0;
    int k1 = tseg[2 * i] - 1, k2 = tseg[2 * i + 1] - 1;
    if (k2 < 0) {
      for (int m = 0; m < nm; m++)
        for (int j = 0; j < ngf; j++) f[j] = f[j] + a0[m + ld * k1] * PICO(t, j, m, 0);
    } else {
      int c = ssit[2 * i] - 1;
      for (int m = 0; m < nm; m++)
        for (int j = 0; j < ngf; j++) f[j] = f[j] + a0[m + ld * k1] * PICO(t, j, m, c);
      c = ssit[2 * i + 1] - 1;
      for (int m = 0; m < nm; m++)
        for (int j = 0; j < ngf; j++)
          f[j] = f[j] + a0[m + ld * k2] * PICO(t, ngf - 1 - j, m, c);
      for (int j = 0; j < ngf; j++) f[j] = 0.5 * f[j];
    }
  }
}

/* b += coeff * mat^T ak: src/calb2.f:3-23. */
static void cal_b(const orc_tables *t, double *b, const double *ak, const double *mat,
                  double coeff) {
  for (int i = 0; i < t->nm; i++) {
    double ans = 0.0;
    for (int j = 0; j < t->nm; j++) ans = ans + M2(mat, j, i) * ak[j];
    b[i] = b[i] + ans * coeff;
  }
}

/* Edge lift to test functions: src/mat1.f:172-209.  e is 0-based. */
static void tri_f(const orc_tables *t, int e, const int *itm, const int *its,
                  const double *fbdy, double *b, double coeff) {
  int nm = t->nm, ngf = t->ngf;
  for (int k = 0; k < 3; k++) {
    int ns1 = itm[3 * e + k] - 1;
    int isg = its[3 * e + k];
    int isg2 = (ngf + 1) * (1 - isg) / 2;
    double sig = (double)isg * coeff * 0.5;
    for (int m = 0; m < nm; m++) {
      double ans = 0.0;
      for (int j = 1; j <= ngf; j++) {
        int j1 = isg2 + isg * j;
        double var = PICO(t, j1 - 1, m, k) * fbdy[(j - 1) + LDF * ns1];
        ans = ans + t->wif[j - 1] * var;
      }
      b[m] = b[m] + ans * sig;
    }
  }
}

/* Cholesky solve with the reference factor: src/dtrsm1.f:47-82. */
void orc_dtrsm4(int m, const double *a, double *b, const double *diag) {
  for (int i = 0; i < m; i++) {
    double temp = b[i];
    for (int k = 0; k < i; k++) temp = temp - M2(a, k, i) * b[k];
    temp = temp * diag[i];
    b[i] = temp;
  }
  for (int k = m - 1; k >= 0; k--) {
    b[k] = b[k] * diag[k];
    for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * M2(a, i, k);
  }
}

/* Projection onto the zero-on-edge-1 basis: src/mat1.f:418-439. */
void orc_reduc_gn(const orc_tables *t, const double *gn, double *gnr) {
  int nc = t->nc;
  for (int j = 0; j <= nc - 1; j++) {
    int c1 = t->cip[2 * j], c2 = t->cip[1 + 2 * j];
    int ind0 = c1 + 1;
    for (int l = 0; l <= nc - 1 - j; l += 2) gnr[c2 + l] = gn[c1 + l + 1] + gn[ind0 - 1];
    for (int l = 1; l <= nc - 1 - j; l += 2) gnr[c2 + l] = gn[c1 + l + 1] - gn[ind0 - 1];
  }
}

/* Expansion back from the reduced basis: src/mat1.f:445-473. */
void orc_reduc_inv_gn(const orc_tables *t, double *gn, const double *gnr) {
  int nc = t->nc, nm = t->nm;
  for (int j = 0; j <= nc - 1; j++) {
    int c1 = t->cip[2 * j], c2 = t->cip[1 + 2 * j];
    int k = c1;
    gn[k] = 0.0;
    for (int l = 0; l <= nc - j - 1; l += 2) gn[k] = gn[k] + gnr[c2 + l];
    for (int l = 1; l <= nc - j - 1; l += 2) gn[k] = gn[k] - gnr[c2 + l];
    for (int l = 0; l <= nc - j - 1; l++) gn[c1 + l + 1] = gnr[c2 + l];
  }
  gn[nm - 1] = 0.0;
}

/* src/mat1.f:478-488. */
static void rot_vec(int nm, double *gn, double *gt, const double *gx, const double *gy,
                    double dnx, double dny) {
  for (int j = 0; j < nm; j++) {
    double a = gx[j] * dnx + gy[j] * dny;
    double b = gx[j] * dny - gy[j] * dnx;
    gn[j] = a;
    gt[j] = b;
  }
}

/* src/nonlin.f:62-86. */
static void spectoxy(const orc_tables *t, double *mff, const double *ak) {
  for (int j = 0; j < t->ng; j++) mff[j] = 0.0;
  for (int m = 0; m < t->nm; m++)
    for (int j = 0; j < t->ng; j++) mff[j] = mff[j] + ak[m] * PI2(t, j, m);
}

/* xytospec + inttri2: src/nonlin.f:35-57, 91-107. */
static void xytospec(const orc_tables *t, double *b, const double *mff) {
  for (int m = 0; m < t->nm; m++) {
    double ans = 0.0;
    for (int i = 0; i < t->ng; i++) ans = ans + t->wi[i] * mff[i] * PI2(t, i, m);
    b[m] = ans;
  }
}

static void multiply(int ng, double *mff, const double *a, const double *b) {
  for (int j = 0; j < ng; j++) mff[j] = a[j] * b[j]; /* src/nonlin.f:111-121 */
}

static void divide(int ng, double *mff, const double *a, const double *b) {
  for (int j = 0; j < ng; j++) mff[j] = a[j] / b[j]; /* src/nonlin.f:123-133 */
}

/* Face-flux helper: fbd(j,f) * norm(c,f) (the loops after every cal_flu,
 * e.g. src/calvarnl_fs.f:350-373). */
static void scale_face(const orc_tables *t, int nf, const double *fbd, const double *norm,
                       int c, double *out) {
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < t->ngf; j++) out[j + LDF * i] = fbd[j + LDF * i] * norm[c + 2 * i];
}

/* ------------------------------------------------------------------------- */
/* Linear RHS: src/calvar_lin.f:7-287.                                        */
/* ------------------------------------------------------------------------- */
static void cal_var_lin(const orc_tables *t, int nf, int ne, int ld, const double *u0,
                        const double *v0, const double *h0, double *u1, double *v1, double *h1,
                        const int *tseg, const int *ssit, const int *itm, const int *its,
                        const double *norm, const int *bdl, const double *vect,
                        const double *normt, double g0, const double *hb) {
  int nm = t->nm, nc = t->nc;
  double hh0 = hb[0];
  double dia1[LDM], dia2[LDM];
  for (int i = 0; i < nm; i++) {
    dia1[i] = 1.0 / M2(t->amm, i, i);
    dia2[i] = 1.0 / M2(t->amml, i, i);
  }
  size_t fsz = (size_t)LDF * nf;
  double *fxx = malloc(sizeof(double) * nm * ne), *fzero = calloc((size_t)nm * ne, sizeof(double));
  double *fbd = calloc(fsz, sizeof(double)), *fbdxx = calloc(fsz, sizeof(double));
  double *fbdxy = calloc(fsz, sizeof(double)), *fbdyx = calloc(fsz, sizeof(double));
  double *fbdyy = calloc(fsz, sizeof(double));
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < nm; j++) fxx[j + nm * i] = -g0 * h0[j + ld * i];
  /* fxx == fyy and fxy == fyx == 0 (src/calvar_lin.f:74-81) */
  orc_cal_flu(t, nf, tseg, fbd, fxx, nm, ssit);
  scale_face(t, nf, fbd, norm, 0, fbdxx);
  orc_cal_flu(t, nf, tseg, fbd, fzero, nm, ssit);
  scale_face(t, nf, fbd, norm, 1, fbdxy);
  orc_cal_flu(t, nf, tseg, fbd, fzero, nm, ssit);
  scale_face(t, nf, fbd, norm, 0, fbdyx);
  orc_cal_flu(t, nf, tseg, fbd, fxx, nm, ssit);
  scale_face(t, nf, fbd, norm, 1, fbdyy);

  for (int i = 0; i < ne; i++) {
    double ax = vect[5 * i], bx = vect[5 * i + 1], ay = vect[5 * i + 2], by = vect[5 * i + 3];
    double detai = vect[5 * i + 4];
    double gx[LDM] = {0}, gy[LDM] = {0}, gxx[LDM], gxy[LDM] = {0}, gyx[LDM] = {0}, gyy[LDM];
    double gn[LDM], gt[LDM], bu[LDM];
    for (int j = 0; j < nm; j++) {
      gxx[j] = fxx[j + nm * i];
      gyy[j] = fxx[j + nm * i];
    }
    cal_b(t, gx, gxx, t->agx, -by * detai);
    cal_b(t, gx, gxx, t->agy, ay * detai);
    cal_b(t, gx, gxy, t->agx, bx * detai);
    cal_b(t, gx, gxy, t->agy, -ax * detai);
    tri_f(t, i, itm, its, fbdxx, gx, detai);
    tri_f(t, i, itm, its, fbdxy, gx, detai);
    double coeff = by * detai;
    cal_b(t, gy, gyx, t->agx, -coeff);
    coeff = -ay * detai;
    cal_b(t, gy, gyx, t->agy, -coeff);
    coeff = -bx * detai;
    cal_b(t, gy, gyy, t->agx, -coeff);
    coeff = ax * detai;
    cal_b(t, gy, gyy, t->agy, -coeff);
    tri_f(t, i, itm, its, fbdyx, gy, detai);
    tri_f(t, i, itm, its, fbdyy, gy, detai);
    if (bdl[i]) {
      double dnx = normt[2 * i], dny = normt[2 * i + 1];
      rot_vec(nm, gn, gt, gx, gy, dnx, dny);
      orc_reduc_gn(t, gn, bu);
      orc_dtrsm4(nm - nc - 1, t->amml, bu, dia2);
      orc_dtrsm4(nm, t->amm, gt, dia1);
      orc_reduc_inv_gn(t, gn, bu);
      rot_vec(nm, gx, gy, gn, gt, dnx, dny);
    } else {
      orc_dtrsm4(nm, t->amm, gx, dia1);
      orc_dtrsm4(nm, t->amm, gy, dia1);
    }
    for (int j = 0; j < nm; j++) {
      u1[j + ld * i] = gx[j];
      v1[j + ld * i] = gy[j];
    }
  }

  /* h (src/calvar_lin.f:224-282) */
  orc_cal_flu(t, nf, tseg, fbd, u0, ld, ssit);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < t->ngf; j++) fbdxx[j + LDF * i] = fbd[j + LDF * i] * norm[2 * i] * hh0;
  orc_cal_flu(t, nf, tseg, fbd, v0, ld, ssit);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < t->ngf; j++)
      fbdyy[j + LDF * i] = fbd[j + LDF * i] * norm[2 * i + 1] * hh0;
  for (int i = 0; i < ne; i++) {
    double ax = vect[5 * i], bx = vect[5 * i + 1], ay = vect[5 * i + 2], by = vect[5 * i + 3];
    double detai = vect[5 * i + 4];
    double b[LDM] = {0}, ak[LDM];
    for (int j = 0; j < nm; j++) ak[j] = u0[j + ld * i] * hh0;
    cal_b(t, b, ak, t->agx, by * detai);
    cal_b(t, b, ak, t->agy, -ay * detai);
    for (int j = 0; j < nm; j++) ak[j] = v0[j + ld * i] * hh0;
    cal_b(t, b, ak, t->agx, -bx * detai);
    cal_b(t, b, ak, t->agy, ax * detai);
    tri_f(t, i, itm, its, fbdxx, b, -detai);
    tri_f(t, i, itm, its, fbdyy, b, -detai);
    orc_dtrsm4(nm, t->amm, b, dia1);
    for (int j = 0; j < nm; j++) h1[j + ld * i] = b[j];
  }
  free(fxx);
  free(fzero);
  free(fbd);
  free(fbdxx);
  free(fbdxy);
  free(fbdyx);
  free(fbdyy);
}

/* ------------------------------------------------------------------------- */
/* Nonlinear RHS: src/calvarnl_fs.f:9-549, with the no-slip                   */
/* (src/calvarnl_ns.f:9-512) and inviscid (src/calvarnl_inv.f:6-498)          */
/* boundary differences.                                                     */
/* ------------------------------------------------------------------------- */
static void cal_var_nl(const orc_tables *t, int variant, int nf, int ne, int ld,
                       const double *u0, const double *v0, const double *h0, double *u1,
                       double *v1, double *h1, const int *tseg, const int *ssit, const int *itm,
                       const int *its, const double *norm, const int *bdl, const double *vect,
                       const double *normt, double mu0, double mb0, const double *fc, double g0,
                       const double *w0, const double *z0, const double *hb) {
  int nm = t->nm, nc = t->nc, ng = t->ng;
  double dia1[LDM], dia2[LDM] = {0};
  for (int i = 0; i < nm; i++) dia1[i] = 1.0 / M2(t->amm, i, i);
  for (int i = 0; i < nm - nc - 1; i++) dia2[i] = 1.0 / M2(t->amml, i, i);
  size_t fsz = (size_t)LDF * nf, esz = (size_t)nm * ne;
  double *fbd = calloc(fsz, 8), *fbdxx = calloc(fsz, 8), *fbdxy = calloc(fsz, 8);
  double *fbdyx = calloc(fsz, 8), *fbdyy = calloc(fsz, 8);
  double *qxx = malloc(esz * 8), *qxy = malloc(esz * 8), *qyx = malloc(esz * 8),
         *qyy = malloc(esz * 8);
  double *adu = malloc(esz * 8), *adv = malloc(esz * 8), *uh = malloc(esz * 8),
         *vh = malloc(esz * 8), *w1 = malloc(esz * 8), *z1 = malloc(esz * 8);
  double *fxx = malloc(esz * 8), *fxy = malloc(esz * 8), *fyx = malloc(esz * 8),
         *fyy = malloc(esz * 8);

  /* A: velocity traces (src/calvarnl_fs.f:88-101) */
  orc_cal_flu(t, nf, tseg, fbd, u0, ld, ssit);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < t->ngf; j++) {
      fbdxx[j + LDF * i] = fbd[j + LDF * i] * norm[2 * i];
      fbdxy[j + LDF * i] = fbd[j + LDF * i] * norm[2 * i + 1];
    }
  orc_cal_flu(t, nf, tseg, fbd, v0, ld, ssit);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < t->ngf; j++) {
      fbdyx[j + LDF * i] = fbd[j + LDF * i] * norm[2 * i];
      fbdyy[j + LDF * i] = fbd[j + LDF * i] * norm[2 * i + 1];
    }

  /* B: velocity gradient q (src/calvarnl_fs.f:115-226) */
  for (int i = 0; i < ne; i++) {
    double uk[LDM], vk[LDM], gx1[LDM], gy1[LDM], gxx[LDM], gxy[LDM], gyx[LDM], gyy[LDM];
    for (int j = 0; j < nm; j++) {
      uk[j] = u0[j + ld * i];
      vk[j] = v0[j + ld * i];
    }
    double ax = vect[5 * i], bx = vect[5 * i + 1], ay = vect[5 * i + 2], by = vect[5 * i + 3];
    double detai = vect[5 * i + 4];
    for (int j = 0; j < nm; j++) gx1[j] = gy1[j] = 0.0;
    cal_b(t, gx1, uk, t->agx, -1.0);
    cal_b(t, gy1, uk, t->agy, -1.0);
    for (int j = 0; j < nm; j++) {
      gxx[j] = (by * gx1[j] - ay * gy1[j]) * detai;
      gxy[j] = (-bx * gx1[j] + ax * gy1[j]) * detai;
    }
    tri_f(t, i, itm, its, fbdxx, gxx, detai);
    tri_f(t, i, itm, its, fbdxy, gxy, detai);
    for (int j = 0; j < nm; j++) gx1[j] = gy1[j] = 0.0;
    cal_b(t, gx1, vk, t->agx, -1.0);
    cal_b(t, gy1, vk, t->agy, -1.0);
    for (int j = 0; j < nm; j++) {
      gyx[j] = (by * gx1[j] - ay * gy1[j]) * detai;
      gyy[j] = (-bx * gx1[j] + ax * gy1[j]) * detai;
    }
    tri_f(t, i, itm, its, fbdyx, gyx, detai);
    tri_f(t, i, itm, its, fbdyy, gyy, detai);
    if (variant == ORC_FS && bdl[i]) {
      /* src/calvarnl_fs.f:165-210 */
      double dnx = normt[2 * i], dny = normt[2 * i + 1];
      double gnx[LDM], gtx[LDM], gny[LDM], gty[LDM], gnn[LDM], gnt[LDM], gtn[LDM], gtt[LDM];
      double bu[LDM];
      rot_vec(nm, gnx, gtx, gxx, gyx, dnx, dny);
      rot_vec(nm, gny, gty, gxy, gyy, dnx, dny);
      rot_vec(nm, gnn, gnt, gnx, gny, dnx, dny);
      rot_vec(nm, gtn, gtt, gtx, gty, dnx, dny);
      orc_reduc_gn(t, gtn, bu);
      orc_dtrsm4(nm, t->amm, gnn, dia1);
      orc_dtrsm4(nm, t->amm, gnt, dia1);
      orc_dtrsm4(nm - nc - 1, t->amml, bu, dia2);
      orc_dtrsm4(nm, t->amm, gtt, dia1);
      orc_reduc_inv_gn(t, gtn, bu);
      rot_vec(nm, gnx, gny, gnn, gnt, dnx, dny);
      rot_vec(nm, gtx, gty, gtn, gtt, dnx, dny);
      rot_vec(nm, gxx, gyx, gnx, gtx, dnx, dny);
      rot_vec(nm, gxy, gyy, gny, gty, dnx, dny);
    } else {
      orc_dtrsm4(nm, t->amm, gxx, dia1);
      orc_dtrsm4(nm, t->amm, gxy, dia1);
      orc_dtrsm4(nm, t->amm, gyx, dia1);
      orc_dtrsm4(nm, t->amm, gyy, dia1);
    }
    for (int j = 0; j < nm; j++) {
      qxx[j + nm * i] = gxx[j];
      qxy[j + nm * i] = gxy[j];
      qyx[j + nm * i] = gyx[j];
      qyy[j + nm * i] = gyy[j];
    }
  }

  /* C: nonlinear terms at quadrature points (src/calvarnl_fs.f:243-332) */
  for (int i = 0; i < ne; i++) {
    double uk[LDM], vk[LDM], hk[LDM], wk[LDM], zk[LDM], gxx[LDM], gxy[LDM], gyx[LDM], gyy[LDM];
    double uux[LDM], vuy[LDM], uvx[LDM], vvy[LDM], uhk[LDM], vhk[LDM];
    double mff[LDG], mff1[LDG], mff2[LDG], mff3[LDG], mff4[LDG];
    for (int j = 0; j < nm; j++) {
      uk[j] = u0[j + ld * i];
      vk[j] = v0[j + ld * i];
      hk[j] = h0[j + ld * i] + hb[j + ld * i];
      wk[j] = w0[j + ld * i] - mb0 * uk[j];
      zk[j] = z0[j + ld * i] - mb0 * vk[j];
      gxx[j] = qxx[j + nm * i];
      gxy[j] = qxy[j + nm * i] - fc[j + ld * i];
      gyx[j] = qyx[j + nm * i] + fc[j + ld * i];
      gyy[j] = qyy[j + nm * i];
    }
    spectoxy(t, mff1, uk);
    spectoxy(t, mff2, vk);
    spectoxy(t, mff3, gxx);
    spectoxy(t, mff4, gxy);
    multiply(ng, mff, mff1, mff3);
    xytospec(t, uux, mff);
    multiply(ng, mff, mff2, mff4);
    xytospec(t, vuy, mff);
    spectoxy(t, mff3, gyx);
    spectoxy(t, mff4, gyy);
    multiply(ng, mff, mff1, mff3);
    xytospec(t, uvx, mff);
    multiply(ng, mff, mff2, mff4);
    xytospec(t, vvy, mff);
    spectoxy(t, mff3, hk);
    multiply(ng, mff, mff1, mff3);
    xytospec(t, uhk, mff);
    multiply(ng, mff, mff2, mff3);
    xytospec(t, vhk, mff);
    spectoxy(t, mff4, wk);
    divide(ng, mff, mff4, mff3);
    xytospec(t, wk, mff);
    spectoxy(t, mff4, zk);
    divide(ng, mff, mff4, mff3);
    xytospec(t, zk, mff);
    if (bdl[i]) {
      if (variant == ORC_NS) {
        /* src/calvarnl_ns.f:247-265 */
        double bu[LDM], bv[LDM];
        orc_reduc_gn(t, uhk, bu);
        orc_reduc_gn(t, vhk, bv);
        orc_dtrsm4(nm - nc - 1, t->amml, bu, dia2);
        orc_dtrsm4(nm - nc - 1, t->amml, bv, dia2);
        orc_reduc_inv_gn(t, uhk, bu);
        orc_reduc_inv_gn(t, vhk, bv);
      } else {
        /* src/calvarnl_fs.f:298-315 (same in calvarnl_inv.f) */
        double dnx = normt[2 * i], dny = normt[2 * i + 1], gn[LDM], gt[LDM], bu[LDM];
        rot_vec(nm, gn, gt, uhk, vhk, dnx, dny);
        orc_reduc_gn(t, gn, bu);
        orc_dtrsm4(nm - nc - 1, t->amml, bu, dia2);
        orc_dtrsm4(nm, t->amm, gt, dia1);
        orc_reduc_inv_gn(t, gn, bu);
        rot_vec(nm, uhk, vhk, gn, gt, dnx, dny);
      }
    } else {
      orc_dtrsm4(nm, t->amm, uhk, dia1);
      orc_dtrsm4(nm, t->amm, vhk, dia1);
    }
    for (int j = 0; j < nm; j++) {
      adu[j + nm * i] = uux[j] + vuy[j];
      adv[j + nm * i] = uvx[j] + vvy[j];
      uh[j + nm * i] = uhk[j];
      vh[j + nm * i] = vhk[j];
      w1[j + nm * i] = wk[j];
      z1[j + nm * i] = zk[j];
    }
  }

  /* D: stress tensor and its traces (src/calvarnl_fs.f:341-373) */
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < nm; j++) {
      fxx[j + nm * i] = mu0 * qxx[j + nm * i] - g0 * h0[j + ld * i];
      fxy[j + nm * i] = mu0 * qxy[j + nm * i];
      fyx[j + nm * i] = mu0 * qyx[j + nm * i];
      fyy[j + nm * i] = mu0 * qyy[j + nm * i] - g0 * h0[j + ld * i];
    }
  orc_cal_flu(t, nf, tseg, fbd, fxx, nm, ssit);
  scale_face(t, nf, fbd, norm, 0, fbdxx);
  orc_cal_flu(t, nf, tseg, fbd, fxy, nm, ssit);
  scale_face(t, nf, fbd, norm, 1, fbdxy);
  orc_cal_flu(t, nf, tseg, fbd, fyx, nm, ssit);
  scale_face(t, nf, fbd, norm, 0, fbdyx);
  orc_cal_flu(t, nf, tseg, fbd, fyy, nm, ssit);
  scale_face(t, nf, fbd, norm, 1, fbdyy);

  /* E: u, v tendencies (src/calvarnl_fs.f:382-478) */
  for (int i = 0; i < ne; i++) {
    double ax = vect[5 * i], bx = vect[5 * i + 1], ay = vect[5 * i + 2], by = vect[5 * i + 3];
    double detai = vect[5 * i + 4];
    double gx[LDM], gy[LDM], gxx[LDM], gxy[LDM], gyx[LDM], gyy[LDM];
    for (int j = 0; j < nm; j++) {
      gx[j] = -adu[j + nm * i] + w1[j + nm * i];
      gy[j] = -adv[j + nm * i] + z1[j + nm * i];
      gxx[j] = fxx[j + nm * i];
      gxy[j] = fxy[j + nm * i];
      gyx[j] = fyx[j + nm * i];
      gyy[j] = fyy[j + nm * i];
    }
    cal_b(t, gx, gxx, t->agx, -by * detai);
    cal_b(t, gx, gxx, t->agy, ay * detai);
    cal_b(t, gx, gxy, t->agx, bx * detai);
    cal_b(t, gx, gxy, t->agy, -ax * detai);
    tri_f(t, i, itm, its, fbdxx, gx, detai);
    tri_f(t, i, itm, its, fbdxy, gx, detai);
    cal_b(t, gy, gyx, t->agx, -by * detai);
    cal_b(t, gy, gyx, t->agy, ay * detai);
    cal_b(t, gy, gyy, t->agx, bx * detai);
    cal_b(t, gy, gyy, t->agy, -ax * detai);
    tri_f(t, i, itm, its, fbdyx, gy, detai);
    tri_f(t, i, itm, its, fbdyy, gy, detai);
    if (bdl[i]) {
      if (variant == ORC_NS) {
        /* src/calvarnl_ns.f:403-421 */
        double bu[LDM], bv[LDM];
        orc_reduc_gn(t, gx, bu);
        orc_reduc_gn(t, gy, bv);
        orc_dtrsm4(nm - nc - 1, t->amml, bu, dia2);
        orc_dtrsm4(nm - nc - 1, t->amml, bv, dia2);
        orc_reduc_inv_gn(t, gx, bu);
        orc_reduc_inv_gn(t, gy, bv);
      } else {
        /* src/calvarnl_fs.f:436-459 */
        double dnx = normt[2 * i], dny = normt[2 * i + 1], gn[LDM], gt[LDM], bu[LDM];
        rot_vec(nm, gn, gt, gx, gy, dnx, dny);
        orc_reduc_gn(t, gn, bu);
        orc_dtrsm4(nm - nc - 1, t->amml, bu, dia2);
        orc_dtrsm4(nm, t->amm, gt, dia1);
        orc_reduc_inv_gn(t, gn, bu);
        rot_vec(nm, gx, gy, gn, gt, dnx, dny);
      }
    } else {
      orc_dtrsm4(nm, t->amm, gx, dia1);
      orc_dtrsm4(nm, t->amm, gy, dia1);
    }
    for (int j = 0; j < nm; j++) {
      u1[j + ld * i] = gx[j];
      v1[j + ld * i] = gy[j];
    }
  }

  /* F, G: eta tendency (src/calvarnl_fs.f:487-544) */
  orc_cal_flu(t, nf, tseg, fbd, uh, nm, ssit);
  scale_face(t, nf, fbd, norm, 0, fbdxx);
  orc_cal_flu(t, nf, tseg, fbd, vh, nm, ssit);
  scale_face(t, nf, fbd, norm, 1, fbdyy);
  for (int i = 0; i < ne; i++) {
    double ax = vect[5 * i], bx = vect[5 * i + 1], ay = vect[5 * i + 2], by = vect[5 * i + 3];
    double detai = vect[5 * i + 4];
    double b[LDM] = {0}, ak[LDM];
    for (int j = 0; j < nm; j++) ak[j] = uh[j + nm * i];
    cal_b(t, b, ak, t->agx, by * detai);
    cal_b(t, b, ak, t->agy, -ay * detai);
    for (int j = 0; j < nm; j++) ak[j] = vh[j + nm * i];
    cal_b(t, b, ak, t->agx, -bx * detai);
    cal_b(t, b, ak, t->agy, ax * detai);
    tri_f(t, i, itm, its, fbdxx, b, -detai);
    tri_f(t, i, itm, its, fbdyy, b, -detai);
    orc_dtrsm4(nm, t->amm, b, dia1);
    for (int j = 0; j < nm; j++) h1[j + ld * i] = b[j];
  }

  free(fbd);
  free(fbdxx);
  free(fbdxy);
  free(fbdyx);
  free(fbdyy);
  free(qxx);
  free(qxy);
  free(qyx);
  free(qyy);
  free(adu);
  free(adv);
  free(uh);
  free(vh);
  free(w1);
  free(z1);
  free(fxx);
  free(fxy);
  free(fyx);
  free(fyy);
}

void orc_cal_var(const orc_tables *t, int variant, int nf, int ne, int ld, const double *u0,
                 const double *v0, const double *h0, double *u1, double *v1, double *h1,
                 const int *tseg, const int *ssit, const int *itm, const int *its,
                 const double *norm, const int *bdl, const double *vect, const double *normt,
                 double mu0, double mb0, const double *fc, double g0, const double *w0,
                 const double *z0, const double *hb) {
  if (variant == ORC_LIN)
    cal_var_lin(t, nf, ne, ld, u0, v0, h0, u1, v1, h1, tseg, ssit, itm, its, norm, bdl, vect,
                normt, g0, hb);
  else
    cal_var_nl(t, variant, nf, ne, ld, u0, v0, h0, u1, v1, h1, tseg, ssit, itm, its, norm, bdl,
               vect, normt, mu0, mb0, fc, g0, w0, z0, hb);
}

/* The driver's RK4 loop: src/shwwnl.f:233-284 (same in shwwnl_adap.f:302-353). */
void orc_rk4(const orc_tables *t, int variant, int nf, int ne, int ld, double *u0, double *v0,
             double *h0, const int *tseg, const int *ssit, const int *itm, const int *its,
             const double *norm, const int *bdl, const double *vect, const double *normt,
             double mu0, double mb0, const double *fc, double g0, const double *w0,
             const double *z0, const double *hb, double dt, int nsteps) {
  int nm = t->nm;
  size_t sz = (size_t)ld * ne;
  double *s1 = calloc(3 * sz, 8), *k = calloc(12 * sz, 8);
  double *u1 = s1, *v1 = s1 + sz, *h1 = s1 + 2 * sz;
  for (int ite = 0; ite < nsteps; ite++) {
    for (int s = 0; s < 4; s++) {
      double *ku = k + (3 * s) * sz, *kv = k + (3 * s + 1) * sz, *kh = k + (3 * s + 2) * sz;
      if (s == 0)
        orc_cal_var(t, variant, nf, ne, ld, u0, v0, h0, ku, kv, kh, tseg, ssit, itm, its, norm,
                    bdl, vect, normt, mu0, mb0, fc, g0, w0, z0, hb);
      else
        orc_cal_var(t, variant, nf, ne, ld, u1, v1, h1, ku, kv, kh, tseg, ssit, itm, its, norm,
                    bdl, vect, normt, mu0, mb0, fc, g0, w0, z0, hb);
      if (s < 3)
        for (int i = 0; i < ne; i++)
          for (int j = 0; j < nm; j++) {
            size_t x = j + (size_t)ld * i;
            if (s < 2) {
              u1[x] = u0[x] + dt * ku[x] * 0.5;
              v1[x] = v0[x] + dt * kv[x] * 0.5;
              h1[x] = h0[x] + dt * kh[x] * 0.5;
            } else {
              u1[x] = u0[x] + dt * ku[x];
              v1[x] = v0[x] + dt * kv[x];
              h1[x] = h0[x] + dt * kh[x];
            }
          }
    }
    double d6 = dt / 6.0;
    for (int i = 0; i < ne; i++)
      for (int j = 0; j < nm; j++) {
        size_t x = j + (size_t)ld * i;
        double *f[3] = {u0, v0, h0};
        for (int c = 0; c < 3; c++) {
          double k1 = k[(0 + c) * sz + x], k2 = k[(3 + c) * sz + x];
          double k3 = k[(6 + c) * sz + x], k4 = k[(9 + c) * sz + x];
          f[c][x] = f[c][x] + d6 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
        }
      }
  }
  free(s1);
  free(k);
}

/* Global integrals: src/energy.f:3-96. */
void orc_energy(const orc_tables *t, int ne, int ld, const double *u0, const double *v0,
                const double *h0, double g0, const double *hb, const int *in, const double *x,
                const double *y, double *ke, double *pe, double *mass) {
  int nm = t->nm, ng = t->ng;
  double k = 0.0, p = 0.0, ms = 0.0;
  for (int i = 0; i < ne; i++) {
    int i1 = in[3 * i] - 1, i2 = in[3 * i + 1] - 1, i3 = in[3 * i + 2] - 1;
    double ax = (x[i2] - x[i1]) * 0.5, bx = (x[i3] - x[i1]) * 0.5;
    double ay = (y[i2] - y[i1]) * 0.5, by = (y[i3] - y[i1]) * 0.5;
    double deta = ax * by - ay * bx;
    for (int j = 0; j < nm; j++) ms = ms + t->ammo[j] * deta * h0[j + ld * i];
    double mu[LDG], mv[LDG], mh[LDG], mhb[LDG], mff[LDG];
    for (int j = 0; j < ng; j++) mu[j] = mv[j] = mh[j] = mhb[j] = 0.0;
    for (int m = 0; m < nm; m++) {
      double uk = u0[m + ld * i], vk = v0[m + ld * i], hk = h0[m + ld * i];
      double hbk = h0[m + ld * i] + hb[m + ld * i];
      for (int j = 0; j < ng; j++) {
        mu[j] = mu[j] + uk * PI2(t, j, m);
        mv[j] = mv[j] + vk * PI2(t, j, m);
        mh[j] = mh[j] + hk * PI2(t, j, m);
        mhb[j] = mhb[j] + hbk * PI2(t, j, m);
      }
    }
    for (int j = 0; j < ng; j++) mff[j] = mhb[j] * (mu[j] * mu[j] + mv[j] * mv[j]);
    double ans = 0.0;
    for (int j = 0; j < ng; j++) ans = ans + t->wi[j] * mff[j] * PI2(t, j, 0);
    k = k + ans * deta * 0.5;
    for (int j = 0; j < ng; j++) mff[j] = mh[j] * mh[j];
    ans = 0.0;
    for (int j = 0; j < ng; j++) ans = ans + t->wi[j] * mff[j] * PI2(t, j, 0);
    p = p + ans * deta * 0.5 * g0;
  }
  *ke = k;
  *pe = p;
  *mass = ms;
}

size_t orc_tables_size(void) { return sizeof(orc_tables); }

/* Edge-jump indicator of one modal field: src/adapt.f:1260-1332
 * (checkconv).  For every face with two elements, the field's traces at
 * the ngf points from both sides (the second reversed), |difference|, and
 * amax(e) = the largest over e's faces and points; then amax /= max_e
 * |a0(1,e)| when that exceeds 1e-6, else amax = 0. */
void orc_checkconv(const orc_tables *t, int nf, int ne, int ld, const double *a0,
                   const int *tseg, const int *ssit, double *amax) {
  int nm = t->nm, ngf = t->ngf;
  double fbdy1[ORC_LDF], fbdy2[ORC_LDF];
  for (int i = 0; i < ne; i++) amax[i] = 0.0;
  for (int i = 0; i < nf; i++) {
    for (int j = 0; j < ngf; j++) fbdy1[j] = fbdy2[j] = 0.0;
    int k1 = tseg[2 * i] - 1, k2 = tseg[2 * i + 1] - 1;
    if (tseg[2 * i + 1] > 0) {
      int c = ssit[2 * i] - 1;
      for (int m = 0; m < nm; m++)
        for (int j = 0; j < ngf; j++) fbdy1[j] = fbdy1[j] + a0[m + ld * k1] * PICO(t, j, m, c);
      c = ssit[2 * i + 1] - 1;
      for (int m = 0; m < nm; m++)
        for (int j = 0; j < ngf; j++)
          fbdy2[j] = fbdy2[j] + a0[m + ld * k2] * PICO(t, ngf - 1 - j, m, c);
      for (int j = 0; j < ngf; j++) {
        double dif = fabs(fbdy2[j] - fbdy1[j]);
        if (amax[k1] < dif) amax[k1] = dif;
        if (amax[k2] < dif) amax[k2] = dif;
      }
    }
  }
  double vmax = fabs(a0[0]);
  for (int i = 1; i < ne; i++) {
    double v = fabs(a0[ld * i]);
    vmax = v > vmax ? v : vmax;
  }
  if (vmax > 1.e-6) {
    for (int i = 0; i < ne; i++) amax[i] = amax[i] / vmax;
  } else {
    for (int i = 0; i < ne; i++) amax[i] = 0.0;
  }
}

/* Refinement decisions: src/adapt.f:1122-1192 (cal_conv).  fconv(e): some
 * field's jump exceeds fac1; fderaff(e): every field's jump is below fac2;
 * deraff: some fderaff; supmax: the largest jump; timedir = supmax <= fac3
 * (false asks the driver to roll back); conv = supmax > fac1 or deraff.
 * amax3 (3*ne, may be NULL) receives the three indicators (u, v, h). */
void orc_cal_conv(const orc_tables *t, int nf, int ne, int ld, const double *u0,
                  const double *v0, const double *h0, const int *tseg, const int *ssit,
                  double fac1, double fac2, double fac3, int *fconv, int *fderaff, int *conv,
                  int *deraff, int *timedir, double *supmax, double *amax3) {
  double *um = malloc(sizeof(double) * 3 * (size_t)(ne > 0 ? ne : 1));
  double *vm = um + ne, *hm = vm + ne;
  for (int i = 0; i < ne; i++) fconv[i] = fderaff[i] = 0;
  *conv = 0;
  *deraff = 0;
  orc_checkconv(t, nf, ne, ld, u0, tseg, ssit, um);
  orc_checkconv(t, nf, ne, ld, v0, tseg, ssit, vm);
  orc_checkconv(t, nf, ne, ld, h0, tseg, ssit, hm);
  double sup = um[0];
  for (int i = 0; i < ne; i++) {
    if (um[i] > fac1 || vm[i] > fac1 || hm[i] > fac1) fconv[i] = 1;
    if (um[i] < fac2 && vm[i] < fac2 && hm[i] < fac2) {
      fderaff[i] = 1;
      *deraff = 1;
    }
    sup = um[i] > sup ? um[i] : sup;
    sup = vm[i] > sup ? vm[i] : sup;
    sup = hm[i] > sup ? hm[i] : sup;
  }
  *timedir = sup > fac3 ? 0 : 1;
  if (sup > fac1) *conv = 1;
  if (*deraff) *conv = 1;
  *supmax = sup;
  if (amax3)
    for (int i = 0; i < 3 * ne; i++) amax3[i] = um[i];
  free(um);
}

