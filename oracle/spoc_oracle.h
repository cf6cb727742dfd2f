/*
 * CPU FP64 restatement of SPOC's RK4 right-hand side -- TEST INFRASTRUCTURE.
 *
 * This is the oracle: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.  The
 * product (spoc_amd/) never links or calls it.
 *
 * Every routine follows the reference loop order and expression order so
 * that, compiled without FMA contraction, it reproduces the reference
 * Fortran bit for bit.  Each function cites the reference file:line it
 * restates.
 *
 * Layout (same as the reference): modal arrays are column-major (ld, ne),
 * element e's mode m at a[m + ld*e]; face arrays are (ORC_LDF, nf).
 * Connectivity is 1-based as in Fortran: tseg[2f+s] in 1..ne (0 = no
 * second element), ssit[2f+s] in 1..3, itm[3e+k] in 1..nf, its[3e+k] = +-1.
 */
#ifndef SPOC_ORACLE_H
#define SPOC_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LDM 45 /* nnmod, include/sp.inc:2 */
#define ORC_LDF 9  /* ngf_max, include/sp.inc:4 */
#define ORC_LDG 64 /* room for the 42-point rule (reference: ng_max = 33) */

enum { ORC_LIN = 0, ORC_FS = 1, ORC_NS = 2, ORC_INV = 3 };

typedef struct {
  int nc, nm, ng, ngf;
  double wi[ORC_LDG], xi[ORC_LDG], yi[ORC_LDG];
  double wif[ORC_LDF], tif[ORC_LDF];
  double pi2[ORC_LDG * ORC_LDM];      /* pi2(q, m)  at [q + LDG*m]           */
  double pico[ORC_LDF * ORC_LDM * 3]; /* pico(j,m,k) at [j + LDF*(m + LDM*k)] */
  double amm[ORC_LDM * ORC_LDM];      /* upper Cholesky factor of the mass matrix */
  double amml[ORC_LDM * ORC_LDM];     /* upper Cholesky factor, reduced basis */
  double agx[ORC_LDM * ORC_LDM], agy[ORC_LDM * ORC_LDM];
  double ammo[ORC_LDM * ORC_LDM];     /* unfactored mass matrix */
  int cip[2 * (ORC_LDM + 1)];         /* cip(s, i) at [s + 2*i], s = 0,1 */
  int info1, info2;                   /* DPOTRF info for amm, amml */
} orc_tables;

size_t orc_tables_size(void);
double orc_po(int n, double x);
double orc_dpo(int n, double x);
int orc_setup(orc_tables *t, int nc, int ng, int ngf, const double *wi, const double *xi,
              const double *yi, const double *wif, const double *tif);

void orc_cal_flu(const orc_tables *t, int nf, const int *tseg, double *fbdy, const double *a0,
                 int ld, const int *ssit);
void orc_dtrsm4(int m, const double *a, double *b, const double *diag);
void orc_reduc_gn(const orc_tables *t, const double *gn, double *gnr);
void orc_reduc_inv_gn(const orc_tables *t, double *gn, const double *gnr);

void orc_cal_var(const orc_tables *t, int variant, int nf, int ne, int ld, const double *u0,
                 const double *v0, const double *h0, double *u1, double *v1, double *h1,
                 const int *tseg, const int *ssit, const int *itm, const int *its,
                 const double *norm, const int *bdl, const double *vect, const double *normt,
                 double mu0, double mb0, const double *fc, double g0, const double *w0,
                 const double *z0, const double *hb);

void orc_rk4(const orc_tables *t, int variant, int nf, int ne, int ld, double *u0, double *v0,
             double *h0, const int *tseg, const int *ssit, const int *itm, const int *its,
             const double *norm, const int *bdl, const double *vect, const double *normt,
             double mu0, double mb0, const double *fc, double g0, const double *w0,
             const double *z0, const double *hb, double dt, int nsteps);

void orc_energy(const orc_tables *t, int ne, int ld, const double *u0, const double *v0,
                const double *h0, double g0, const double *hb, const int *in, const double *x,
                const double *y, double *ke, double *pe, double *mass);

/* checkconv and cal_conv (src/adapt.f:1260-1332, 1122-1192); flags as 0/1 ints. */
void orc_checkconv(const orc_tables *t, int nf, int ne, int ld, const double *a0,
                   const int *tseg, const int *ssit, double *amax);
void orc_cal_conv(const orc_tables *t, int nf, int ne, int ld, const double *u0,
                  const double *v0, const double *h0, const int *tseg, const int *ssit,
                  double fac1, double fac2, double fac3, int *fconv, int *fderaff, int *conv,
                  int *deraff, int *timedir, double *supmax, double *amax3);

#ifdef __cplusplus
}
#endif
#endif
