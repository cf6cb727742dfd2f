/* A stand-in for the five C-ABI entry points the Fortran cal_var shim calls
 * (spoc_ctx_create, spoc_mesh_set, spoc_fields_set, spoc_rhs,
 * spoc_last_error).  It records how often each is called and copies of the
 * operands it receives, so that a CPU test can check when the shim
 * re-uploads the mesh and the fields.  Test infrastructure only. */
#include <stdlib.h>
#include <string.h>

static int n_ctx, n_mesh, n_fields, n_rhs;
static int last_ne, last_nf;
static int *last_itm;
static double *last_vect, *last_w0;

int spoc_ctx_create(void **ctx, int device, int nc, int ng, int ngf, const double *wi,
                    const double *wif, const double *pi2, int ld_g, const double *pico,
                    int ld_f, int ld_m, const double *amm, const double *amml, const double *agx,
                    const double *agy, const double *ammo) {
  static int dummy;
  (void)device; (void)nc; (void)ng; (void)ngf; (void)wi; (void)wif; (void)pi2; (void)ld_g;
  (void)pico; (void)ld_f; (void)ld_m; (void)amm; (void)amml; (void)agx; (void)agy; (void)ammo;
  *ctx = &dummy;
  n_ctx++;
  return 0;
}

int spoc_mesh_set(void *ctx, int ne, int nf, const int *itm, const int *its, const int *tseg,
                  const int *ssit, const double *norm, const double *vect, const double *normt,
                  const int *bdl, int nn, const int *in, const double *xgr, const double *ygr) {
  (void)ctx; (void)its; (void)tseg; (void)ssit; (void)norm; (void)normt; (void)bdl; (void)nn;
  (void)in; (void)xgr; (void)ygr;
  n_mesh++;
  last_ne = ne;
  last_nf = nf;
  free(last_itm);
  free(last_vect);
  last_itm = malloc(sizeof(int) * 3 * (size_t)ne);
  last_vect = malloc(sizeof(double) * 5 * (size_t)ne);
  memcpy(last_itm, itm, sizeof(int) * 3 * (size_t)ne);
  memcpy(last_vect, vect, sizeof(double) * 5 * (size_t)ne);
  return 0;
}

int spoc_fields_set(void *ctx, int ld, const double *fc, const double *w0, const double *z0,
                    const double *hb, double mu0, double mb0, double g0) {
  (void)ctx; (void)fc; (void)z0; (void)hb; (void)mu0; (void)mb0; (void)g0;
  n_fields++;
  free(last_w0);
  last_w0 = malloc(sizeof(double) * (size_t)ld * last_ne);
  memcpy(last_w0, w0, sizeof(double) * (size_t)ld * last_ne);
  return 0;
}

int spoc_rhs(void *ctx, int variant, int precision, int ld, const double *u0, const double *v0,
             const double *h0, double *u1, double *v1, double *h1) {
  (void)ctx; (void)variant; (void)precision; (void)u0; (void)v0; (void)h0;
  n_rhs++;
  memset(u1, 0, sizeof(double) * (size_t)ld * last_ne);
  memset(v1, 0, sizeof(double) * (size_t)ld * last_ne);
  memset(h1, 0, sizeof(double) * (size_t)ld * last_ne);
  return 0;
}

const char *spoc_last_error(void) { return ""; }

/* read back by the test */
void stub_counts(int *out) {
  out[0] = n_ctx;
  out[1] = n_mesh;
  out[2] = n_fields;
  out[3] = n_rhs;
}
int stub_ne(void) { return last_ne; }
const int *stub_itm(void) { return last_itm; }
const double *stub_vect(void) { return last_vect; }
const double *stub_w0(void) { return last_w0; }
