// Host side of the C-ABI (include/spoc_abi.h): contexts, table and mesh
// ingest, device-resident RK4 driver, timing.  Compiled without FMA
// contraction so the host-side table builder reproduces the reference
// driver's set-up arithmetic exactly.
#include "../../include/spoc_abi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "error.h"
#include "f32_types.h"
#include "launch.h"

using namespace spoc;

namespace spoc {
int adapt_launch(hipStream_t st, int nc, int ngf, int ne, int nmem, int nch,
                 const double *const *a, const int4 *side, const double *pico,
                 const int *off, const int *cst, const int *cof,
                 double *work, double *T, double *amax, bool decide, double fac1, double fac2,
                 int *flags);
int adapt_chunk();
int nonfinite_launch(hipStream_t st, size_t n, const double *a, int *flag);
size_t energy_scratch_doubles();
int energy_scratch(int device, double **out);
struct F64Plan;
int f64_plan_create(F64Plan **out, int nc, int ng, int ngf);
void f64_plan_destroy(F64Plan *p);
int f64_mesh_set(F64Plan *p, int ne, int nf, const int *itm, const int *its, const int *bdl,
                 const double *vect, const double *normt, const int *dtseg, const int *dssit,
                 const double *dnorm, double *dface);
int f64_fields_set(F64Plan *p, hipStream_t s, const DevTables &tab, const double *fc,
                   const double *w0, const double *z0, const double *hb, const double *hh0e,
                   double mu0, double mb0, double g0);
int f64_rhs(F64Plan *p, hipStream_t s, const DevTables &tab, int variant,
            const double *const *x, double *const *k, TimerHook hook, void *user);
int f64_rk4_begin(F64Plan *p, hipStream_t s, const DevTables &tab, const double *const *u0);
int f64_rk4_step(F64Plan *p, hipStream_t s, const DevTables &tab, int variant, double dt,
                 TimerHook hook, void *user);
int f64_rk4_end(F64Plan *p, hipStream_t s, const DevTables &tab, double *const *u0);
int f64_kernels_per_rhs(const F64Plan *p, int variant);
void locality_order(int ne, int nf, const int *itm, const int *tseg, const int *ssit,
                    const double *vect, const std::vector<int> &mem_e0, bool use_hilbert,
                    std::vector<int> &perm);
void face_order(int ne, int nf, const int *itm, const std::vector<int> &perm,
                std::vector<int> &fperm);
double block_share(int nf, const int *tseg, const int *inv);
int energy_launch(hipStream_t st, int ne, int nm, int ng, const double *u, const double *v,
                  const double *h, const double *hb, const double *pi2, const double *wi,
                  const double *ammo1, const double *deta, double g0, int nr, const int *off,
                  const int *order, double *scratch, double *out);
struct F32Plan;
int f32_plan_create(F32Plan **out, int nc, int ng, int ngf, const double *amm, const double *amml,
                    const double *agx, const double *agy, const double *pi2, const double *pico,
                    const double *wi, const double *wif, int ld_g, int ld_f, int ld_m,
                    const double *dU);
void f32_plan_destroy(F32Plan *p);
int f32_mesh_set(F32Plan *p, int ne, int nf, const int *itm, const int *its, const int *tseg,
                 const int *ssit, const double *norm, const double *vect, const double *normt,
                 const int *bdl);
int f32_fields_set(F32Plan *p, int ld, const double *fc, const double *w0, const double *z0,
                   const double *hb, const double *hh0e, const double *pi2, int ld_g, double mu0,
                   double mb0, double g0);
int f32_rhs(F32Plan *p, hipStream_t s, int variant, const double *const *x, double *const *k,
            TimerHook hook, void *user);
int f32_rk4_begin(F32Plan *p, hipStream_t s, int variant, double *const *u0);
int f32_rk4_step(F32Plan *p, hipStream_t s, int variant, double dt, TimerHook hook, void *user);
int f32_rk4_end(F32Plan *p, hipStream_t s, int variant, double *const *u0);
int f32_rk4_persist(F32Plan *p, hipStream_t s, int variant, double dt, int nsteps,
                    const unsigned **abort, TimerHook hook, void *user, bool test_abort);
int f32_autotune(F32Plan *p, hipStream_t s, int variant, int *out);
void f32_set_tiles(F32Plan *p, int variant, int t0, int t1, int autotune);
void f32_set_flux(F32Plan *p, double alpha);
double f32_flux(const F32Plan *p);
int f32_adj_rhs(F32Plan *p, hipStream_t s, int variant, const double *const *x,
                const double *const *lam, double *const *y, double *hb);
int f32_adj_run(F32Plan *p, hipStream_t s, int variant, double dt, int nsteps, int ncheck,
                const double *const *x0, double *const *lam, double *hb);
}  // namespace spoc

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace

// (the other translation units' way to fail(): error.h)
int spoc::report(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      return fail(SPOC_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                \
  } while (0)

// ------------------------------------------------------------------------
// Operator tables from the quadrature rule, in the order of the driver's
// set-up (src/shwwnl.f:127-156).
// ------------------------------------------------------------------------
double legendre(int n, double x) {  // src/legendre2.f:5-30
  double p0 = 1.0, p1 = x, p = p0;
  if (n == 0) return p0;
  if (n == 1) return p1;
  for (int i = 2; i <= n; i++) {
    double a0 = (double)i, a2 = (double)(2 * i - 1), a3 = (double)(i - 1);
    p = (a2 * x * p1 - a3 * p0) / a0;
    p0 = p1;
    p1 = p;
  }
  return p;
}

double dlegendre(int n, double x) {  // src/legendre2.f:32-62
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0, d = d0;
  if (n == 0) return d0;
  if (n == 1) return d1;
  for (int i = 2; i <= n; i++) {
    double a0 = (double)i, a2 = (double)(2 * i - 1), a3 = (double)(i - 1);
    d = (a2 * x * d1 + a2 * p1 - a3 * d0) / a0;
    double p2 = (a2 * x * p1 - a3 * p0) / a0;
    p0 = p1;
    p1 = p2;
    d0 = d1;
    d1 = d;
  }
  return d;
}

// Upper Cholesky in the column order of the unblocked LAPACK routine the
// reference's DPOTRF takes for these sizes (src/solve.f:108-113, 383-403).
int cholesky_upper(int n, double *a, int ld) {
  for (int j = 0; j < n; j++) {
    double dot = 0.0;
    for (int i = 0; i < j; i++) dot = dot + a[i + ld * j] * a[i + ld * j];
    double ajj = a[j + ld * j] - dot;
    if (ajj <= 0.0) {
      a[j + ld * j] = ajj;
      return j + 1;
    }
    ajj = std::sqrt(ajj);
    a[j + ld * j] = ajj;
    if (j + 1 < n) {
      if (j > 0)
        for (int c = j + 1; c < n; c++) {
          double t = 0.0;
          for (int i = 0; i < j; i++) t = t + a[i + ld * c] * a[i + ld * j];
          a[j + ld * c] = a[j + ld * c] + (-1.0) * t;
        }
      double r = 1.0 / ajj;
      for (int c = j + 1; c < n; c++) a[j + ld * c] = r * a[j + ld * c];
    }
  }
  return 0;
}

struct Prof {
  bool on = false;
  std::vector<hipEvent_t> ev;   // pool
  size_t used = 0;
  std::vector<std::pair<int, size_t>> marks;  // (slot, event index)
};

}  // namespace

struct spoc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nc = 0, ng = 0, ngf = 0, nm = 0;
  std::vector<double> h_wi, h_ammo1;
  // device tables
  double *dtab = nullptr;
  DevTables tab{};
  // mesh
  int ne = 0, nf = 0;
  void *dmesh = nullptr;
  DevMesh mesh{};
  bool have_nodes = false;
  // fields
  double *dfields = nullptr;  // fc, w0, z0, hb
  Physics phys{};
  bool have_fields = false;
  // state and work
  double *dstate = nullptr;  // u0 x3, x x3, acc x3, k x3, work x8
  double *dface = nullptr;
  double *den = nullptr;     // energy scratch
  double *dadapt = nullptr;  // refinement indicator scratch (adapt_launch)
  int *dchunk = nullptr;     // its reduction chunks: starts (nch + 1), members' first (nmem + 1)
  int4 *dside = nullptr;     // its face per element side: (k1, c1, k2, c2), 0-based
  int nch = 0;
  // element offsets of the meshes run together: {0, ne} for one mesh, or the
  // members of an ensemble (spoc_ensemble_mesh_set); device copy holds
  // {0, ne, mem_e0...}
  std::vector<int> mem_e0;
  int *dmem = nullptr;
  double *dadj = nullptr;    // adjoint state, reference layout (nm, ne) x 3, then hb's adjoint
  bool have_adj = false, have_adj_hb = false;
  int ncheck = 0;            // checkpoint interval of the nonlinear adjoint (0: automatic)
  bool have_state = false;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  double last_ms = 0.0;
  Prof prof;
  std::vector<std::string> knames;
  std::vector<double> kms;
  std::vector<int> kcount;
  F32Plan *f32 = nullptr;
  F64Plan *f64 = nullptr;
  std::vector<double> h_pi2;  // pi2(ng, nm) for projecting fields onto quadrature points
  // `steps` RK4 steps captured as a graph, valid while its key matches
  hipGraphExec_t graph = nullptr;
  struct GraphKey {
    int variant, precision;
    double dt;
    long version;
    int steps;
    bool operator==(const GraphKey &o) const {
      return variant == o.variant && precision == o.precision && dt == o.dt &&
             version == o.version && steps == o.steps;
    }
  } graph_key{-1, -1, 0.0, -1, 0};
  long version = 0;     // bumped by every mesh/field change
  int use_graph = 1;    // spoc_set_graph
  int use_persist = 1;  // spoc_set_persistent (2: with the abort word raised)
  QueueBound queue;     // dispatches queued ahead of the host in spoc_rk4_steps
  int *dflag = nullptr; // device word of the non-finite test
  // internal element numbering (spoc_set_ordering): perm[internal] = the
  // caller's element, iperm its inverse; both empty when they coincide.
  // dorder: device copy of iperm (energy sums run in the caller's order)
  int ordering = SPOC_ORDER_AUTO;
  // pinned host staging of the host <-> device copies: room for 4 fields of
  // (nm, ne); completion events of the copies (fields x pieces)
  double *stage = nullptr;
  size_t stage_cap = 0;
  hipEvent_t fev[4 * 4] = {};
  std::vector<int> perm, iperm;
  int *dorder = nullptr;
};

namespace {

// RK4 steps per captured graph (see spoc_rk4_steps)
constexpr int kGraphSteps = 8;

// Destroys the captured graph once every launch of it has finished.
void destroy_graph(spoc_ctx *c) {
  if (!c->graph) return;
  (void)hipStreamSynchronize(c->stream);
  (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
}

void drop_graph(spoc_ctx *c) {
  destroy_graph(c);
  c->version++;
}

// RK4 kernel dispatches of one step: per right-hand side, the FP64 path runs
// a face and an element pass (two of each for the nonlinear variants), the
// mixed path one element kernel (two for the nonlinear variants).
int step_dispatches(int variant, int precision) {
  const int per_rhs = precision == SPOC_FP64 ? (variant == SPOC_LIN ? 2 : 4)
                                             : (variant == SPOC_LIN ? 1 : 2);
  return 4 * per_rhs;
}

void free_mesh(spoc_ctx *c) {
  (void)hipStreamSynchronize(c->stream);  // (copies may still target what is freed)
  drop_graph(c);
  if (c->dmesh) (void)hipFree(c->dmesh);
  if (c->dmem) (void)hipFree(c->dmem);
  c->dmem = nullptr;
  if (c->dadj) (void)hipFree(c->dadj);
  c->dadj = nullptr;
  c->have_adj = c->have_adj_hb = false;
  if (c->dstate) (void)hipFree(c->dstate);
  if (c->dface) (void)hipFree(c->dface);
  if (c->den) (void)hipFree(c->den);
  if (c->dadapt) (void)hipFree(c->dadapt);
  c->dadapt = nullptr;
  if (c->dchunk) (void)hipFree(c->dchunk);
  c->dchunk = nullptr;
  if (c->dside) (void)hipFree(c->dside);
  c->dside = nullptr;
  if (c->dorder) (void)hipFree(c->dorder);
  c->dorder = nullptr;
  if (c->stage) (void)hipHostFree(c->stage);
  c->stage = nullptr;
  c->stage_cap = 0;
  c->perm.clear();
  c->iperm.clear();
  if (c->dfields) (void)hipFree(c->dfields);
  c->dmesh = nullptr;
  c->dstate = nullptr;
  c->dface = nullptr;
  c->den = nullptr;
  c->dfields = nullptr;
  c->have_state = c->have_fields = false;
}

double *state_ptr(spoc_ctx *c, int slot) { return c->dstate + (size_t)slot * c->nm * c->ne; }

void prof_hook(void *user, int slot, hipStream_t st) {
  spoc_ctx *c = (spoc_ctx *)user;
  if (!c->prof.on) return;
  if (c->prof.used >= c->prof.ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    c->prof.ev.push_back(e);
  }
  size_t i = c->prof.used++;
  (void)hipEventRecord(c->prof.ev[i], st);
  c->prof.marks.push_back({slot, i});
}

int f32_fail(int rc) {
  if (rc == -1)
    return fail(SPOC_EINVAL,
                "FP32 path: a wall face is not local edge 1 of its element (prepare the mesh "
                "with the CM corner fix, src/chmal.f) or the shape is unsupported");
  if (rc == -3) return fail(SPOC_ESTATE, "FP32 path: mesh and fields must be set first");
  hipError_t e = hipGetLastError();
  return fail(SPOC_EHIP, "FP32 kernel launch failed: %s", hipGetErrorString(e));
}

// Host threads for the repacking between the caller's (ld, ne) arrays and
// the packed (nm, ne) layout, which is bound by host memory bandwidth (one
// thread does not fill it): up to 16.
int host_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::min<unsigned>(16u, hw ? hw : 1u);
}

}  // namespace

extern "C" {

const char *spoc_last_error(void) { return g_err.c_str(); }
int spoc_version(void) { return 1; }

int spoc_tables_build(int nc, int ng, int ngf, const double *wi, const double *xi,
                      const double *yi, const double *wif, const double *tif, int ld_g,
                      int ld_f, int ld_m, double *pi2, double *pico, double *amm, double *amml,
                      double *agx, double *agy, double *ammo) {
  return spoc::guarded([&]() -> int {
    (void)wif;
    int nm = nm_of(nc);
    if (nc < 1 || nc > 8 || ng < 1 || ngf < 1 || ld_g < ng || ld_f < ngf || ld_m < nm)
      return fail(SPOC_EINVAL, "spoc_tables_build: bad sizes nc=%d ng=%d ngf=%d", nc, ng, ngf);
    // basis phi_m = P_a(xi) P_b(eta), m = ip(nc, a) + b (src/nonlin.f:19-28)
    int m = 0;
    for (int a = 0; a <= nc; a++)
      for (int b = 0; b <= nc - a; b++, m++) {
        for (int q = 0; q < ng; q++) pi2[q + (size_t)ld_g * m] = legendre(a, xi[q]) * legendre(b, yi[q]);
        for (int j = 0; j < ngf; j++) {
          double t = tif[j];
          pico[j + (size_t)ld_f * (m + (size_t)ld_m * 0)] = legendre(a, t) * legendre(b, -1.0);
          pico[j + (size_t)ld_f * (m + (size_t)ld_m * 1)] = legendre(a, -t) * legendre(b, t);
          pico[j + (size_t)ld_f * (m + (size_t)ld_m * 2)] = legendre(a, -1.0) * legendre(b, -t);
        }
      }
    // mass and derivative matrices by quadrature (src/mat1.f:90-142)
    std::vector<double> M0((size_t)ld_m * ld_m, 0.0);
    for (double *mat : {M0.data(), agx, agy})
      for (int i = 0; i < ld_m * ld_m; i++) mat[i] = 0.0;
    int i1 = 0;
    for (int k1 = 0; k1 <= nc; k1++)
      for (int k2 = 0; k2 <= nc - k1; k2++, i1++) {
        int i2 = 0;
        for (int k3 = 0; k3 <= nc; k3++)
          for (int k4 = 0; k4 <= nc - k3; k4++, i2++) {
            double s0 = 0.0, sx = 0.0, sy = 0.0;
            for (int q = 0; q < ng; q++) {
              double x = xi[q], y = yi[q];
              s0 = s0 + wi[q] * (legendre(k1, x) * legendre(k2, y) * legendre(k3, x) * legendre(k4, y));
              sx = sx + wi[q] * (legendre(k1, x) * legendre(k2, y) * dlegendre(k3, x) * legendre(k4, y));
              sy = sy + wi[q] * (legendre(k1, x) * legendre(k2, y) * legendre(k3, x) * dlegendre(k4, y));
            }
            M0[i1 + (size_t)ld_m * i2] = s0;
            agx[i1 + (size_t)ld_m * i2] = sx;
            agy[i1 + (size_t)ld_m * i2] = sy;
          }
      }
    // reduced-basis mass matrix (src/mat1.f:285-368)
    for (int i = 0; i < ld_m * ld_m; i++) amml[i] = 0.0;
    const int js[4] = {0, 1, 0, 1}, ls[4] = {0, 0, 1, 1};
    for (int pass = 0; pass < 4; pass++)
      for (int i = 0; i <= nc - 1; i++)
        for (int j = js[pass]; j <= nc - 1 - i; j += 2)
          for (int k = 0; k <= nc - 1; k++)
            for (int l = ls[pass]; l <= nc - 1 - k; l += 2) {
              int r = ip_of(nc - 1, i) + j, c = ip_of(nc - 1, k) + l;
              int ri = ip_of(nc, i) + j + 1, ci = ip_of(nc, k) + l + 1;
              int r0 = ip_of(nc, i), c0 = ip_of(nc, k);
              double a22 = M0[ri + (size_t)ld_m * ci], a02 = M0[r0 + (size_t)ld_m * ci];
              double a20 = M0[ri + (size_t)ld_m * c0], a00 = M0[r0 + (size_t)ld_m * c0];
              double v = pass == 0   ? a22 + a02 + a20 + a00
                         : pass == 1 ? a22 - a02 + a20 - a00
                         : pass == 2 ? a22 + a02 - a20 - a00
                                     : a22 - a02 - a20 + a00;
              amml[r + (size_t)ld_m * c] = v;
            }
    for (int i = 0; i < ld_m * ld_m; i++) {
      ammo[i] = M0[i];
      amm[i] = M0[i];
    }
    int info1 = cholesky_upper(nm, amm, ld_m);
    int info2 = cholesky_upper(nm - nc - 1, amml, ld_m);
    if (info1 || info2)
      return fail(SPOC_ESINGULAR,
                  "mass matrix not positive definite (p=%d with a %d-point rule: need >= %d points)",
                  nc, ng, nm);
    return SPOC_OK;
  });
}

int spoc_ctx_create(spoc_ctx **out, int device, int nc, int ng, int ngf, const double *wi,
                    const double *wif, const double *pi2, int ld_g, const double *pico,
                    int ld_f, int ld_m, const double *amm, const double *amml,
                    const double *agx, const double *agy, const double *ammo) {
  return spoc::guarded([&]() -> int {
    if (!out) return fail(SPOC_EINVAL, "spoc_ctx_create: null ctx pointer");
    *out = nullptr;
    int nm = nm_of(nc);
    bool ok = false;
#define SPOC_OKSHAPE(NC_, NG_, NGF_) ok = ok || (nc == NC_ && ng == NG_ && ngf == NGF_);
    SPOC_FOR_EACH_SHAPE(SPOC_OKSHAPE)
#undef SPOC_OKSHAPE
    if (!ok)
      return fail(SPOC_EINVAL,
                  "unsupported (p, triangle points, edge points) = (%d, %d, %d); supported: "
                  "(2,6,3) (3,12,4) (4,25,5) (5,25,6) (6,33,7) (7,42,8)",
                  nc, ng, ngf);
    if (ld_g < ng || ld_f < ngf || ld_m < nm) return fail(SPOC_EINVAL, "leading dimension too small");
    for (int i = 0; i < nm; i++)
      if (!(amm[i + (size_t)ld_m * i] > 0.0))
        return fail(SPOC_ESINGULAR, "amm factor has a non-positive diagonal (DPOTRF failed?)");
    HIPCHK(hipSetDevice(device));
    spoc_ctx *c = new spoc_ctx();
    c->device = device;
    c->nc = nc;
    c->ng = ng;
    c->ngf = ngf;
    c->nm = nm;
    int nl = nm - nc - 1;
    // host-side layout conversion
    std::vector<double> agxT(nm * nm), agyT(nm * nm), p2(ng * nm), p2T(nm * ng),
        pc(ngf * nm * 3), pcT(nm * ngf * 3), U(nm * nm, 0.0), UL(nm * nm, 0.0), d1(nm), d2(nm, 0.0),
        w(ng), wf(ngf), a1(nm);
    for (int j = 0; j < nm; j++)
      for (int i = 0; i < nm; i++) {
        agxT[i + nm * j] = agx[j + (size_t)ld_m * i];
        agyT[i + nm * j] = agy[j + (size_t)ld_m * i];
      }
    for (int m = 0; m < nm; m++)
      for (int q = 0; q < ng; q++) {
        p2[q + ng * m] = pi2[q + (size_t)ld_g * m];
        p2T[m + nm * q] = pi2[q + (size_t)ld_g * m];
      }
    for (int k = 0; k < 3; k++)
      for (int m = 0; m < nm; m++)
        for (int j = 0; j < ngf; j++) {
          double v = pico[j + (size_t)ld_f * (m + (size_t)ld_m * k)];
          pc[j + ngf * (m + nm * k)] = v;
          pcT[m + nm * (j + ngf * k)] = v;
        }
    for (int i = 0; i < nm; i++)
      for (int k = 0; k <= i; k++) U[k + nm * i] = amm[k + (size_t)ld_m * i];
    for (int i = 0; i < nl; i++)
      for (int k = 0; k <= i; k++) UL[k + nm * i] = amml[k + (size_t)ld_m * i];
    for (int i = 0; i < nm; i++) d1[i] = 1.0 / amm[i + (size_t)ld_m * i];
    for (int i = 0; i < nl; i++) d2[i] = 1.0 / amml[i + (size_t)ld_m * i];
    for (int q = 0; q < ng; q++) w[q] = wi[q];
    for (int j = 0; j < ngf; j++) wf[j] = wif[j];
    for (int i = 0; i < nm; i++) a1[i] = ammo[i];
    c->h_wi = w;
    c->h_ammo1 = a1;
    std::vector<std::vector<double> *> parts = {&agxT, &agyT, &p2, &p2T, &pc, &pcT, &U,
                                                &UL,   &d1,   &d2, &w,   &wf,  &a1};
    size_t tot = 0;
    for (auto *p : parts) tot += (p->size() + 31) / 32 * 32;
    hipError_t he = hipStreamCreate(&c->stream);
    if (he == hipSuccess) he = hipMalloc(&c->dtab, tot * sizeof(double));
    if (he != hipSuccess) {
      delete c;
      return fail(SPOC_EHIP, "device allocation failed: %s", hipGetErrorString(he));
    }
    std::vector<double> host(tot, 0.0);
    std::vector<const double *> dptr;
    size_t off = 0;
    for (auto *p : parts) {
      std::memcpy(host.data() + off, p->data(), p->size() * sizeof(double));
      dptr.push_back(c->dtab + off);
      off += (p->size() + 31) / 32 * 32;
    }
    he = hipMemcpy(c->dtab, host.data(), tot * sizeof(double), hipMemcpyHostToDevice);
    if (he != hipSuccess) {
      (void)hipFree(c->dtab);
      delete c;
      return fail(SPOC_EHIP, "table upload failed: %s", hipGetErrorString(he));
    }
    c->tab = DevTables{dptr[0], dptr[1], dptr[2], dptr[3], dptr[4],  dptr[5], dptr[6],
                       dptr[7], dptr[8], dptr[9], dptr[10], dptr[11], dptr[12]};
    (void)hipEventCreate(&c->t0);
    (void)hipEventCreate(&c->t1);
    if (hipMalloc((void **)&c->dflag, 256) != hipSuccess) {
      spoc_ctx_destroy(c);
      return fail(SPOC_EHIP, "device allocation failed");
    }
    c->h_pi2.assign((size_t)ng * nm, 0.0);
    for (int m = 0; m < nm; m++)
      for (int q = 0; q < ng; q++) c->h_pi2[q + (size_t)ng * m] = pi2[q + (size_t)ld_g * m];
    if (f32_plan_create(&c->f32, nc, ng, ngf, amm, amml, agx, agy, pi2, pico, wi, wif, ld_g, ld_f,
                        ld_m, c->tab.U) != 0) {
      spoc_ctx_destroy(c);
      return fail(SPOC_EHIP, "FP32 operator upload failed");
    }
    f64_plan_create(&c->f64, nc, ng, ngf);
    *out = c;
    return SPOC_OK;
  });
}

int spoc_ctx_destroy(spoc_ctx *c) {
  return spoc::guarded([&]() -> int {
    if (!c) return SPOC_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_mesh(c);
    f32_plan_destroy(c->f32);
    f64_plan_destroy(c->f64);
    if (c->dtab) (void)hipFree(c->dtab);
    if (c->dflag) (void)hipFree(c->dflag);
    if (c->stage) (void)hipHostFree(c->stage);
    for (auto e : c->fev)
      if (e) (void)hipEventDestroy(e);
    for (auto e : c->prof.ev) (void)hipEventDestroy(e);
    if (c->t0) (void)hipEventDestroy(c->t0);
    if (c->t1) (void)hipEventDestroy(c->t1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SPOC_OK;
  });
}

static int mesh_set(spoc_ctx *c, int ne, int nf, const int *itm, const int *its,
                    const int *tseg, const int *ssit, const double *norm, const double *vect,
                    const double *normt, const int *bdl, int nn, const int *in,
                    const double *xgr, const double *ygr, const std::vector<int> &mem_e0) {
  if (!c) return fail(SPOC_EINVAL, "null context");
  if (ne <= 0 || nf <= 0) return fail(SPOC_EINVAL, "spoc_mesh_set: ne=%d nf=%d", ne, nf);
  // validate connectivity on the host before any kernel indexes with it
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < 3; k++) {
      int f = itm[3 * e + k];
      if (f < 1 || f > nf) return fail(SPOC_EINVAL, "itm(%d,%d)=%d out of range", k + 1, e + 1, f);
      if (its[3 * e + k] != 1 && its[3 * e + k] != -1)
        return fail(SPOC_EINVAL, "its(%d,%d)=%d is not +-1", k + 1, e + 1, its[3 * e + k]);
    }
  for (int f = 0; f < nf; f++) {
    int a = tseg[2 * f], b = tseg[2 * f + 1];
    if (a < 1 || a > ne || b < 0 || b > ne)
      return fail(SPOC_EINVAL, "tseg(:,%d)=(%d,%d) out of range", f + 1, a, b);
    int s1 = ssit[2 * f], s2 = ssit[2 * f + 1];
    if (s1 < 1 || s1 > 3 || (b > 0 && (s2 < 1 || s2 > 3)))
      return fail(SPOC_EINVAL, "ssit(:,%d)=(%d,%d) out of range", f + 1, s1, s2);
  }
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_mesh(c);
  int nm = c->nm, ngf = c->ngf;
  // the internal numbering: elements along a space-filling curve of their
  // centroids inside each member, faces in order of first appearance; every
  // operand is renumbered here, and upload, download, energies and the
  // refinement indicator map back to the caller's numbering
  std::vector<int> pitm, pits, ptseg, pssit, pbdl, pin;
  std::vector<double> pnorm, pvect, pnormt;
  if (c->ordering != SPOC_ORDER_INPUT) {
    std::vector<int> perm, fperm;
    locality_order(ne, nf, itm, tseg, ssit, vect, mem_e0, c->ordering != SPOC_ORDER_ZCURVE,
                   perm);
    face_order(ne, nf, itm, perm, fperm);
    bool same = true;
    for (int e = 0; e < ne && same; e++) same = perm[e] == e;
    for (int f = 0; f < nf && same; f++) same = fperm[f] == f;
    std::vector<int> iperm(ne);
    for (int e = 0; e < ne; e++) iperm[perm[e]] = e;
    // automatic: keep the caller's numbering when it already puts at least
    // as many face neighbours in one block of 16 as the curve does
    if (!same && c->ordering == SPOC_ORDER_AUTO &&
        block_share(nf, tseg, nullptr) >= block_share(nf, tseg, iperm.data()))
      same = true;
    if (!same) {
      std::vector<int> ifperm(nf);
      for (int f = 0; f < nf; f++) ifperm[fperm[f]] = f;
      pitm.resize(3 * (size_t)ne);
      pits.resize(3 * (size_t)ne);
      pvect.resize(5 * (size_t)ne);
      pnormt.resize(2 * (size_t)ne);
      pbdl.resize(ne);
      if (in) pin.resize(3 * (size_t)ne);
      for (int en = 0; en < ne; en++) {
        const int e = perm[en];
        for (int k = 0; k < 3; k++) {
          pitm[3 * en + k] = ifperm[itm[3 * e + k] - 1] + 1;
          pits[3 * en + k] = its[3 * e + k];
          if (in) pin[3 * en + k] = in[3 * e + k];
        }
        for (int i = 0; i < 5; i++) pvect[5 * en + i] = vect[5 * e + i];
        for (int i = 0; i < 2; i++) pnormt[2 * en + i] = normt[2 * e + i];
        pbdl[en] = bdl[e];
      }
      ptseg.resize(2 * (size_t)nf);
      pssit.resize(2 * (size_t)nf);
      pnorm.resize(2 * (size_t)nf);
      for (int fn = 0; fn < nf; fn++) {
        const int f = fperm[fn];
        for (int i = 0; i < 2; i++) {  // sides keep their order (the face sums do)
          const int t = tseg[2 * f + i];
          ptseg[2 * fn + i] = t == 0 ? 0 : iperm[t - 1] + 1;
          pssit[2 * fn + i] = ssit[2 * f + i];
          pnorm[2 * fn + i] = norm[2 * f + i];
        }
      }
      itm = pitm.data();
      its = pits.data();
      tseg = ptseg.data();
      ssit = pssit.data();
      norm = pnorm.data();
      vect = pvect.data();
      normt = pnormt.data();
      bdl = pbdl.data();
      if (in) in = pin.data();
      c->perm = std::move(perm);
      c->iperm = std::move(iperm);
    }
  }
  std::vector<double> deta(ne, 0.0);
  c->have_nodes = (in && xgr && ygr && nn > 0);
  if (c->have_nodes) {
    for (int e = 0; e < ne; e++) {  // src/energy.f:37-42
      int i1 = in[3 * e] - 1, i2 = in[3 * e + 1] - 1, i3 = in[3 * e + 2] - 1;
      if (i1 < 0 || i2 < 0 || i3 < 0 || i1 >= nn || i2 >= nn || i3 >= nn)
        return fail(SPOC_EINVAL, "in(:,%d) out of range", e + 1);
      double ax = (xgr[i2] - xgr[i1]) * 0.5, bx = (xgr[i3] - xgr[i1]) * 0.5;
      double ay = (ygr[i2] - ygr[i1]) * 0.5, by = (ygr[i3] - ygr[i1]) * 0.5;
      deta[e] = ax * by - ay * bx;
    }
  }
  size_t ibytes = sizeof(int) * ((size_t)3 * ne * 2 + (size_t)2 * nf * 2 + ne);
  size_t dbytes = sizeof(double) * ((size_t)2 * nf + (size_t)5 * ne + (size_t)2 * ne + ne);
  char *blob = nullptr;
  HIPCHK(hipMalloc((void **)&blob, dbytes + ibytes + 256));
  c->dmesh = blob;
  double *d = (double *)blob;
  double *dnorm = d, *dvect = dnorm + 2 * (size_t)nf, *dnormt = dvect + 5 * (size_t)ne,
         *ddeta = dnormt + 2 * (size_t)ne;
  int *ip = (int *)(ddeta + ne);
  int *ditm = ip, *dits = ditm + 3 * (size_t)ne, *dtseg = dits + 3 * (size_t)ne,
      *dssit = dtseg + 2 * (size_t)nf, *dbdl = dssit + 2 * (size_t)nf;
  std::vector<int> b(ne), ss(2 * (size_t)nf);
  for (int e = 0; e < ne; e++) b[e] = bdl[e] != 0;
  for (size_t i = 0; i < 2 * (size_t)nf; i++) ss[i] = ssit[i];
  for (int f = 0; f < nf; f++)
    if (tseg[2 * f + 1] == 0) ss[2 * f + 1] = 1;  // unused on walls; keep in range
  HIPCHK(hipMemcpy(dnorm, norm, sizeof(double) * 2 * nf, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dvect, vect, sizeof(double) * 5 * ne, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dnormt, normt, sizeof(double) * 2 * ne, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ddeta, deta.data(), sizeof(double) * ne, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ditm, itm, sizeof(int) * 3 * ne, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dits, its, sizeof(int) * 3 * ne, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dtseg, tseg, sizeof(int) * 2 * nf, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dssit, ss.data(), sizeof(int) * 2 * nf, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dbdl, b.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
  c->ne = ne;
  c->nf = nf;
  c->mesh = DevMesh{ne, nf, ditm, dits, dtseg, dssit, dnorm, dvect, dnormt, dbdl, ddeta};
  // state: u0(3) x(3) acc(3) k(3) work(8)
  HIPCHK(hipMalloc((void **)&c->dstate, sizeof(double) * (size_t)20 * nm * ne));
  HIPCHK(hipMemset(c->dstate, 0, sizeof(double) * (size_t)20 * nm * ne));
  HIPCHK(hipMalloc((void **)&c->dface, sizeof(double) * (size_t)nf * 6 * ngf));
  const int nmem = (int)mem_e0.size() - 1;
  HIPCHK(hipMalloc((void **)&c->den, sizeof(double) * ((size_t)ne * (2 + nm) + 3 * (nmem + 1) + 8)));
  // refinement indicator: reduction chunks of consecutive elements inside
  // one member; scratch = partial and member maxima (5 nch + 5 nmem),
  // indicators (3 ne), flags (2 ne ints)
  {
    std::vector<int> cst, cof;
    for (int r = 0; r < nmem; r++) {
      cof.push_back((int)cst.size());
      for (int e = mem_e0[r]; e < mem_e0[r + 1]; e += adapt_chunk()) cst.push_back(e);
    }
    c->nch = (int)cst.size();
    cof.push_back(c->nch);
    cst.push_back(ne);
    cst.insert(cst.end(), cof.begin(), cof.end());
    HIPCHK(hipMalloc((void **)&c->dchunk, sizeof(int) * cst.size()));
    HIPCHK(hipMemcpy(c->dchunk, cst.data(), sizeof(int) * cst.size(), hipMemcpyHostToDevice));
    std::vector<int4> side(3 * (size_t)ne);
    for (size_t t = 0; t < side.size(); t++) {
      const int i = itm[t] - 1, k2 = tseg[2 * i + 1] - 1;
      side[t] = make_int4(tseg[2 * i] - 1, ssit[2 * i] - 1, k2, k2 >= 0 ? ssit[2 * i + 1] - 1 : 0);
    }
    HIPCHK(hipMalloc((void **)&c->dside, sizeof(int4) * side.size()));
    HIPCHK(hipMemcpy(c->dside, side.data(), sizeof(int4) * side.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void **)&c->dadapt,
                     sizeof(double) * (5 * (size_t)c->nch + 5 * (size_t)nmem + 4 * (size_t)ne +
                                       9 * (size_t)ngf * ne + 8)));
  }
  std::vector<int> off = {0, ne};
  off.insert(off.end(), mem_e0.begin(), mem_e0.end());
  HIPCHK(hipMalloc((void **)&c->dmem, sizeof(int) * off.size()));
  HIPCHK(hipMemcpy(c->dmem, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
  c->mem_e0 = mem_e0;
  if (!c->iperm.empty()) {
    HIPCHK(hipMalloc((void **)&c->dorder, sizeof(int) * ne));
    HIPCHK(hipMemcpy(c->dorder, c->iperm.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMalloc((void **)&c->dadj, sizeof(double) * (size_t)4 * nm * ne));
  HIPCHK(hipMalloc((void **)&c->dfields, sizeof(double) * ((size_t)4 * nm * ne + ne)));
  HIPCHK(hipMemset(c->dfields, 0, sizeof(double) * ((size_t)4 * nm * ne + ne)));
  if (f32_mesh_set(c->f32, ne, nf, itm, its, tseg, ssit, norm, vect, normt, bdl))
    return fail(SPOC_EHIP, "FP32 mesh upload failed");
  if (f64_mesh_set(c->f64, ne, nf, itm, its, b.data(), vect, normt, dtseg, dssit, dnorm, c->dface))
    return fail(SPOC_EHIP, "FP64 mesh upload failed");
  return SPOC_OK;
}

int spoc_mesh_set(spoc_ctx *c, int ne, int nf, const int *itm, const int *its, const int *tseg,
                  const int *ssit, const double *norm, const double *vect, const double *normt,
                  const int *bdl, int nn, const int *in, const double *xgr, const double *ygr) {
  return spoc::guarded([&]() -> int {
    return mesh_set(c, ne, nf, itm, its, tseg, ssit, norm, vect, normt, bdl, nn, in, xgr, ygr,
                    {0, ne});
  });
}

int spoc_ensemble_mesh_set(spoc_ctx *c, int nmem, const int *ne_m, const int *nf_m,
                           const int *nn_m, const int *itm, const int *its, const int *tseg,
                           const int *ssit, const double *norm, const double *vect,
                           const double *normt, const int *bdl, const int *in,
                           const double *xgr, const double *ygr) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    if (nmem < 1 || !ne_m || !nf_m) return fail(SPOC_EINVAL, "spoc_ensemble_mesh_set: nmem=%d", nmem);
    const bool nodes = nn_m && in && xgr && ygr;
    long ne = 0, nf = 0, nn = 0;
    std::vector<int> e0(nmem + 1, 0);
    for (int m = 0; m < nmem; m++) {
      if (ne_m[m] <= 0 || nf_m[m] <= 0 || (nodes && nn_m[m] <= 0))
        return fail(SPOC_EINVAL, "member %d: ne=%d nf=%d", m, ne_m[m], nf_m[m]);
      ne += ne_m[m];
      nf += nf_m[m];
      if (nodes) nn += nn_m[m];
      e0[m + 1] = (int)ne;
      if (ne > (1L << 30) || nf > (1L << 30)) return fail(SPOC_EINVAL, "ensemble too large");
    }
    // union mesh: member m's 1-based element, face and node numbers shifted by
    // the totals of the members before it (0, a wall, stays 0)
    std::vector<int> uitm(3 * ne), utseg(2 * nf), uin(nodes ? 3 * ne : 0);
    long eo = 0, fo = 0, no = 0;
    for (int m = 0; m < nmem; m++) {
      for (long i = 3 * eo; i < 3 * (eo + ne_m[m]); i++) {
        uitm[i] = itm[i] + (int)fo;
        if (nodes) uin[i] = in[i] + (int)no;
      }
      for (long i = 2 * fo; i < 2 * (fo + nf_m[m]); i++)
        utseg[i] = tseg[i] == 0 ? 0 : tseg[i] + (int)eo;
      eo += ne_m[m];
      fo += nf_m[m];
      if (nodes) no += nn_m[m];
    }
    return mesh_set(c, (int)ne, (int)nf, uitm.data(), its, utseg.data(), ssit, norm, vect, normt,
                    bdl, nodes ? (int)nn : 0, nodes ? uin.data() : nullptr, nodes ? xgr : nullptr,
                    nodes ? ygr : nullptr, e0);
  });
}

// The caller's element of internal element e.
static inline size_t caller_elem(const spoc_ctx *c, int e) {
  return c->perm.empty() ? (size_t)e : (size_t)c->perm[e];
}

// Pinned staging for n fields of (nm, ne) doubles, and the per-field events.
static int ensure_stage(spoc_ctx *c, int n) {
  const size_t want = (size_t)n * c->nm * c->ne;
  if (c->stage_cap < want) {
    // (copies of an earlier call may still read or write the old staging)
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->stage) (void)hipHostFree(c->stage);
    c->stage = nullptr;
    c->stage_cap = 0;
    const size_t cap = std::max(want, (size_t)4 * c->nm * c->ne);
    HIPCHK(hipHostMalloc((void **)&c->stage, cap * sizeof(double), hipHostMallocDefault));
    c->stage_cap = cap;
  }
  for (auto &e : c->fev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return SPOC_OK;
}

// Host (ld, ne) arrays in the caller's numbering -> device (nm, ne) arrays in
// the internal numbering.  Each host thread owns a slice of the elements:
// it packs its slice of every field into pinned staging and queues that
// slice's copy at once, so the copies run while the other slices are being
// packed (the stream takes copies from several threads; their order does
// not matter, the slices do not overlap).
static int upload_fields(spoc_ctx *c, int ld, const double *const *src, double *const *dst,
                         int n) {
  const int nm = c->nm, ne = c->ne;
  const size_t fsz = (size_t)nm * ne;
  int rc = ensure_stage(c, n);
  if (rc) return rc;
  // the staging may still feed an earlier call's copies
  HIPCHK(hipStreamSynchronize(c->stream));
  const int nt = ne < 4096 ? 1 : host_threads();
  // a small mesh whose fields land side by side on the device goes in one
  // copy (its fixed cost is most of the time there)
  bool one = nt == 1;
  for (int f = 0; f < n && one; f++) one = src[f] && dst[f] == dst[0] + f * fsz;
  std::atomic<int> bad{0};
  auto work = [&](int t) {
    const int e0 = (int)((long long)ne * t / nt), e1 = (int)((long long)ne * (t + 1) / nt);
    for (int f = 0; f < n && e1 > e0; f++) {
      if (!src[f]) continue;
      double *st = c->stage + f * fsz;
      const double *s = src[f];
      for (int e = e0; e < e1; e++)
        std::memcpy(st + (size_t)e * nm, s + (size_t)ld * caller_elem(c, e), sizeof(double) * nm);
      if (one) continue;
      if (hipMemcpyAsync(dst[f] + (size_t)e0 * nm, st + (size_t)e0 * nm,
                         sizeof(double) * nm * (e1 - e0), hipMemcpyHostToDevice,
                         c->stream) != hipSuccess)
        bad = 1;
    }
    if (one && hipMemcpyAsync(dst[0], c->stage, sizeof(double) * fsz * n, hipMemcpyHostToDevice,
                              c->stream) != hipSuccess)
      bad = 1;
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto &t : th) t.join();
  // (no wait for the copies: the stream orders everything after them, the
  // staging is reused only by later copies on the same stream or after the
  // wait at the top of the next upload, and it is freed only after a wait)
  if (bad) {
    (void)hipStreamSynchronize(c->stream);
    return fail(SPOC_EHIP, "host-to-device copy failed");
  }
  return SPOC_OK;
}

// Device (nm, ne) arrays -> host (ld, ne) arrays.  Every field is copied in
// kPieces pieces of elements (one on a small mesh), each with its
// completion event, all queued at once; each host thread owns a slice of the elements and unpacks it from
// every field as soon as the piece holding it has arrived, so the unpacking
// (the bound: it writes 28 of every 45 doubles of the caller's arrays)
// runs while the later pieces are still being copied.
static int download_fields(spoc_ctx *c, int ld, double *const *dst, double *const *src, int n) {
  const int nm = c->nm, ne = c->ne;
  // (a small mesh, unpacked by one thread, gains nothing from pieces: one
  // copy per field)
  const int kPieces = ne < 4096 ? 1 : 4;
  const size_t fsz = (size_t)nm * ne;
  int rc = ensure_stage(c, n);
  if (rc) return rc;
  auto cut = [&](int k, int parts) { return (int)((long long)ne * k / parts); };
  // (a small mesh whose fields lie side by side on the device: one copy,
  // and every field's event marks its end)
  bool one = kPieces == 1;
  for (int f = 0; f < n && one; f++) one = src[f] == src[0] + f * fsz;
  if (one) {
    HIPCHK(hipMemcpyAsync(c->stage, src[0], sizeof(double) * fsz * n, hipMemcpyDeviceToHost,
                          c->stream));
    for (int f = 0; f < n; f++) HIPCHK(hipEventRecord(c->fev[f], c->stream));
  }
  for (int f = 0; f < n && !one; f++)
    for (int p = 0; p < kPieces; p++) {
      const size_t e0 = cut(p, kPieces), e1 = cut(p + 1, kPieces);
      if (e1 > e0)
        HIPCHK(hipMemcpyAsync(c->stage + f * fsz + e0 * nm, src[f] + e0 * nm,
                              sizeof(double) * nm * (e1 - e0), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipEventRecord(c->fev[f * kPieces + p], c->stream));
    }
  const int nt = ne < 4096 ? 1 : host_threads();
  std::atomic<int> bad{0};
  auto work = [&](int t) {
    const int e0 = cut(t, nt), e1 = cut(t + 1, nt);
    for (int f = 0; f < n && e1 > e0; f++) {
      const double *st = c->stage + f * fsz;
      double *d = dst[f];
      // the pieces that hold [e0, e1)
      for (int p = 0; p < kPieces; p++) {
        const int a = std::max(e0, cut(p, kPieces)), b = std::min(e1, cut(p + 1, kPieces));
        if (a >= b) continue;
        if (hipEventSynchronize(c->fev[f * kPieces + p]) != hipSuccess) {
          bad = 1;
          return;
        }
        for (int e = a; e < b; e++)
          std::memcpy(d + (size_t)ld * caller_elem(c, e), st + (size_t)e * nm,
                      sizeof(double) * nm);
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto &t : th) t.join();
  if (bad) return fail(SPOC_EHIP, "device-to-host copy failed");
  return SPOC_OK;
}

int spoc_fields_set(spoc_ctx *c, int ld, const double *fc, const double *w0, const double *z0,
                    const double *hb, double mu0, double mb0, double g0) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->dmesh) return fail(SPOC_ESTATE, "spoc_fields_set: set the mesh first");
    if (ld < c->nm || !hb) return fail(SPOC_EINVAL, "spoc_fields_set: ld=%d, hb=%p", ld, hb);
    (void)hipSetDevice(c->device);
    drop_graph(c);
    size_t n = (size_t)c->nm * c->ne;
    double *dst[4] = {c->dfields, c->dfields + n, c->dfields + 2 * n, c->dfields + 3 * n};
    const double *src[4] = {fc, w0, z0, hb};
    HIPCHK(hipMemset(c->dfields, 0, sizeof(double) * 3 * n));
    int rc = upload_fields(c, ld, src, dst, 4);
    if (rc) return rc;
    // hh0 = hb(1,1) of each element's own mesh (src/calvar_lin.f:65)
    std::vector<double> hh0e(c->ne);
    for (size_t m = 0; m + 1 < c->mem_e0.size(); m++)
      for (int e = c->mem_e0[m]; e < c->mem_e0[m + 1]; e++)
        hh0e[e] = hb[(size_t)ld * c->mem_e0[m]];
    HIPCHK(hipMemcpy(c->dfields + 4 * n, hh0e.data(), sizeof(double) * c->ne, hipMemcpyHostToDevice));
    c->phys = Physics{mu0, mb0, g0, hb[0], c->dfields + 4 * n};
    // the mixed path's set-up reads host arrays element by element: hand it
    // the fields in the internal numbering
    std::vector<double> internal;
    const double *f4[4] = {fc, w0, z0, hb};
    int ld4 = ld;
    if (!c->perm.empty()) {
      internal.assign(4 * n, 0.0);
      for (int f = 0; f < 4; f++) {
        if (!f4[f]) continue;
        double *d = internal.data() + f * n;
        for (int e = 0; e < c->ne; e++)
          std::memcpy(d + (size_t)c->nm * e, f4[f] + (size_t)ld * caller_elem(c, e),
                      sizeof(double) * c->nm);
        f4[f] = d;
      }
      ld4 = c->nm;
    }
    if (f32_fields_set(c->f32, ld4, f4[0], f4[1], f4[2], f4[3], hh0e.data(), c->h_pi2.data(),
                       c->ng, mu0, mb0, g0))
      return fail(SPOC_EHIP, "FP32 field upload failed");
    if (f64_fields_set(c->f64, c->stream, c->tab, c->dfields, c->dfields + n, c->dfields + 2 * n,
                       c->dfields + 3 * n, c->dfields + 4 * n, mu0, mb0, g0))
      return fail(SPOC_EHIP, "FP64 field set-up failed");
    c->have_fields = true;
    return SPOC_OK;
  });
}

int spoc_state_upload(spoc_ctx *c, int ld, const double *u0, const double *v0,
                      const double *h0) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->dmesh) return fail(SPOC_ESTATE, "spoc_state_upload: set the mesh first");
    if (ld < c->nm || !u0 || !v0 || !h0) return fail(SPOC_EINVAL, "spoc_state_upload: bad args");
    (void)hipSetDevice(c->device);
    const double *src[3] = {u0, v0, h0};
    double *dst[3] = {state_ptr(c, 0), state_ptr(c, 1), state_ptr(c, 2)};
    int rc = upload_fields(c, ld, src, dst, 3);
    if (!rc) c->have_state = true;
    return rc;
  });
}

int spoc_state_download(spoc_ctx *c, int ld, double *u0, double *v0, double *h0) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_state) return fail(SPOC_ESTATE, "spoc_state_download: no state");
    if (ld < c->nm || !u0 || !v0 || !h0) return fail(SPOC_EINVAL, "spoc_state_download: bad args");
    (void)hipSetDevice(c->device);
    double *dst[3] = {u0, v0, h0};
    double *src[3] = {state_ptr(c, 0), state_ptr(c, 1), state_ptr(c, 2)};
    return download_fields(c, ld, dst, src, 3);
  });
}

static int check_ready(spoc_ctx *c, int variant) {
  if (!c) return fail(SPOC_EINVAL, "null context");
  if (!c->dmesh) return fail(SPOC_ESTATE, "no mesh: call spoc_mesh_set first");
  if (!c->have_fields) return fail(SPOC_ESTATE, "no fields: call spoc_fields_set first");
  if (variant < SPOC_LIN || variant > SPOC_INV) return fail(SPOC_EINVAL, "bad variant %d", variant);
  return SPOC_OK;
}

int spoc_rhs(spoc_ctx *c, int variant, int precision, int ld, const double *u0,
             const double *v0, const double *h0, double *u1, double *v1, double *h1) {
  return spoc::guarded([&]() -> int {
    int rc = check_ready(c, variant);
    if (rc) return rc;
    if (ld < c->nm) return fail(SPOC_EINVAL, "ld=%d < nm=%d", ld, c->nm);
    (void)hipSetDevice(c->device);
    // stage input goes to the x slots (3..5); tendencies come back from k (9..11)
    const double *src[3] = {u0, v0, h0};
    double *xs[3] = {state_ptr(c, 3), state_ptr(c, 4), state_ptr(c, 5)};
    rc = upload_fields(c, ld, src, xs, 3);
    if (rc) return rc;
    const double *x[3] = {xs[0], xs[1], xs[2]};
    double *ks[3] = {state_ptr(c, 9), state_ptr(c, 10), state_ptr(c, 11)};
    if (precision != SPOC_FP32_MFMA && f32_flux(c->f32) != 0.0)
      return fail(SPOC_EINVAL, "the Lax-Friedrichs flux option runs on the FP32-MFMA path only");
    if (precision == SPOC_FP32_MFMA) {
      rc = f32_rhs(c->f32, c->stream, variant, x, ks, nullptr, nullptr);
      if (rc) return f32_fail(rc);
    } else if (precision == SPOC_FP64) {
      rc = f64_rhs(c->f64, c->stream, c->tab, variant, x, ks, nullptr, nullptr);
      if (rc) return fail(SPOC_EHIP, "FP64 kernel launch failed: %s",
                          hipGetErrorString(hipGetLastError()));
    } else {
      return fail(SPOC_EINVAL, "unknown precision %d", precision);
    }
    double *dst[3] = {u1, v1, h1};
    return download_fields(c, ld, dst, ks, 3);
  });
}

int spoc_rk4_steps(spoc_ctx *c, int variant, int precision, double dt, int nsteps) {
  return spoc::guarded([&]() -> int {
    int rc = check_ready(c, variant);
    if (rc) return rc;
    if (!c->have_state) return fail(SPOC_ESTATE, "no state: call spoc_state_upload first");
    if (nsteps < 0) return fail(SPOC_EINVAL, "nsteps=%d", nsteps);
    if (precision != SPOC_FP64 && precision != SPOC_FP32_MFMA)
      return fail(SPOC_EINVAL, "unknown precision %d", precision);
    if (precision != SPOC_FP32_MFMA && f32_flux(c->f32) != 0.0)
      return fail(SPOC_EINVAL, "the Lax-Friedrichs flux option runs on the FP32-MFMA path only");
    (void)hipSetDevice(c->device);
    c->prof.used = 0;
    c->prof.marks.clear();
    const double *u0[3] = {state_ptr(c, 0), state_ptr(c, 1), state_ptr(c, 2)};
    double *uw[3] = {state_ptr(c, 0), state_ptr(c, 1), state_ptr(c, 2)};
    // one RK4 step (4 right-hand sides with their fused stage updates)
    auto step = [&](TimerHook hook) -> int {
      if (precision == SPOC_FP32_MFMA) {
        int r = f32_rk4_step(c->f32, c->stream, variant, dt, hook, c);
        return r ? f32_fail(r) : SPOC_OK;
      }
      int r = f64_rk4_step(c->f64, c->stream, c->tab, variant, dt, hook, c);
      return r ? fail(SPOC_EHIP, "FP64 kernel launch failed: %s",
                      hipGetErrorString(hipGetLastError()))
               : SPOC_OK;
    };
    // an error leaves nothing queued behind it
    auto bail = [&](int r) {
      (void)hipStreamSynchronize(c->stream);
      c->queue.reset();
      return r;
    };
    c->queue.reset();
    const int per_step = precision == SPOC_FP64 ? 4 * f64_kernels_per_rhs(c->f64, variant)
                                                : step_dispatches(variant, precision);
    HIPCHK(hipEventRecord(c->t0, c->stream));
    if (precision == SPOC_FP32_MFMA && (rc = f32_rk4_begin(c->f32, c->stream, variant, uw)))
      return bail(f32_fail(rc));
    if (precision == SPOC_FP64 && f64_rk4_begin(c->f64, c->stream, c->tab, u0))
      return bail(fail(SPOC_EHIP, "FP64 state conversion failed"));
    // small linear meshes on the mixed path: every step in one persistent
    // launch (f32_rk4_persist says where it applies)
    const unsigned *abort_word = nullptr;
    int persist = 1;
    if (precision == SPOC_FP32_MFMA && c->use_persist &&
        (persist = f32_rk4_persist(c->f32, c->stream, variant, dt, nsteps, &abort_word,
                                   c->prof.on ? prof_hook : nullptr, c, c->use_persist == 2)) < 0)
      return bail(f32_fail(persist));
    const bool graph = c->use_graph && !c->prof.on && nsteps >= 2;
    if (persist == 0) {
      // (the steps ran above)
    } else if (graph) {
      // the step's launches are identical every step: replay them as a graph
      // of kGraphSteps steps (one when fewer are asked for), which costs less
      // per launch than one graph per step on small meshes; the steps left
      // over run eagerly
      const int per = nsteps >= kGraphSteps ? kGraphSteps : 1;
      spoc_ctx::GraphKey key{variant, precision, dt, c->version, per};
      if (!c->graph || !(c->graph_key == key)) {
        destroy_graph(c);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess)
          return bail(fail(SPOC_EHIP, "graph capture: %s", hipGetErrorString(e)));
        int r = SPOC_OK;
        for (int k = 0; k < per && !r; k++) r = step(nullptr);
        e = hipStreamEndCapture(c->stream, &g);
        if (r) {
          if (g) (void)hipGraphDestroy(g);
          return bail(r);
        }
        if (e != hipSuccess) return bail(fail(SPOC_EHIP, "graph capture: %s", hipGetErrorString(e)));
        e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          c->graph = nullptr;
          return bail(fail(SPOC_EHIP, "graph instantiate: %s", hipGetErrorString(e)));
        }
        c->graph_key = key;
      }
      for (int it = 0; it + per <= nsteps; it += per) {
        hipError_t e = c->queue.admit(per * per_step);
        if (e == hipSuccess) e = hipGraphLaunch(c->graph, c->stream);
        if (e == hipSuccess) e = c->queue.mark(c->stream);
        if (e != hipSuccess) return bail(fail(SPOC_EHIP, "graph launch: %s", hipGetErrorString(e)));
      }
      for (int it = nsteps - nsteps % per; it < nsteps; it++) {
        hipError_t e = c->queue.admit(per_step);
        if (e != hipSuccess) return bail(fail(SPOC_EHIP, "queue wait: %s", hipGetErrorString(e)));
        if ((rc = step(nullptr))) return bail(rc);
        if ((e = c->queue.mark(c->stream)) != hipSuccess)
          return bail(fail(SPOC_EHIP, "queue mark: %s", hipGetErrorString(e)));
      }
    } else {
      for (int it = 0; it < nsteps; it++) {
        hipError_t e = c->queue.admit(per_step);
        if (e != hipSuccess) return bail(fail(SPOC_EHIP, "queue wait: %s", hipGetErrorString(e)));
        if ((rc = step(prof_hook))) return bail(rc);
        if ((e = c->queue.mark(c->stream)) != hipSuccess)
          return bail(fail(SPOC_EHIP, "queue mark: %s", hipGetErrorString(e)));
      }
    }
    if (precision == SPOC_FP32_MFMA && (rc = f32_rk4_end(c->f32, c->stream, variant, uw)))
      return bail(f32_fail(rc));
    if (precision == SPOC_FP64 && f64_rk4_end(c->f64, c->stream, c->tab, uw))
      return bail(fail(SPOC_EHIP, "FP64 state conversion failed"));
    HIPCHK(hipEventRecord(c->t1, c->stream));
    // the driver's testnan: is every coefficient of the new state finite?
    if (nonfinite_launch(c->stream, (size_t)3 * c->nm * c->ne, state_ptr(c, 0), c->dflag))
      return bail(fail(SPOC_EHIP, "non-finite test launch failed"));
    int bad = 0;
    unsigned aborted = 0;
    HIPCHK(hipMemcpyAsync(&bad, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (abort_word)
      HIPCHK(hipMemcpyAsync(&aborted, abort_word, sizeof(unsigned), hipMemcpyDeviceToHost,
                            c->stream));
    HIPCHK(hipEventSynchronize(c->t1));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->queue.reset();
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->t0, c->t1));
    c->last_ms = ms;
    // per-kernel totals from the hook events: slot k -> k+1 brackets kernel k
    c->knames.clear();
    c->kms.clear();
    c->kcount.clear();
    if (c->prof.on) {
      std::vector<double> tot(16, 0.0);
      std::vector<int> cnt(16, 0);
      for (size_t i = 0; i + 1 < c->prof.marks.size(); i++) {
        int s0 = c->prof.marks[i].first, s1 = c->prof.marks[i + 1].first;
        if (s1 != s0 + 1) continue;
        float m = 0.f;
        if (hipEventElapsedTime(&m, c->prof.ev[c->prof.marks[i].second],
                                c->prof.ev[c->prof.marks[i + 1].second]) == hipSuccess) {
          tot[s0] += m;
          cnt[s0]++;
        }
      }
      const char *lin[] = {"face_lin", "elem_lin", "wall_lin"};
      const char *nl[] = {"face_uv", "elem_q", "wall_q", "elem_c", "wall_c", "face_nl2", "elem_eg",
                          "wall_eg"};
      int nk = variant == SPOC_LIN ? 3 : 8;
      if (precision == SPOC_FP32_MFMA) {
        const char *l32[] = {persist == 0 ? "lin_persistent" : "lin_f64"};
        const char *n32[] = {"bc_f32", "eg_f32f64"};
        nk = variant == SPOC_LIN ? 1 : 2;
        for (int k = 0; k < nk; k++) {
          c->knames.push_back(variant == SPOC_LIN ? l32[k] : n32[k]);
          c->kms.push_back(tot[k]);
          c->kcount.push_back(cnt[k]);
        }
      } else {
        for (int k = 0; k < nk; k++) {
          c->knames.push_back(variant == SPOC_LIN ? lin[k] : nl[k]);
          c->kms.push_back(tot[k]);
          c->kcount.push_back(cnt[k]);
        }
      }
    }
    if (aborted)
      return fail(SPOC_EHIP, "persistent RK4 kernel: a wait for a neighbouring block ran out (not "
                             "every wave of the launch was resident); the state is not advanced "
                             "consistently: call spoc_set_persistent(ctx, 0) and run again");
    if (bad)
      return fail(SPOC_ENAN, "the state holds a NaN or an infinity after %d RK4 steps (dt=%g)",
                  nsteps, dt);
    return SPOC_OK;
  });
}

static int adj_fail(int rc) {
  if (rc == -1) return fail(SPOC_EINVAL, "adjoint: the nonlinear variants need the base state");
  return f32_fail(rc);
}

int spoc_rhs_adjoint(spoc_ctx *c, int variant, int ld, const double *u0, const double *v0,
                     const double *h0, const double *au, const double *av, const double *ah,
                     double *ru, double *rv, double *rh, double *rhb) {
  return spoc::guarded([&]() -> int {
    int rc = check_ready(c, variant);
    if (rc) return rc;
    if (ld < c->nm || !au || !av || !ah || !ru || !rv || !rh)
      return fail(SPOC_EINVAL, "spoc_rhs_adjoint: bad arguments");
    if (variant != SPOC_LIN && (!u0 || !v0 || !h0))
      return fail(SPOC_EINVAL, "spoc_rhs_adjoint: the nonlinear variants need the state (u0,v0,h0)");
    if (f32_flux(c->f32) != 0.0)
      return fail(SPOC_EINVAL, "the adjoint is of the central flux: spoc_set_flux(ctx, 0) first");
    (void)hipSetDevice(c->device);
    size_t n = (size_t)c->nm * c->ne;
    const double *src[3] = {au, av, ah};
    double *lam[3] = {c->dadj, c->dadj + n, c->dadj + 2 * n};
    double *ys[3] = {state_ptr(c, 9), state_ptr(c, 10), state_ptr(c, 11)};
    double *xs[3] = {state_ptr(c, 12), state_ptr(c, 13), state_ptr(c, 14)};
    if ((rc = upload_fields(c, ld, src, lam, 3))) return rc;
    if (variant != SPOC_LIN) {
      const double *xsrc[3] = {u0, v0, h0};
      if ((rc = upload_fields(c, ld, xsrc, xs, 3))) return rc;
    }
    double *hb = c->dadj + 3 * n;
    rc = f32_adj_rhs(c->f32, c->stream, variant, variant != SPOC_LIN ? xs : nullptr, lam, ys,
                     variant != SPOC_LIN ? hb : nullptr);
    if (rc) return adj_fail(rc);
    c->have_adj = c->have_adj_hb = false;  // the adjoint slots now hold this call's data
    double *dst[3] = {ru, rv, rh};
    if ((rc = download_fields(c, ld, dst, ys, 3))) return rc;
    if (rhb) {
      if (variant == SPOC_LIN) {
        for (int e = 0; e < c->ne; e++)
          for (int m = 0; m < c->nm; m++) rhb[m + (size_t)ld * e] = 0.0;
      } else {
        double *d1[1] = {rhb}, *s1[1] = {hb};
        if ((rc = download_fields(c, ld, d1, s1, 1))) return rc;
      }
    }
    return SPOC_OK;
  });
}

int spoc_adjoint_upload(spoc_ctx *c, int ld, const double *au, const double *av,
                        const double *ah) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->dmesh) return fail(SPOC_ESTATE, "spoc_adjoint_upload: set the mesh first");
    if (ld < c->nm || !au || !av || !ah) return fail(SPOC_EINVAL, "spoc_adjoint_upload: bad args");
    (void)hipSetDevice(c->device);
    size_t n = (size_t)c->nm * c->ne;
    const double *src[3] = {au, av, ah};
    double *dst[3] = {c->dadj, c->dadj + n, c->dadj + 2 * n};
    int rc = upload_fields(c, ld, src, dst, 3);
    if (!rc) c->have_adj = true;
    c->have_adj_hb = false;
    return rc;
  });
}

int spoc_adjoint_download(spoc_ctx *c, int ld, double *au, double *av, double *ah) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_adj) return fail(SPOC_ESTATE, "spoc_adjoint_download: no adjoint state");
    if (ld < c->nm || !au || !av || !ah) return fail(SPOC_EINVAL, "spoc_adjoint_download: bad args");
    (void)hipSetDevice(c->device);
    size_t n = (size_t)c->nm * c->ne;
    double *dst[3] = {au, av, ah};
    double *src[3] = {c->dadj, c->dadj + n, c->dadj + 2 * n};
    return download_fields(c, ld, dst, src, 3);
  });
}

int spoc_adjoint_hb_download(spoc_ctx *c, int ld, double *ahb) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_adj_hb)
      return fail(SPOC_ESTATE, "spoc_adjoint_hb_download: no bathymetry adjoint (run "
                               "spoc_rk4_adjoint_steps with a nonlinear variant first)");
    if (ld < c->nm || !ahb) return fail(SPOC_EINVAL, "spoc_adjoint_hb_download: bad args");
    (void)hipSetDevice(c->device);
    size_t n = (size_t)c->nm * c->ne;
    double *dst[1] = {ahb}, *src[1] = {c->dadj + 3 * n};
    return download_fields(c, ld, dst, src, 1);
  });
}

int spoc_set_checkpoint(spoc_ctx *c, int every) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    if (every < 0) return fail(SPOC_EINVAL, "checkpoint interval %d", every);
    c->ncheck = every;
    return SPOC_OK;
  });
}

int spoc_rk4_adjoint_steps(spoc_ctx *c, int variant, double dt, int nsteps) {
  return spoc::guarded([&]() -> int {
    int rc = check_ready(c, variant);
    if (rc) return rc;
    if (!c->have_adj) return fail(SPOC_ESTATE, "no adjoint state: call spoc_adjoint_upload first");
    if (variant != SPOC_LIN && !c->have_state)
      return fail(SPOC_ESTATE, "the nonlinear adjoint needs the initial state: spoc_state_upload");
    if (nsteps < 0) return fail(SPOC_EINVAL, "nsteps=%d", nsteps);
    if (f32_flux(c->f32) != 0.0)
      return fail(SPOC_EINVAL, "the adjoint is of the central flux: spoc_set_flux(ctx, 0) first");
    (void)hipSetDevice(c->device);
    drop_graph(c);
    size_t n = (size_t)c->nm * c->ne;
    double *lam[3] = {c->dadj, c->dadj + n, c->dadj + 2 * n};
    const double *x0[3] = {state_ptr(c, 0), state_ptr(c, 1), state_ptr(c, 2)};
    HIPCHK(hipEventRecord(c->t0, c->stream));
    rc = f32_adj_run(c->f32, c->stream, variant, dt, nsteps, c->ncheck, x0, lam,
                     variant != SPOC_LIN ? c->dadj + 3 * n : nullptr);
    if (rc) return adj_fail(rc);
    HIPCHK(hipEventRecord(c->t1, c->stream));
    HIPCHK(hipEventSynchronize(c->t1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->t0, c->t1));
    c->last_ms = ms;
    c->have_adj_hb = variant != SPOC_LIN;
    return SPOC_OK;
  });
}

int spoc_autotune(spoc_ctx *c, int variant, int *tiles) {
  return spoc::guarded([&]() -> int {
    int rc = check_ready(c, variant);
    if (rc) return rc;
    (void)hipSetDevice(c->device);
    drop_graph(c);
    rc = f32_autotune(c->f32, c->stream, variant, tiles);
    return rc ? f32_fail(rc) : SPOC_OK;
  });
}

int spoc_set_tiles(spoc_ctx *c, int variant, int tile0, int tile1, int autotune) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    if (variant < SPOC_LIN || variant > SPOC_INV) return fail(SPOC_EINVAL, "variant=%d", variant);
    for (int t : {tile0, tile1}) {
      const int wg = t & 15, form = t >> 4;
      if (t != 0 && (t < 0 || (wg != 2 && wg != 4 && wg != 8) || form >= TILE_FORMS))
        return fail(SPOC_EINVAL,
                    "tile %d: 2, 4 or 8 waves per workgroup plus 16 x a kernel form 0..%d "
                    "(0: automatic)", t, TILE_FORMS - 1);
    }
    drop_graph(c);
    f32_set_tiles(c->f32, variant, tile0, tile1, autotune);
    return SPOC_OK;
  });
}

int spoc_set_flux(spoc_ctx *c, double alpha) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    if (!(alpha >= 0.0) || alpha > 100.0) return fail(SPOC_EINVAL, "flux alpha=%g", alpha);
    drop_graph(c);
    f32_set_flux(c->f32, alpha);
    return SPOC_OK;
  });
}

int spoc_set_ordering(spoc_ctx *c, int mode) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    if (mode != SPOC_ORDER_INPUT && mode != SPOC_ORDER_HILBERT && mode != SPOC_ORDER_ZCURVE &&
        mode != SPOC_ORDER_AUTO)
      return fail(SPOC_EINVAL, "ordering mode %d (SPOC_ORDER_INPUT 0, SPOC_ORDER_HILBERT 1, "
                               "SPOC_ORDER_ZCURVE 2, SPOC_ORDER_AUTO 3)", mode);
    c->ordering = mode;
    return SPOC_OK;
  });
}

int spoc_set_graph(spoc_ctx *c, int on) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    c->use_graph = on != 0;
    if (!on) drop_graph(c);
    return SPOC_OK;
  });
}

int spoc_set_persistent(spoc_ctx *c, int on) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    if (on < 0 || on > 2) return fail(SPOC_EINVAL, "spoc_set_persistent: on=%d (0, 1 or 2)", on);
    c->use_persist = on;
    return SPOC_OK;
  });
}

int spoc_sync(spoc_ctx *c) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return SPOC_OK;
  });
}

// (ke, pe, mass) of nr element ranges given by the device offsets off.
static int energies(spoc_ctx *c, int nr, const int *off, double *r) {
  (void)hipSetDevice(c->device);
  size_t n = (size_t)c->nm * c->ne;
  double *out = c->den + (size_t)c->ne * (2 + c->nm);
  int rc = energy_launch(c->stream, c->ne, c->nm, c->ng, state_ptr(c, 0), state_ptr(c, 1),
                         state_ptr(c, 2), c->dfields + 3 * n, c->tab.pi2, c->tab.wi,
                         c->tab.ammo1, c->mesh.deta, c->phys.g0, nr, off, c->dorder, c->den,
                         out);
  if (rc) return fail(SPOC_EHIP, "energy kernel launch failed");
  HIPCHK(hipMemcpyAsync(r, out, sizeof(double) * 3 * nr, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SPOC_OK;
}

int spoc_energy(spoc_ctx *c, double *ke, double *pe, double *mass) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_state || !c->have_fields)
      return fail(SPOC_ESTATE, "spoc_energy: need mesh, fields and state");
    if (!c->have_nodes) return fail(SPOC_ESTATE, "spoc_energy: mesh was set without node arrays");
    (void)hipSetDevice(c->device);
    double r[3];
    int rc = energies(c, 1, c->dmem, r);
    if (rc) return rc;
    *ke = r[0];
    *pe = r[1];
    *mass = r[2];
    if (std::isnan(r[0])) return fail(SPOC_ENAN, "spoc_energy: the kinetic energy is NaN");
    return SPOC_OK;
  });
}

// The refinement indicator of the resident state (adapt_launch).  Layout of
// c->dadapt: partial and member maxima (5 nch + 5 nmem doubles; the
// per-member supmax and deraff at 5 nch + 3 nmem), normalised indicators
// (3 ne), flags (2 ne ints), traces (9 ngf ne).
struct AdaptBufs {
  double *work, *res, *amax, *T;
  int *flags;
};
static AdaptBufs adapt_bufs(spoc_ctx *c) {
  const size_t nmem = c->mem_e0.size() - 1, ne = c->ne, nch = c->nch;
  AdaptBufs b;
  b.work = c->dadapt;
  b.res = b.work + 5 * nch + 3 * nmem;
  b.amax = b.work + 5 * nch + 5 * nmem;
  b.flags = (int *)(b.amax + 3 * ne);
  b.T = b.amax + 4 * ne;
  return b;
}

static int adapt_run(spoc_ctx *c, bool decide, double fac1, double fac2) {
  const int nmem = (int)c->mem_e0.size() - 1;
  AdaptBufs b = adapt_bufs(c);
  const double *a[3] = {state_ptr(c, 0), state_ptr(c, 1), state_ptr(c, 2)};
  const int rc = adapt_launch(c->stream, c->nc, c->ngf, c->ne, nmem, c->nch, a, c->dside,
                              c->tab.pico, c->dmem + 2, c->dchunk,
                              c->dchunk + c->nch + 1, b.work, b.T, b.amax, decide, fac1, fac2,
                              b.flags);
  if (rc == -1) return fail(SPOC_EINVAL, "refinement indicator: no kernel for this shape");
  if (rc) return fail(SPOC_EHIP, "refinement indicator kernel launch failed");
  return SPOC_OK;
}

int spoc_checkconv(spoc_ctx *c, int field, double *amax) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_state) return fail(SPOC_ESTATE, "spoc_checkconv: no state");
    if (field < 0 || field > 2 || !amax)
      return fail(SPOC_EINVAL, "spoc_checkconv: field=%d (0 u, 1 v, 2 h)", field);
    (void)hipSetDevice(c->device);
    int rc = adapt_run(c, false, 0.0, 0.0);
    if (rc) return rc;
    const double *d = adapt_bufs(c).amax + (size_t)field * c->ne;
    std::vector<double> tmp(c->perm.empty() ? 0 : c->ne);
    double *h = c->perm.empty() ? amax : tmp.data();
    HIPCHK(hipMemcpyAsync(h, d, sizeof(double) * c->ne, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t e = 0; e < tmp.size(); e++) amax[c->perm[e]] = tmp[e];
    return SPOC_OK;
  });
}

int spoc_cal_conv(spoc_ctx *c, double fac1, double fac2, double fac3, int *fconv, int *fderaff,
                  int *conv, int *deraff, int *timedir, double *supmax, double *amax3) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_state) return fail(SPOC_ESTATE, "spoc_cal_conv: no state");
    if (!fconv || !fderaff || !conv || !deraff || !timedir)
      return fail(SPOC_EINVAL, "spoc_cal_conv: null output");
    (void)hipSetDevice(c->device);
    int rc = adapt_run(c, true, fac1, fac2);
    if (rc) return rc;
    const int nmem = (int)c->mem_e0.size() - 1;
    const size_t ne = c->ne;
    AdaptBufs b = adapt_bufs(c);
    std::vector<double> r(2 * (size_t)nmem);
    const bool perm = !c->perm.empty();
    std::vector<int> fl(perm ? 2 * ne : 0);
    std::vector<double> am(perm && amax3 ? 3 * ne : 0);
    int *hf0 = perm ? fl.data() : fconv, *hf1 = perm ? fl.data() + ne : fderaff;
    double *ha = perm ? am.data() : amax3;
    HIPCHK(hipMemcpyAsync(r.data(), b.res, sizeof(double) * r.size(), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(hf0, b.flags, sizeof(int) * ne, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hf1, b.flags + ne, sizeof(int) * ne, hipMemcpyDeviceToHost, c->stream));
    if (amax3)
      HIPCHK(hipMemcpyAsync(ha, b.amax, sizeof(double) * 3 * ne, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (perm)  // back to the caller's numbering
      for (size_t e = 0; e < ne; e++) {
        const size_t o = c->perm[e];
        fconv[o] = fl[e];
        fderaff[o] = fl[ne + e];
        if (amax3)
          for (int f = 0; f < 3; f++) amax3[f * ne + o] = am[f * ne + e];
      }
    for (int m = 0; m < nmem; m++) {  // src/adapt.f:1180-1189
      const double sup = r[m];
      deraff[m] = r[nmem + m] > 0.0;
      timedir[m] = sup > fac3 ? 0 : 1;
      conv[m] = (sup > fac1 || deraff[m]) ? 1 : 0;
      if (supmax) supmax[m] = sup;
    }
    return SPOC_OK;
  });
}

int spoc_energy_noalloc(spoc_ctx *c, double *out_dev) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_state || !c->have_fields)
      return fail(SPOC_ESTATE, "spoc_energy_noalloc: need mesh, fields and state");
    if (!c->have_nodes) return fail(SPOC_ESTATE, "spoc_energy_noalloc: mesh set without nodes");
    if (!out_dev) return fail(SPOC_EINVAL, "spoc_energy_noalloc: null output");
    const size_t need = (size_t)c->ne * (2 + (size_t)c->nm), cap = energy_scratch_doubles();
    if (need > cap)
      return fail(SPOC_EINVAL,
                  "spoc_energy_noalloc: %d elements of %d modes need %zu doubles of scratch, more "
                  "than the %zu it holds (ne (nm + 2) <= %zu); use spoc_energy",
                  c->ne, c->nm, need, cap, cap);
    HIPCHK(hipSetDevice(c->device));
    double *scr = nullptr;
    const int r = energy_scratch(c->device, &scr);
    if (r == -1) return fail(SPOC_EINVAL, "spoc_energy_noalloc: device %d out of range", c->device);
    if (r) return fail(SPOC_EHIP, "spoc_energy_noalloc: hipGetSymbolAddress failed");
    const size_t n = (size_t)c->nm * c->ne;
    const int nmem = (int)c->mem_e0.size() - 1;
    const int rc = energy_launch(c->stream, c->ne, c->nm, c->ng, state_ptr(c, 0), state_ptr(c, 1),
                                 state_ptr(c, 2), c->dfields + 3 * n, c->tab.pi2, c->tab.wi,
                                 c->tab.ammo1, c->mesh.deta, c->phys.g0, nmem, c->dmem + 2,
                                 c->dorder, scr, out_dev);
    if (rc) return fail(SPOC_EHIP, "spoc_energy_noalloc: kernel launch failed");
    return SPOC_OK;
  });
}

int spoc_ensemble_energy(spoc_ctx *c, int nmem, double *ke, double *pe, double *mass) {
  return spoc::guarded([&]() -> int {
    if (!c || !c->have_state || !c->have_fields)
      return fail(SPOC_ESTATE, "spoc_ensemble_energy: need mesh, fields and state");
    if (!c->have_nodes) return fail(SPOC_ESTATE, "spoc_ensemble_energy: mesh set without nodes");
    if (nmem != (int)c->mem_e0.size() - 1)
      return fail(SPOC_EINVAL, "spoc_ensemble_energy: nmem=%d, the mesh has %d members", nmem,
                  (int)c->mem_e0.size() - 1);
    std::vector<double> r(3 * (size_t)nmem);
    int rc = energies(c, nmem, c->dmem + 2, r.data());
    if (rc) return rc;
    int nan_member = -1;
    for (int m = 0; m < nmem; m++) {
      ke[m] = r[3 * m];
      pe[m] = r[3 * m + 1];
      mass[m] = r[3 * m + 2];
      if (nan_member < 0 && std::isnan(ke[m])) nan_member = m;
    }
    if (nan_member >= 0)
      return fail(SPOC_ENAN, "spoc_ensemble_energy: the kinetic energy of member %d is NaN",
                  nan_member);
    return SPOC_OK;
  });
}

int spoc_set_profiling(spoc_ctx *c, int on) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    c->prof.on = on != 0;
    return SPOC_OK;
  });
}

int spoc_last_timing(spoc_ctx *c, double *total_ms, int *nkernels, char *names, int len,
                     double *ms, int *launches) {
  return spoc::guarded([&]() -> int {
    if (!c) return fail(SPOC_EINVAL, "null context");
    if (total_ms) *total_ms = c->last_ms;
    if (!nkernels) return SPOC_OK;
    int cap = *nkernels, n = (int)c->knames.size();
    *nkernels = n;
    int pos = 0;
    for (int i = 0; i < n && i < cap; i++) {
      if (ms) ms[i] = c->kms[i];
      if (launches) launches[i] = c->kcount[i];
      if (names) {
        int l = (int)c->knames[i].size();
        if (pos + l + 1 > len) break;
        std::memcpy(names + pos, c->knames[i].c_str(), l + 1);
        pos += l + 1;
      }
    }
    return SPOC_OK;
  });
}

}  // extern "C"
