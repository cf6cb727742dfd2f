// Host side of the plain FP64 path: the element-blocked copies of the mesh,
// fields and state (f64_types.h), the wall-element list, work arrays and the
// launch sequence.  The C-ABI keeps packed (nm, ne) arrays on the device;
// this path converts at the start and the end of each spoc_rk4_steps call
// and around each spoc_rhs.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "common.h"
#include "f64_types.h"

namespace spoc {

struct F64Plan {
  int nc, ng, ngf, nm;
  int ne = 0, nf = 0, nblk = 0, nwall = 0;
  void *dmesh = nullptr;    // geometry, integer rows, wall list
  double *dwork = nullptr;  // every blocked array below
  F64Mesh mesh{};
  double *u0[3], *x[3], *acc[3], *k[3], *fld[4], *Q[4], *F[4], *S[2], *UH[2], *G[3], *MF, *WS;
  double *face = nullptr;
  const double *hh0e = nullptr;
  double mu0 = 0.0, mb0 = 0.0, g0 = 0.0;
  bool have_fields = false;
};

namespace {

int call(const F64Plan *p, const F64Args &a, int phase, const double *in, double *out) {
#define SPOC_CALL(NC_, NG_, NGF_) \
  if (p->nc == NC_ && p->ng == NG_ && p->ngf == NGF_) return rhs_f64_##NC_##_##NG_##_##NGF_(a, phase, in, out);
  SPOC_FOR_EACH_SHAPE(SPOC_CALL)
#undef SPOC_CALL
  return -1;
}

void free_mesh(F64Plan *p) {
  if (p->dmesh) (void)hipFree(p->dmesh);
  if (p->dwork) (void)hipFree(p->dwork);
  p->dmesh = nullptr;
  p->dwork = nullptr;
  p->have_fields = false;
  p->ne = p->nf = p->nblk = p->nwall = 0;
}

F64Args args(const F64Plan *p, hipStream_t s, const DevTables &tab, int variant, int stage,
             double dt) {
  F64Args a{};
  a.variant = variant;
  a.stage = stage;
  a.dt = dt;
  a.dt6 = dt / 6.0;
  a.stream = s;
  a.tab = tab;
  a.mesh = p->mesh;
  a.mu0 = p->mu0;
  a.mb0 = p->mb0;
  a.g0 = p->g0;
  a.hh0e = p->hh0e;
  a.hb = p->fld[3];
  a.w0 = p->fld[1];
  a.z0 = p->fld[2];
  a.fc = p->fld[0];
  for (int c = 0; c < 3; c++) {
    a.x[c] = stage <= 0 ? p->u0[c] : p->x[c];
    a.u0[c] = p->u0[c];
    a.xo[c] = p->x[c];
    a.acc[c] = p->acc[c];
    a.k[c] = p->k[c];
    a.G[c] = p->G[c];
  }
  for (int c = 0; c < 4; c++) {
    a.Q[c] = p->Q[c];
    a.F[c] = p->F[c];
  }
  for (int c = 0; c < 2; c++) {
    a.S[c] = p->S[c];
    a.UH[c] = p->UH[c];
  }
  a.MF = p->MF;
  a.WS = p->WS;
  a.face = p->face;
  return a;
}

}  // namespace

int f64_plan_create(F64Plan **out, int nc, int ng, int ngf) {
  F64Plan *p = new F64Plan();
  p->nc = nc;
  p->ng = ng;
  p->ngf = ngf;
  p->nm = nm_of(nc);
  *out = p;
  return 0;
}

void f64_plan_destroy(F64Plan *p) {
  if (!p) return;
  free_mesh(p);
  delete p;
}

// Host arrays in the internal numbering (1-based faces as the C-ABI takes
// them); device arrays: tseg, ssit, norm and the face-value buffer (nf x 6
// x ngf doubles).
int f64_mesh_set(F64Plan *p, int ne, int nf, const int *itm, const int *its, const int *bdl,
                 const double *vect, const double *normt, const int *dtseg, const int *dssit,
                 const double *dnorm, double *dface) {
  free_mesh(p);
  const int nblk = (ne + F64_WB - 1) / F64_WB, npad = nblk * F64_WB, nm = p->nm, ng = p->ng;
  std::vector<double> geo((size_t)G_ROWS * npad, 0.0);
  std::vector<int> ib((size_t)I_ROWS * npad, 0), wall;
  for (int e = 0; e < npad; e++)
    for (int k = 0; k < 3; k++) ib[f64_at(e, I_ITS + k, I_ROWS)] = 1;
  for (int e = 0; e < ne; e++) {
    for (int r = 0; r < 5; r++) geo[f64_at(e, G_AX + r, G_ROWS)] = vect[5 * (size_t)e + r];
    geo[f64_at(e, G_NX, G_ROWS)] = normt[2 * (size_t)e];
    geo[f64_at(e, G_NY, G_ROWS)] = normt[2 * (size_t)e + 1];
    for (int k = 0; k < 3; k++) {
      ib[f64_at(e, I_ITM + k, I_ROWS)] = itm[3 * (size_t)e + k] - 1;
      ib[f64_at(e, I_ITS + k, I_ROWS)] = its[3 * (size_t)e + k];
    }
    ib[f64_at(e, I_BDL, I_ROWS)] = bdl[e] != 0;
    if (bdl[e] != 0) wall.push_back(e);
  }
  const size_t gb = geo.size() * sizeof(double), ibb = ib.size() * sizeof(int);
  const size_t wb = (wall.size() + 1) * sizeof(int);
  if (hipMalloc(&p->dmesh, gb + ibb + wb + 256) != hipSuccess) return -2;
  char *m = (char *)p->dmesh;
  double *dgeo = (double *)m;
  int *dib = (int *)(m + gb), *dwall = (int *)(m + gb + ibb);
  if (hipMemcpy(dgeo, geo.data(), gb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dib, ib.data(), ibb, hipMemcpyHostToDevice) != hipSuccess ||
      (!wall.empty() &&
       hipMemcpy(dwall, wall.data(), wall.size() * sizeof(int), hipMemcpyHostToDevice) !=
           hipSuccess))
    return -2;
  // rows: u0, x, acc, k (3 each), fields (4), Q (4), F (4), S (2), UH (2),
  // G (3) of nm rows; MF (9 ng), WS (8 ng)
  const size_t row = (size_t)npad, vec = row * nm;
  const size_t total = vec * (12 + 4 + 15) + row * 17 * ng;
  if (hipMalloc((void **)&p->dwork, total * sizeof(double) + 256) != hipSuccess) return -2;
  if (hipMemset(p->dwork, 0, total * sizeof(double)) != hipSuccess) return -2;
  double *w = p->dwork;
  auto take = [&](size_t n) {
    double *r = w;
    w += n;
    return r;
  };
  for (int c = 0; c < 3; c++) p->u0[c] = take(vec);
  for (int c = 0; c < 3; c++) p->x[c] = take(vec);
  for (int c = 0; c < 3; c++) p->acc[c] = take(vec);
  for (int c = 0; c < 3; c++) p->k[c] = take(vec);
  for (int c = 0; c < 4; c++) p->fld[c] = take(vec);
  for (int c = 0; c < 4; c++) p->Q[c] = take(vec);
  for (int c = 0; c < 4; c++) p->F[c] = take(vec);
  for (int c = 0; c < 2; c++) p->S[c] = take(vec);
  for (int c = 0; c < 2; c++) p->UH[c] = take(vec);
  for (int c = 0; c < 3; c++) p->G[c] = take(vec);
  p->MF = take(row * 9 * ng);
  p->WS = take(row * 8 * ng);
  p->ne = ne;
  p->nf = nf;
  p->nblk = nblk;
  p->nwall = (int)wall.size();
  p->face = dface;
  p->mesh = F64Mesh{ne, nf, nblk, p->nwall, dtseg, dssit, dnorm, dgeo, dib, dwall};
  return 0;
}

// Fields from the packed (nm, ne) device arrays fc, w0, z0, hb (any may be
// null: zero), hh0e [ne] on the device.
int f64_fields_set(F64Plan *p, hipStream_t s, const DevTables &tab, const double *fc,
                   const double *w0, const double *z0, const double *hb, const double *hh0e,
                   double mu0, double mb0, double g0) {
  if (!p->dwork) return -3;
  const F64Args a = args(p, s, tab, 0, -1, 0.0);
  const double *src[4] = {fc, w0, z0, hb};
  for (int c = 0; c < 4; c++) {
    if (!src[c]) continue;
    int rc = call(p, a, 0, src[c], p->fld[c]);
    if (rc) return rc;
  }
  if (hipStreamSynchronize(s) != hipSuccess) return -2;
  p->hh0e = hh0e;
  p->mu0 = mu0;
  p->mb0 = mb0;
  p->g0 = g0;
  p->have_fields = true;
  return 0;
}

// Phase 0 or 1 (layout change) of the three fields in[f] -> out[f]: one
// launch when both sides hold them at equal strides, else one per field.
static int layout3(F64Plan *p, F64Args a, int phase, const double *const *in,
                   double *const *out) {
  auto stride = [](const void *p0, const void *p1, const void *p2) -> size_t {
    const uintptr_t a0 = (uintptr_t)p0, a1 = (uintptr_t)p1, a2 = (uintptr_t)p2;
    if (a1 <= a0 || a2 - a1 != a1 - a0 || (a1 - a0) % sizeof(double)) return 0;
    return (a1 - a0) / sizeof(double);
  };
  const size_t si = stride(in[0], in[1], in[2]), so = stride(out[0], out[1], out[2]);
  if (si && so) {
    a.nf = 3;
    a.sin = si;
    a.sout = so;
    return call(p, a, phase, in[0], out[0]);
  }
  for (int c = 0; c < 3; c++) {
    const int rc = call(p, a, phase, in[c], out[c]);
    if (rc) return rc;
  }
  return 0;
}

// Tendencies of the packed device state x[3] into k[3] (packed).
int f64_rhs(F64Plan *p, hipStream_t s, const DevTables &tab, int variant,
            const double *const *x, double *const *k, TimerHook hook, void *user) {
  if (!p->dwork || !p->have_fields) return -3;
  F64Args a = args(p, s, tab, variant, -1, 0.0);
  int rc;
  if ((rc = layout3(p, a, 0, x, p->x))) return rc;
  for (int c = 0; c < 3; c++) a.x[c] = p->x[c];
  a.hook = hook;
  a.hook_user = user;
  if ((rc = call(p, a, 2, nullptr, nullptr))) return rc;
  return layout3(p, a, 1, p->k, k);
}

// The RK4 loop in three parts, so that the caller can capture the step in a
// graph: begin (packed state -> blocked), one step (4 stages), end.
int f64_rk4_begin(F64Plan *p, hipStream_t s, const DevTables &tab,
                  const double *const *u0) {
  if (!p->dwork || !p->have_fields) return -3;
  const F64Args a = args(p, s, tab, 0, -1, 0.0);
  return layout3(p, a, 0, u0, p->u0);
}

int f64_rk4_step(F64Plan *p, hipStream_t s, const DevTables &tab, int variant, double dt,
                 TimerHook hook, void *user) {
  for (int st = 0; st < 4; st++) {
    F64Args a = args(p, s, tab, variant, st, dt);
    a.hook = hook;
    a.hook_user = user;
    int rc = call(p, a, 2, nullptr, nullptr);
    if (rc) return rc;
  }
  return 0;
}

int f64_rk4_end(F64Plan *p, hipStream_t s, const DevTables &tab, double *const *u0) {
  const F64Args a = args(p, s, tab, 0, -1, 0.0);
  return layout3(p, a, 1, p->u0, u0);
}

// Kernels per right-hand side (for the dispatch bound of the RK4 driver).
int f64_kernels_per_rhs(const F64Plan *p, int variant) {
  const bool walls = p->nwall > 0;
  if (variant == 0) return 2 + walls;
  return 5 + (walls ? 2 + (variant == 1) : 0);
}

}  // namespace spoc
