// Host side of the mixed-precision MFMA path: operators of the
// mass-orthonormal basis (formed in FP64 from the reference tables, stored
// as MFMA operand images in FP32 and FP64), per-element mesh data, device
// buffers and the launch sequence.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "f32_types.h"
#include "launch.h"

namespace spoc {

#define SPOC_DECL_F32(NC_, NG_, NGF_)                                                        \
  int rhs_f32_##NC_##_##NG_##_##NGF_(const F32Args &a, int phase, const double *in, double *out); \
  int adj_f64_##NC_##_##NG_##_##NGF_(const F32Args &a, int phase, const double *in, double *out);
SPOC_FOR_EACH_SHAPE(SPOC_DECL_F32)
#undef SPOC_DECL_F32

namespace {
int up(int x, int m) { return (x + m - 1) / m * m; }
// floats of the quadrature-point values: blocks x k-steps x 64 lanes x 4
size_t qf_floats(int nepad, int ng) { return (size_t)(nepad / 16) * ((ng + 3) / 4) * 256; }

// Inverse of an upper-triangular matrix (column-major, ld), FP64.
std::vector<double> upper_inverse(int n, const double *U, int ld) {
  std::vector<double> X((size_t)n * n, 0.0);
  for (int c = 0; c < n; c++) {
    // solve U x = e_c by back substitution
    for (int i = n - 1; i >= 0; i--) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = i + 1; k < n; k++) s -= U[i + (size_t)ld * k] * X[k + (size_t)n * c];
      X[i + (size_t)n * c] = s / U[i + (size_t)ld * i];
    }
  }
  return X;
}

// FP32 image of the R x K operator op(i, k), rows permuted by sigma_row.
template <class F>
void img32(float *dst, int R, int K, F op) {
  const int RB = cdiv(R, 16), KQ = cdiv(K, 16);
  for (int rb = 0; rb < RB; rb++)
    for (int q = 0; q < KQ; q++)
      for (int l = 0; l < 64; l++)
        for (int t = 0; t < 4; t++) {
          int i = sigma_row(16 * rb + (l & 15)), k = 4 * (4 * q + t) + (l >> 4);
          dst[((size_t)(rb * KQ + q) * 64 + l) * 4 + t] = (i < R && k < K) ? (float)op(i, k) : 0.0f;
        }
}

// FP64 image of op(i, k), rows in order.
template <class F>
void img64(double *dst, int R, int K, F op) {
  const int RB = cdiv(R, 16), KH = cdiv(K, 8);
  for (int rb = 0; rb < RB; rb++)
    for (int h = 0; h < KH; h++)
      for (int l = 0; l < 64; l++)
        for (int t = 0; t < 2; t++) {
          int i = 16 * rb + (l & 15), k = 4 * (2 * h + t) + (l >> 4);
          dst[((size_t)(rb * KH + h) * 64 + l) * 2 + t] = (i < R && k < K) ? op(i, k) : 0.0;
        }
}
}  // namespace

struct F32Plan {
  int nc, ng, ngf, nm, nl;
  float *dpack32 = nullptr, *dgl32 = nullptr;
  double *dpack64 = nullptr, *dgl64 = nullptr, *dpack64t = nullptr;
  F32Ops ops{};
  // mesh
  int ne = 0, nepad = 0;
  bool walls_on_edge1 = true;
  void *dmesh = nullptr;
  F32Mesh mesh{};
  // fields
  float *dfld = nullptr;
  double *dhh0e = nullptr;
  F32Fields fld{};
  bool have_fields = false;
  // state/work
  void *dstate = nullptr;
  F32State st{};
  // workgroup tiles per variant: [0] lin or bc kernel, [1] eg kernel;
  // tuned = picked by timing on this mesh (0: not yet), forced = set by the
  // caller (0: automatic)
  int tuned[4][2] = {};
  int forced[4][2] = {};
  bool autotune = true;
  double lf_alpha = 0.0;  // Lax-Friedrichs option (0: central flux)
  // persistent linear RK4: its abort word and per-block stage counts
  // (PERSIST_SYNC_WORDS, allocated at its first use)
  unsigned *dsync = nullptr;
};

void f32_plan_destroy(F32Plan *p);

static int call(const F32Plan *p, const F32Args &a, int phase, const double *in, double *out) {
#define SPOC_CALL(NC_, NG_, NGF_) \
  if (p->nc == NC_ && p->ng == NG_ && p->ngf == NGF_) return rhs_f32_##NC_##_##NG_##_##NGF_(a, phase, in, out);
  SPOC_FOR_EACH_SHAPE(SPOC_CALL)
#undef SPOC_CALL
  return -1;
}

static int call_adj(const F32Plan *p, const F32Args &a, int phase, const double *in, double *out) {
#define SPOC_CALL(NC_, NG_, NGF_) \
  if (p->nc == NC_ && p->ng == NG_ && p->ngf == NGF_) return adj_f64_##NC_##_##NG_##_##NGF_(a, phase, in, out);
  SPOC_FOR_EACH_SHAPE(SPOC_CALL)
#undef SPOC_CALL
  return -1;
}

int f32_plan_create(F32Plan **out, int nc, int ng, int ngf, const double *amm, const double *amml,
                    const double *agx, const double *agy, const double *pi2, const double *pico,
                    const double *wi, const double *wif, int ld_g, int ld_f, int ld_m,
                    const double *dU) {
  F32Plan *p = new F32Plan();
  p->nc = nc;
  p->ng = ng;
  p->ngf = ngf;
  int nm = nm_of(nc), nl = nm - nc - 1, ne3 = 3 * ngf;
  p->nm = nm;
  p->nl = nl;
  // FP64 building blocks
  std::vector<double> Ui = upper_inverse(nm, amm, ld_m);  // U^{-1}
  auto UI = [&](int a, int b) { return Ui[a + (size_t)nm * b]; };
  std::vector<double> ULi = upper_inverse(nl, amml, ld_m);
  auto AGX = [&](int a, int b) { return agx[a + (size_t)ld_m * b]; };
  auto AGY = [&](int a, int b) { return agy[a + (size_t)ld_m * b]; };
  auto PI2 = [&](int q, int m) { return pi2[q + (size_t)ld_g * m]; };
  auto PICO = [&](int j, int m, int k) { return pico[j + (size_t)ld_f * (m + (size_t)ld_m * k)]; };
  // Dx = U^-T agx^T U^-1: Dx(i,j) = sum_{a,b} UI(a,i) agx(b,a) UI(b,j)
  std::vector<double> T1((size_t)nm * nm), DX((size_t)nm * nm), DY((size_t)nm * nm);
  for (int pass = 0; pass < 2; pass++) {
    // T1(a, j) = sum_b agx(b, a) UI(b, j)
    for (int a = 0; a < nm; a++)
      for (int j = 0; j < nm; j++) {
        double s = 0.0;
        for (int b = 0; b < nm; b++) s += (pass ? AGY(b, a) : AGX(b, a)) * UI(b, j);
        T1[a + (size_t)nm * j] = s;
      }
    std::vector<double> &D = pass ? DY : DX;
    for (int i = 0; i < nm; i++)
      for (int j = 0; j < nm; j++) {
        double s = 0.0;
        for (int a = 0; a < nm; a++) s += UI(a, i) * T1[a + (size_t)nm * j];
        D[i + (size_t)nm * j] = s;
      }
  }
  // PSI(q, m) = sum_a pi2(q, a) UI(a, m)
  std::vector<double> PSI((size_t)ng * nm);
  for (int q = 0; q < ng; q++)
    for (int m = 0; m < nm; m++) {
      double s = 0.0;
      for (int a = 0; a < nm; a++) s += PI2(q, a) * UI(a, m);
      PSI[q + (size_t)ng * m] = s;
    }
  // TR((k,j), m) = sum_a pico(j, a, k) UI(a, m)
  // LF(i, (k,j)) = sum_a UI(a, i) wif(j) pico(j, a, k)
  std::vector<double> TR((size_t)ne3 * nm), LF((size_t)nm * ne3);
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < ngf; j++)
      for (int m = 0; m < nm; m++) {
        double s = 0.0, t = 0.0;
        for (int a = 0; a < nm; a++) {
          s += PICO(j, a, k) * UI(a, m);
          t += UI(a, m) * wif[j] * PICO(j, a, k);
        }
        TR[(k * ngf + j) + (size_t)ne3 * m] = s;
        LF[m + (size_t)nm * (k * ngf + j)] = t;
      }
  // wall projector: G = U R^T Ml^{-1} R U^T with R the reduction of
  // src/mat1.f:418-439 and Ml = R amm R^T (the reference's amml)
  std::vector<double> R((size_t)nl * nm, 0.0);
  for (int j = 0; j <= nc - 1; j++) {
    int c1 = ip_of(nc, j), c2 = ip_of(nc - 1, j);
    for (int l = 0; l <= nc - 1 - j; l++) {
      R[(c2 + l) + (size_t)nl * (c1 + l + 1)] = 1.0;
      R[(c2 + l) + (size_t)nl * c1] = (l % 2 == 0) ? 1.0 : -1.0;
    }
  }
  // Z = U R^T (nm x nl); Ml^{-1} = ULi ULi^T
  std::vector<double> Z((size_t)nm * nl, 0.0), MLi((size_t)nl * nl, 0.0), GL((size_t)nm * nm);
  for (int i = 0; i < nm; i++)
    for (int c = 0; c < nl; c++) {
      double s = 0.0;
      for (int k = i; k < nm; k++) s += amm[i + (size_t)ld_m * k] * R[c + (size_t)nl * k];
      Z[i + (size_t)nm * c] = s;
    }
  for (int a = 0; a < nl; a++)
    for (int b = 0; b < nl; b++) {
      double s = 0.0;
      for (int k = 0; k < nl; k++) s += ULi[a + (size_t)nl * k] * ULi[b + (size_t)nl * k];
      MLi[a + (size_t)nl * b] = s;
    }
  for (int i = 0; i < nm; i++)
    for (int j = 0; j < nm; j++) {
      double s = 0.0;
      for (int a = 0; a < nl; a++) {
        double t = 0.0;
        for (int b = 0; b < nl; b++) t += MLi[a + (size_t)nl * b] * Z[j + (size_t)nm * b];
        s += Z[i + (size_t)nm * a] * t;
      }
      GL[i * nm + j] = s - (i == j ? 1.0 : 0.0);  // row-major
    }
  // device images (layouts in f32_types.h)
  const Pack32 P = pack32(nm, ng, ne3);
  const Pack64 Q = pack64(nm, ne3);
  std::vector<float> h32(P.total, 0.0f), g32(img32_size(nm, nm), 0.0f);
  std::vector<double> h64(Q.total, 0.0), g64(img64_size(nm, nm), 0.0);
  auto dx = [&](int i, int k) { return DX[i + (size_t)nm * k]; };
  auto dy = [&](int i, int k) { return DY[i + (size_t)nm * k]; };
  auto lf = [&](int i, int k) { return LF[i + (size_t)nm * k]; };
  auto tr = [&](int i, int k) { return TR[i + (size_t)ne3 * k]; };
  auto psi = [&](int i, int k) { return PSI[i + (size_t)ng * k]; };
  auto psiw = [&](int i, int k) { return wi[k] * PSI[k + (size_t)ng * i]; };
  auto gl = [&](int i, int k) { return GL[(size_t)i * nm + k]; };
  img32(h32.data() + P.dx, nm, nm, dx);
  img32(h32.data() + P.dy, nm, nm, dy);
  img32(h32.data() + P.lf, nm, ne3, lf);
  img32(h32.data() + P.tr, ne3, nm, tr);
  img32(h32.data() + P.psi, ng, nm, psi);
  img32(h32.data() + P.psiw, nm, ng, psiw);
  for (int q = 0; q < ng; q++) {
    h32[P.xq + q] = (float)PI2(q, nc + 1);  // P_1(xi) = xi
    h32[P.yq + q] = (float)PI2(q, 1);       // P_1(eta) = eta
  }
  for (int m = 0; m < nm; m++) h32[P.tail + m] = (float)psi(ng - 1, m);
  img32(g32.data(), nm, nm, gl);
  img64(h64.data() + Q.dx, nm, nm, dx);
  img64(h64.data() + Q.dy, nm, nm, dy);
  img64(h64.data() + Q.lf, nm, ne3, lf);
  img64(h64.data() + Q.tr, ne3, nm, tr);
  img64(g64.data(), nm, nm, gl);
  const Pack64T QT = pack64t(nm, ng, ne3);
  std::vector<double> h64t(QT.total, 0.0);
  img64(h64t.data() + QT.dxt, nm, nm, [&](int i, int k) { return dx(k, i); });
  img64(h64t.data() + QT.dyt, nm, nm, [&](int i, int k) { return dy(k, i); });
  img64(h64t.data() + QT.lft, ne3, nm, [&](int i, int k) { return lf(k, i); });
  img64(h64t.data() + QT.trt, nm, ne3, [&](int i, int k) { return tr(k, i); });
  img64(h64t.data() + QT.psi, ng, nm, psi);
  img64(h64t.data() + QT.psiwt, ng, nm, [&](int i, int k) { return wi[i] * PSI[i + (size_t)ng * k]; });
  img64(h64t.data() + QT.psit, nm, ng, [&](int i, int k) { return psi(k, i); });
  for (int q = 0; q < ng; q++) {
    h64t[QT.xq + q] = PI2(q, nc + 1);
    h64t[QT.yq + q] = PI2(q, 1);
  }
  auto put = [](auto **d, const auto &h) {
    size_t b = h.size() * sizeof(h[0]);
    return hipMalloc(d, b) == hipSuccess &&
           hipMemcpy(*d, h.data(), b, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!put(&p->dpack32, h32) || !put(&p->dgl32, g32) || !put(&p->dpack64, h64) ||
      !put(&p->dgl64, g64) || !put(&p->dpack64t, h64t)) {
    f32_plan_destroy(p);
    return -2;
  }
  p->ops = F32Ops{p->dpack32, p->dpack64, p->dpack64t, p->dgl32, p->dgl64, dU};
  *out = p;
  return 0;
}

static void free_mesh(F32Plan *p) {
  if (p->dmesh) (void)hipFree(p->dmesh);
  if (p->dfld) (void)hipFree(p->dfld);
  if (p->dhh0e) (void)hipFree(p->dhh0e);
  p->dhh0e = nullptr;
  if (p->dstate) (void)hipFree(p->dstate);
  p->dmesh = p->dfld = nullptr;
  p->dstate = nullptr;
  p->have_fields = false;
}

void f32_plan_destroy(F32Plan *p) {
  if (!p) return;
  free_mesh(p);
  for (void *d : {(void *)p->dpack32, (void *)p->dgl32, (void *)p->dpack64, (void *)p->dgl64,
                  (void *)p->dpack64t, (void *)p->dsync})
    if (d) (void)hipFree(d);
  delete p;
}

int f32_mesh_set(F32Plan *p, int ne, int nf, const int *itm, const int *its, const int *tseg,
                 const int *ssit, const double *norm, const double *vect, const double *normt,
                 const int *bdl) {
  (void)nf;
  free_mesh(p);
  std::memset(p->tuned, 0, sizeof p->tuned);
  int nm = p->nm, ng = p->ng, ngf = p->ngf;
  int nepad = up(ne, 16);
  p->ne = ne;
  p->nepad = nepad;
  std::vector<float> geo((size_t)ne * GEO, 0.0f);
  std::vector<double> geo64((size_t)ne * GEO, 0.0);
  std::vector<int> nbr((size_t)ne * 3), nbe((size_t)ne * 3), b(ne);
  p->walls_on_edge1 = true;
  for (int e = 0; e < ne; e++) {
    float *g = geo.data() + (size_t)e * GEO;
    for (int i = 0; i < 5; i++) g[i] = (float)vect[5 * e + i];
    for (int k = 0; k < 3; k++) {
      int f = itm[3 * e + k] - 1;
      g[5 + k] = (float)norm[2 * f];
      g[8 + k] = (float)norm[2 * f + 1];
      g[11 + k] = (float)its[3 * e + k];
      int a = tseg[2 * f] - 1, c = tseg[2 * f + 1] - 1;
      if (c < 0) {
        nbr[3 * e + k] = -1;
        nbe[3 * e + k] = 0;
        if (k != 0) p->walls_on_edge1 = false;
      } else if (a == e) {
        nbr[3 * e + k] = c;
        nbe[3 * e + k] = ssit[2 * f + 1] - 1;
      } else {
        nbr[3 * e + k] = a;
        nbe[3 * e + k] = ssit[2 * f] - 1;
      }
    }
    g[14] = (float)normt[2 * e];
    g[15] = (float)normt[2 * e + 1];
    double *g64 = geo64.data() + (size_t)e * GEO;
    for (int i = 0; i < 5; i++) g64[i] = vect[5 * e + i];
    for (int k = 0; k < 3; k++) {
      int f = itm[3 * e + k] - 1;
      g64[5 + k] = norm[2 * f];
      g64[8 + k] = norm[2 * f + 1];
      g64[11 + k] = (double)its[3 * e + k];
    }
    g64[14] = normt[2 * e];
    g64[15] = normt[2 * e + 1];
    b[e] = bdl[e] != 0;
  }
  size_t bytes = geo64.size() * 8 + geo.size() * 4 + (size_t)ne * 3 * 4 * 2 + (size_t)ne * 4 + 1024;
  if (hipMalloc(&p->dmesh, bytes) != hipSuccess) return -2;
  char *m = (char *)p->dmesh;
  double *dgeo64 = (double *)m;
  float *dgeo = (float *)(dgeo64 + geo64.size());
  int *dnbr = (int *)(dgeo + geo.size()), *dnbe = dnbr + (size_t)ne * 3, *dbdl = dnbe + (size_t)ne * 3;
  if (hipMemcpy(dgeo64, geo64.data(), geo64.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dgeo, geo.data(), geo.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dnbr, nbr.data(), nbr.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dnbe, nbe.data(), nbe.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dbdl, b.data(), b.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return -2;
  p->mesh = F32Mesh{ne, dgeo, dgeo64, dnbr, dnbe, dbdl};
  // state / work (element-blocked, see f32_types.h)
  const int ne3 = 3 * ngf;
  const size_t n64 = (size_t)nepad * nm, nt = (size_t)nepad * ne3;
  // FP64: c0 (3), ys (9), k (3), lam xbar aw avol (12), aF aUH aL (8), xb (3), hbbar (1),
  // gvol (3); trL (4 nt), abar (4 nt), abar2 (2 nt), btr (2 nt).  FP32: trA x2, trS (7 nt)
  const size_t b64 = (42 * n64 + 12 * nt) * 8, b32 = (7 * nt) * 4;
  if (hipMalloc(&p->dstate, b64 + b32 + 256) != hipSuccess) return -2;
  if (hipMemset(p->dstate, 0, b64 + b32 + 256) != hipSuccess) return -2;
  double *s64 = (double *)p->dstate;
  float *s32 = (float *)(s64 + 42 * n64 + 12 * nt);
  F32State &st = p->st;
  for (int f = 0; f < 3; f++) {
    st.c0[f] = s64 + f * n64;
    for (int s = 0; s < 3; s++) st.ys[s][f] = s64 + (3 + 3 * s + f) * n64;
    st.k[f] = s64 + (12 + f) * n64;
  }
  st.trL[0] = s64 + 15 * n64;
  st.trL[1] = st.trL[0] + 2 * nt;
  for (int f = 0; f < 3; f++) {
    double *a = s64 + 15 * n64 + 4 * nt;
    st.lam[f] = a + f * n64;
    st.xbar[f] = a + (3 + f) * n64;
    st.aw[f] = a + (6 + f) * n64;
    st.avol[f] = a + (9 + f) * n64;
  }
  {
    double *a = s64 + 27 * n64 + 4 * nt;
    for (int f = 0; f < 4; f++) st.aF[f] = a + f * n64;
    for (int f = 0; f < 2; f++) {
      st.aUH[f] = a + (4 + f) * n64;
      st.aL[f] = a + (6 + f) * n64;
    }
    for (int f = 0; f < 3; f++) st.xb[f] = a + (8 + f) * n64;
    st.hbbar = a + 11 * n64;
    st.abar = a + 12 * n64;
    st.abar2 = st.abar + 4 * nt;
    st.btr = st.abar2 + 2 * nt;
  }
  for (int f = 0; f < 3; f++) st.gvol[f] = (float *)(s64 + 39 * n64 + 12 * nt) + f * n64;
  st.trA[0] = s32;
  st.trA[1] = s32 + 2 * nt;
  st.trS = s32 + 4 * nt;
  // fields buffer: quadrature-point values (nepad/16 blocks x KG k-steps x
  // 64 lanes x 4 floats), then fc3 (ne*3)
  const size_t nq = qf_floats(nepad, ng);
  if (hipMalloc(&p->dfld, (nq + (size_t)ne * 3) * 4 + 256) != hipSuccess) return -2;
  if (hipMalloc(&p->dhh0e, 2 * (size_t)ne * sizeof(double)) != hipSuccess) return -2;
  if (hipMemset(p->dfld, 0, (nq + (size_t)ne * 3) * 4 + 256) != hipSuccess) return -2;
  return 0;
}

int f32_fields_set(F32Plan *p, int ld, const double *fc, const double *w0, const double *z0,
                   const double *hb, const double *hh0e, const double *pi2, int ld_g, double mu0,
                   double mb0, double g0) {
  int ne = p->ne, nepad = p->nepad, nm = p->nm, ng = p->ng, nc = p->nc;
  const size_t nq = qf_floats(nepad, ng);
  const int kg = (ng + 3) / 4;
  std::vector<float> h(nq + (size_t)ne * 3, 0.0f);
  // word f of point q of element e (see F32Fields::qf)
  auto at = [&](int e, int q, int f) -> float & {
    const size_t ks = q >> 2, l = (size_t)(q & 3) * 16 + (e & 15);
    return h[(((size_t)(e >> 4) * kg + ks) * 64 + l) * 4 + f];
  };
  for (int e = 0; e < nepad; e++)
    for (int q = 0; q < 4 * kg; q++) at(e, q, 0) = 1.0f;  // finite hb where nothing is read
  const double *src[3] = {hb, w0, z0};
  std::vector<double> hcore(ne);  // the nonlinear core's depth (core_depth)
  for (int e = 0; e < ne; e++) hcore[e] = hb[(size_t)ld * e];
  for (int e = 0; e < ne; e++)
    for (int q = 0; q < ng; q++) {
      for (int f = 0; f < 3; f++) {
        if (!src[f]) {
          at(e, q, f) = 0.0f;
          continue;
        }
        double s = 0.0;
        for (int m = 0; m < nm; m++) s += pi2[q + (size_t)ld_g * m] * src[f][m + (size_t)ld * e];
        at(e, q, f) = (float)s;
      }
      double s = 0.0;  // hb - hcore (FP32 mass flux)
      for (int m = 0; m < nm; m++)
        s += pi2[q + (size_t)ld_g * m] * (hb[m + (size_t)ld * e] - (m == 0 ? hcore[e] : 0.0));
      at(e, q, 3) = (float)s;
    }
  if (fc)
    for (int e = 0; e < ne; e++) {
      h[nq + 3 * (size_t)e] = (float)fc[0 + (size_t)ld * e];
      h[nq + 3 * (size_t)e + 1] = (float)fc[1 + (size_t)ld * e];
      h[nq + 3 * (size_t)e + 2] = (float)fc[nc + 1 + (size_t)ld * e];
    }
  if (hipMemcpy(p->dfld, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->dhh0e, hh0e, (size_t)ne * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->dhh0e + ne, hcore.data(), (size_t)ne * sizeof(double),
                hipMemcpyHostToDevice) != hipSuccess)
    return -2;
  p->fld = F32Fields{p->dfld, p->dfld + nq, p->dhh0e, p->dhh0e + ne, (float)mu0, (float)mb0, g0};
  p->have_fields = true;
  return 0;
}

static F32Args args(F32Plan *p, hipStream_t s, int variant, int stage, double dt) {
  F32Args a{};
  a.variant = variant;
  for (int k = 0; k < 2; k++)
    a.tile[k] = p->forced[variant & 3][k] ? p->forced[variant & 3][k]
              : p->tuned[variant & 3][k] ? p->tuned[variant & 3][k] : 4;
  a.stage = stage;
  a.dt = dt;
  a.lf_alpha = p->lf_alpha;
  a.stream = s;
  a.ops = p->ops;
  a.mesh = p->mesh;
  a.fld = p->fld;
  a.st = p->st;
  return a;
}

static int rhs_kernels(F32Plan *p, const F32Args &a, TimerHook hook, void *user) {
  int rc;
  if (hook) hook(user, 0, a.stream);
  if (a.variant == 0) {
    rc = call(p, a, 3, nullptr, nullptr);
    if (hook) hook(user, 1, a.stream);
    return rc;
  }
  rc = call(p, a, 4, nullptr, nullptr);
  if (rc) return rc;
  if (hook) hook(user, 1, a.stream);
  rc = call(p, a, 5, nullptr, nullptr);
  if (hook) hook(user, 2, a.stream);
  return rc;
}

// Time each element kernel of `variant` in every compiled form (register
// budget, E+G split) at every workgroup size in TILE_CHOICES on the current
// mesh (tendency evaluations, which leave the RK4 state untouched) and keep
// the fastest, coded as waves + 16 x form.  Results do not depend on the
// choice: it only changes how element blocks map to waves and workgroups
// and how the compiler allocates registers.
int f32_ready(const F32Plan *p);
int f32_autotune(F32Plan *p, hipStream_t s, int variant, int *out) {
  int rc = f32_ready(p);
  if (rc) return rc;
  // timed as a middle RK4 stage with dt = 0 (the path a step runs: RK
  // update and next-stage traces); it only writes work arrays that every
  // RK4 start and every right-hand side evaluation set up afresh
  F32Args a = args(p, s, variant, 1, 0.0);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -2;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -2;
  }
  // ms of `reps` launches of tile code `code` after one untimed launch
  auto time_tile = [&](int k, int phase, int code, int reps, float *ms) -> int {
    a.tile[k] = code;
    int r = call(p, a, phase, nullptr, nullptr);  // warm-up
    if (r) return r;
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < reps && !r; i++) r = call(p, a, phase, nullptr, nullptr);
    (void)hipEventRecord(e1, s);
    if (r) return r;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms, e0, e1) != hipSuccess)
      return -2;
    return 0;
  };
  const int nk = variant == 0 ? 1 : 2;
  for (int k = 0; k < nk && !rc; k++) {
    const int phase = variant == 0 ? 3 : 4 + k;
    // every candidate once, then the three fastest again, in turns (four
    // rounds, the best of each kept), so that noise or a clock that settles
    // during the first pass does not decide between close candidates
    constexpr int kFinal = 3, kRounds = 4;
    std::vector<std::pair<float, int>> cand;
    for (int code = 0; code < 16 * TILE_FORMS && !rc; code++) {
      const int wg = code & 15;
      if (wg != TILE_CHOICES[0] && wg != TILE_CHOICES[1] && wg != TILE_CHOICES[2]) continue;
      float ms = 0.f;
      if ((rc = time_tile(k, phase, code, 3, &ms))) break;
      cand.push_back({ms, code});
    }
    if (rc) break;
    std::sort(cand.begin(), cand.end());
    const int nf = std::min<int>(kFinal, (int)cand.size());
    std::vector<float> best(nf, 1e30f);
    for (int round = 0; round < kRounds && !rc; round++)
      for (int i = 0; i < nf && !rc; i++) {
        float ms = 0.f;
        if ((rc = time_tile(k, phase, cand[i].second, 4, &ms))) break;
        best[i] = std::min(best[i], ms);
      }
    if (rc) break;
    int pick = cand.empty() ? 4 : cand[0].second;
    float pbest = 1e30f;
    for (int i = 0; i < nf; i++)
      if (best[i] < pbest) {
        pbest = best[i];
        pick = cand[i].second;
      }
    a.tile[k] = pick;
    p->tuned[variant & 3][k] = pick;
  }
  if (nk == 1) p->tuned[variant & 3][1] = 4;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (out) {
    out[0] = a.tile[0];
    out[1] = nk == 2 ? a.tile[1] : 0;
  }
  return rc;
}

void f32_set_flux(F32Plan *p, double alpha) { p->lf_alpha = alpha; }
double f32_flux(const F32Plan *p) { return p->lf_alpha; }

void f32_set_tiles(F32Plan *p, int variant, int t0, int t1, int autotune) {
  p->forced[variant & 3][0] = t0;
  p->forced[variant & 3][1] = t1;
  p->autotune = autotune != 0;
}

static int maybe_tune(F32Plan *p, hipStream_t s, int variant) {
  if (!p->autotune || p->forced[variant & 3][0] || p->tuned[variant & 3][0]) return 0;
  return f32_autotune(p, s, variant, nullptr);
}

int f32_ready(const F32Plan *p) {
  if (!p || !p->dmesh) return -3;
  if (!p->have_fields) return -3;
  if (!p->walls_on_edge1) return -1;
  return 0;
}

// Tendencies of the reference-layout FP64 state x[3] (device, [e][nm]) into
// k[3] (device, same layout).
// Phase 0 or 1 (basis change) of the three fields in[f] -> out[f]: one
// launch when both sides hold them at equal strides, else one per field.
static int basis3(const F32Plan *p, F32Args a, int phase, const double *const *in,
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
  for (int f = 0; f < 3; f++) {
    const int rc = call(p, a, phase, in[f], out[f]);
    if (rc) return rc;
  }
  return 0;
}

int f32_rhs(F32Plan *p, hipStream_t s, int variant, const double *const *x, double *const *k,
            TimerHook hook, void *user) {
  int rc = f32_ready(p);
  if (rc || (rc = maybe_tune(p, s, variant))) return rc;
  F32Args a = args(p, s, variant, -1, 0.0);
  if ((rc = basis3(p, a, 0, x, p->st.c0))) return rc;
  if ((rc = call(p, a, 2, nullptr, nullptr))) return rc;
  if ((rc = rhs_kernels(p, a, hook, user))) return rc;
  return basis3(p, a, 1, p->st.k, k);
}

// The RK4 loop in three parts so that the caller can capture the step in a
// graph: begin (reference layout -> psi, stage-0 traces), one step (4
// stages), end (psi -> reference layout).
int f32_rk4_begin(F32Plan *p, hipStream_t s, int variant, double *const *u0) {
  int rc = f32_ready(p);
  if (rc || (rc = maybe_tune(p, s, variant))) return rc;
  F32Args a = args(p, s, variant, -1, 0.0);
  if ((rc = basis3(p, a, 0, u0, p->st.c0))) return rc;
  return call(p, a, 2, nullptr, nullptr);
}

int f32_rk4_step(F32Plan *p, hipStream_t s, int variant, double dt, TimerHook hook, void *user) {
  F32Args a = args(p, s, variant, -1, dt);
  for (int st = 0; st < 4; st++) {
    a.stage = st;
    int rc = rhs_kernels(p, a, hook, user);
    if (rc) return rc;
  }
  return 0;
}

// nsteps linear RK4 steps in the persistent kernel, one launch per
// PERSIST_STEPS steps, after f32_rk4_begin (which made the stage-0 traces).
// Returns 1, launching nothing, where it does not apply: a nonlinear
// variant, the Lax-Friedrichs flux, a mesh of more than persist_blocks(nc)
// blocks.  *abort: the device word that a wave raises when its wait for a
// neighbouring block runs out (every wave then leaves the launch); the caller reads it
// once the stream has drained.  hook (may be null) brackets the launches.
// test_abort: the abort word starts raised, so that every wave leaves at
// its first wait (the failure path, forced from the host).
namespace {
// Persistent launches on one device run one after another, whatever plan or
// stream they come from: two at once could each hold part of the device and
// wait for each other's blocks (until the bounded waits ended both with an
// error).  persist_done[device]: the last one's end.
constexpr int kMaxDevices = 64;
std::mutex persist_mu;
hipEvent_t persist_done[kMaxDevices] = {};
}  // namespace

int f32_rk4_persist(F32Plan *p, hipStream_t s, int variant, double dt, int nsteps,
                    const unsigned **abort, TimerHook hook, void *user, bool test_abort) {
  if (variant != 0 || p->lf_alpha != 0.0 || (p->ne + 15) / 16 > persist_blocks(p->nc) ||
      nsteps < 1)
    return 1;
  if (!p->dsync &&
      hipMalloc((void **)&p->dsync, PERSIST_SYNC_WORDS * sizeof(unsigned)) != hipSuccess) {
    p->dsync = nullptr;
    return -2;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -2;
  std::lock_guard<std::mutex> lock(persist_mu);
  hipEvent_t &done = persist_done[dev];
  if (!done && hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess) {
    done = nullptr;
    return -2;
  }
  if (hipStreamWaitEvent(s, done, 0) != hipSuccess) return -2;
  F32Args a = args(p, s, variant, 0, dt);
  a.sync = p->dsync;
  a.abort0 = test_abort ? 1u : 0u;
  if (hook) hook(user, 0, s);
  for (int n = 0; n < nsteps; n += PERSIST_STEPS) {
    a.steps = std::min(PERSIST_STEPS, nsteps - n);
    int rc = call(p, a, 6, nullptr, nullptr);
    if (rc) return rc;  // (1 at the first launch: the grid does not fit)
  }
  if (hook) hook(user, 1, s);
  if (hipEventRecord(done, s) != hipSuccess) return -2;
  *abort = p->dsync + 1;
  return 0;
}

int f32_rk4_end(F32Plan *p, hipStream_t s, int variant, double *const *u0) {
  F32Args a = args(p, s, variant, -1, 0.0);
  return basis3(p, a, 1, p->st.c0, u0);
}

// Adjoint (transpose of the linearised right-hand side and of the RK4
// step), FP64.  Phases of the adjoint launcher: 0/1 basis change of adjoint
// variables in/out, 2 w = dt/6 lam, 3/4/5 the passes of R^T (the linear
// variant needs 3 and 4), 6 base-state traces.
static int adj_rhs_passes(F32Plan *p, F32Args &a) {
  int rc;
  if (a.variant != 0 && (rc = call_adj(p, a, 6, nullptr, nullptr))) return rc;
  if ((rc = call_adj(p, a, 3, nullptr, nullptr)) || (rc = call_adj(p, a, 4, nullptr, nullptr)))
    return rc;
  if (a.variant != 0) rc = call_adj(p, a, 5, nullptr, nullptr);
  return rc;
}

// One adjoint right-hand side: y = J_R(x)^T lam (reference coordinates,
// device arrays [e][nm]); x may be null for the linear variant; hb, if not
// null, receives J_{R,hb}(x)^T lam.
int f32_adj_rhs(F32Plan *p, hipStream_t s, int variant, const double *const *x,
                const double *const *lam, double *const *y, double *hb) {
  int rc = f32_ready(p);
  if (rc) return rc;
  F32Args a = args(p, s, variant, -1, 0.0);
  if (variant != 0) {
    if (!x) return -1;
    for (int f = 0; f < 3; f++)
      if ((rc = call(p, a, 0, x[f], p->st.xb[f]))) return rc;
    if (hipMemsetAsync(p->st.hbbar, 0, (size_t)p->nepad * p->nm * sizeof(double), s) !=
        hipSuccess)
      return -2;
  }
  for (int f = 0; f < 3; f++)
    if ((rc = call_adj(p, a, 0, lam[f], p->st.aw[f]))) return rc;
  if ((rc = adj_rhs_passes(p, a))) return rc;
  for (int f = 0; f < 3; f++)
    if ((rc = call_adj(p, a, 1, p->st.k[f], y[f]))) return rc;
  if (hb && variant != 0 && (rc = call_adj(p, a, 1, p->st.hbbar, hb))) return rc;
  return 0;
}

// nsteps adjoint RK4 steps backwards from lam (reference coordinates, in
// place), linearised along the trajectory that starts at x0 (reference
// coordinates; unused by the linear variant), computed by the
// mixed-precision forward path.  ncheck >= 1: checkpoints every `ncheck`
// steps, then per segment the step states are recomputed, then per step its
// four stage inputs.  ncheck <= 0 (automatic): when every step's four stage
// inputs fit in half the device's free memory (at 49,952 triangles and
// p = 5, 100 MB per step) the forward sweep keeps them all and nothing is
// recomputed; otherwise checkpoints every 50 steps.  hb (may be null)
// receives the accumulated adjoint of the bathymetry.
int f32_adj_run(F32Plan *p, hipStream_t s, int variant, double dt, int nsteps, int ncheck,
                const double *const *x0, double *const *lam, double *hb) {
  int rc = f32_ready(p);
  if (rc) return rc;
  F32Args a = args(p, s, variant, -1, dt);
  const size_t n64 = (size_t)p->nepad * p->nm, sbytes = 3 * n64 * sizeof(double);
  double *mem = nullptr;
  int nseg = 0;
  bool keep_all = false;  // every step's four stage inputs from the forward sweep
  if (variant != 0) {
    if (!x0) return -1;
    if (ncheck < 1) {
      size_t freeb = 0, totalb = 0;
      if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return -2;
      keep_all = nsteps > 0 && sbytes * 4 * (size_t)nsteps <= freeb / 2;
      ncheck = keep_all ? std::max(nsteps, 1) : 50;
    }
    nseg = (nsteps + ncheck - 1) / ncheck;
    // keep_all: 4 stage inputs per step; else checkpoints (nseg), one
    // segment's step states (ncheck) and the 4 stage inputs of one step
    const size_t nstates = keep_all ? 4 * (size_t)nsteps : (size_t)(nseg + ncheck + 4);
    if (hipMalloc(&mem, sbytes * nstates) != hipSuccess) return -2;
  }
  auto state = [&](size_t i) { return mem + i * 3 * n64; };
  auto copy = [&](double *dst, const double *src) {
    return hipMemcpyAsync(dst, src, sbytes, hipMemcpyDeviceToDevice, s) == hipSuccess ? 0 : -2;
  };
  double *ckpt0 = state(0), *seg0 = state(nseg), *base0 = state(nseg + ncheck);
  // keep_all: stage input k of step n at state(4 n + k)
  // dispatches queued ahead of the host, bounded per step: a forward step
  // queues at most 8 kernels and a copy, a backward step the 4 stage inputs'
  // recomputation (a copy, the trace pass, 3 right-hand sides, 4 copies) and
  // 4 adjoint stages of at most 4 passes
  QueueBound qb;
  auto fence = [&](int n) { return qb.admit(n) == hipSuccess ? 0 : -2; };
  auto mark = [&]() { return qb.mark(s) == hipSuccess ? 0 : -2; };
  do {
    if (variant != 0) {
      // forward sweep: checkpoints at the start of every segment
      for (int f = 0; f < 3 && !rc; f++) rc = call(p, a, 0, x0[f], p->st.c0[f]);
      if (!rc) rc = maybe_tune(p, s, variant);
      if (rc) break;
      F32Args fa = args(p, s, variant, -1, dt);
      if ((rc = call(p, fa, 2, nullptr, nullptr))) break;
      for (int n = 0; n < nsteps && !rc; n++) {
        rc = fence(13);
        if (keep_all) {
          // stage input 0 is the step's start state, 1..3 the stored Y2..Y4
          if (!rc) rc = copy(state(4 * (size_t)n), p->st.c0[0]);
          for (int st = 0; st < 4 && !rc; st++) {
            fa.stage = st;
            rc = rhs_kernels(p, fa, nullptr, nullptr);
            if (!rc && st < 3) rc = copy(state(4 * (size_t)n + st + 1), p->st.ys[st][0]);
          }
        } else {
          if (!rc && n % ncheck == 0)
            rc = copy(ckpt0 + (size_t)(n / ncheck) * 3 * n64, p->st.c0[0]);
          for (int st = 0; st < 4 && !rc; st++) {
            fa.stage = st;
            rc = rhs_kernels(p, fa, nullptr, nullptr);
          }
        }
        if (!rc) rc = mark();
      }
      if (rc) break;
      if (hipMemsetAsync(p->st.hbbar, 0, n64 * sizeof(double), s) != hipSuccess) {
        rc = -2;
        break;
      }
    }
    for (int f = 0; f < 3 && !rc; f++) rc = call_adj(p, a, 0, lam[f], p->st.lam[f]);
    if (rc || (rc = call_adj(p, a, 2, nullptr, nullptr))) break;
    for (int sg = (variant != 0 ? nseg - 1 : 0); sg >= 0 && !rc; sg--) {
      const int n0 = variant != 0 ? sg * ncheck : 0;
      const int n1 = variant != 0 ? std::min(nsteps, n0 + ncheck) : nsteps;
      F32Args fa = args(p, s, variant, -1, dt);
      if (variant != 0 && !keep_all) {
        // the segment's step states
        rc = copy(p->st.c0[0], ckpt0 + (size_t)sg * 3 * n64);
        if (!rc) rc = call(p, fa, 2, nullptr, nullptr);
        for (int n = n0; n < n1 && !rc; n++) {
          rc = fence(9);
          if (!rc) rc = copy(seg0 + (size_t)(n - n0) * 3 * n64, p->st.c0[0]);
          for (int st = 0; st < 4 && !rc && n + 1 < n1; st++) {
            fa.stage = st;
            rc = rhs_kernels(p, fa, nullptr, nullptr);
          }
          if (!rc) rc = mark();
        }
      }
      for (int n = n1 - 1; n >= n0 && !rc; n--) {
        if ((rc = fence(32))) break;
        double *base = keep_all ? state(4 * (size_t)n) : base0;
        if (variant != 0 && !keep_all) {
          // the four stage inputs of step n
          rc = copy(p->st.c0[0], seg0 + (size_t)(n - n0) * 3 * n64);
          if (!rc) rc = call(p, fa, 2, nullptr, nullptr);
          for (int st = 0; st < 4 && !rc; st++) {
            // (the first stage's input is the step's start state itself)
            rc = copy(base0 + (size_t)st * 3 * n64, st == 0 ? p->st.c0[0] : p->st.ys[st - 1][0]);
            if (st < 3 && !rc) {
              fa.stage = st;
              rc = rhs_kernels(p, fa, nullptr, nullptr);
            }
          }
        }
        for (int as = 0; as < 4 && !rc; as++) {
          a.stage = as;
          if (variant != 0)
            for (int f = 0; f < 3; f++) a.st.xb[f] = base + (size_t)(3 - as) * 3 * n64 + f * n64;
          rc = adj_rhs_passes(p, a);
        }
        if (!rc) rc = mark();
      }
    }
    for (int f = 0; f < 3 && !rc; f++) rc = call_adj(p, a, 1, p->st.lam[f], lam[f]);
    if (!rc && hb && variant != 0) rc = call_adj(p, a, 1, p->st.hbbar, hb);
  } while (false);
  if (rc) (void)hipStreamSynchronize(s);  // nothing left queued behind an error
  if (mem) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(mem);
  }
  return rc;
}

}  // namespace spoc
