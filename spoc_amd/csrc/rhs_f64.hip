// Plain FP64 right-hand side for gfx950.
//
// Every value is produced by the same sequence of IEEE operations as the
// reference Fortran (no FMA contraction: this file is compiled with
// -ffp-contract=off), so the device result equals the reference's bit for
// bit.  The parallel structure is the GPU's own: elements are processed in
// tiles of E per 256-thread workgroup with every per-element vector staged in
// LDS; one thread owns one (element, output index) dot product and runs it
// in the reference's summation order; face traces are a separate pass over
// faces because the reference accumulates both sides of a face into one sum
// (src/calflu.f:41-62).
//
// Kernels per right-hand side:
//   linear     : face_lin  -> elem_lin (+ RK4 stage update)
//   nonlinear  : face_uv   -> elem_bc (velocity gradient q, quadrature terms)
//                face_nl2  -> elem_eg (u, v, eta tendencies + RK4 update)
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

namespace spoc {
namespace f64 {

constexpr int NT = 256;  // threads per workgroup

template <int NC, int NG, int NGF>
struct Shape {
  static constexpr int NM = nm_of(NC);
  static constexpr int NL = NM - NC - 1;          // reduced (zero on edge 1) basis
  static constexpr int NMP = NM | 1;              // odd LDS stride: fewer bank conflicts
  static constexpr int NGP = NG | 1;
  static constexpr int E = NM <= 15 ? 32 : 16;    // elements per workgroup
};

// ---------------------------------------------------------------------------
// Per-element building blocks, each the reference routine's arithmetic.
// ---------------------------------------------------------------------------

// ans_i = sum_j mat(j, i) * x(j), j ascending (src/calb2.f:14-18).
template <int NM>
__device__ __forceinline__ double dot_col(const double *__restrict__ matT, int i,
                                          const double *x) {
  double ans = 0.0;
#pragma unroll
  for (int j = 0; j < NM; j++) ans = ans + matT[i + NM * j] * x[j];
  return ans;
}

// Edge lift of one face field to mode i (src/mat1.f:184-207).  fb holds the
// element's 3 faces, NGF values each, for this field.
template <int NM, int NGF>
__device__ __forceinline__ double tri_f(double b, const double *__restrict__ picoT,
                                        const double *__restrict__ wif, int i,
                                        const double *fb, const int *its, double coeff) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int isg = its[k];
    int isg2 = (NGF + 1) * (1 - isg) / 2;
    double sig = (double)isg * coeff * 0.5;
    double ans = 0.0;
#pragma unroll
    for (int j = 1; j <= NGF; j++) {
      int j1 = isg2 + isg * j;
      double var = picoT[i + NM * ((j1 - 1) + NGF * k)] * fb[k * NGF + (j - 1)];
      ans = ans + wif[j - 1] * var;
    }
    b = b + ans * sig;
  }
  return b;
}

// Cholesky solve U^T U x = b with b in registers (src/dtrsm1.f:58-75).
template <int M, int LD>
__device__ __forceinline__ void dtrsm4_reg(double (&b)[M > 0 ? M : 1],
                                           const double *__restrict__ U,
                                           const double *__restrict__ dia) {
#pragma unroll
  for (int i = 0; i < M; i++) {
    double temp = b[i];
#pragma unroll
    for (int k = 0; k < i; k++) temp = temp - U[k + LD * i] * b[k];
    b[i] = temp * dia[i];
  }
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {
    b[k] = b[k] * dia[k];
#pragma unroll
    for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * U[i + LD * k];
  }
}

// Same solve on an LDS/private vector (boundary-element path).
template <int LD>
__device__ void dtrsm4_mem(int m, double *b, const double *__restrict__ U,
                           const double *__restrict__ dia) {
  for (int i = 0; i < m; i++) {
    double temp = b[i];
    for (int k = 0; k < i; k++) temp = temp - U[k + LD * i] * b[k];
    b[i] = temp * dia[i];
  }
  for (int k = m - 1; k >= 0; k--) {
    b[k] = b[k] * dia[k];
    for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * U[i + LD * k];
  }
}

// Projection onto the zero-on-edge-1 basis (src/mat1.f:418-439).
template <int NC>
__device__ void reduc_gn(const double *gn, double *gnr) {
  for (int j = 0; j <= NC - 1; j++) {
    int c1 = ip_of(NC, j), c2 = ip_of(NC - 1, j);
    for (int l = 0; l <= NC - 1 - j; l += 2) gnr[c2 + l] = gn[c1 + l + 1] + gn[c1];
    for (int l = 1; l <= NC - 1 - j; l += 2) gnr[c2 + l] = gn[c1 + l + 1] - gn[c1];
  }
}

// Expansion back to the full basis (src/mat1.f:445-473).
template <int NC>
__device__ void reduc_inv_gn(double *gn, const double *gnr) {
  for (int j = 0; j <= NC - 1; j++) {
    int c1 = ip_of(NC, j), c2 = ip_of(NC - 1, j);
    double s = 0.0;
    for (int l = 0; l <= NC - j - 1; l += 2) s = s + gnr[c2 + l];
    for (int l = 1; l <= NC - j - 1; l += 2) s = s - gnr[c2 + l];
    gn[c1] = s;
    for (int l = 0; l <= NC - j - 1; l++) gn[c1 + l + 1] = gnr[c2 + l];
  }
  gn[nm_of(NC) - 1] = 0.0;
}

// Free-slip projection of a vector field (gx, gy) on a wall element:
// rotate to (n, t), solve the normal part in the reduced basis and the
// tangential part with the full mass matrix, rotate back
// (src/calvarnl_fs.f:447-459, src/calvar_lin.f:184-197).
template <int NC, int NM, int NL>
__device__ void wall_rot_solve(double *gx, double *gy, int sx, double dnx, double dny,
                               const DevTables &T) {
  double gn[NM], gt[NM], bu[NL > 0 ? NL : 1];
  for (int j = 0; j < NM; j++) {
    gn[j] = gx[j * sx] * dnx + gy[j * sx] * dny;
    gt[j] = gx[j * sx] * dny - gy[j * sx] * dnx;
  }
  reduc_gn<NC>(gn, bu);
  dtrsm4_mem<NM>(NL, bu, T.UL, T.dia2);
  dtrsm4_mem<NM>(NM, gt, T.U, T.dia1);
  reduc_inv_gn<NC>(gn, bu);
  for (int j = 0; j < NM; j++) {
    gx[j * sx] = gn[j] * dnx + gt[j] * dny;
    gy[j * sx] = gn[j] * dny - gt[j] * dnx;
  }
}

// No-slip projection: both components in the reduced basis
// (src/calvarnl_ns.f:253-264, 409-421).
template <int NC, int NM, int NL>
__device__ void wall_red_solve(double *g, int sx, const DevTables &T) {
  double v[NM], bu[NL > 0 ? NL : 1];
  for (int j = 0; j < NM; j++) v[j] = g[j * sx];
  reduc_gn<NC>(v, bu);
  dtrsm4_mem<NM>(NL, bu, T.UL, T.dia2);
  reduc_inv_gn<NC>(v, bu);
  for (int j = 0; j < NM; j++) g[j * sx] = v[j];
}

// Free-slip projection of the velocity-gradient tensor on a wall element
// (src/calvarnl_fs.f:171-204).  q[c] are the 4 components xx, xy, yx, yy.
template <int NC, int NM, int NL>
__device__ void wall_tensor_solve(double *qxx, double *qxy, double *qyx, double *qyy,
                                  double dnx, double dny, const DevTables &T) {
  double gnn[NM], gnt[NM], gtn[NM], gtt[NM], bu[NL > 0 ? NL : 1];
  for (int j = 0; j < NM; j++) {
    double gnx = qxx[j] * dnx + qyx[j] * dny, gtx = qxx[j] * dny - qyx[j] * dnx;
    double gny = qxy[j] * dnx + qyy[j] * dny, gty = qxy[j] * dny - qyy[j] * dnx;
    gnn[j] = gnx * dnx + gny * dny;
    gnt[j] = gnx * dny - gny * dnx;
    gtn[j] = gtx * dnx + gty * dny;
    gtt[j] = gtx * dny - gty * dnx;
  }
  reduc_gn<NC>(gtn, bu);
  dtrsm4_mem<NM>(NM, gnn, T.U, T.dia1);
  dtrsm4_mem<NM>(NM, gnt, T.U, T.dia1);
  dtrsm4_mem<NM>(NL, bu, T.UL, T.dia2);
  dtrsm4_mem<NM>(NM, gtt, T.U, T.dia1);
  reduc_inv_gn<NC>(gtn, bu);
  for (int j = 0; j < NM; j++) {
    double gnx = gnn[j] * dnx + gnt[j] * dny, gny = gnn[j] * dny - gnt[j] * dnx;
    double gtx = gtn[j] * dnx + gtt[j] * dny, gty = gtn[j] * dny - gtt[j] * dnx;
    qxx[j] = gnx * dnx + gtx * dny;
    qyx[j] = gnx * dny - gtx * dnx;
    qxy[j] = gny * dnx + gty * dny;
    qyy[j] = gny * dny - gty * dnx;
  }
}

// ---------------------------------------------------------------------------
// Face traces (src/calflu.f:6-71), one thread per face.  Each input field's
// trace is multiplied by a component of the face normal (and for the linear
// eta equation by hh0), as the loops after every cal_flu call do.
// ---------------------------------------------------------------------------
struct FaceSpec {
  const double *in[6];  // modal inputs (NM, ne)
  double pre[6];        // value fed to the trace is pre*a when haspre
  int haspre[6];
  int nout;
  int src[6];           // input field of each output
  int comp[6];          // normal component of each output (0: n_x, 1: n_y)
  double post[6];       // extra factor after the normal when haspost
  int haspost[6];
  const double *hpost;  // if set: per-element factor (of the face's first element) in place of post
};

template <int NC, int NGF, int NIN>
__global__ __launch_bounds__(NT) void face_kernel(int nf, const int *__restrict__ tseg,
                                                  const int *__restrict__ ssit,
                                                  const double *__restrict__ norm,
                                                  const double *__restrict__ pico, FaceSpec spec,
                                                  double *__restrict__ out) {
  constexpr int NM = nm_of(NC);
  int f = blockIdx.x * NT + threadIdx.x;
  if (f >= nf) return;
  int k1 = tseg[2 * f] - 1, k2 = tseg[2 * f + 1] - 1;
  double acc[NIN][NGF];
#pragma unroll
  for (int c = 0; c < NIN; c++)
#pragma unroll
    for (int j = 0; j < NGF; j++) acc[c][j] = 0.0;
  int s1 = (k2 < 0) ? 0 : ssit[2 * f] - 1;
  for (int m = 0; m < NM; m++) {
    const double *pc = pico + NGF * (m + NM * s1);
#pragma unroll
    for (int c = 0; c < NIN; c++) {
      double a = spec.in[c][m + NM * k1];
      if (spec.haspre[c]) a = spec.pre[c] * a;
#pragma unroll
      for (int j = 0; j < NGF; j++) acc[c][j] = acc[c][j] + a * pc[j];
    }
  }
  if (k2 >= 0) {
    int s2 = ssit[2 * f + 1] - 1;
    for (int m = 0; m < NM; m++) {
      const double *pc = pico + NGF * (m + NM * s2);
#pragma unroll
      for (int c = 0; c < NIN; c++) {
        double a = spec.in[c][m + NM * k2];
        if (spec.haspre[c]) a = spec.pre[c] * a;
#pragma unroll
        for (int j = 0; j < NGF; j++) acc[c][j] = acc[c][j] + a * pc[NGF - 1 - j];
      }
    }
#pragma unroll
    for (int c = 0; c < NIN; c++)
#pragma unroll
      for (int j = 0; j < NGF; j++) acc[c][j] = 0.5 * acc[c][j];
  }
  double n[2] = {norm[2 * f], norm[2 * f + 1]};
  for (int o = 0; o < spec.nout; o++) {
    double *dst = out + ((size_t)f * spec.nout + o) * NGF;
    int c = spec.src[o];
    double nn = n[spec.comp[o]], post = spec.hpost ? spec.hpost[k1] : spec.post[o];
    int haspost = spec.haspost[o];
#pragma unroll
    for (int cc = 0; cc < NIN; cc++)
      if (cc == c) {
#pragma unroll
        for (int j = 0; j < NGF; j++) {
          double v = acc[cc][j] * nn;
          if (haspost) v = v * post;
          dst[j] = v;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Tile helpers.
// ---------------------------------------------------------------------------

// Copy NF modal fields of the tile's elements into LDS [field][e][NMP].
template <int NM, int NMP, int E>
__device__ __forceinline__ void load_fields(double *lds, const double *const *src, int nfld,
                                            int e0, int ne) {
  for (int idx = threadIdx.x; idx < nfld * E * NM; idx += NT) {
    int c = idx / (E * NM), r = idx % (E * NM), e = r / NM, m = r % NM;
    double v = 0.0;
    if (e0 + e < ne) v = src[c][(size_t)(e0 + e) * NM + m];
    lds[(c * E + e) * NMP + m] = v;
  }
}

// Gather the tile's face values: fb[((e*3 + k)*NOUT + o)*NGF + j].
template <int NGF, int E, int NOUT>
__device__ __forceinline__ void load_faces(double *fb, const double *__restrict__ face,
                                           const int *__restrict__ itm, int e0, int ne) {
  for (int idx = threadIdx.x; idx < E * 3 * NOUT * NGF; idx += NT) {
    int j = idx % NGF, r = idx / NGF, o = r % NOUT, ek = r / NOUT, k = ek % 3, e = ek / 3;
    double v = 0.0;
    if (e0 + e < ne) {
      int f = itm[3 * (e0 + e) + k] - 1;
      v = face[((size_t)f * NOUT + o) * NGF + j];
    }
    fb[idx] = v;
  }
}

// RK4 stage update of one value (src/shwwnl.f:242-284), stage s = 0..3.
// acc accumulates ((k1 + 2k2) + 2k3) + k4 in the reference's order.
__device__ __forceinline__ void rk_update(int s, double k, double dt, double dt6, double *u0,
                                          double *x, double *acc) {
  if (s == 0) {
    *acc = k;
    *x = *u0 + dt * k * 0.5;
  } else if (s == 1) {
    *acc = *acc + 2.0 * k;
    *x = *u0 + dt * k * 0.5;
  } else if (s == 2) {
    *acc = *acc + 2.0 * k;
    *x = *u0 + dt * k;
  } else {
    double a = *acc + k;
    *u0 = *u0 + dt6 * a;
  }
}

// Where an element kernel sends its tendencies.
struct StageOut {
  double *k[3];        // raw tendencies (spoc_rhs), or null
  double *u0[3];       // RK4 base state (read; written at stage 3)
  double *x[3];        // next stage input
  double *acc[3];      // RK4 accumulator
  int stage;           // 0..3, or -1 for "tendencies only"
  double dt, dt6;
};

template <int NM>
__device__ __forceinline__ void emit(const StageOut &so, int c, size_t g, double kval) {
  if (so.stage < 0) {
    so.k[c][g] = kval;
  } else {
    rk_update(so.stage, kval, so.dt, so.dt6, &so.u0[c][g], &so.x[c][g], &so.acc[c][g]);
  }
}

// ---------------------------------------------------------------------------
// Linear right-hand side element kernel (src/calvar_lin.f:115-282).
// Face inputs per face: fbdxx, fbdyy (of fxx = -g0*h0), fbdx, fbdy (of u0*hh0,
// v0*hh0).  fxy = fyx = 0, whose terms add only zeros in the reference.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF>
__global__ __launch_bounds__(NT) void elem_lin(DevMesh M, DevTables T, Physics P,
                                               const double *__restrict__ u,
                                               const double *__restrict__ v,
                                               const double *__restrict__ h,
                                               const double *__restrict__ face, StageOut so) {
  using S = Shape<NC, NG, NGF>;
  constexpr int NM = S::NM, NMP = S::NMP, E = S::E, NL = S::NL;
  __shared__ double sf[3 * E * NMP];        // fxx, u0, v0
  __shared__ double sg[3 * E * NMP];        // gx, gy, b (tendencies before the solve)
  __shared__ double sfb[E * 3 * 4 * NGF];
  __shared__ double svect[E * 5];
  __shared__ int sits[E * 3], sbdl[E];
  const int e0 = blockIdx.x * E, ne = M.ne;
  {
    const double *src[3] = {h, u, v};
    load_fields<NM, NMP, E>(sf, src, 3, e0, ne);
  }
  load_faces<NGF, E, 4>(sfb, face, M.itm, e0, ne);
  for (int idx = threadIdx.x; idx < E * 5; idx += NT)
    svect[idx] = (e0 + idx / 5 < ne) ? M.vect[(size_t)e0 * 5 + idx] : 1.0;
  for (int idx = threadIdx.x; idx < E * 3; idx += NT)
    sits[idx] = (e0 + idx / 3 < ne) ? M.its[(size_t)e0 * 3 + idx] : 1;
  for (int idx = threadIdx.x; idx < E; idx += NT) sbdl[idx] = (e0 + idx < ne) ? M.bdl[e0 + idx] : 0;
  __syncthreads();
  // fxx(j) = -g0 * h0(j) (src/calvar_lin.f:74-81)
  for (int idx = threadIdx.x; idx < E * NM; idx += NT) {
    int e = idx / NM, m = idx % NM;
    sf[e * NMP + m] = -P.g0 * sf[e * NMP + m];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < E * NM; idx += NT) {
    int e = idx / NM, i = idx % NM;
    const double *vc = svect + 5 * e;
    double ax = vc[0], bx = vc[1], ay = vc[2], by = vc[3], detai = vc[4];
    const double *fx = sf + e * NMP;
    const double *fb = sfb + e * 3 * 4 * NGF;
    double Fx[3 * NGF], Fy[3 * NGF], Fu[3 * NGF], Fv[3 * NGF];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < NGF; j++) {
        Fx[k * NGF + j] = fb[(k * 4 + 0) * NGF + j];
        Fy[k * NGF + j] = fb[(k * 4 + 1) * NGF + j];
        Fu[k * NGF + j] = fb[(k * 4 + 2) * NGF + j];
        Fv[k * NGF + j] = fb[(k * 4 + 3) * NGF + j];
      }
    double sx = dot_col<NM>(T.agxT, i, fx), sy = dot_col<NM>(T.agyT, i, fx);
    // u: cal_b(gx,gxx,agx,-by*detai), cal_b(gx,gxx,agy,ay*detai), tri_f x2
    double gx = 0.0;
    gx = gx + sx * (-by * detai);
    gx = gx + sy * (ay * detai);
    gx = tri_f<NM, NGF>(gx, T.picoT, T.wif, i, Fx, sits + 3 * e, detai);
    // v: cal_b(gy,gyy,agx,-(-bx*detai)), cal_b(gy,gyy,agy,-(ax*detai)), tri_f
    double gy = 0.0;
    gy = gy + sx * (-(-bx * detai));
    gy = gy + sy * (-(ax * detai));
    gy = tri_f<NM, NGF>(gy, T.picoT, T.wif, i, Fy, sits + 3 * e, detai);
    // h (src/calvar_lin.f:244-274)
    const double H = P.hh0e[e0 + e < ne ? e0 + e : 0];
    double ak[NM];
#pragma unroll
    for (int j = 0; j < NM; j++) ak[j] = sf[(E + e) * NMP + j] * H;
    double b = 0.0;
    b = b + dot_col<NM>(T.agxT, i, ak) * (by * detai);
    b = b + dot_col<NM>(T.agyT, i, ak) * (-ay * detai);
#pragma unroll
    for (int j = 0; j < NM; j++) ak[j] = sf[(2 * E + e) * NMP + j] * H;
    b = b + dot_col<NM>(T.agxT, i, ak) * (-bx * detai);
    b = b + dot_col<NM>(T.agyT, i, ak) * (ax * detai);
    b = tri_f<NM, NGF>(b, T.picoT, T.wif, i, Fu, sits + 3 * e, -detai);
    b = tri_f<NM, NGF>(b, T.picoT, T.wif, i, Fv, sits + 3 * e, -detai);
    sg[e * NMP + i] = gx;
    sg[(E + e) * NMP + i] = gy;
    sg[(2 * E + e) * NMP + i] = b;
  }
  __syncthreads();
  // Mass solves: one thread per (element, field); wall elements' velocity
  // pair goes through the free-slip projection in a serial pass.
  for (int idx = threadIdx.x; idx < 3 * E; idx += NT) {
    int c = idx / E, e = idx % E;
    if (c < 2 && sbdl[e]) continue;
    double r[NM];
    double *g = sg + (c * E + e) * NMP;
#pragma unroll
    for (int j = 0; j < NM; j++) r[j] = g[j];
    dtrsm4_reg<NM, NM>(r, T.U, T.dia1);
#pragma unroll
    for (int j = 0; j < NM; j++) g[j] = r[j];
  }
  for (int e = threadIdx.x; e < E; e += NT)
    if (sbdl[e]) {
      const double *nt = M.normt + 2 * (size_t)(e0 + e);
      wall_rot_solve<NC, NM, NL>(sg + e * NMP, sg + (E + e) * NMP, 1, nt[0], nt[1], T);
    }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 3 * E * NM; idx += NT) {
    int c = idx / (E * NM), r = idx % (E * NM), e = r / NM, m = r % NM;
    if (e0 + e >= ne) continue;
    emit<NM>(so, c, (size_t)(e0 + e) * NM + m, sg[(c * E + e) * NMP + m]);
  }
}

// ---------------------------------------------------------------------------
// Nonlinear stage B + C (src/calvarnl_fs.f:115-332 and the ns/inv variants):
// velocity gradient q with its wall projection, then the quadrature-point
// products and the mass-flux solve.  Outputs per element: F = mu0*q - g0*h
// (4 modal fields), the source terms -adu + w1 and -adv + z1, and uh, vh.
// Face inputs: fbdxx, fbdxy, fbdyx, fbdyy of the velocity traces.
// ---------------------------------------------------------------------------
struct BCOut {
  double *F[4];
  double *src[2];
  double *uh[2];
};

template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void elem_bc(DevMesh M, DevTables T, Physics P,
                                              const double *__restrict__ u,
                                              const double *__restrict__ v,
                                              const double *__restrict__ h,
                                              const double *__restrict__ hb,
                                              const double *__restrict__ w0,
                                              const double *__restrict__ z0,
                                              const double *__restrict__ fc,
                                              const double *__restrict__ face, BCOut out) {
  using S = Shape<NC, NG, NGF>;
  constexpr int NM = S::NM, NMP = S::NMP, E = S::E, NL = S::NL, NGP = S::NGP;
  // modal fields: 0 u, 1 v, 2 qxx, 3 qxy, 4 qyx, 5 qyy, 6 h, 7 hb, 8 w0, 9 z0, 10 fc
  __shared__ double sm[11 * E * NMP];
  __shared__ double sq[8 * E * NGP];        // products at quadrature points
  __shared__ double sfb[E * 3 * 4 * NGF];
  __shared__ double svect[E * 5];
  __shared__ int sits[E * 3], sbdl[E];
  const int e0 = blockIdx.x * E, ne = M.ne;
  {
    const double *src[2] = {u, v};
    load_fields<NM, NMP, E>(sm, src, 2, e0, ne);
    const double *src2[5] = {h, hb, w0, z0, fc};
    load_fields<NM, NMP, E>(sm + 6 * E * NMP, src2, 5, e0, ne);
  }
  load_faces<NGF, E, 4>(sfb, face, M.itm, e0, ne);
  for (int idx = threadIdx.x; idx < E * 5; idx += NT)
    svect[idx] = (e0 + idx / 5 < ne) ? M.vect[(size_t)e0 * 5 + idx] : 1.0;
  for (int idx = threadIdx.x; idx < E * 3; idx += NT)
    sits[idx] = (e0 + idx / 3 < ne) ? M.its[(size_t)e0 * 3 + idx] : 1;
  for (int idx = threadIdx.x; idx < E; idx += NT) sbdl[idx] = (e0 + idx < ne) ? M.bdl[e0 + idx] : 0;
  __syncthreads();
  auto fld = [&](int c, int e) { return sm + (c * E + e) * NMP; };
  // B: q = weak gradient of u and v (src/calvarnl_fs.f:129-158)
  for (int idx = threadIdx.x; idx < 2 * E * NM; idx += NT) {
    int c = idx / (E * NM), r = idx % (E * NM), e = r / NM, i = r % NM;
    const double *vc = svect + 5 * e;
    double ax = vc[0], bx = vc[1], ay = vc[2], by = vc[3], detai = vc[4];
    const double *a = fld(c, e);
    double gx1 = 0.0, gy1 = 0.0;
    gx1 = gx1 + dot_col<NM>(T.agxT, i, a) * (-1.0);
    gy1 = gy1 + dot_col<NM>(T.agyT, i, a) * (-1.0);
    double gA = (by * gx1 - ay * gy1) * detai;
    double gB = (-bx * gx1 + ax * gy1) * detai;
    const double *fb = sfb + e * 3 * 4 * NGF;
    double FA[3 * NGF], FB[3 * NGF];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < NGF; j++) {
        FA[k * NGF + j] = fb[(k * 4 + 2 * c) * NGF + j];
        FB[k * NGF + j] = fb[(k * 4 + 2 * c + 1) * NGF + j];
      }
    gA = tri_f<NM, NGF>(gA, T.picoT, T.wif, i, FA, sits + 3 * e, detai);
    gB = tri_f<NM, NGF>(gB, T.picoT, T.wif, i, FB, sits + 3 * e, detai);
    fld(2 + 2 * c, e)[i] = gA;      // qxx or qyx
    fld(3 + 2 * c, e)[i] = gB;      // qxy or qyy
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 4 * E; idx += NT) {
    int c = idx / E, e = idx % E;
    if (VAR == 1 && sbdl[e]) continue;
    double r[NM];
    double *g = fld(2 + c, e);
#pragma unroll
    for (int j = 0; j < NM; j++) r[j] = g[j];
    dtrsm4_reg<NM, NM>(r, T.U, T.dia1);
#pragma unroll
    for (int j = 0; j < NM; j++) g[j] = r[j];
  }
  if (VAR == 1) {
    for (int e = threadIdx.x; e < E; e += NT)
      if (sbdl[e]) {
        const double *nt = M.normt + 2 * (size_t)(e0 + e);
        wall_tensor_solve<NC, NM, NL>(fld(2, e), fld(3, e), fld(4, e), fld(5, e), nt[0], nt[1],
                                      T);
      }
  }
  __syncthreads();
  // F = mu0*q - g0*h (src/calvarnl_fs.f:341-348); modal combinations of
  // stage C (src/calvarnl_fs.f:245-255) overwrite h, hb, w0, z0 slots.
  for (int idx = threadIdx.x; idx < E * NM; idx += NT) {
    int e = idx / NM, m = idx % NM;
    if (e0 + e < ne) {
      size_t g = (size_t)(e0 + e) * NM + m;
      double h0v = fld(6, e)[m];
      out.F[0][g] = P.mu0 * fld(2, e)[m] - P.g0 * h0v;
      out.F[1][g] = P.mu0 * fld(3, e)[m];
      out.F[2][g] = P.mu0 * fld(4, e)[m];
      out.F[3][g] = P.mu0 * fld(5, e)[m] - P.g0 * h0v;
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < E * NM; idx += NT) {
    int e = idx / NM, m = idx % NM;
    double uk = fld(0, e)[m], vk = fld(1, e)[m], f = fld(10, e)[m];
    double hk = fld(6, e)[m] + fld(7, e)[m];
    double wk = fld(8, e)[m] - P.mb0 * uk;
    double zk = fld(9, e)[m] - P.mb0 * vk;
    fld(3, e)[m] = fld(3, e)[m] - f;   // gxy = qxy - fc
    fld(4, e)[m] = fld(4, e)[m] + f;   // gyx = qyx + fc
    fld(6, e)[m] = hk;
    fld(8, e)[m] = wk;
    fld(9, e)[m] = zk;
  }
  __syncthreads();
  // C: values at quadrature points and their products (src/nonlin.f:62-86,
  // 111-133; src/calvarnl_fs.f:257-296)
  for (int idx = threadIdx.x; idx < E * NG; idx += NT) {
    int e = idx / NG, q = idx % NG;
    double mf[9];
    const int cs[9] = {0, 1, 2, 3, 4, 5, 6, 8, 9};
#pragma unroll
    for (int c = 0; c < 9; c++) mf[c] = 0.0;
    for (int m = 0; m < NM; m++) {
      double p = T.pi2[q + NG * m];
#pragma unroll
      for (int c = 0; c < 9; c++) mf[c] = mf[c] + fld(cs[c], e)[m] * p;
    }
    // mf: 0 u, 1 v, 2 gxx, 3 gxy, 4 gyx, 5 gyy, 6 hk, 7 wk, 8 zk
    double *o = sq + e * NGP + q;
    o[0 * E * NGP] = mf[0] * mf[2];   // uux
    o[1 * E * NGP] = mf[1] * mf[3];   // vuy
    o[2 * E * NGP] = mf[0] * mf[4];   // uvx
    o[3 * E * NGP] = mf[1] * mf[5];   // vvy
    o[4 * E * NGP] = mf[0] * mf[6];   // uhk
    o[5 * E * NGP] = mf[1] * mf[6];   // vhk
    o[6 * E * NGP] = mf[7] / mf[6];   // wk
    o[7 * E * NGP] = mf[8] / mf[6];   // zk
  }
  __syncthreads();
  // xytospec of the 8 products (src/nonlin.f:35-57, 91-107).  Results reuse
  // LDS slots: 2 -> src_x, 3 -> src_y, 4 -> uhk, 5 -> vhk.
  for (int idx = threadIdx.x; idx < E * NM; idx += NT) {
    int e = idx / NM, m = idx % NM;
    double r[8];
#pragma unroll
    for (int c = 0; c < 8; c++) r[c] = 0.0;
    for (int q = 0; q < NG; q++) {
      double p = T.pi2T[m + NM * q], w = T.wi[q];
#pragma unroll
      for (int c = 0; c < 8; c++) r[c] = r[c] + w * sq[(c * E + e) * NGP + q] * p;
    }
    double adu = r[0] + r[1], adv = r[2] + r[3];
    fld(2, e)[m] = -adu + r[6];
    fld(3, e)[m] = -adv + r[7];
    fld(4, e)[m] = r[4];
    fld(5, e)[m] = r[5];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 2 * E; idx += NT) {
    int c = idx / E, e = idx % E;
    if (sbdl[e]) continue;
    double r[NM];
    double *g = fld(4 + c, e);
#pragma unroll
    for (int j = 0; j < NM; j++) r[j] = g[j];
    dtrsm4_reg<NM, NM>(r, T.U, T.dia1);
#pragma unroll
    for (int j = 0; j < NM; j++) g[j] = r[j];
  }
  for (int e = threadIdx.x; e < E; e += NT)
    if (sbdl[e]) {
      if (VAR == 2) {
        wall_red_solve<NC, NM, NL>(fld(4, e), 1, T);
        wall_red_solve<NC, NM, NL>(fld(5, e), 1, T);
      } else {
        const double *nt = M.normt + 2 * (size_t)(e0 + e);
        wall_rot_solve<NC, NM, NL>(fld(4, e), fld(5, e), 1, nt[0], nt[1], T);
      }
    }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 4 * E * NM; idx += NT) {
    int c = idx / (E * NM), r = idx % (E * NM), e = r / NM, m = r % NM;
    if (e0 + e >= ne) continue;
    size_t g = (size_t)(e0 + e) * NM + m;
    double val = fld(2 + c, e)[m];
    if (c < 2) out.src[c][g] = val;
    else out.uh[c - 2][g] = val;
  }
}

// ---------------------------------------------------------------------------
// Nonlinear stages E + G (src/calvarnl_fs.f:382-544): u, v tendencies from
// the stress tensor, sources and wall projection; eta tendency from the mass
// fluxes.  Face inputs per face: fbdxx, fbdxy, fbdyx, fbdyy (of F) and fbdx,
// fbdy (of uh, vh).
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void elem_eg(DevMesh M, DevTables T, Physics P,
                                              const double *const *__restrict__ Fp,
                                              const double *__restrict__ srcx,
                                              const double *__restrict__ srcy,
                                              const double *__restrict__ uh,
                                              const double *__restrict__ vh,
                                              const double *__restrict__ face, StageOut so) {
  using S = Shape<NC, NG, NGF>;
  constexpr int NM = S::NM, NMP = S::NMP, E = S::E, NL = S::NL;
  // 0..3 F, 4 uh, 5 vh ; outputs 0 gx, 1 gy, 2 b
  __shared__ double sm[6 * E * NMP];
  __shared__ double sg[3 * E * NMP];
  __shared__ double sfb[E * 3 * 6 * NGF];
  __shared__ double svect[E * 5];
  __shared__ int sits[E * 3], sbdl[E];
  const int e0 = blockIdx.x * E, ne = M.ne;
  {
    const double *src[6] = {Fp[0], Fp[1], Fp[2], Fp[3], uh, vh};
    load_fields<NM, NMP, E>(sm, src, 6, e0, ne);
    const double *src2[2] = {srcx, srcy};
    load_fields<NM, NMP, E>(sg, src2, 2, e0, ne);
  }
  load_faces<NGF, E, 6>(sfb, face, M.itm, e0, ne);
  for (int idx = threadIdx.x; idx < E * 5; idx += NT)
    svect[idx] = (e0 + idx / 5 < ne) ? M.vect[(size_t)e0 * 5 + idx] : 1.0;
  for (int idx = threadIdx.x; idx < E * 3; idx += NT)
    sits[idx] = (e0 + idx / 3 < ne) ? M.its[(size_t)e0 * 3 + idx] : 1;
  for (int idx = threadIdx.x; idx < E; idx += NT) sbdl[idx] = (e0 + idx < ne) ? M.bdl[e0 + idx] : 0;
  __syncthreads();
  auto fld = [&](int c, int e) { return sm + (c * E + e) * NMP; };
  for (int idx = threadIdx.x; idx < 3 * E * NM; idx += NT) {
    int c = idx / (E * NM), r = idx % (E * NM), e = r / NM, i = r % NM;
    const double *vc = svect + 5 * e;
    double ax = vc[0], bx = vc[1], ay = vc[2], by = vc[3], detai = vc[4];
    const double *fb = sfb + e * 3 * 6 * NGF;
    double FA[3 * NGF], FB[3 * NGF];
    int oa = (c == 2) ? 4 : 2 * c, ob = oa + 1;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < NGF; j++) {
        FA[k * NGF + j] = fb[(k * 6 + oa) * NGF + j];
        FB[k * NGF + j] = fb[(k * 6 + ob) * NGF + j];
      }
    double g;
    if (c < 2) {
      // gx = -adu + w1 ; cal_b x4 ; tri_f x2  (src/calvarnl_fs.f:393-429)
      const double *A = fld(2 * c, e), *B = fld(2 * c + 1, e);
      g = sg[(c * E + e) * NMP + i];
      g = g + dot_col<NM>(T.agxT, i, A) * (-by * detai);
      g = g + dot_col<NM>(T.agyT, i, A) * (ay * detai);
      g = g + dot_col<NM>(T.agxT, i, B) * (bx * detai);
      g = g + dot_col<NM>(T.agyT, i, B) * (-ax * detai);
      g = tri_f<NM, NGF>(g, T.picoT, T.wif, i, FA, sits + 3 * e, detai);
      g = tri_f<NM, NGF>(g, T.picoT, T.wif, i, FB, sits + 3 * e, detai);
    } else {
      // eta (src/calvarnl_fs.f:509-536)
      const double *A = fld(4, e), *B = fld(5, e);
      g = 0.0;
      g = g + dot_col<NM>(T.agxT, i, A) * (by * detai);
      g = g + dot_col<NM>(T.agyT, i, A) * (-ay * detai);
      g = g + dot_col<NM>(T.agxT, i, B) * (-bx * detai);
      g = g + dot_col<NM>(T.agyT, i, B) * (ax * detai);
      g = tri_f<NM, NGF>(g, T.picoT, T.wif, i, FA, sits + 3 * e, -detai);
      g = tri_f<NM, NGF>(g, T.picoT, T.wif, i, FB, sits + 3 * e, -detai);
    }
    sg[(c * E + e) * NMP + i] = g;  // only this thread reads or writes this slot
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 3 * E; idx += NT) {
    int c = idx / E, e = idx % E;
    if (c < 2 && sbdl[e]) continue;
    double r[NM];
    double *g = sg + (c * E + e) * NMP;
#pragma unroll
    for (int j = 0; j < NM; j++) r[j] = g[j];
    dtrsm4_reg<NM, NM>(r, T.U, T.dia1);
#pragma unroll
    for (int j = 0; j < NM; j++) g[j] = r[j];
  }
  for (int e = threadIdx.x; e < E; e += NT)
    if (sbdl[e]) {
      if (VAR == 2) {
        wall_red_solve<NC, NM, NL>(sg + e * NMP, 1, T);
        wall_red_solve<NC, NM, NL>(sg + (E + e) * NMP, 1, T);
      } else {
        const double *nt = M.normt + 2 * (size_t)(e0 + e);
        wall_rot_solve<NC, NM, NL>(sg + e * NMP, sg + (E + e) * NMP, 1, nt[0], nt[1], T);
      }
    }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 3 * E * NM; idx += NT) {
    int c = idx / (E * NM), r = idx % (E * NM), e = r / NM, m = r % NM;
    if (e0 + e >= ne) continue;
    emit<NM>(so, c, (size_t)(e0 + e) * NM + m, sg[(c * E + e) * NMP + m]);
  }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF>
int launch_rhs(const RhsArgs &a) {
  using S = Shape<NC, NG, NGF>;
  constexpr int NM = S::NM, E = S::E;
  const DevMesh &M = a.mesh;
  int nfb = (M.nf + NT - 1) / NT, neb = (M.ne + E - 1) / E;
  hipStream_t st = a.stream;
  Physics P = a.phys;
  StageOut so;
  for (int c = 0; c < 3; c++) {
    so.k[c] = a.k[c];
    so.u0[c] = a.u0[c];
    so.x[c] = a.xout[c];
    so.acc[c] = a.acc[c];
  }
  so.stage = a.stage;
  so.dt = a.dt;
  so.dt6 = a.dt / 6.0;
  if (a.variant == 0) {
    FaceSpec fs = {};
    fs.in[0] = a.x[2];
    fs.pre[0] = -P.g0;
    fs.haspre[0] = 1;
    fs.in[1] = a.x[0];
    fs.in[2] = a.x[1];
    fs.nout = 4;
    int src[4] = {0, 0, 1, 2}, comp[4] = {0, 1, 0, 1};
    for (int o = 0; o < 4; o++) {
      fs.src[o] = src[o];
      fs.comp[o] = comp[o];
      fs.post[o] = P.hh0;
      fs.haspost[o] = o >= 2;
    }
    fs.hpost = P.hh0e;
    a.timer(0, st);
    face_kernel<NC, NGF, 3><<<nfb, NT, 0, st>>>(M.nf, M.tseg, M.ssit, M.norm, a.tab.pico, fs,
                                                  a.face);
    a.timer(1, st);
    elem_lin<NC, NG, NGF><<<neb, NT, 0, st>>>(M, a.tab, P, a.x[0], a.x[1], a.x[2], a.face, so);
    a.timer(2, st);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  // nonlinear
  FaceSpec fa = {};
  fa.in[0] = a.x[0];
  fa.in[1] = a.x[1];
  fa.nout = 4;
  {
    int src[4] = {0, 0, 1, 1}, comp[4] = {0, 1, 0, 1};
    for (int o = 0; o < 4; o++) {
      fa.src[o] = src[o];
      fa.comp[o] = comp[o];
    }
  }
  a.timer(0, st);
  face_kernel<NC, NGF, 2><<<nfb, NT, 0, st>>>(M.nf, M.tseg, M.ssit, M.norm, a.tab.pico, fa,
                                                a.face);
  BCOut bo;
  for (int c = 0; c < 4; c++) bo.F[c] = a.work[c];
  bo.src[0] = a.work[4];
  bo.src[1] = a.work[5];
  bo.uh[0] = a.work[6];
  bo.uh[1] = a.work[7];
  a.timer(1, st);
#define SPOC_BC(V)                                                                        \
  elem_bc<NC, NG, NGF, V><<<neb, NT, 0, st>>>(M, a.tab, P, a.x[0], a.x[1], a.x[2], a.hb, \
                                              a.w0, a.z0, a.fc, a.face, bo)
  if (a.variant == 1) SPOC_BC(1);
  else if (a.variant == 2) SPOC_BC(2);
  else SPOC_BC(3);
#undef SPOC_BC
  FaceSpec f2 = {};
  for (int c = 0; c < 6; c++) {
    f2.in[c] = a.work[c < 4 ? c : c + 2];
    f2.src[c] = c;
    f2.comp[c] = (c == 0 || c == 2 || c == 4) ? 0 : 1;
  }
  f2.nout = 6;
  a.timer(2, st);
  face_kernel<NC, NGF, 6><<<nfb, NT, 0, st>>>(M.nf, M.tseg, M.ssit, M.norm, a.tab.pico, f2,
                                                a.face);
  a.timer(3, st);
#define SPOC_EG(V)                                                                         \
  elem_eg<NC, NG, NGF, V><<<neb, NT, 0, st>>>(M, a.tab, P, a.Fptr, a.work[4], a.work[5], \
                                              a.work[6], a.work[7], a.face, so)
  if (a.variant == 1) SPOC_EG(1);
  else if (a.variant == 2) SPOC_EG(2);
  else SPOC_EG(3);
#undef SPOC_EG
  a.timer(4, st);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace f64

// One translation unit per shape (the build compiles this file once per
// shape with SPOC_NC/SPOC_NG/SPOC_NGF defined) keeps compile times parallel.
#if defined(SPOC_NC)
int rhs_f64_shape(const RhsArgs &a, int nc, int ng, int ngf);
#define SPOC_CAT3(a, b, c, d) a##_##b##_##c##_##d
#define SPOC_NAME(a, b, c, d) SPOC_CAT3(a, b, c, d)
int SPOC_NAME(rhs_f64, SPOC_NC, SPOC_NG, SPOC_NGF)(const RhsArgs &a) {
  return f64::launch_rhs<SPOC_NC, SPOC_NG, SPOC_NGF>(a);
}
#endif

}  // namespace spoc
