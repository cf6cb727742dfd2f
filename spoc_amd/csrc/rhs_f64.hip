// Plain FP64 right-hand side for gfx950.
//
// Every value is produced by the same sequence of IEEE operations as the
// reference Fortran (no FMA contraction: this file is compiled with
// -ffp-contract=off), so the device result equals the reference's bit for
// bit.  Bit-exactness fixes the order of each output's sums, not who
// computes it or where the data lives, so the work is laid out for the GPU:
//
//  - one element per lane: arrays are element-blocked by 64 (f64_types.h),
//    so a wave reads and writes each row of its 64 elements as one
//    contiguous segment, and each lane runs its element's sequential sums
//    from its own registers;
//  - the operator tables are read at wave-uniform addresses, as scalar loads
//    into SGPRs (the table pointers are separate noalias kernel arguments);
//  - each mass-matrix solve is a lane's triangular solve on its element's
//    vector in registers, so no lane waits for another;
//  - an element kernel's waves split the element's outputs (the q
//    components, the three tendencies), so every kernel has several waves
//    per block of 64 elements;
//  - the wall projections (the free-slip and no-slip sequences of the
//    reference), which only wall elements need, run afterwards in a small
//    kernel over the list of wall elements; the element kernels leave those
//    elements' raw vectors in a work array;
//  - face traces are a separate pass (one thread per face and field),
//    because the reference accumulates both sides of a face into one sum
//    (src/calflu.f:41-62).
//
// Kernels per right-hand side:
//   linear    : face (eta, u, v) -> elem_lin -> wall_lin
//   nonlinear : face (u, v) -> elem_q -> wall_q (free slip) -> elem_c ->
//               wall_c -> face (F, uh, vh) -> elem_eg -> wall_eg
#include <hip/hip_runtime.h>

#include "common.h"
#include "f64_types.h"

namespace spoc {
namespace f64 {

constexpr int WB = F64_WB;

// A compiler barrier: no instruction moves across it, and memory counts as
// changed, so table values read from LDS before it are read again after it
// rather than kept in registers.  Placed after each output of a fully
// unrolled loop, it keeps the reads of later outputs from being hoisted and
// earlier reads from being reused (either would run out of registers).
__device__ __forceinline__ void fence() {
  __builtin_amdgcn_sched_barrier(0);
  asm volatile("" ::: "memory");
}

template <int NC, int NG, int NGF>
struct Sh {
  static constexpr int NM = nm_of(NC);
  static constexpr int NL = NM - NC - 1;  // reduced (zero on edge 1) basis
  static constexpr int NE3 = 3 * NGF;
};

__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// ---------------------------------------------------------------------------
// Per-lane building blocks, each the reference routine's arithmetic.
// ---------------------------------------------------------------------------
template <int NM>
__device__ __forceinline__ void vload(double (&v)[NM], const double *p, int e) {
#pragma unroll
  for (int m = 0; m < NM; m++) v[m] = p[f64_at(e, m, NM)];
}

template <int NM>
__device__ __forceinline__ void vstore(double *p, int e, const double (&v)[NM]) {
#pragma unroll
  for (int m = 0; m < NM; m++) p[f64_at(e, m, NM)] = v[m];
}

// ans = sum_j mat(j, i) x(j), j ascending (src/calb2.f:14-18); row i of the
// LDS image holds mat(0..NM-1, i) contiguously, read at a uniform address.
template <int NM>
__device__ __forceinline__ double dot_col(const double *mrow, const double (&x)[NM]) {
  double ans = 0.0;
#pragma unroll
  for (int j = 0; j < NM; j++) ans = ans + mrow[j] * x[j];
  return ans;
}

// The three faces' values of face output o for this lane's element:
// fb[k NGF + j].
template <int NGF>
__device__ __forceinline__ void face_vals(double (&fb)[3 * NGF], const double *__restrict__ face,
                                          int nout, int o, const int (&fk)[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < NGF; j++) fb[k * NGF + j] = face[((size_t)fk[k] * nout + o) * NGF + j];
}

// Edge lift of one face field to mode i (src/mat1.f:184-207): for each local
// edge k the face points are taken forward (its = 1) or reversed (its = -1).
// prow = P + 3 NGF i, the LDS image of pico(., i, k) (P[(3 i + k) NGF + j]);
// the lane picks the point by its own address.
template <int NGF>
__device__ __forceinline__ double tri_f(double b, const double *prow,
                                        const double *__restrict__ wif,
                                        const double (&fb)[3 * NGF], const int (&its)[3],
                                        double coeff) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int isg = its[k];
    const double sig = (double)isg * coeff * 0.5;
    const double *pk = prow + k * NGF;
    double ans = 0.0;
#pragma unroll
    for (int j = 1; j <= NGF; j++) {
      const double var = pk[isg > 0 ? j - 1 : NGF - j] * fb[k * NGF + (j - 1)];
      ans = ans + wif[j - 1] * var;
    }
    b = b + ans * sig;
  }
  return b;
}

// Cholesky solve U^T U x = b with b in registers (src/dtrsm1.f:58-75).
template <int M, int LD>
__device__ __forceinline__ void dtrsm4_reg(double (&b)[M], const double *__restrict__ U,
                                           const double *__restrict__ dia) {
#pragma unroll
  for (int i = 0; i < M; i++) {
    double temp = b[i];
#pragma unroll
    for (int k = 0; k < i; k++) temp = temp - U[k + LD * i] * b[k];
    b[i] = temp * dia[i];
    fence();
  }
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {
    b[k] = b[k] * dia[k];
#pragma unroll
    for (int i = 0; i < k; i++) b[i] = b[i] - b[k] * U[i + LD * k];
    fence();
  }
}

// Two independent solves of dtrsm4_reg, row by row side by side: each keeps
// its own order of operations, and the two dependent chains interleave (the
// wall kernels run one lane per wall element, so their time is the chains'
// latency).
template <int M1, int M2, int LD>
__device__ __forceinline__ void dtrsm4_pair(double (&b1)[M1], const double *__restrict__ U1,
                                            const double *__restrict__ d1, double (&b2)[M2],
                                            const double *__restrict__ U2,
                                            const double *__restrict__ d2) {
  constexpr int M = M1 > M2 ? M1 : M2;
#pragma unroll
  for (int i = 0; i < M; i++) {
    if (i < M1) {
      double temp = b1[i];
#pragma unroll
      for (int k = 0; k < i; k++) temp = temp - U1[k + LD * i] * b1[k];
      b1[i] = temp * d1[i];
    }
    if (i < M2) {
      double temp = b2[i];
#pragma unroll
      for (int k = 0; k < i; k++) temp = temp - U2[k + LD * i] * b2[k];
      b2[i] = temp * d2[i];
    }
    fence();
  }
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {
    if (k < M1) {
      b1[k] = b1[k] * d1[k];
#pragma unroll
      for (int i = 0; i < k; i++) b1[i] = b1[i] - b1[k] * U1[i + LD * k];
    }
    if (k < M2) {
      b2[k] = b2[k] * d2[k];
#pragma unroll
      for (int i = 0; i < k; i++) b2[i] = b2[i] - b2[k] * U2[i + LD * k];
    }
    fence();
  }
}

// Projection onto the zero-on-edge-1 basis (src/mat1.f:418-439).
template <int NC, int NL>
__device__ __forceinline__ void reduc_gn(const double *gn, double (&gnr)[NL]) {
#pragma unroll
  for (int j = 0; j <= NC - 1; j++) {
    const int c1 = ip_of(NC, j), c2 = ip_of(NC - 1, j);
#pragma unroll
    for (int l = 0; l <= NC - 1 - j; l += 2) gnr[c2 + l] = gn[c1 + l + 1] + gn[c1];
#pragma unroll
    for (int l = 1; l <= NC - 1 - j; l += 2) gnr[c2 + l] = gn[c1 + l + 1] - gn[c1];
  }
}

// Expansion back to the full basis (src/mat1.f:445-473).
template <int NC, int NL>
__device__ __forceinline__ void reduc_inv_gn(double *gn, const double (&gnr)[NL]) {
#pragma unroll
  for (int j = 0; j <= NC - 1; j++) {
    const int c1 = ip_of(NC, j), c2 = ip_of(NC - 1, j);
    double s = 0.0;
#pragma unroll
    for (int l = 0; l <= NC - j - 1; l += 2) s = s + gnr[c2 + l];
#pragma unroll
    for (int l = 1; l <= NC - j - 1; l += 2) s = s - gnr[c2 + l];
    gn[c1] = s;
#pragma unroll
    for (int l = 0; l <= NC - j - 1; l++) gn[c1 + l + 1] = gnr[c2 + l];
  }
  gn[nm_of(NC) - 1] = 0.0;
}

// Free-slip projection of a vector field (gx, gy) on a wall element: rotate
// to (n, t), solve the normal part in the reduced basis and the tangential
// part with the full mass matrix, rotate back (src/calvarnl_fs.f:447-459,
// src/calvar_lin.f:184-197).  In place; U, UL: the factors, column-major
// with leading dimension NM.
template <int NC, int NM, int NL>
__device__ __forceinline__ void wall_rot_solve(double (&gx)[NM], double (&gy)[NM], double dnx,
                                               double dny, const double *U, const double *dia1,
                                               const double *UL, const double *dia2) {
  double bu[NL];
#pragma unroll
  for (int j = 0; j < NM; j++) {
    const double gn = gx[j] * dnx + gy[j] * dny, gt = gx[j] * dny - gy[j] * dnx;
    gx[j] = gn;
    gy[j] = gt;
  }
  reduc_gn<NC, NL>(gx, bu);
  dtrsm4_pair<NL, NM, NM>(bu, UL, dia2, gy, U, dia1);
  reduc_inv_gn<NC, NL>(gx, bu);
#pragma unroll
  for (int j = 0; j < NM; j++) {
    const double gn = gx[j], gt = gy[j];
    gx[j] = gn * dnx + gt * dny;
    gy[j] = gn * dny - gt * dnx;
  }
}

// No-slip projection: the field in the reduced basis
// (src/calvarnl_ns.f:253-264, 409-421).
template <int NC, int NM, int NL>
__device__ __forceinline__ void wall_red_solve(double (&g)[NM], const double *UL,
                                               const double *dia2) {
  double bu[NL];
  reduc_gn<NC, NL>(g, bu);
  dtrsm4_reg<NL, NM>(bu, UL, dia2);
  reduc_inv_gn<NC, NL>(g, bu);
}

// Free-slip projection of the velocity-gradient tensor on a wall element
// (src/calvarnl_fs.f:171-204), in place: q = xx, xy, yx, yy hold gnn, gnt,
// gtn, gtt between the two rotations.
template <int NC, int NM, int NL>
__device__ __forceinline__ void wall_tensor_solve(double (&qxx)[NM], double (&qxy)[NM],
                                                  double (&qyx)[NM], double (&qyy)[NM],
                                                  double dnx, double dny, const double *U,
                                                  const double *dia1, const double *UL,
                                                  const double *dia2) {
  double bu[NL];
#pragma unroll
  for (int j = 0; j < NM; j++) {
    const double gnx = qxx[j] * dnx + qyx[j] * dny, gtx = qxx[j] * dny - qyx[j] * dnx;
    const double gny = qxy[j] * dnx + qyy[j] * dny, gty = qxy[j] * dny - qyy[j] * dnx;
    qxx[j] = gnx * dnx + gny * dny;  // gnn
    qxy[j] = gnx * dny - gny * dnx;  // gnt
    qyx[j] = gtx * dnx + gty * dny;  // gtn
    qyy[j] = gtx * dny - gty * dnx;  // gtt
  }
  reduc_gn<NC, NL>(qyx, bu);
  dtrsm4_pair<NM, NM, NM>(qxx, U, dia1, qxy, U, dia1);
  dtrsm4_pair<NL, NM, NM>(bu, UL, dia2, qyy, U, dia1);
  reduc_inv_gn<NC, NL>(qyx, bu);
#pragma unroll
  for (int j = 0; j < NM; j++) {
    const double gnn = qxx[j], gnt = qxy[j], gtn = qyx[j], gtt = qyy[j];
    const double gnx = gnn * dnx + gnt * dny, gny = gnn * dny - gnt * dnx;
    const double gtx = gtn * dnx + gtt * dny, gty = gtn * dny - gtt * dnx;
    qxx[j] = gnx * dnx + gtx * dny;
    qyx[j] = gnx * dny - gtx * dnx;
    qxy[j] = gny * dnx + gty * dny;
    qyy[j] = gny * dny - gty * dnx;
  }
}

// Pointer c of three (a select, not an index: a runtime index into the
// kernel argument would copy the whole argument block to scratch).
template <class T>
__device__ __forceinline__ T pick(int c, T a, T b, T d) {
  return c == 0 ? a : (c == 1 ? b : d);
}

// Where field c's RK4 stage update goes.
struct Out {
  double *k, *u0, *xo, *acc;
};
__device__ __forceinline__ Out out_of(const F64Args &A, int c) {
  return Out{pick(c, A.k[0], A.k[1], A.k[2]), pick(c, A.u0[0], A.u0[1], A.u0[2]),
             pick(c, A.xo[0], A.xo[1], A.xo[2]), pick(c, A.acc[0], A.acc[1], A.acc[2])};
}

// RK4 stage update of one value (src/shwwnl.f:242-284), stage s = 0..3;
// acc accumulates ((k1 + 2k2) + 2k3) + k4 in the reference's order.  Stage
// -1: the tendency itself.
__device__ __forceinline__ void emit(const F64Args &A, const Out &o, size_t g, double k) {
  const int s = A.stage;
  if (s < 0) {
    o.k[g] = k;
  } else if (s == 0) {
    o.acc[g] = k;
    o.xo[g] = o.u0[g] + A.dt * k * 0.5;
  } else if (s == 1) {
    o.acc[g] = o.acc[g] + 2.0 * k;
    o.xo[g] = o.u0[g] + A.dt * k * 0.5;
  } else if (s == 2) {
    o.acc[g] = o.acc[g] + 2.0 * k;
    o.xo[g] = o.u0[g] + A.dt * k;
  } else {
    const double a = o.acc[g] + k;
    o.u0[g] = o.u0[g] + A.dt6 * a;
  }
}

template <int NM>
__device__ __forceinline__ void emit_vec(const F64Args &A, int c, int e, const double (&g)[NM]) {
  const Out o = out_of(A, c);
#pragma unroll
  for (int m = 0; m < NM; m++) emit(A, o, f64_at(e, m, NM), g[m]);
}

// Per-lane element data.
struct Elem {
  int e;
  bool valid, wall;
  double ax, bx, ay, by, detai, nx, ny;
  int fk[3], its[3];
};

__device__ __forceinline__ Elem elem_of(const F64Mesh &M, int e) {
  Elem q;
  q.e = e;
  q.valid = e < M.ne;
  const double *g = M.geo;
  q.ax = g[f64_at(e, G_AX, G_ROWS)];
  q.bx = g[f64_at(e, G_BX, G_ROWS)];
  q.ay = g[f64_at(e, G_AY, G_ROWS)];
  q.by = g[f64_at(e, G_BY, G_ROWS)];
  q.detai = g[f64_at(e, G_DETAI, G_ROWS)];
  q.nx = g[f64_at(e, G_NX, G_ROWS)];
  q.ny = g[f64_at(e, G_NY, G_ROWS)];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    q.fk[k] = M.ib[f64_at(e, I_ITM + k, I_ROWS)];
    q.its[k] = M.ib[f64_at(e, I_ITS + k, I_ROWS)];
  }
  q.wall = M.ib[f64_at(e, I_BDL, I_ROWS)] != 0;
  return q;
}

// Operator images in LDS, staged by each workgroup: agx and agy with row i
// holding column i of the Fortran matrix (the terms of one output in order),
// pico as P[(3 i + k) NGF + j], the mass-matrix factor U (column-major, as the
// solves walk it) and 1 / U_ii.
template <int NM, int NGF>
struct OpsLds {
  static constexpr int AX = 0, AY = NM * NM, P = 2 * NM * NM, UU = P + 3 * NM * NGF,
                       DIA = UU + NM * NM, SIZE = DIA + NM;
};

template <int NM, int NGF>
__device__ __forceinline__ void stage_ops(double *lds, const DevTables &T, int nthreads) {
  using L = OpsLds<NM, NGF>;
  for (int t = threadIdx.x; t < NM * NM; t += nthreads) {
    const int i = t / NM, j = t % NM;
    lds[L::AX + t] = T.agxT[i + NM * j];
    lds[L::AY + t] = T.agyT[i + NM * j];
    lds[L::UU + t] = T.U[t];
  }
  for (int t = threadIdx.x; t < 3 * NM * NGF; t += nthreads) {
    const int j = t % NGF, k = (t / NGF) % 3, i = t / (3 * NGF);
    lds[L::P + t] = T.picoT[i + NM * (j + NGF * k)];
  }
  for (int t = threadIdx.x; t < NM; t += nthreads) lds[L::DIA + t] = T.dia1[t];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Face traces (src/calflu.f:6-71): one thread per (face, field); the field is
// blockIdx.y.  Each field's trace is multiplied by a component of the face
// normal (and for the linear eta equation by hh0), as the loops after every
// cal_flu call do.
// ---------------------------------------------------------------------------
struct FaceSpec {
  const double *in[6];  // blocked modal inputs
  double pre[6];        // the trace is of pre * a where haspre
  int haspre[6];
  int nout;             // outputs per face
  int nof[6];           // outputs of each field (1 or 2)
  int out[6][2];        // their output indices
  int comp[6][2];       // normal component of each (0: n_x, 1: n_y)
  int haspost[6][2];    // times hpost of the face's first element
  const double *hpost;
};

template <int NC, int NGF>
__global__ __launch_bounds__(256) void face_kernel(int nf, const int *__restrict__ tseg,
                                                   const int *__restrict__ ssit,
                                                   const double *__restrict__ norm,
                                                   const double *__restrict__ pico,
                                                   FaceSpec spec, double *__restrict__ out) {
  constexpr int NM = nm_of(NC);
  __shared__ double sp[3 * NM * NGF];
  for (int i = threadIdx.x; i < 3 * NM * NGF; i += 256) sp[i] = pico[i];
  __syncthreads();
  const int f = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (f >= nf) return;
  const int k1 = tseg[2 * f] - 1, k2 = tseg[2 * f + 1] - 1;
  const int s1 = (k2 < 0) ? 0 : ssit[2 * f] - 1;
  const double *in = spec.in[0];
  bool haspre = spec.haspre[0] != 0;
  double pre = spec.pre[0];
  int nof = spec.nof[0], o0 = spec.out[0][0], o1 = spec.out[0][1], c0 = spec.comp[0][0],
      c1 = spec.comp[0][1], h0 = spec.haspost[0][0], h1 = spec.haspost[0][1];
#pragma unroll
  for (int t = 1; t < 6; t++)  // (selects, not an index into the argument)
    if (c == t) {
      in = spec.in[t];
      haspre = spec.haspre[t] != 0;
      pre = spec.pre[t];
      nof = spec.nof[t];
      o0 = spec.out[t][0];
      o1 = spec.out[t][1];
      c0 = spec.comp[t][0];
      c1 = spec.comp[t][1];
      h0 = spec.haspost[t][0];
      h1 = spec.haspost[t][1];
    }
  double acc[NGF];
#pragma unroll
  for (int j = 0; j < NGF; j++) acc[j] = 0.0;
#pragma unroll 4
  for (int m = 0; m < NM; m++) {
    double a = in[f64_at(k1, m, NM)];
    if (haspre) a = pre * a;
    const double *pc = sp + NGF * (m + NM * s1);
#pragma unroll
    for (int j = 0; j < NGF; j++) acc[j] = acc[j] + a * pc[j];
  }
  if (k2 >= 0) {
    const int s2 = ssit[2 * f + 1] - 1;
#pragma unroll 4
    for (int m = 0; m < NM; m++) {
      double a = in[f64_at(k2, m, NM)];
      if (haspre) a = pre * a;
      const double *pc = sp + NGF * (m + NM * s2);
#pragma unroll
      for (int j = 0; j < NGF; j++) acc[j] = acc[j] + a * pc[NGF - 1 - j];
    }
#pragma unroll
    for (int j = 0; j < NGF; j++) acc[j] = 0.5 * acc[j];
  }
  const double nxf = norm[2 * f], nyf = norm[2 * f + 1];
  for (int t = 0; t < nof; t++) {
    double *dst = out + ((size_t)f * spec.nout + (t ? o1 : o0)) * NGF;
    const double nn = (t ? c1 : c0) ? nyf : nxf;
    const bool post = (t ? h1 : h0) != 0;
    const double pv = post ? spec.hpost[k1] : 1.0;
#pragma unroll
    for (int j = 0; j < NGF; j++) {
      double v = acc[j] * nn;
      if (post) v = v * pv;
      dst[j] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Passes over one output vector w (the element's row of a blocked work
// array), one output at a time: the loop over outputs stays rolled, so only
// the pass's input vector and a few accumulators are live.
// ---------------------------------------------------------------------------
// w[i] = (w[i] + dot(agx_i, x) c1) + dot(agy_i, x) c2; w[i] starts from 0.0
// when init (the reference's "b = 0" followed by two cal_b calls).
template <int NM>
__device__ __forceinline__ void vol_pass(double *w, int e, const double (&x)[NM],
                                         const double *AX, const double *AY, double c1,
                                         double c2, bool init) {
#pragma unroll 2
  for (int i = 0; i < NM; i++) {
    const size_t o = f64_at(e, i, NM);
    double r = init ? 0.0 : w[o];
    r = r + dot_col<NM>(AX + NM * i, x) * c1;
    r = r + dot_col<NM>(AY + NM * i, x) * c2;
    w[o] = r;
  }
}

// w[i] = tri_f(w[i], ...) for the face values fb.
template <int NM, int NGF>
__device__ __forceinline__ void lift_pass(double *w, int e, const double (&fb)[3 * NGF],
                                          const int (&its)[3], const double *P,
                                          const double *__restrict__ wif, double coeff) {
#pragma unroll 2
  for (int i = 0; i < NM; i++) {
    const size_t o = f64_at(e, i, NM);
    w[o] = tri_f<NGF>(w[o], P + 3 * NGF * i, wif, fb, its, coeff);
  }
}

// Solve the mass matrix for the output vector w and send it on: the RK4
// stage update (emit) or, with store, back into w.
template <int NM>
__device__ __forceinline__ void solve_out(const F64Args &A, int c, int e, double *w,
                                          const double *U, const double *dia, bool store) {
  double g[NM];
  vload(g, w, e);
  dtrsm4_reg<NM, NM>(g, U, dia);
  if (store)
    vstore(w, e, g);
  else
    emit_vec(A, c, e, g);
}

// ---------------------------------------------------------------------------
// Linear right-hand side (src/calvar_lin.f:115-282).  Wave 0: u tendency,
// wave 1: v, wave 2: eta, each built in G[c].  Face outputs per face: fbdxx,
// fbdyy (of fxx = -g0 h), fbdx, fbdy (of hh0 u, hh0 v); fxy = fyx = 0, whose
// terms add only zeros in the reference.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF>
__global__ __launch_bounds__(192) void elem_lin(F64Args A, const double *__restrict__ wif) {
  constexpr int NM = Sh<NC, NG, NGF>::NM, NE3 = Sh<NC, NG, NGF>::NE3;
  using L = OpsLds<NM, NGF>;
  __shared__ double lds[L::SIZE];
  stage_ops<NM, NGF>(lds, A.tab, 192);
  const double *AX = lds + L::AX, *AY = lds + L::AY, *P = lds + L::P;
  const int c = wave_id(), e = blockIdx.x * WB + (threadIdx.x & (WB - 1));
  const Elem q = elem_of(A.mesh, e);
  const double ax = q.ax, bx = q.bx, ay = q.ay, by = q.by, detai = q.detai;
  double *w = pick(c, A.G[0], A.G[1], A.G[2]);
  if (c < 2) {
    double fx[NM];
    vload(fx, A.x[2], e);
#pragma unroll
    for (int m = 0; m < NM; m++) fx[m] = -A.g0 * fx[m];  // fxx (src/calvar_lin.f:74-81)
    // gx: cal_b(agx, -by detai), cal_b(agy, ay detai); gy: cal_b(agx,
    // -(-bx detai)), cal_b(agy, -(ax detai))
    const double c1 = c == 0 ? -by * detai : -(-bx * detai);
    const double c2 = c == 0 ? ay * detai : -(ax * detai);
    vol_pass(w, e, fx, AX, AY, c1, c2, true);
    double fb[NE3];
    face_vals<NGF>(fb, A.face, 4, c, q.fk);
    lift_pass<NM, NGF>(w, e, fb, q.its, P, wif, detai);
  } else {
    const double H = A.hh0e[q.valid ? e : 0];
    double ak[NM];
    vload(ak, A.x[0], e);
#pragma unroll
    for (int m = 0; m < NM; m++) ak[m] = ak[m] * H;
    vol_pass(w, e, ak, AX, AY, by * detai, -ay * detai, true);
    vload(ak, A.x[1], e);
#pragma unroll
    for (int m = 0; m < NM; m++) ak[m] = ak[m] * H;
    vol_pass(w, e, ak, AX, AY, -bx * detai, ax * detai, false);
    double fb[NE3];
    face_vals<NGF>(fb, A.face, 4, 2, q.fk);
    lift_pass<NM, NGF>(w, e, fb, q.its, P, wif, -detai);
    face_vals<NGF>(fb, A.face, 4, 3, q.fk);
    lift_pass<NM, NGF>(w, e, fb, q.its, P, wif, -detai);
  }
  // The stage update below writes the next stage input in place of this
  // stage's, which the block's other waves read above: wait for them.
  __syncthreads();
  // the velocity pair of a wall element stays in G for wall_vel
  if (!q.valid || (c < 2 && q.wall)) return;
  solve_out<NM>(A, c, e, w, lds + L::UU, lds + L::DIA, false);
}

// ---------------------------------------------------------------------------
// Nonlinear stage B (src/calvarnl_fs.f:115-204 and the ns/inv variants): the
// velocity gradient q.  Wave r builds q component r (0 xx, 1 xy of u; 2 yx,
// 3 yy of v) in Q[r] and, except for free-slip wall elements (wall_q),
// solves it.  Face outputs: fbdxx, fbdxy, fbdyx, fbdyy of the velocity
// traces.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(256) void elem_q(F64Args A, const double *__restrict__ wif) {
  constexpr int NM = Sh<NC, NG, NGF>::NM, NE3 = Sh<NC, NG, NGF>::NE3;
  using L = OpsLds<NM, NGF>;
  __shared__ double lds[L::SIZE];
  stage_ops<NM, NGF>(lds, A.tab, 256);
  const double *AX = lds + L::AX, *AY = lds + L::AY, *P = lds + L::P;
  const int r = wave_id(), e = blockIdx.x * WB + (threadIdx.x & (WB - 1));
  const Elem q = elem_of(A.mesh, e);
  const double ax = q.ax, bx = q.bx, ay = q.ay, by = q.by, detai = q.detai;
  double *w = r < 2 ? (r == 0 ? A.Q[0] : A.Q[1]) : (r == 2 ? A.Q[2] : A.Q[3]);
  {
    double a[NM];
    vload(a, r < 2 ? A.x[0] : A.x[1], e);
#pragma unroll 2
    for (int i = 0; i < NM; i++) {
      double gx1 = 0.0, gy1 = 0.0;
      gx1 = gx1 + dot_col<NM>(AX + NM * i, a) * (-1.0);
      gy1 = gy1 + dot_col<NM>(AY + NM * i, a) * (-1.0);
      w[f64_at(e, i, NM)] =
          (r & 1) ? (-bx * gx1 + ax * gy1) * detai : (by * gx1 - ay * gy1) * detai;
    }
  }
  double fb[NE3];
  face_vals<NGF>(fb, A.face, 4, r, q.fk);
  lift_pass<NM, NGF>(w, e, fb, q.its, P, wif, detai);
  if (!q.valid || (VAR == 1 && q.wall)) return;
  solve_out<NM>(A, 0, e, w, lds + L::UU, lds + L::DIA, true);
}

// ---------------------------------------------------------------------------
// Nonlinear stage C (src/calvarnl_fs.f:245-332; src/nonlin.f:35-133): F =
// mu0 q - g0 h (src/calvarnl_fs.f:341-348); the nine
// fields' values at the quadrature points (spectoxy); their products and
// quotients, times the weights; the projections onto the modes (xytospec);
// the sources -adu + w1, -adv + z1 and the mass fluxes uh, vh, which are
// solved here except on wall elements (wall_c).  The block's four waves
// split each phase; LDS: pi2 with row q holding pi2(q, 0..NM-1), U, 1/U_ii.
// ---------------------------------------------------------------------------
template <int NM, int NG>
struct QuadLds {
  static constexpr int PQ = 0, UU = NG * NM, DIA = UU + NM * NM, SIZE = DIA + NM;
};

template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(256) void elem_c(F64Args A, const double *__restrict__ wi) {
  constexpr int NM = Sh<NC, NG, NGF>::NM;
  // modes per pass of the source projections: each pass re-reads the weighted
  // products, so fewer passes where the registers allow (14 at p >= 6: elem_c
  // -3% at p = 6; at p = 4 and 5 it would halve the resident waves)
  constexpr int MC = NM >= 28 ? 14 : 7;
  using L = QuadLds<NM, NG>;
  __shared__ double lds[L::SIZE];
  for (int t = threadIdx.x; t < NG * NM; t += 256) lds[L::PQ + t] = A.tab.pi2T[t];
  for (int t = threadIdx.x; t < NM * NM; t += 256) lds[L::UU + t] = A.tab.U[t];
  for (int t = threadIdx.x; t < NM; t += 256) lds[L::DIA + t] = A.tab.dia1[t];
  __syncthreads();
  const double *PQ = lds + L::PQ;  // row q: pi2(q, 0..NM-1)
  const int w = wave_id(), e = blockIdx.x * WB + (threadIdx.x & (WB - 1));
  const Elem q = elem_of(A.mesh, e);
  const double mu0 = A.mu0, g0 = A.g0, mb0 = A.mb0;
  // F = mu0 q - g0 h (elementwise; wave w the component w)
  {
    const double *Qw = w < 2 ? (w == 0 ? A.Q[0] : A.Q[1]) : (w == 2 ? A.Q[2] : A.Q[3]);
    double *Fw = w < 2 ? (w == 0 ? A.F[0] : A.F[1]) : (w == 2 ? A.F[2] : A.F[3]);
    const bool gh = w == 0 || w == 3;
#pragma unroll 4
    for (int m = 0; m < NM; m++) {
      const size_t i = f64_at(e, m, NM);
      Fw[i] = gh ? mu0 * Qw[i] - g0 * A.x[2][i] : mu0 * Qw[i];
    }
  }
  // the nine fields at the quadrature points (wave w: fields w, w+4, w+8):
  // u, v, gxx, gxy = qxy - fc, gyx = qyx + fc, gyy, hk = h + hb,
  // wk = w0 - mb0 u, zk = z0 - mb0 v
  for (int c = w; c < 9; c += 4) {
    // field c = X op Y: op 0 none, 1 minus, 2 plus, 3 minus mb0 times
    const double *X, *Y = nullptr;
    int op = 0;
    switch (c) {
      case 0: X = A.x[0]; break;
      case 1: X = A.x[1]; break;
      case 2: X = A.Q[0]; break;
      case 3: X = A.Q[1]; Y = A.fc; op = 1; break;
      case 4: X = A.Q[2]; Y = A.fc; op = 2; break;
      case 5: X = A.Q[3]; break;
      case 6: X = A.x[2]; Y = A.hb; op = 2; break;
      case 7: X = A.w0; Y = A.x[0]; op = 3; break;
      default: X = A.z0; Y = A.x[1]; op = 3; break;
    }
    double v[NM];
    vload(v, X, e);
    if (op) {
      double y[NM];
      vload(y, Y, e);
      if (op == 1) {
#pragma unroll
        for (int m = 0; m < NM; m++) v[m] = v[m] - y[m];
      } else if (op == 2) {
#pragma unroll
        for (int m = 0; m < NM; m++) v[m] = v[m] + y[m];
      } else {
#pragma unroll
        for (int m = 0; m < NM; m++) v[m] = v[m] - mb0 * y[m];
      }
    }
#pragma unroll 2
    for (int qq = 0; qq < NG; qq++) {
      const double *row = PQ + NM * qq;
      double mf = 0.0;
#pragma unroll
      for (int m = 0; m < NM; m++) mf = mf + v[m] * row[m];
      A.MF[f64_at(e, c * NG + qq, 9 * NG)] = mf;
    }
  }
  __syncthreads();  // (global memory written by the block's waves: visible after)
  // products at the points, times the weights (wave w: points w, w+4, ...)
  for (int qq = w; qq < NG; qq += 4) {
    double mf[9];
#pragma unroll
    for (int c = 0; c < 9; c++) mf[c] = A.MF[f64_at(e, c * NG + qq, 9 * NG)];
    double o[8];
    o[0] = mf[0] * mf[2];  // uux
    o[1] = mf[1] * mf[3];  // vuy
    o[2] = mf[0] * mf[4];  // uvx
    o[3] = mf[1] * mf[5];  // vvy
    o[4] = mf[0] * mf[6];  // uhk
    o[5] = mf[1] * mf[6];  // vhk
    o[6] = mf[7] / mf[6];  // wk
    o[7] = mf[8] / mf[6];  // zk
    const double wq = wi[qq];
#pragma unroll
    for (int c = 0; c < 8; c++) A.WS[f64_at(e, c * NG + qq, 8 * NG)] = wq * o[c];
  }
  __syncthreads();
  // projections onto the modes: wave 0 the u source -adu + w1 (products 0,
  // 1, 6), wave 1 the v source -adv + z1 (2, 3, 7), waves 2 and 3 the mass
  // fluxes uh (4) and vh (5), solved here except on wall elements (wall_c)
  if (w < 2) {
    const int p0 = w == 0 ? 0 : 2, p1 = p0 + 1, p2 = w == 0 ? 6 : 7;
    double *Sw = w == 0 ? A.S[0] : A.S[1];
    for (int m0 = 0; m0 < NM; m0 += MC) {
      double r[3][MC];
#pragma unroll
      for (int t = 0; t < 3; t++)
#pragma unroll
        for (int mm = 0; mm < MC; mm++) r[t][mm] = 0.0;
#pragma unroll 3
      for (int qq = 0; qq < NG; qq++) {
        const double ws[3] = {A.WS[f64_at(e, p0 * NG + qq, 8 * NG)],
                              A.WS[f64_at(e, p1 * NG + qq, 8 * NG)],
                              A.WS[f64_at(e, p2 * NG + qq, 8 * NG)]};
        const double *row = PQ + NM * qq;
#pragma unroll
        for (int mm = 0; mm < MC; mm++) {
          const double p = row[m0 + mm < NM ? m0 + mm : NM - 1];
#pragma unroll
          for (int t = 0; t < 3; t++) r[t][mm] = r[t][mm] + ws[t] * p;
        }
      }
#pragma unroll
      for (int mm = 0; mm < MC; mm++)
        if (m0 + mm < NM) {
          const double ad = r[0][mm] + r[1][mm];
          Sw[f64_at(e, m0 + mm, NM)] = -ad + r[2][mm];
        }
    }
  } else {
    const int pc = w == 2 ? 4 : 5;
    double *Uw = w == 2 ? A.UH[0] : A.UH[1];
    double g[NM];
#pragma unroll
    for (int m = 0; m < NM; m++) g[m] = 0.0;
#pragma unroll 2
    for (int qq = 0; qq < NG; qq++) {
      const double ws = A.WS[f64_at(e, pc * NG + qq, 8 * NG)];
      const double *row = PQ + NM * qq;
#pragma unroll
      for (int m = 0; m < NM; m++) g[m] = g[m] + ws * row[m];
    }
    if (!q.wall) dtrsm4_reg<NM, NM>(g, lds + L::UU, lds + L::DIA);  // (else wall_c)
    vstore(Uw, e, g);
  }
}

// ---------------------------------------------------------------------------
// Nonlinear stages E + G (src/calvarnl_fs.f:382-544): waves 0 and 1 build the
// u and v tendencies from the stress tensor F and the sources, wave 2 the eta
// tendency from the mass fluxes, in G[c]; then the solve and the RK4 stage
// update, except for the velocity pair of wall elements (wall_vel).  Face
// outputs per face: fbdxx, fbdxy, fbdyx, fbdyy (of F) and fbdx, fbdy (of uh,
// vh).
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(192) void elem_eg(F64Args A, const double *__restrict__ wif) {
  constexpr int NM = Sh<NC, NG, NGF>::NM, NE3 = Sh<NC, NG, NGF>::NE3;
  using L = OpsLds<NM, NGF>;
  __shared__ double lds[L::SIZE];
  stage_ops<NM, NGF>(lds, A.tab, 192);
  const double *AX = lds + L::AX, *AY = lds + L::AY, *P = lds + L::P;
  const int c = wave_id(), e = blockIdx.x * WB + (threadIdx.x & (WB - 1));
  const Elem q = elem_of(A.mesh, e);
  const double ax = q.ax, bx = q.bx, ay = q.ay, by = q.by, detai = q.detai;
  double *w = pick(c, A.G[0], A.G[1], A.G[2]);
  // the volume passes: u, v (gx = -adu + w1, then cal_b x4,
  // src/calvarnl_fs.f:393-429) or eta (b = 0, cal_b x4,
  // src/calvarnl_fs.f:509-536)
  const double *A1 = pick<const double *>(c, A.F[0], A.F[2], A.UH[0]);
  const double *A2 = pick<const double *>(c, A.F[1], A.F[3], A.UH[1]);
  const double c1 = c < 2 ? -by * detai : by * detai, c2 = c < 2 ? ay * detai : -ay * detai;
  const double c3 = c < 2 ? bx * detai : -bx * detai, c4 = c < 2 ? -ax * detai : ax * detai;
  if (c < 2) {
    const double *src = c == 0 ? A.S[0] : A.S[1];
#pragma unroll 4
    for (int i = 0; i < NM; i++) w[f64_at(e, i, NM)] = src[f64_at(e, i, NM)];
  }
  {
    double a[NM];
    vload(a, A1, e);
    vol_pass(w, e, a, AX, AY, c1, c2, c == 2);
    vload(a, A2, e);
    vol_pass(w, e, a, AX, AY, c3, c4, false);
  }
  // the face lifts: F's faces with detai, or the mass fluxes' with -detai
  const double cf = c < 2 ? detai : -detai;
  double fb[NE3];
  face_vals<NGF>(fb, A.face, 6, c < 2 ? 2 * c : 4, q.fk);
  lift_pass<NM, NGF>(w, e, fb, q.its, P, wif, cf);
  face_vals<NGF>(fb, A.face, 6, c < 2 ? 2 * c + 1 : 5, q.fk);
  lift_pass<NM, NGF>(w, e, fb, q.its, P, wif, cf);
  if (!q.valid || (c < 2 && q.wall)) return;  // (wall_vel)
  solve_out<NM>(A, c, e, w, lds + L::UU, lds + L::DIA, false);
}

// ---------------------------------------------------------------------------
// Wall elements, one lane each, vectors in registers: the reference's
// projection sequences on the raw vectors the element kernels left.  The
// factors are staged in LDS (U, 1/U_ii, UL, 1/UL_ii).
// ---------------------------------------------------------------------------
template <int NM>
__device__ __forceinline__ const double *stage_factors(double *lds, const DevTables &T) {
  for (int t = threadIdx.x; t < NM * NM; t += 64) {
    lds[t] = T.U[t];
    lds[NM * NM + t] = T.UL[t];
  }
  for (int t = threadIdx.x; t < NM; t += 64) {
    lds[2 * NM * NM + t] = T.dia1[t];
    lds[2 * NM * NM + NM + t] = T.dia2[t];
  }
  __syncthreads();
  return lds;
}

// the velocity pair of the linear variant and of E+G: free slip (lin, fs,
// inv) or no slip (ns), then the RK4 stage update
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(64) void wall_vel(F64Args A) {
  constexpr int NM = Sh<NC, NG, NGF>::NM, NL = Sh<NC, NG, NGF>::NL;
  __shared__ double lds[2 * NM * NM + 2 * NM];
  const double *U = stage_factors<NM>(lds, A.tab), *UL = U + NM * NM;
  const double *dia1 = U + 2 * NM * NM, *dia2 = dia1 + NM;
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= A.mesh.nwall) return;
  const int e = A.mesh.wall[t];
  if (VAR == 2) {  // each component alone (both at once: 3x slower at p = 7, registers)
#pragma unroll 1
    for (int c = 0; c < 2; c++) {
      double g[NM];
      vload(g, c == 0 ? A.G[0] : A.G[1], e);
      wall_red_solve<NC, NM, NL>(g, UL, dia2);
      emit_vec(A, c, e, g);
    }
  } else {
    double gx[NM], gy[NM];
    vload(gx, A.G[0], e);
    vload(gy, A.G[1], e);
    const double nx = A.mesh.geo[f64_at(e, G_NX, G_ROWS)], ny = A.mesh.geo[f64_at(e, G_NY, G_ROWS)];
    wall_rot_solve<NC, NM, NL>(gx, gy, nx, ny, U, dia1, UL, dia2);
    emit_vec(A, 0, e, gx);
    emit_vec(A, 1, e, gy);
  }
}

// the velocity-gradient tensor of free-slip wall elements (stage B)
template <int NC, int NG, int NGF>
__global__ __launch_bounds__(64) void wall_q(F64Args A) {
  constexpr int NM = Sh<NC, NG, NGF>::NM, NL = Sh<NC, NG, NGF>::NL;
  __shared__ double lds[2 * NM * NM + 2 * NM];
  const double *U = stage_factors<NM>(lds, A.tab), *UL = U + NM * NM;
  const double *dia1 = U + 2 * NM * NM, *dia2 = dia1 + NM;
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= A.mesh.nwall) return;
  const int e = A.mesh.wall[t];
  double q0[NM], q1[NM], q2[NM], q3[NM];
  vload(q0, A.Q[0], e);
  vload(q1, A.Q[1], e);
  vload(q2, A.Q[2], e);
  vload(q3, A.Q[3], e);
  const double nx = A.mesh.geo[f64_at(e, G_NX, G_ROWS)], ny = A.mesh.geo[f64_at(e, G_NY, G_ROWS)];
  wall_tensor_solve<NC, NM, NL>(q0, q1, q2, q3, nx, ny, U, dia1, UL, dia2);
  vstore(A.Q[0], e, q0);
  vstore(A.Q[1], e, q1);
  vstore(A.Q[2], e, q2);
  vstore(A.Q[3], e, q3);
}

// the mass fluxes of wall elements (stage C)
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(64) void wall_c(F64Args A) {
  constexpr int NM = Sh<NC, NG, NGF>::NM, NL = Sh<NC, NG, NGF>::NL;
  __shared__ double lds[2 * NM * NM + 2 * NM];
  const double *U = stage_factors<NM>(lds, A.tab), *UL = U + NM * NM;
  const double *dia1 = U + 2 * NM * NM, *dia2 = dia1 + NM;
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= A.mesh.nwall) return;
  const int e = A.mesh.wall[t];
  double u[NM], v[NM];
  vload(u, A.UH[0], e);
  vload(v, A.UH[1], e);
  if (VAR == 2) {
    wall_red_solve<NC, NM, NL>(u, UL, dia2);
    wall_red_solve<NC, NM, NL>(v, UL, dia2);
  } else {
    const double nx = A.mesh.geo[f64_at(e, G_NX, G_ROWS)], ny = A.mesh.geo[f64_at(e, G_NY, G_ROWS)];
    wall_rot_solve<NC, NM, NL>(u, v, nx, ny, U, dia1, UL, dia2);
  }
  vstore(A.UH[0], e, u);
  vstore(A.UH[1], e, v);
}

// ---------------------------------------------------------------------------
// Layout changes between the packed (NM, ne) arrays of the C-ABI and the
// blocked arrays of this path (padding elements read as zero).
// ---------------------------------------------------------------------------
// (blockIdx.y: the field, at in + y * sin, out + y * sout)
template <int NM>
__global__ __launch_bounds__(256) void to_blocked(int ne, int nblk, const double *__restrict__ in,
                                                  double *__restrict__ out, size_t sin,
                                                  size_t sout) {
  in += blockIdx.y * sin;
  out += blockIdx.y * sout;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)nblk * NM * WB) return;
  const int c = (int)(idx % WB), r = (int)((idx / WB) % NM), b = (int)(idx / ((size_t)WB * NM));
  const int e = b * WB + c;
  out[idx] = e < ne ? in[(size_t)e * NM + r] : 0.0;
}

template <int NM>
__global__ __launch_bounds__(256) void from_blocked(int ne, const double *__restrict__ in,
                                                    double *__restrict__ out, size_t sin,
                                                    size_t sout) {
  in += blockIdx.y * sin;
  out += blockIdx.y * sout;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)ne * NM) return;
  const int e = (int)(idx / NM), r = (int)(idx % NM);
  out[idx] = in[f64_at(e, r, NM)];
}

// ---------------------------------------------------------------------------
// Host launcher.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF>
int launch(const F64Args &a, int phase, const double *in, double *out) {
  constexpr int NM = nm_of(NC);
  const F64Mesh &M = a.mesh;
  hipStream_t st = a.stream;
  const unsigned nf = a.nf > 1 ? (unsigned)a.nf : 1u;
  if (phase == 0) {
    const size_t n = (size_t)M.nblk * NM * WB;
    to_blocked<NM><<<dim3((unsigned)((n + 255) / 256), nf), 256, 0, st>>>(M.ne, M.nblk, in, out,
                                                                          a.sin, a.sout);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  if (phase == 1) {
    const size_t n = (size_t)M.ne * NM;
    from_blocked<NM><<<dim3((unsigned)((n + 255) / 256), nf), 256, 0, st>>>(M.ne, in, out, a.sin,
                                                                            a.sout);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  const DevTables &T = a.tab;
  const dim3 gf2((M.nf + 255) / 256, 2), gf3((M.nf + 255) / 256, 3), gf6((M.nf + 255) / 256, 6);
  const unsigned nb = (unsigned)M.nblk, nw = (unsigned)((M.nwall + 63) / 64);
  if (a.variant == 0) {
    FaceSpec fs = {};
    fs.in[0] = a.x[2];
    fs.pre[0] = -a.g0;
    fs.haspre[0] = 1;
    fs.in[1] = a.x[0];
    fs.in[2] = a.x[1];
    fs.nout = 4;
    fs.nof[0] = 2;
    fs.out[0][0] = 0;
    fs.comp[0][0] = 0;
    fs.out[0][1] = 1;
    fs.comp[0][1] = 1;
    fs.nof[1] = 1;
    fs.out[1][0] = 2;
    fs.comp[1][0] = 0;
    fs.haspost[1][0] = 1;
    fs.nof[2] = 1;
    fs.out[2][0] = 3;
    fs.comp[2][0] = 1;
    fs.haspost[2][0] = 1;
    fs.hpost = a.hh0e;
    a.timer(0);
    face_kernel<NC, NGF><<<gf3, 256, 0, st>>>(M.nf, M.tseg, M.ssit, M.norm, T.pico, fs, a.face);
    a.timer(1);
    elem_lin<NC, NG, NGF><<<nb, 192, 0, st>>>(a, T.wif);
    a.timer(2);
    if (nw) wall_vel<NC, NG, NGF, 0><<<nw, 64, 0, st>>>(a);
    a.timer(3);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  FaceSpec fa = {};
  fa.in[0] = a.x[0];
  fa.in[1] = a.x[1];
  fa.nout = 4;
  for (int c = 0; c < 2; c++) {
    fa.nof[c] = 2;
    fa.out[c][0] = 2 * c;
    fa.comp[c][0] = 0;
    fa.out[c][1] = 2 * c + 1;
    fa.comp[c][1] = 1;
  }
  a.timer(0);
  face_kernel<NC, NGF><<<gf2, 256, 0, st>>>(M.nf, M.tseg, M.ssit, M.norm, T.pico, fa, a.face);
  a.timer(1);
#define SPOC_V(V, STMT) \
  if (a.variant == V) { STMT; }
#define SPOC_Q(V) elem_q<NC, NG, NGF, V><<<nb, 256, 0, st>>>(a, T.wif)
  SPOC_V(1, SPOC_Q(1)) SPOC_V(2, SPOC_Q(2)) SPOC_V(3, SPOC_Q(3))
  a.timer(2);
  if (a.variant == 1 && nw) wall_q<NC, NG, NGF><<<nw, 64, 0, st>>>(a);
  a.timer(3);
#define SPOC_C(V) elem_c<NC, NG, NGF, V><<<nb, 256, 0, st>>>(a, T.wi)
  SPOC_V(1, SPOC_C(1)) SPOC_V(2, SPOC_C(2)) SPOC_V(3, SPOC_C(3))
  a.timer(4);
  if (nw) {
    SPOC_V(1, (wall_c<NC, NG, NGF, 1><<<nw, 64, 0, st>>>(a)))
    SPOC_V(2, (wall_c<NC, NG, NGF, 2><<<nw, 64, 0, st>>>(a)))
    SPOC_V(3, (wall_c<NC, NG, NGF, 3><<<nw, 64, 0, st>>>(a)))
  }
  FaceSpec f2 = {};
  const double *fin[6] = {a.F[0], a.F[1], a.F[2], a.F[3], a.UH[0], a.UH[1]};
  for (int c = 0; c < 6; c++) {
    f2.in[c] = fin[c];
    f2.nof[c] = 1;
    f2.out[c][0] = c;
    f2.comp[c][0] = c & 1;
  }
  f2.nout = 6;
  a.timer(5);
  face_kernel<NC, NGF><<<gf6, 256, 0, st>>>(M.nf, M.tseg, M.ssit, M.norm, T.pico, f2, a.face);
  a.timer(6);
#define SPOC_EG(V) elem_eg<NC, NG, NGF, V><<<nb, 192, 0, st>>>(a, T.wif)
  SPOC_V(1, SPOC_EG(1)) SPOC_V(2, SPOC_EG(2)) SPOC_V(3, SPOC_EG(3))
  a.timer(7);
  if (nw) {
    SPOC_V(1, (wall_vel<NC, NG, NGF, 1><<<nw, 64, 0, st>>>(a)))
    SPOC_V(2, (wall_vel<NC, NG, NGF, 2><<<nw, 64, 0, st>>>(a)))
    SPOC_V(3, (wall_vel<NC, NG, NGF, 3><<<nw, 64, 0, st>>>(a)))
  }
  a.timer(8);
#undef SPOC_V
#undef SPOC_Q
#undef SPOC_C
#undef SPOC_EG
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace f64

// One translation unit per shape (the build compiles this file once per
// shape with SPOC_NC/SPOC_NG/SPOC_NGF defined) keeps compile times parallel.
#if defined(SPOC_NC)
#define SPOC_CAT3(a, b, c, d) a##_##b##_##c##_##d
#define SPOC_NAME(a, b, c, d) SPOC_CAT3(a, b, c, d)
int SPOC_NAME(rhs_f64, SPOC_NC, SPOC_NG, SPOC_NGF)(const F64Args &a, int phase, const double *in,
                                                   double *out) {
  return f64::launch<SPOC_NC, SPOC_NG, SPOC_NGF>(a, phase, in, out);
}
#endif

}  // namespace spoc
