// Discrete adjoint of the RK4 step of the linear right-hand side
// (src/calvar_lin.f driven by src/shwwnl.f:233-284), FP64 on MFMA, in the
// mass-orthonormal coordinates of the mixed-precision path.
//
// The forward right-hand side of one element is
//   G = P_wall( Dvol(x) + LF * F(avg(T x)) )
// with T the edge traces, avg the face average with the neighbour's
// reversed traces (own trace on a wall edge) and F the pointwise face
// factors.  Its transpose is evaluated in two passes, because the transpose
// of the face average moves data between neighbours:
//   adj_lift:   lam' = P_wall lam;  volume part Dvol^T lam' -> avol;
//               A = F^T (LF^T lam') at the element's own edge points -> abar
//   adj_gather: avg^T A, which is again the face average (own + neighbour
//               reversed, own alone on a wall edge), then T^T, plus avol.
// P_wall (I + n n^T (G - I), or G - I on both components for no slip) is
// symmetric because G is an orthogonal projector, so it is its own
// transpose.  The second pass ends in the adjoint RK4 stage (below).
#include <hip/hip_runtime.h>

#include "mfma.h"

namespace spoc {
namespace f32 {

// Adjoint of one RK4 step x1 = x + dt/6 (k1 + 2 k2 + 2 k3 + k4),
// k_i = R(x_i), x2 = x + dt/2 k1, x3 = x + dt/2 k2, x4 = x + dt k3, run as
// four adjoint stages s = 0..3 (= k4, k3, k2, k1).  Stage s gets w (the
// adjoint of k_{4-s}), computes y = R^T w and updates, with lam the adjoint
// of x1 and xbar its running sum:
//   s=0: xbar = lam + y;  w = dt/3 lam + dt y
//   s=1: xbar += y;       w = dt/3 lam + dt/2 y
//   s=2: xbar += y;       w = dt/6 lam + dt/2 y
//   s=3: lam = xbar + y;  w = dt/6 lam   (the next, earlier step's first input)
__device__ __forceinline__ void adj_rk_value(const F32Args &A, int f, size_t idx, double y) {
  double *lam = A.st.lam[f] + idx, *xb = A.st.xbar[f] + idx, *w = A.st.aw[f] + idx;
  const double dt = A.dt;
  switch (A.stage) {
    case 0:
      *xb = *lam + y;
      *w = (dt / 3.0) * *lam + dt * y;
      break;
    case 1:
      *xb = *xb + y;
      *w = (dt / 3.0) * *lam + (dt * 0.5) * y;
      break;
    case 2:
      *xb = *xb + y;
      *w = (dt / 6.0) * *lam + (dt * 0.5) * y;
      break;
    default: {
      double l = *xb + y;
      *lam = l;
      *w = (dt / 6.0) * l;
    }
  }
}

// Pass 1 of R^T (VAR 0: src/calvar_lin.f; the nonlinear variants' linear
// core takes the same path with the flux projection).
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void adj_lift_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, RBM = Sh::RBM, RBE = Sh::RBE;
  constexpr Pack64T P = pack64t(NM, NE3);
  constexpr int N16 = P.trt / 2;  // Dx^T, Dy^T, LF^T in 16-byte words
  __shared__ floatx4 lds[N16];
  const int b0 = first_block();
  if (b0 >= nblk) return;
  stage_lds(lds, reinterpret_cast<const floatx4 *>(A.ops.pack64t), N16);
  __syncthreads();
  const int eb = b0 + (threadIdx.x >> 6);
  if (eb >= nblk) return;
  const double *O = reinterpret_cast<const double *>(lds);
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.0;
  const double g0 = A.fld.g0, hh0 = valid ? A.fld.hh0e[e] : 0.0;
  const size_t mb = (size_t)e0 * NM;
  const bool wall = valid && A.mesh.bdl[e] != 0;
  const bool anyw = __ballot(wall) != 0;
  double lu[KM], lv[KM], lh[KM];
  bload<NM>(A.st.aw[0] + mb, g, col, valid, lu);
  bload<NM>(A.st.aw[1] + mb, g, col, valid, lv);
  bload<NM>(A.st.aw[2] + mb, g, col, valid, lh);
  if (anyw) wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gg[14], gg[15], lu, lv);
  // volume: eta from the u, v rows; (u, v) from the eta row
  {
    double p0[KM], p1[KM], q[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) {
      p0[s] = di * (-by * lu[s] + bx * lv[s]);
      p1[s] = di * (ay * lu[s] - ax * lv[s]);
      q[s] = di * lh[s];
    }
    double bu[KM], bv[KM];
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double adx[KM], ady[KM];
      afrag64<NM>(O + P.dxt, rb, adx);
      afrag64<NM>(O + P.dyt, rb, ady);
      doublex4 H = chain64(ady, p1, chain64(adx, p0, zero64()));
      doublex4 X0 = chain64(adx, q, zero64()), X1 = chain64(ady, q, zero64());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int s = 4 * rb + r, m = drow(rb, r, g);
        if (valid && m < NM) A.st.avol[2][mb + m * 16 + col] = -g0 * H[r];
        if (s < KM) {
          bu[s] = by * X0[r] - ay * X1[r];
          bv[s] = -bx * X0[r] + ax * X1[r];
        }
      }
    }
    if (VAR != 0 && anyw) wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gg[14], gg[15], bu, bv);
#pragma unroll
    for (int s = 0; s < KM; s++) {
      const int m = 4 * s + g;
      if (valid && m < NM) {
        A.st.avol[0][mb + m * 16 + col] = hh0 * bu[s];
        A.st.avol[1][mb + m * 16 + col] = hh0 * bv[s];
      }
    }
  }
  // edge parts: transposed lift, then the transposed face factors
#pragma unroll
  for (int s = 0; s < KM; s++) {
    lu[s] *= di;
    lv[s] *= di;
    lh[s] *= di;
  }
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double al[KM];
    afrag64<NM>(O + P.lft, rb, al);
    doublex4 Fx = chain64(al, lu, zero64()), Fy = chain64(al, lv, zero64());
    doublex4 Fh = chain64(al, lh, zero64());
    if (valid) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int kj = drow(rb, r, g);
        if (kj < NE3) {
          const int k = kj / NGF;
          const double sg = gg[11 + k];
          tr_put<NE3, 2>(A.st.abar, eb, col, 0, kj,
                         0.5 * sg * (-g0) * (gg[5 + k] * Fx[r] + gg[8 + k] * Fy[r]));
          tr_put<NE3, 2>(A.st.abar, eb, col, 1, kj, -0.5 * sg * Fh[r]);
        }
      }
    }
  }
}

// Pass 2 of R^T and the adjoint RK4 stage (stage < 0: y -> k).
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void adj_gather_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, KE = Sh::KE, RBM = Sh::RBM;
  constexpr Pack64T P = pack64t(NM, NE3);
  constexpr int N16 = (P.total - P.trt) / 2;
  __shared__ floatx4 lds[N16];
  const int b0 = first_block();
  if (b0 >= nblk) return;
  stage_lds(lds, reinterpret_cast<const floatx4 *>(A.ops.pack64t + P.trt), N16);
  __syncthreads();
  const int eb = b0 + (threadIdx.x >> 6);
  if (eb >= nblk) return;
  const double *O = reinterpret_cast<const double *>(lds);
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double hh0 = valid ? A.fld.hh0e[e] : 0.0;
  const size_t mb = (size_t)e0 * NM;
  const bool wall = valid && A.mesh.bdl[e] != 0;
  const bool anyw = __ballot(wall) != 0;
  double th[KE], tu[KE], tv[KE];
#pragma unroll
  for (int s = 0; s < KE; s++) {
    const int kj = 4 * s + g;
    th[s] = tu[s] = tv[s] = 0.0;
    if (kj < NE3 && valid) {
      const int k = kj / NGF, j = kj % NGF;
      th[s] = face_avg<NGF, 2>(A.st.abar, A.mesh, e, k, j, 0);
      const double sa = hh0 * face_avg<NGF, 2>(A.st.abar, A.mesh, e, k, j, 1);
      tu[s] = sa * gg[5 + k];
      tv[s] = sa * gg[8 + k];
    }
  }
  double yu[KM], yv[KM], yh[KM];
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    double at[KE];
    afrag64<NE3>(O, rb, at);
    doublex4 Th = chain64(at, th, zero64()), Tu = chain64(at, tu, zero64());
    doublex4 Tv = chain64(at, tv, zero64());
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (4 * rb + r < KM) {
        yh[4 * rb + r] = Th[r];
        yu[4 * rb + r] = Tu[r];
        yv[4 * rb + r] = Tv[r];
      }
  }
  if (VAR != 0 && anyw) wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gg[14], gg[15], yu, yv);
#pragma unroll
  for (int s = 0; s < KM; s++) {
    const int m = 4 * s + g;
    if (valid && m < NM) {
      const size_t idx = mb + m * 16 + col;
      const double y[3] = {A.st.avol[0][idx] + yu[s], A.st.avol[1][idx] + yv[s],
                           A.st.avol[2][idx] + yh[s]};
#pragma unroll
      for (int f = 0; f < 3; f++) {
        if (A.stage < 0) A.st.k[f][idx] = y[f];
        else adj_rk_value(A, f, idx, y[f]);
      }
    }
  }
}

// Adjoint variables transform with U^-T: lam_c = U^-T lam_a on the way in
// (forward substitution with the lower factor U^T), lam_a = U^T lam_c out.
template <int NM>
__global__ void adj_to_psi_kernel(int ne, const double *__restrict__ U, const double *__restrict__ a,
                                  double *__restrict__ c) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double x[NM];
#pragma unroll
  for (int i = 0; i < NM; i++) {
    double s = a[(size_t)e * NM + i];
#pragma unroll
    for (int k = 0; k < i; k++) s -= U[k + NM * i] * x[k];
    x[i] = s / U[i + NM * i];
  }
  double *cb = c + (size_t)(e & ~15) * NM + (e & 15);
#pragma unroll
  for (int i = 0; i < NM; i++) cb[i * 16] = x[i];
}

template <int NM>
__global__ void adj_from_psi_kernel(int ne, const double *__restrict__ U,
                                    const double *__restrict__ c, double *__restrict__ a) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * NM) return;
  int e = t / NM, i = t % NM;
  const double *cb = c + (size_t)(e & ~15) * NM + (e & 15);
  double s = 0.0;
  for (int k = 0; k <= i; k++) s += U[k + NM * i] * cb[k * 16];
  a[(size_t)e * NM + i] = s;
}

// w = dt/6 lam: first input of a backward sweep.
template <int NM>
__global__ void adj_start_kernel(int n, double dt6, const double *__restrict__ lu,
                                 const double *__restrict__ lv, const double *__restrict__ lh,
                                 double *__restrict__ wu, double *__restrict__ wv,
                                 double *__restrict__ wh) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  wu[t] = dt6 * lu[t];
  wv[t] = dt6 * lv[t];
  wh[t] = dt6 * lh[t];
}

template <int NC, int NG, int NGF>
int adj_launch(const F32Args &A, int phase, const double *in, double *out) {
  constexpr int NM = nm_of(NC);
  const int ne = A.mesh.ne, nblk = cdiv(ne, 16);
  const int grid = 8 * cdiv(cdiv(nblk, NW), 8);
  hipStream_t st = A.stream;
  const int v = A.variant;
  switch (phase) {
    case 0:  // adjoint into psi coordinates
      adj_to_psi_kernel<NM><<<(ne + 127) / 128, 128, 0, st>>>(ne, A.ops.U, in, out);
      break;
    case 1:  // adjoint back to the reference coordinates
      adj_from_psi_kernel<NM><<<(ne * NM + 255) / 256, 256, 0, st>>>(ne, A.ops.U, in, out);
      break;
    case 2: {  // w = dt/6 lam
      const int n = nblk * 16 * NM;
      adj_start_kernel<NM><<<(n + 255) / 256, 256, 0, st>>>(
          n, A.dt / 6.0, A.st.lam[0], A.st.lam[1], A.st.lam[2], A.st.aw[0], A.st.aw[1], A.st.aw[2]);
      break;
    }
    case 3:
      if (v == 0) adj_lift_kernel<NC, NG, NGF, 0><<<grid, NT, 0, st>>>(A, nblk);
      else return -1;
      break;
    case 4:
      if (v == 0) adj_gather_kernel<NC, NG, NGF, 0><<<grid, NT, 0, st>>>(A, nblk);
      else return -1;
      break;
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace f32

#if defined(SPOC_NC)
#define SPOC_CAT4(a, b, c, d) a##_##b##_##c##_##d
#define SPOC_NAME4(a, b, c, d) SPOC_CAT4(a, b, c, d)
int SPOC_NAME4(adj_f64, SPOC_NC, SPOC_NG, SPOC_NGF)(const F32Args &a, int phase, const double *in,
                                                  double *out) {
  return f32::adj_launch<SPOC_NC, SPOC_NG, SPOC_NGF>(a, phase, in, out);
}
#endif

}  // namespace spoc
