// Discrete adjoint of the RK4 step (src/shwwnl.f:233-284) of the right-hand
// side of every variant, FP64 on MFMA, in the mass-orthonormal coordinates
// of the mixed-precision path.
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
//
// The nonlinear variants add the transpose of the B+C and E+G stages
// (src/calvarnl_fs.f:115-544 and the ns/inv variants) linearised at the
// stage's base state xb: adj_lift also transposes the E+G volume and lift
// terms; adj_mid gathers the face adjoints, applies the transposed
// products at the quadrature points (with the base values recomputed from
// xb and its traces btr) and the transposed velocity gradient; adj_final
// gathers the gradient's face adjoints and runs the RK4 stage.
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

// A fragments of k-steps s0 .. s0+3 (s0 a multiple of 4) of row block rb.
template <int K>
__device__ __forceinline__ void afrag64_4(const double *img, int rb, int s0, double (&a)[4]) {
  constexpr int KH = cdiv(K, 8);
  const doublex2 *p = reinterpret_cast<const doublex2 *>(img) + (size_t)rb * KH * 64 + lane();
  doublex2 v0 = p[(s0 / 2) * 64];
  doublex2 v1 = (s0 / 2 + 1 < KH) ? p[(s0 / 2 + 1) * 64] : doublex2{0.0, 0.0};
  a[0] = v0[0];
  a[1] = v0[1];
  a[2] = v1[0];
  a[3] = v1[1];
}

// Free-slip wall projection of the velocity-gradient tensor
// (src/calvarnl_fs.f:165-204) in FP64, B fragments in place: the t-n
// component goes through the reduced basis.  Symmetric, so it is also the
// projection of the tensor's adjoint.
template <int NM>
__device__ __forceinline__ void wall_tensor64(const double *__restrict__ gl, bool wall, double nx,
                                              double ny, double (&q0)[cdiv(NM, 4)],
                                              double (&q1)[cdiv(NM, 4)],
                                              double (&q2)[cdiv(NM, 4)],
                                              double (&q3)[cdiv(NM, 4)]) {
  constexpr int KM = cdiv(NM, 4), RBM = cdiv(NM, 16);
  if (!wall) nx = ny = 0.0;
  double v[KM];
#pragma unroll
  for (int s = 0; s < KM; s++) v[s] = (q0[s] * ny - q2[s] * nx) * nx + (q1[s] * ny - q3[s] * nx) * ny;
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    double ag[KM];
    afrag64<NM>(gl, rb, ag);
    doublex4 d = chain64(ag, v, zero64());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int s = 4 * rb + r;
      if (s < KM) {
        q0[s] += nx * ny * d[r];
        q2[s] -= nx * nx * d[r];
        q1[s] += ny * ny * d[r];
        q3[s] -= nx * ny * d[r];
      }
    }
  }
}

// Pass 1 of R^T (VAR 0: src/calvar_lin.f; the nonlinear variants' linear
// core takes the same path with the flux projection, plus the transposed
// E+G terms).
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void adj_lift_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, RBM = Sh::RBM, RBE = Sh::RBE;
  constexpr Pack64T P = pack64t(NM, NG, NE3);
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
  const double g0 = A.fld.g0, hh0 = valid ? core_depth<VAR>(A.fld, e) : 0.0;
  const size_t mb = (size_t)e0 * NM;
  const bool wall = valid && A.mesh.bdl[e] != 0;
  const bool anyw = __ballot(wall) != 0;
  double lu[KM], lv[KM], lh[KM];
  bload<NM>(A.st.aw[0] + mb, g, col, valid, lu);
  bload<NM>(A.st.aw[1] + mb, g, col, valid, lv);
  bload<NM>(A.st.aw[2] + mb, g, col, valid, lh);
  if (anyw) wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gg[14], gg[15], lu, lv);
  if (VAR != 0) {
    // the (u, v) sources enter the tendency as they are: keep their adjoint;
    // transposed E+G volume terms of the viscous flux F = mu q
    double pu[KM], pv[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) {
      const int m = 4 * s + g;
      if (valid && m < NM) {
        A.st.aL[0][mb + m * 16 + col] = lu[s];
        A.st.aL[1][mb + m * 16 + col] = lv[s];
      }
      pu[s] = di * lu[s];
      pv[s] = di * lv[s];
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double adx[KM], ady[KM];
      afrag64<NM>(O + P.dxt, rb, adx);
      afrag64<NM>(O + P.dyt, rb, ady);
      doublex4 o0 = chain64(adx, pu, zero64()), o1 = chain64(ady, pu, zero64());
      doublex4 o2 = chain64(adx, pv, zero64()), o3 = chain64(ady, pv, zero64());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = drow(rb, r, g);
        if (valid && m < NM) {
          const size_t idx = mb + m * 16 + col;
          A.st.aF[0][idx] = -by * o0[r] + ay * o1[r];
          A.st.aF[1][idx] = bx * o0[r] - ax * o1[r];
          A.st.aF[2][idx] = -by * o2[r] + ay * o3[r];
          A.st.aF[3][idx] = bx * o2[r] - ax * o3[r];
        }
      }
    }
  }
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
        if (VAR != 0) {  // the same vector is the adjoint of the raw mass flux u (h + hb - H0)
          A.st.aUH[0][mb + m * 16 + col] = bu[s];
          A.st.aUH[1][mb + m * 16 + col] = bv[s];
        }
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
          tr_put<NE3, 4>(A.st.abar, eb, col, 0, kj,
                         0.5 * sg * (-g0) * (gg[5 + k] * Fx[r] + gg[8 + k] * Fy[r]));
          tr_put<NE3, 4>(A.st.abar, eb, col, 1, kj, -0.5 * sg * Fh[r]);
          if (VAR != 0) {  // the viscous-flux traces of the u and v equations
            tr_put<NE3, 4>(A.st.abar, eb, col, 2, kj, 0.5 * sg * Fx[r]);
            tr_put<NE3, 4>(A.st.abar, eb, col, 3, kj, 0.5 * sg * Fy[r]);
          }
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
  constexpr Pack64T P = pack64t(NM, NG, NE3);
  constexpr int N16 = (P.total_lin - P.trt) / 2;
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
  const double hh0 = valid ? core_depth<VAR>(A.fld, e) : 0.0;
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
      th[s] = face_avg<NGF, 4>(A.st.abar, A.mesh, e, k, j, 0);
      const double sa = hh0 * face_avg<NGF, 4>(A.st.abar, A.mesh, e, k, j, 1);
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

// Velocity traces of the base state xb (FP64) for the base gradient q.
template <int NC, int NG, int NGF>
__global__ __launch_bounds__(NT) void adj_btr_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, RBE = Sh::RBE;
  const int eb = first_block() + (threadIdx.x >> 6);
  if (eb >= nblk) return;
  const int l = lane(), g = l >> 4, col = l & 15, e = eb * 16 + col;
  const bool valid = e < A.mesh.ne;
  const size_t mb = (size_t)eb * 16 * NM;
  double bu[KM], bv[KM];
  bload<NM>(A.st.xb[0] + mb, g, col, valid, bu);
  bload<NM>(A.st.xb[1] + mb, g, col, valid, bv);
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double at[KM];
    afrag64<NM>(A.ops.pack64 + Sh::P64.tr, rb, at);
    doublex4 Tu = chain64(at, bu, zero64()), Tv = chain64(at, bv, zero64());
    if (valid) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int kj = drow(rb, r, g);
        if (kj < NE3) {
          tr_put<NE3, 2>(A.st.btr, eb, col, 0, kj, Tu[r]);
          tr_put<NE3, 2>(A.st.btr, eb, col, 1, kj, Tv[r]);
        }
      }
    }
  }
}

// Pass 2 of R^T for the nonlinear variants (FP64, operators read from
// global memory), in two kernels.  adj_mid1: the face adjoints of pass 1
// (eta trace, mass-flux trace, viscous-flux traces) gathered and lifted,
// added in place to the element-local work arrays of pass 1: avol (the
// partial result), aUH (the mass-flux adjoints) and aF, which then holds
// the adjoint of the gradient q, q-bar = mu F-bar + T^T (traces).
// adj_mid2: the base state's values at the quadrature points (u, v, eta and
// the base gradient q, recomputed from xb and btr as the forward B stage
// computes them); the transposed pointwise products of
// src/calvarnl_fs.f:245-296 and src/nonlin.f:35-133
//   gx = -(u qxx + v (qxy - f)) + (w0 - mb0 u) / (eta + hb)
//   gy = -(u (qyx + f) + v qyy) + (z0 - mb0 v) / (eta + hb)
//   uh = u (eta + hb),  vh = v (eta + hb);
// projection back to modes (PSI^T), which also gives the adjoint of hb at
// the quadrature points (hb enters only through eta + hb); the wall
// projection of q-bar and the transposed gradient of src/calvarnl_fs.f:129-158.
// Writes the partial result to avol and the gradient's face adjoints to
// abar2.  (One kernel holding all of it needed about 660 registers per lane
// at p = 5 and spilled; the hand-off through the work arrays it already
// reads costs less.)
// (adj_mid1 fits two waves per SIMD in 256 registers up to p = 6: 4% off
// the adjoint step at p = 5 against one wave with the registers it asks for)
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NC <= 6 ? 2 : 1))) void adj_mid1_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, KE = Sh::KE, KG = Sh::KG, RBM = Sh::RBM,
                RBE = Sh::RBE, RBG = Sh::RBG;
  constexpr Pack64 P = Sh::P64;
  constexpr Pack64T T = pack64t(NM, NG, NE3);
  const int eb = first_block() + (threadIdx.x >> 6);
  if (eb >= nblk) return;
  const double *O = A.ops.pack64, *OT = A.ops.pack64t;
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.0;
  const double hh0 = valid ? core_depth<VAR>(A.fld, e) : 0.0;
  const double mu0 = A.fld.mu0, mb0 = A.fld.mb0;
  const size_t mb = (size_t)e0 * NM;
  const bool wall = valid && A.mesh.bdl[e] != 0;
  const bool anyw = __ballot(wall) != 0;
  const double nwx = wall ? gg[14] : 0.0, nwy = wall ? gg[15] : 0.0;

  // face adjoints of pass 1
  double yu[KM], yv[KM], yh[KM], uhr[KM], vhr[KM];
  double q0[KM], q1[KM], q2[KM], q3[KM];  // adjoint of the gradient q (B fragments)
  {
    double th[KE], su[KE], sv[KE], t0x[KE], t0y[KE], t1x[KE], t1y[KE];
#pragma unroll
    for (int s = 0; s < KE; s++) {
      const int kj = 4 * s + g;
      th[s] = su[s] = sv[s] = t0x[s] = t0y[s] = t1x[s] = t1y[s] = 0.0;
      if (kj < NE3 && valid) {
        const int k = kj / NGF, j = kj % NGF;
        const double nfx = gg[5 + k], nfy = gg[8 + k];
        th[s] = face_avg<NGF, 4>(A.st.abar, A.mesh, e, k, j, 0);
        const double xs = face_avg<NGF, 4>(A.st.abar, A.mesh, e, k, j, 1);
        const double a0 = mu0 * face_avg<NGF, 4>(A.st.abar, A.mesh, e, k, j, 2);
        const double a1 = mu0 * face_avg<NGF, 4>(A.st.abar, A.mesh, e, k, j, 3);
        su[s] = nfx * xs;
        sv[s] = nfy * xs;
        t0x[s] = nfx * a0;
        t0y[s] = nfy * a0;
        t1x[s] = nfx * a1;
        t1y[s] = nfy * a1;
      }
    }
    double du[KM], dv[KM];
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double at[KE];
      afrag64<NE3>(OT + T.trt, rb, at);
      doublex4 Th = chain64(at, th, zero64()), Du = chain64(at, su, zero64());
      doublex4 Dv = chain64(at, sv, zero64()), Q0 = chain64(at, t0x, zero64());
      doublex4 Q1 = chain64(at, t0y, zero64()), Q2 = chain64(at, t1x, zero64());
      doublex4 Q3 = chain64(at, t1y, zero64());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int s = 4 * rb + r;
        if (s < KM) {
          yh[s] = Th[r];
          du[s] = Du[r];
          dv[s] = Dv[r];
          q0[s] = Q0[r];
          q1[s] = Q1[r];
          q2[s] = Q2[r];
          q3[s] = Q3[r];
        }
      }
    }
    // the mass flux of both the linear core and the B stage is wall-projected
    if (anyw) wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gg[14], gg[15], du, dv);
#pragma unroll
    for (int s = 0; s < KM; s++) {
      const int m = 4 * s + g;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0, h0 = 0.0, h1 = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0,
             f3 = 0.0;
      if (valid && m < NM) {
        const size_t idx = mb + m * 16 + col;
        v0 = A.st.avol[0][idx];
        v1 = A.st.avol[1][idx];
        v2 = A.st.avol[2][idx];
        h0 = A.st.aUH[0][idx];
        h1 = A.st.aUH[1][idx];
        f0 = A.st.aF[0][idx];
        f1 = A.st.aF[1][idx];
        f2 = A.st.aF[2][idx];
        f3 = A.st.aF[3][idx];
      }
      yu[s] = v0 + hh0 * du[s];
      yv[s] = v1 + hh0 * dv[s];
      yh[s] += v2;
      uhr[s] = h0 + du[s];
      vhr[s] = h1 + dv[s];
      q0[s] += mu0 * f0;
      q1[s] += mu0 * f1;
      q2[s] += mu0 * f2;
      q3[s] += mu0 * f3;
    }
  }
#pragma unroll
  for (int s = 0; s < KM; s++) {
    const int m = 4 * s + g;
    if (valid && m < NM) {
      const size_t idx = mb + m * 16 + col;
      A.st.avol[0][idx] = yu[s];
      A.st.avol[1][idx] = yv[s];
      A.st.avol[2][idx] = yh[s];
      A.st.aUH[0][idx] = uhr[s];
      A.st.aUH[1][idx] = vhr[s];
      A.st.aF[0][idx] = q0[s];
      A.st.aF[1][idx] = q1[s];
      A.st.aF[2][idx] = q2[s];
      A.st.aF[3][idx] = q3[s];
    }
  }
}

template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void adj_mid2_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, KE = Sh::KE, KG = Sh::KG, RBM = Sh::RBM,
                RBE = Sh::RBE, RBG = Sh::RBG;
  constexpr Pack64 P = Sh::P64;
  constexpr Pack64T T = pack64t(NM, NG, NE3);
  const int eb = first_block() + (threadIdx.x >> 6);
  if (eb >= nblk) return;
  const double *O = A.ops.pack64, *OT = A.ops.pack64t;
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.0;
  const double hh0 = valid ? core_depth<VAR>(A.fld, e) : 0.0;
  const double mu0 = A.fld.mu0, mb0 = A.fld.mb0;
  const size_t mb = (size_t)e0 * NM;
  const bool wall = valid && A.mesh.bdl[e] != 0;
  const bool anyw = __ballot(wall) != 0;
  const double nwx = wall ? gg[14] : 0.0, nwy = wall ? gg[15] : 0.0;

  double uhr[KM], vhr[KM];
  bload<NM>(A.st.aUH[0] + mb, g, col, valid, uhr);
  bload<NM>(A.st.aUH[1] + mb, g, col, valid, vhr);
  double lsu[KM], lsv[KM];
  bload<NM>(A.st.aL[0] + mb, g, col, valid, lsu);
  bload<NM>(A.st.aL[1] + mb, g, col, valid, lsv);

  // (2) base state and its gradient
  double bu[KM], bv[KM], bh[KM];
  bload<NM>(A.st.xb[0] + mb, g, col, valid, bu);
  bload<NM>(A.st.xb[1] + mb, g, col, valid, bv);
  bload<NM>(A.st.xb[2] + mb, g, col, valid, bh);
  double Qb0[KM], Qb1[KM], Qb2[KM], Qb3[KM];
  {
    double fb0[KE], fb1[KE], fb2[KE], fb3[KE];
#pragma unroll
    for (int s = 0; s < KE; s++) {
      const int kj = 4 * s + g;
      fb0[s] = fb1[s] = fb2[s] = fb3[s] = 0.0;
      if (kj < NE3 && valid) {
        const int k = kj / NGF, j = kj % NGF;
        const double sg = 0.5 * gg[11 + k], nx = sg * gg[5 + k], ny = sg * gg[8 + k];
        const double fu = face_avg<NGF, 2>(A.st.btr, A.mesh, e, k, j, 0);
        const double fv = face_avg<NGF, 2>(A.st.btr, A.mesh, e, k, j, 1);
        fb0[s] = fu * nx;
        fb1[s] = fu * ny;
        fb2[s] = fv * nx;
        fb3[s] = fv * ny;
      }
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double adx[KM], ady[KM], al[KE];
      afrag64<NM>(O + P.dx, rb, adx);
      afrag64<NM>(O + P.dy, rb, ady);
      afrag64<NE3>(O + P.lf, rb, al);
      doublex4 Au = chain64(adx, bu, zero64()), Bu = chain64(ady, bu, zero64());
      doublex4 Av = chain64(adx, bv, zero64()), Bv = chain64(ady, bv, zero64());
      doublex4 L0 = chain64(al, fb0, zero64()), L1 = chain64(al, fb1, zero64());
      doublex4 L2 = chain64(al, fb2, zero64()), L3 = chain64(al, fb3, zero64());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int s = 4 * rb + r;
        if (s < KM) {
          Qb0[s] = di * (-by * Au[r] + ay * Bu[r] + L0[r]);
          Qb1[s] = di * (bx * Au[r] - ax * Bu[r] + L1[r]);
          Qb2[s] = di * (-by * Av[r] + ay * Bv[r] + L2[r]);
          Qb3[s] = di * (bx * Av[r] - ax * Bv[r] + L3[r]);
        }
      }
    }
    if (VAR == 1 && anyw) wall_tensor64<NM>(A.ops.gl64, wall, nwx, nwy, Qb0, Qb1, Qb2, Qb3);
  }

  // (3) quadrature points
  doublex4 au[RBM], av[RBM], ah[RBM], aq0[RBM], aq1[RBM], aq2[RBM], aq3[RBM];
#pragma unroll
  for (int rb = 0; rb < RBM; rb++)
    au[rb] = av[rb] = ah[rb] = aq0[rb] = aq1[rb] = aq2[rb] = aq3[rb] = zero64();
  const float *fc3 = A.fld.fc3 + (size_t)(valid ? e : 0) * 3;
  const double fc0 = fc3[0], fc1 = fc3[1], fc2 = fc3[2];
#pragma unroll
  for (int qb = 0; qb < RBG; qb++) {
    double ap[KM], aw[KM];
    afrag64<NM>(OT + T.psi, qb, ap);
    afrag64<NM>(OT + T.psiwt, qb, aw);
    doublex4 Vu = chain64(ap, bu, zero64()), Vv = chain64(ap, bv, zero64());
    doublex4 Vh = chain64(ap, bh, zero64());
    doublex4 Vq0 = chain64(ap, Qb0, zero64()), Vq1 = chain64(ap, Qb1, zero64());
    doublex4 Vq2 = chain64(ap, Qb2, zero64()), Vq3 = chain64(ap, Qb3, zero64());
    doublex4 Gx = chain64(aw, lsu, zero64()), Gy = chain64(aw, lsv, zero64());
    doublex4 Uh = chain64(aw, uhr, zero64()), Wh = chain64(aw, vhr, zero64());
    double pu[4], pv[4], ph[4], p0[4], p1[4], p2[4], p3[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int q = drow(qb, r, g);
      pu[r] = pv[r] = ph[r] = p0[r] = p1[r] = p2[r] = p3[r] = 0.0;
      if (q < NG && valid) {
        const float *F = A.fld.qf + (((size_t)eb * KG + 4 * qb + r) * 64 + l) * 4;
        const double f = fc0 + fc1 * OT[T.yq + q] + fc2 * OT[T.xq + q];
        const double hk = Vh[r] + (double)F[0], hf = Vh[r] + (double)F[3];
        const double ihk = 1.0 / hk;
        const double u = Vu[r], v = Vv[r], gx = Gx[r], gy = Gy[r];
        pu[r] = gx * (-Vq0[r] - mb0 * ihk) - gy * (Vq2[r] + f) + Uh[r] * hf;
        pv[r] = -gx * (Vq1[r] - f) + gy * (-Vq3[r] - mb0 * ihk) + Wh[r] * hf;
        ph[r] = -(gx * ((double)F[1] - mb0 * u) + gy * ((double)F[2] - mb0 * v)) *
                    ihk * ihk +
                Uh[r] * u + Wh[r] * v;
        p0[r] = -u * gx;
        p1[r] = -v * gx;
        p2[r] = -u * gy;
        p3[r] = -v * gy;
      }
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double a4[4];
      afrag64_4<NG>(OT + T.psit, rb, 4 * qb, a4);
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (4 * qb + r < KG) {
          au[rb] = mfma64(a4[r], pu[r], au[rb]);
          av[rb] = mfma64(a4[r], pv[r], av[rb]);
          ah[rb] = mfma64(a4[r], ph[r], ah[rb]);
          aq0[rb] = mfma64(a4[r], p0[r], aq0[rb]);
          aq1[rb] = mfma64(a4[r], p1[r], aq1[rb]);
          aq2[rb] = mfma64(a4[r], p2[r], aq2[rb]);
          aq3[rb] = mfma64(a4[r], p3[r], aq3[rb]);
        }
    }
  }

  // the partial result and q-bar of adj_mid1
  double yu[KM], yv[KM], yh[KM], q0[KM], q1[KM], q2[KM], q3[KM];
  bload<NM>(A.st.avol[0] + mb, g, col, valid, yu);
  bload<NM>(A.st.avol[1] + mb, g, col, valid, yv);
  bload<NM>(A.st.avol[2] + mb, g, col, valid, yh);
  bload<NM>(A.st.aF[0] + mb, g, col, valid, q0);
  bload<NM>(A.st.aF[1] + mb, g, col, valid, q1);
  bload<NM>(A.st.aF[2] + mb, g, col, valid, q2);
  bload<NM>(A.st.aF[3] + mb, g, col, valid, q3);
  // (4) combine; the quadrature part of the eta adjoint is also hb's
#pragma unroll
  for (int s = 0; s < KM; s++) {
    const int rb = s >> 2, r = s & 3, m = 4 * s + g;
    yu[s] += au[rb][r];
    yv[s] += av[rb][r];
    yh[s] += ah[rb][r];
    q0[s] += aq0[rb][r];
    q1[s] += aq1[rb][r];
    q2[s] += aq2[rb][r];
    q3[s] += aq3[rb][r];
    if (valid && m < NM) A.st.hbbar[mb + m * 16 + col] += ah[rb][r];
  }
  if (VAR == 1 && anyw) wall_tensor64<NM>(A.ops.gl64, wall, nwx, nwy, q0, q1, q2, q3);
  // transposed gradient: volume ...
  {
    double p0[KM], p1[KM], p2[KM], p3[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) {
      p0[s] = di * (-by * q0[s] + bx * q1[s]);
      p1[s] = di * (ay * q0[s] - ax * q1[s]);
      p2[s] = di * (-by * q2[s] + bx * q3[s]);
      p3[s] = di * (ay * q2[s] - ax * q3[s]);
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double adx[KM], ady[KM];
      afrag64<NM>(OT + T.dxt, rb, adx);
      afrag64<NM>(OT + T.dyt, rb, ady);
      doublex4 U = chain64(ady, p1, chain64(adx, p0, zero64()));
      doublex4 V = chain64(ady, p3, chain64(adx, p2, zero64()));
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int s = 4 * rb + r;
        if (s < KM) {
          yu[s] += U[r];
          yv[s] += V[r];
        }
      }
    }
  }
  // ... and its face terms at the element's own edge points
#pragma unroll
  for (int s = 0; s < KM; s++) {
    q0[s] *= di;
    q1[s] *= di;
    q2[s] *= di;
    q3[s] *= di;
  }
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double al[KM];
    afrag64<NM>(OT + T.lft, rb, al);
    doublex4 F0 = chain64(al, q0, zero64()), F1 = chain64(al, q1, zero64());
    doublex4 F2 = chain64(al, q2, zero64()), F3 = chain64(al, q3, zero64());
    if (valid) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int kj = drow(rb, r, g);
        if (kj < NE3) {
          const int k = kj / NGF;
          const double sg = 0.5 * gg[11 + k], nx = sg * gg[5 + k], ny = sg * gg[8 + k];
          tr_put<NE3, 2>(A.st.abar2, eb, col, 0, kj, nx * F0[r] + ny * F1[r]);
          tr_put<NE3, 2>(A.st.abar2, eb, col, 1, kj, nx * F2[r] + ny * F3[r]);
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < KM; s++) {
    const int m = 4 * s + g;
    if (valid && m < NM) {
      const size_t idx = mb + m * 16 + col;
      A.st.avol[0][idx] = yu[s];
      A.st.avol[1][idx] = yv[s];
      A.st.avol[2][idx] = yh[s];
    }
  }
}

// Pass 3 (nonlinear variants): the gradient's face adjoints, then the
// adjoint RK4 stage (stage < 0: y -> k).
template <int NC, int NG, int NGF>
__global__ __launch_bounds__(NT) void adj_final_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, KE = Sh::KE, RBM = Sh::RBM;
  constexpr Pack64T T = pack64t(NM, NG, NE3);
  const int eb = first_block() + (threadIdx.x >> 6);
  if (eb >= nblk) return;
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const size_t mb = (size_t)e0 * NM;
  double tu[KE], tv[KE];
#pragma unroll
  for (int s = 0; s < KE; s++) {
    const int kj = 4 * s + g;
    tu[s] = tv[s] = 0.0;
    if (kj < NE3 && valid) {
      const int k = kj / NGF, j = kj % NGF;
      tu[s] = face_avg<NGF, 2>(A.st.abar2, A.mesh, e, k, j, 0);
      tv[s] = face_avg<NGF, 2>(A.st.abar2, A.mesh, e, k, j, 1);
    }
  }
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    double at[KE];
    afrag64<NE3>(A.ops.pack64t + T.trt, rb, at);
    doublex4 U = chain64(at, tu, zero64()), V = chain64(at, tv, zero64());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = drow(rb, r, g);
      if (valid && m < NM) {
        const size_t idx = mb + m * 16 + col;
        const double y[3] = {A.st.avol[0][idx] + U[r], A.st.avol[1][idx] + V[r],
                             A.st.avol[2][idx]};
#pragma unroll
        for (int f = 0; f < 3; f++) {
          if (A.stage < 0) A.st.k[f][idx] = y[f];
          else adj_rk_value(A, f, idx, y[f]);
        }
      }
    }
  }
}

// Adjoint variables transform with U^-T: lam_c = U^-T lam_a on the way in
// (forward substitution with the lower factor U^T), lam_a = U^T lam_c out.
// Both run one thread per element, 64 elements per workgroup, with U and
// the element-major side staged in LDS as the forward basis change does.
template <int NM>
__global__ __launch_bounds__(PSI_WG) void adj_to_psi_kernel(int ne, const double *__restrict__ U,
                                                            const double *__restrict__ a,
                                                            double *__restrict__ c) {
  __shared__ double Us[NM * NM];
  __shared__ double T[PSI_WG * (NM + 1)];
  const int e0 = blockIdx.x * PSI_WG, n = min(PSI_WG, ne - e0);
  for (int i = threadIdx.x; i < NM * NM; i += PSI_WG) Us[i] = U[i];
  const double *src = a + (size_t)e0 * NM;
  for (int t = threadIdx.x; t < n * NM; t += PSI_WG) T[(t / NM) * (NM + 1) + t % NM] = src[t];
  __syncthreads();
  const int e = e0 + threadIdx.x;
  if (e >= ne) return;
  double x[NM];
#pragma unroll
  for (int i = 0; i < NM; i++) {
    double s = T[threadIdx.x * (NM + 1) + i];
#pragma unroll
    for (int k = 0; k < i; k++) s -= Us[k + NM * i] * x[k];
    x[i] = s / Us[i + NM * i];
  }
  double *cb = c + (size_t)(e & ~15) * NM + (e & 15);
#pragma unroll
  for (int i = 0; i < NM; i++) cb[i * 16] = x[i];
}

template <int NM>
__global__ __launch_bounds__(PSI_WG) void adj_from_psi_kernel(int ne, const double *__restrict__ U,
                                                              const double *__restrict__ c,
                                                              double *__restrict__ a) {
  __shared__ double Us[NM * NM];
  __shared__ double T[PSI_WG * (NM + 1)];
  const int e0 = blockIdx.x * PSI_WG, n = min(PSI_WG, ne - e0);
  for (int i = threadIdx.x; i < NM * NM; i += PSI_WG) Us[i] = U[i];
  __syncthreads();
  const int e = e0 + threadIdx.x;
  if (e < ne) {
    double x[NM];
    const double *cb = c + (size_t)(e & ~15) * NM + (e & 15);
#pragma unroll
    for (int k = 0; k < NM; k++) x[k] = cb[k * 16];
#pragma unroll
    for (int i = 0; i < NM; i++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k <= i; k++) s += Us[k + NM * i] * x[k];
      T[threadIdx.x * (NM + 1) + i] = s;
    }
  }
  __syncthreads();
  double *dst = a + (size_t)e0 * NM;
  for (int t = threadIdx.x; t < n * NM; t += PSI_WG) dst[t] = T[(t / NM) * (NM + 1) + t % NM];
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
      adj_to_psi_kernel<NM><<<(ne + PSI_WG - 1) / PSI_WG, PSI_WG, 0, st>>>(ne, A.ops.U, in, out);
      break;
    case 1:  // adjoint back to the reference coordinates
      adj_from_psi_kernel<NM><<<(ne + PSI_WG - 1) / PSI_WG, PSI_WG, 0, st>>>(ne, A.ops.U, in, out);
      break;
    case 2: {  // w = dt/6 lam
      const int n = nblk * 16 * NM;
      adj_start_kernel<NM><<<(n + 255) / 256, 256, 0, st>>>(
          n, A.dt / 6.0, A.st.lam[0], A.st.lam[1], A.st.lam[2], A.st.aw[0], A.st.aw[1], A.st.aw[2]);
      break;
    }
    case 3:  // pass 1
      if (v == 0) adj_lift_kernel<NC, NG, NGF, 0><<<grid, NT, 0, st>>>(A, nblk);
      else if (v == 1) adj_lift_kernel<NC, NG, NGF, 1><<<grid, NT, 0, st>>>(A, nblk);
      else if (v == 2) adj_lift_kernel<NC, NG, NGF, 2><<<grid, NT, 0, st>>>(A, nblk);
      else adj_lift_kernel<NC, NG, NGF, 3><<<grid, NT, 0, st>>>(A, nblk);
      break;
    case 4:  // pass 2 (the last pass of the linear variant)
      if (v == 0) adj_gather_kernel<NC, NG, NGF, 0><<<grid, NT, 0, st>>>(A, nblk);
      else if (v == 1) {
        adj_mid1_kernel<NC, NG, NGF, 1><<<grid, NT, 0, st>>>(A, nblk);
        adj_mid2_kernel<NC, NG, NGF, 1><<<grid, NT, 0, st>>>(A, nblk);
      } else if (v == 2) {
        adj_mid1_kernel<NC, NG, NGF, 2><<<grid, NT, 0, st>>>(A, nblk);
        adj_mid2_kernel<NC, NG, NGF, 2><<<grid, NT, 0, st>>>(A, nblk);
      } else {
        adj_mid1_kernel<NC, NG, NGF, 3><<<grid, NT, 0, st>>>(A, nblk);
        adj_mid2_kernel<NC, NG, NGF, 3><<<grid, NT, 0, st>>>(A, nblk);
      }
      break;
    case 5:  // pass 3 (nonlinear)
      if (v == 0) return -1;
      adj_final_kernel<NC, NG, NGF><<<grid, NT, 0, st>>>(A, nblk);
      break;
    case 6:  // base-state traces (nonlinear)
      adj_btr_kernel<NC, NG, NGF><<<grid, NT, 0, st>>>(A, nblk);
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
