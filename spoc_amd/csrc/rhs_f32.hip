// Mixed-precision right-hand side for gfx950: FP32 MFMA
// (v_mfma_f32_16x16x4_f32) for the quadrature of the nonlinear, viscous,
// Coriolis and forcing terms; FP64 MFMA (v_mfma_f64_16x16x4_f64) for the
// linear gravity-wave core, the wall projections of the tendencies and the
// traces of the stage input; FP64 RK4 state.
//
// Coordinates: the mass-orthonormal basis psi = U^{-T} phi, so each step of
// the reference's cal_var is a small dense product
//   out[row, element] = Op[row, k] * X[k, element]
// with the operator on the MFMA A side and 16 elements on the N side.  One
// wavefront owns a block of 16 elements.  Element operands stay in
// registers from load to store: a product's result is already the B operand
// of the next product (layout in f32_types.h), and the per-element scalars
// (Jacobian terms, normals, quadrature-point products) are lane-local
// because an MFMA column is one element.  The operator images of a kernel
// are staged in LDS once per workgroup.
//
// Edge coupling: every element stores the traces of its fields at the
// Gauss-Legendre points of its three edges; the consumer averages its own
// trace with the neighbour's (reversed along the shared edge), as the
// reference's face average does (src/calflu.f:41-62).  Both sides of a face
// read the same two stored numbers, so the face flux is single-valued.
//
// Kernels per right-hand side:  linear: lin_kernel (all FP64);  nonlinear:
// bc_kernel (every FP32 volume term: weak gradient q, quadrature-point
// products, mass flux and their divergence) then eg_kernel (the FP64 linear
// core, volume and face terms, with the flux traces' face values lifted
// along; wall projection, RK4 stage, traces of the next stage input).
#include <hip/hip_runtime.h>

#include "mfma.h"

namespace spoc {
namespace f32 {

// Kernel forms.  Each element kernel is compiled in three forms, which the
// autotuner times on the caller's mesh with each workgroup size (a tile code
// is waves per workgroup + 16 x form).  The forms differ in the register
// budget, given as resident waves per SIMD (512 registers per lane shared by
// the resident waves), and for E+G in the split of a block's work:
//  - split: each block of 16 elements gets two waves, one for the u and v
//    equations and one for eta; at p = 5 and 6 either half fits in 128
//    registers where the whole needs about 165, so four waves share a SIMD
//    instead of three (measured: E+G 240 -> 234 us at p = 6 on 199,362
//    triangles, 52 -> 46 us at p = 5 on 49,952);
//  - form 0 is the measured default: B+C at 3 waves (2 at p = 7, whose 36
//    modes spill at 3), E+G split at 4 waves at p = 5 and 6 and whole at 3
//    (p <= 4) or 2 (p = 7), the linear kernel whole at 3;
//  - form 1 trades waves for registers (B+C one wave fewer; E+G the other
//    split: whole at 3 or 2 waves where form 0 splits, split where form 0
//    does not) or, for the linear kernel, splits its block like E+G, which
//    halves the dependent chain of each wave on meshes too small to fill
//    the chip;
//  - form 2 asks for one wave more than form 0 (B+C and the linear kernel)
//    or gives the default E+G split more registers (one wave fewer).
constexpr int NFORMS = TILE_FORMS;
constexpr bool eg_split0(int nc) { return nc == 5 || nc == 6; }
constexpr int bc_waves(int nc, int form) {
  return form == 1 ? (nc >= 7 ? 1 : 2) : form == 2 ? (nc >= 7 ? 3 : 4) : (nc >= 7 ? 2 : 3);
}
constexpr bool eg_split(int nc, int form) { return form == 1 ? !eg_split0(nc) : eg_split0(nc); }
constexpr int eg_waves(int nc, int form) {
  return form == 2 ? (nc >= 7 ? 1 : eg_split0(nc) ? 3 : 2)
                   : (nc >= 7 ? 2 : eg_split(nc, form) ? 4 : 3);
}
constexpr bool lin_split(int form) { return form == 1; }
constexpr int lin_waves(int form) { return form == 2 ? 4 : 3; }

// Stage input of field f: the step's start state c0 at the first stage (and
// for a bare right-hand side, stage < 0), else what the previous stage
// stored in ys[stage - 1].  (A select on the launch-uniform stage, not a
// branch.)
__device__ __forceinline__ const double *stage_in(const F32Args &A, int stage, int f) {
  return stage <= 0 ? A.st.c0[f] : A.st.ys[stage - 1][f];
}

// Linear gravity-wave core (VAR 0: the whole linear right-hand side,
// src/calvar_lin.f; nonlinear variants: the -g grad(eta) terms and the
// -div(H0 u) term (H0 = core_depth), src/calvarnl_fs.f:298-321, 436-471 and the ns/inv
// equivalents): volume terms from the FP64 stage input, face terms from
// the FP64 traces trL.  dx, dy, lf: FP64 images of Dx, Dy and LF.  G[f]
// receives the three tendencies (u, v, eta) in the result layout.  FLUX
// (nonlinear variants): the face values of the FP32 flux traces trS (the
// viscous flux for u and v, the remaining mass flux for eta) join the
// core's own before the lift, which is linear, so one FP64 product lifts
// both, and their gathers issue with the core's.  ROLE: 0 all three
// equations, 1 those of u and v only (G[0], G[1]), 2 that of eta only (G[2]).
// RES: the stage input comes from registers (xin, B fragments, zero past the
// mesh and past NM) instead of memory.
template <int NC, int NG, int NGF, int VAR, bool LF = false, bool FLUX = false, int ROLE = 0,
          bool RES = false>
__device__ __forceinline__ void core64(const F32Args &A, int stage, const double *dx,
                                       const double *dy, const double *lf, const Nbr3 &N, int eb,
                                       doublex4 (&G)[3][S<NC, NG, NGF>::RBM],
                                       const double (*xin)[S<NC, NG, NGF>::KM] = nullptr) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, KE = Sh::KE, RBM = Sh::RBM;
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.0;
  const double g0 = A.fld.g0, hh0 = (valid ? 1.0 : 0.0) * core_depth<VAR>(A.fld, valid ? e : 0);
  const size_t mb = (size_t)e0 * NM;
  // (a select, not an index: a runtime index into the kernel argument would
  // copy the whole argument block to scratch)
  const double *trL = (stage > 0 && (stage & 1)) ? A.st.trL[1] : A.st.trL[0];
  // local Lax-Friedrichs option: F.n_out gains (c/2) |n| (q_own - q_nbr) for
  // q = eta, u, v with c = alpha sqrt(g hh0); lifted as (c/4) |n| (q_nbr -
  // q_own) (a wall point is its own neighbour: no jump).  A compile-time
  // switch (the launch picks the instance from alpha != 0): a runtime branch
  // in each face iteration would keep the loads of successive faces apart.
  constexpr bool lfo = LF;
  const double lfc = lfo ? A.lf_alpha * sqrt(g0 * A.fld.hh0e[valid ? e : 0]) : 0.0;
  const float *trA = (stage > 0 && (stage & 1)) ? A.st.trA[1] : A.st.trA[0];
  if (ROLE != 2) {  // u, v: -g grad(eta), volume and face terms
    double bh[KM], fx[KE], fy[KE];
    {
      if (RES) {
#pragma unroll
        for (int s = 0; s < KM; s++) bh[s] = xin[2][s];
      } else {
        bload<NM>(stage_in(A, stage, 2) + mb, g, col, valid, bh);
      }
#pragma unroll
      for (int s = 0; s < KM; s++) bh[s] *= -g0;
    }
    {
#pragma unroll
      for (int s = 0; s < KE; s++) {
        const int kj = 4 * s + g, kc = kj < NE3 ? kj : NE3 - 1, k = kc / NGF;
        const bool ok = valid && kj < NE3;
        const FacePt fp = face_pt<NGF>(N, e, kc);
        const double fh = -g0 * face_mean<NGF, 2>(trL, fp, 0);
        double px = 0.0, py = 0.0;
        if (lfo) {
          const double nx = gg[5 + k], ny = gg[8 + k], w = 0.25 * lfc * sqrt(nx * nx + ny * ny);
          px = w * (double)(tr_at<NE3, 2>(trA, fp.eb, 0, fp.kb) - tr_at<NE3, 2>(trA, fp.ea, 0, fp.ka));
          py = w * (double)(tr_at<NE3, 2>(trA, fp.eb, 1, fp.kb) - tr_at<NE3, 2>(trA, fp.ea, 1, fp.ka));
        }
        double sx = 0.0, sy = 0.0;
        if (FLUX) {
          sx = 0.5 * gg[11 + k] * (double)face_mean<NGF, 3>(A.st.trS, fp, 0);
          sy = 0.5 * gg[11 + k] * (double)face_mean<NGF, 3>(A.st.trS, fp, 1);
        }
        fx[s] = ok ? 0.5 * fh * gg[11 + k] * gg[5 + k] + px + sx : 0.0;
        fy[s] = ok ? 0.5 * fh * gg[11 + k] * gg[8 + k] + py + sy : 0.0;
      }
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      doublex4 Gx, Gy;
      {
        double adx[KM], ady[KM];
        afrag64<NM>(dx, rb, adx);
        afrag64<NM>(dy, rb, ady);
        doublex4 Ax = chain64(adx, bh, zero64()), Ay = chain64(ady, bh, zero64());
#pragma unroll
        for (int r = 0; r < 4; r++) {
          Gx[r] = -by * Ax[r] + ay * Ay[r];
          Gy[r] = bx * Ax[r] - ax * Ay[r];
        }
      }
      {
        double al[KE];
        afrag64<NE3>(lf, rb, al);
        Gx = chain64(al, fx, Gx);
        Gy = chain64(al, fy, Gy);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        G[0][rb][r] = di * Gx[r];
        G[1][rb][r] = di * Gy[r];
      }
    }
  }
  phase_fence();
  if (ROLE != 1) {  // eta: -div(H0 u); the nonlinear variants project the flux on walls
    double bu[KM], bv[KM], fh[KE];
    {
      if (RES) {
#pragma unroll
        for (int s = 0; s < KM; s++) {
          bu[s] = xin[0][s];
          bv[s] = xin[1][s];
        }
      } else {
        bload<NM>(stage_in(A, stage, 0) + mb, g, col, valid, bu);
        bload<NM>(stage_in(A, stage, 1) + mb, g, col, valid, bv);
      }
#pragma unroll
      for (int s = 0; s < KM; s++) {
        bu[s] *= hh0;
        bv[s] *= hh0;
      }
      if (VAR != 0) {
        const bool wall = A.mesh.bdl[valid ? e : 0] != 0 && valid;
        if (__ballot(wall)) wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gg[14], gg[15], bu, bv);
      }
#pragma unroll
      for (int s = 0; s < KM; s++) {
        double u = bu[s], v = bv[s];
        bu[s] = by * u - bx * v;
        bv[s] = -ay * u + ax * v;
      }
    }
    {
#pragma unroll
      for (int s = 0; s < KE; s++) {
        const int kj = 4 * s + g, kc = kj < NE3 ? kj : NE3 - 1, k = kc / NGF;
        const bool ok = valid && kj < NE3;
        const FacePt fp = face_pt<NGF>(N, e, kc);
        double v = -0.5 * gg[11 + k] * face_mean<NGF, 2>(trL, fp, 1);
        if (FLUX) v -= 0.5 * gg[11 + k] * (double)face_mean<NGF, 3>(A.st.trS, fp, 2);
        if (lfo) {
          const double nx = gg[5 + k], ny = gg[8 + k];
          v += 0.25 * lfc * sqrt(nx * nx + ny * ny) *
               (tr_at<NE3, 2>(trL, fp.eb, 0, fp.kb) - tr_at<NE3, 2>(trL, fp.ea, 0, fp.ka));
        }
        fh[s] = ok ? v : 0.0;
      }
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      doublex4 Gh = zero64();
      {
        double adx[KM], ady[KM];
        afrag64<NM>(dx, rb, adx);
        afrag64<NM>(dy, rb, ady);
        Gh = chain64(adx, bu, Gh);
        Gh = chain64(ady, bv, Gh);
      }
      {
        double al[KE];
        afrag64<NE3>(lf, rb, al);
        Gh = chain64(al, fh, Gh);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) G[2][rb][r] = di * Gh[r];
    }
  }
}

// Traces of a new stage input (B fragments, FP64), written to trL[w]:
// T_eta (field 0) ...
template <int NC, int NG, int NGF>
__device__ __forceinline__ void traces_h(const F32Args &A, const double *O, int eb, int w,
                                         const double (&xh)[S<NC, NG, NGF>::KM]) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, RBE = Sh::RBE;
  const int l = lane(), g = l >> 4, col = l & 15;
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double at[KM];
    afrag64<NM>(O + Sh::P64.tr, rb, at);
    doublex4 Th = chain64(at, xh, zero64());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int kj = drow(rb, r, g);
      if (kj < NE3) tr_put<NE3, 2>(w ? A.st.trL[1] : A.st.trL[0], eb, col, 0, kj, Th[r]);
    }
  }
}

// ... and S = hh0 (T_u, T_v).n_face (field 1), with the mass flux of wall
// elements projected first as the core projects it; trA[w] <- (T_u, T_v) in
// FP32 for the nonlinear gradient and the Lax-Friedrichs option.
template <int NC, int NG, int NGF, int VAR>
__device__ __forceinline__ void traces_uv(const F32Args &A, const double *O, int eb, int w,
                                          const double (&xu)[S<NC, NG, NGF>::KM],
                                          const double (&xv)[S<NC, NG, NGF>::KM]) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, RBE = Sh::RBE;
  const int l = lane(), g = l >> 4, col = l & 15, e = eb * 16 + col;
  const bool valid = e < A.mesh.ne;
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double hh0 = (valid ? 1.0 : 0.0) * core_depth<VAR>(A.fld, valid ? e : 0);
  double *trL = w ? A.st.trL[1] : A.st.trL[0];
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double at[KM];
    afrag64<NM>(O + Sh::P64.tr, rb, at);
    doublex4 Tu = chain64(at, xu, zero64()), Tv = chain64(at, xv, zero64());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int kj = drow(rb, r, g);
      if (kj < NE3) {
        int k = kj / NGF;
        if (VAR != 0 || A.lf_alpha != 0.0) {
          float *trA = w ? A.st.trA[1] : A.st.trA[0];
          tr_put<NE3, 2>(trA, eb, col, 0, kj, (float)Tu[r]);
          tr_put<NE3, 2>(trA, eb, col, 1, kj, (float)Tv[r]);
        }
        tr_put<NE3, 2>(trL, eb, col, 1, kj, hh0 * (Tu[r] * gg[5 + k] + Tv[r] * gg[8 + k]));
      }
    }
  }
  if (VAR == 0) return;
  const bool wall = A.mesh.bdl[valid ? e : 0] != 0 && valid;
  if (!__ballot(wall)) return;
  // wall elements: add the traces of (projected flux - flux) / hh0 to the
  // values this lane stored above
  double du[KM], dv[KM];
#pragma unroll
  for (int s = 0; s < KM; s++) {
    du[s] = xu[s];
    dv[s] = xv[s];
  }
  wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gg[14], gg[15], du, dv);
#pragma unroll
  for (int s = 0; s < KM; s++) {
    du[s] -= xu[s];
    dv[s] -= xv[s];
  }
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double at[KM];
    afrag64<NM>(O + Sh::P64.tr, rb, at);
    doublex4 Cu = chain64(at, du, zero64()), Cv = chain64(at, dv, zero64());
    if (wall) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int kj = drow(rb, r, g);
        if (kj < NE3) {
          int k = kj / NGF;
          double *t = trL + ((size_t)eb * (2 * NE3) + NE3 + kj) * 16 + col;
          *t += hh0 * (Cu[r] * gg[5 + k] + Cv[r] * gg[8 + k]);
        }
      }
    }
  }
}

// RK4 stage update (src/shwwnl.f:242-284) of field f, FP64: x holds the
// stage's tendency (B fragments) and leaves as the next stage input.  Stages
// 0-2 store their stage results Y2 = c0 + dt/2 k1, Y3 = c0 + dt/2 k2 and
// Y4 = c0 + dt k3 in ys[0..2] and keep no running sum: the last stage forms
// c0 + dt/6 k1 + dt/3 k2 + dt/3 k3 as c0 + ((Y2 - c0) + 2 (Y3 - c0) +
// (Y4 - c0)) / 3 and writes the step's result to c0, which is also the next
// step's first stage input (stage_in).  Per step that is 7 field reads and 4
// writes (c0 at every stage, ys[0..2] at the last; ys[s] at stage s, c0 at
// the last), of which two reads are the stage inputs of stages 0 and 3, so
// only those of stages 1 and 2 come on top: 13 field passes, against 16 for
// a running sum (6 reads and 7 writes, and the stage inputs of stages 1-3
// on top).  Each difference
// Y - c0 is exact up to the rounding of Y itself, so the step's result
// differs from the reference's order of additions (c0 + dt/6 (((k1 + 2 k2)
// + 2 k3) + k4)) by a few FP64 roundings of c0.  stage < 0: the tendency
// goes to k and x leaves as 0.  Elements past the mesh (the padding of the
// last block) take a zero tendency, so their stored values stay 0; rows past
// NM belong to the next block and are not stored.  The stage is uniform over
// the launch, so its branches are scalar and every load of a field issues
// at once.
template <int NM, int KM>
__device__ __forceinline__ void update(const F32Args &A, int stage, int f, int eb,
                                       double (&x)[KM]) {
  const int g = lane() >> 4, col = lane() & 15;
  const bool valid = eb * 16 + col < A.mesh.ne;
  const size_t mb = (size_t)eb * 16 * NM;
  const double dt = A.dt;
  int ix[KM];
#pragma unroll
  for (int s = 0; s < KM; s++) {
    const int m = 4 * s + g;
    ix[s] = (m < NM ? m : NM - 1) * 16 + col;
    x[s] = valid ? x[s] : 0.0;
  }
  // the stored results are read by a later launch only (non-temporal: they
  // do not displace this launch's neighbour gathers from L2)
  auto store = [&](double *p, const double(&v)[KM]) {
#pragma unroll
    for (int s = 0; s < KM; s++)
      if (4 * s + g < NM) __builtin_nontemporal_store(v[s], p + mb + ix[s]);
  };
  if (stage < 0) {
    store(A.st.k[f], x);
#pragma unroll
    for (int s = 0; s < KM; s++) x[s] = 0.0;
    return;
  }
  double c[KM], y[KM];
#pragma unroll
  for (int s = 0; s < KM; s++) c[s] = A.st.c0[f][mb + ix[s]];
  if (stage == 3) {
    double p[KM], q[KM], r[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) p[s] = __builtin_nontemporal_load(A.st.ys[0][f] + mb + ix[s]);
#pragma unroll
    for (int s = 0; s < KM; s++) q[s] = __builtin_nontemporal_load(A.st.ys[1][f] + mb + ix[s]);
#pragma unroll
    for (int s = 0; s < KM; s++) r[s] = __builtin_nontemporal_load(A.st.ys[2][f] + mb + ix[s]);
    const double third = 1.0 / 3.0, w = dt / 6.0;
#pragma unroll
    for (int s = 0; s < KM; s++) {
      const double d = ((p[s] - c[s]) + 2.0 * (q[s] - c[s])) + (r[s] - c[s]);
      y[s] = c[s] + (third * d + w * x[s]);
    }
    store(A.st.c0[f], y);
  } else {
    const double h = stage == 2 ? dt : 0.5 * dt;
#pragma unroll
    for (int s = 0; s < KM; s++) y[s] = c[s] + h * x[s];
    store(A.st.ys[stage][f], y);
  }
#pragma unroll
  for (int s = 0; s < KM; s++) x[s] = y[s];
}

// Epilogue of the element kernels: wall projection of the (u, v) tendency
// (free slip for VAR 0, 1, 3 and no slip for VAR 2, src/calvar_lin.f and
// src/calvarnl_*.f at the end of cal_var), then either the tendencies
// (stage < 0) or the RK4 stage and the traces of the next stage input.
template <int NC, int NG, int NGF, int VAR, int ROLE = 0>
__device__ __forceinline__ void finish64(const F32Args &A, int stage, const double *O, int eb,
                                         doublex4 (&G)[3][S<NC, NG, NGF>::RBM]) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, KM = Sh::KM;
  const int col = lane() & 15, e = eb * 16 + col;
  const bool valid = e < A.mesh.ne;
  const int w = (stage + 1) & 1;
  if (ROLE != 1) {
    double xh[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) xh[s] = G[2][s >> 2][s & 3];
    update<NM>(A, stage, 2, eb, xh);
    if (stage >= 0) traces_h<NC, NG, NGF>(A, O, eb, w, xh);
  }
  if (ROLE == 2) return;
  phase_fence();
  double xu[KM], xv[KM];
#pragma unroll
  for (int s = 0; s < KM; s++) {
    xu[s] = G[0][s >> 2][s & 3];
    xv[s] = G[1][s >> 2][s & 3];
  }
  const bool wall = A.mesh.bdl[valid ? e : 0] != 0 && valid;
  if (__ballot(wall)) {
    const double *gw = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
    wall_pair64<NM, VAR == 2>(A.ops.gl64, wall, gw[14], gw[15], xu, xv);
  }
  update<NM>(A, stage, 0, eb, xu);
  update<NM>(A, stage, 1, eb, xv);
  if (stage >= 0) traces_uv<NC, NG, NGF, VAR>(A, O, eb, w, xu, xv);
}

// ---------------------------------------------------------------------------
// Traces of the first stage input c0 (before the first stage).
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void prep_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, KM = Sh::KM;
  constexpr Pack64 P = Sh::P64;
  __shared__ floatx4 lds[P.dx / 2];  // LF, TR
  const int b0 = first_block();
  if (b0 >= nblk) return;
  stage_lds(lds, reinterpret_cast<const floatx4 *>(A.ops.pack64), P.dx / 2);
  __syncthreads();
  const int eb = wave_block(b0);
  if (eb >= nblk) return;
  const double *O = reinterpret_cast<const double *>(lds);
  const int l = lane(), g = l >> 4, col = l & 15, e = eb * 16 + col;
  const bool valid = e < A.mesh.ne;
  const size_t mb = (size_t)eb * 16 * NM;
  double x[3][KM];
#pragma unroll
  for (int f = 0; f < 3; f++) bload<NM>(A.st.c0[f] + mb, g, col, valid, x[f]);
  traces_h<NC, NG, NGF>(A, O, eb, 0, x[2]);
  traces_uv<NC, NG, NGF, VAR>(A, O, eb, 0, x[0], x[1]);
}

// ---------------------------------------------------------------------------
// Linear right-hand side (src/calvar_lin.f:7-287) + RK4 stage, all FP64.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int WG, bool LF, int FORM>
__global__ __launch_bounds__(64 * WG) __attribute__((amdgpu_waves_per_eu(lin_waves(FORM)))) void lin_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr Pack64 P = Sh::P64;
  constexpr bool SPLIT = lin_split(FORM);
  __shared__ floatx4 lds[P.total / 2];
  const int b0 = first_block<SPLIT ? WG / 2 : WG>();
  if (b0 >= nblk) return;
  stage_lds<WG>(lds, reinterpret_cast<const floatx4 *>(A.ops.pack64), P.total / 2);
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int eb = b0 + (SPLIT ? wv >> 1 : wv);
  if (eb >= nblk) return;
  const double *O = reinterpret_cast<const double *>(lds);
  const int e = eb * 16 + (lane() & 15);
  const Nbr3 N = nbr3(A.mesh, e < A.mesh.ne ? e : 0);
  doublex4 G[3][Sh::RBM];
  if (!SPLIT) {
    core64<NC, NG, NGF, 0, LF>(A, A.stage, O + P.dx, O + P.dy, O + P.lf, N, eb, G);
    phase_fence();
    finish64<NC, NG, NGF, 0>(A, A.stage, O, eb, G);
  } else if (wv & 1) {  // the eta equation
    core64<NC, NG, NGF, 0, LF, false, 2>(A, A.stage, O + P.dx, O + P.dy, O + P.lf, N, eb, G);
    phase_fence();
    finish64<NC, NG, NGF, 0, 2>(A, A.stage, O, eb, G);
  } else {  // the u and v equations
    core64<NC, NG, NGF, 0, LF, false, 1>(A, A.stage, O + P.dx, O + P.dy, O + P.lf, N, eb, G);
    phase_fence();
    finish64<NC, NG, NGF, 0, 1>(A, A.stage, O, eb, G);
  }
}

// ---------------------------------------------------------------------------
// Persistent linear RK4 for small meshes: A.steps RK4 steps of the linear
// right-hand side in one launch.  Each wave keeps its element block for the
// whole launch, and with it the block's state in registers: the step's
// start c, the stage input x and the running sum d of the stage increments
// (the RK4 stage update of `update`, with Y2 - c, 2 (Y3 - c) and Y4 - c
// added as each stage forms them, in the order the last stage adds them
// there, so the results are the same bit for bit).  Only the traces of each
// stage input go through memory, since a stage reads its neighbours'; before
// a stage a wave waits for the blocks around its own to finish the previous
// one (stage_done / wait_nbrs), and c0 is written once, at the end.  The
// operator images are staged in LDS once.  Meant for meshes of at most
// persist_blocks(NC) blocks (where it is faster than the per-stage
// launches), far fewer waves than the chip holds at one wave per SIMD, so
// every wave is resident and a wait for another block can end; the register
// budget is that of one wave per SIMD.
// ---------------------------------------------------------------------------
constexpr unsigned kSpinLimit = 1u << 20;

// RK4 stage update of field f with the block's state in registers (see
// `update` for the arithmetic): x holds the stage's tendency and leaves as
// the stage result; c (the step's start) and d (the running sum) advance;
// the last stage of the launch stores the step's result to c0.
template <int NM, int KM>
__device__ __forceinline__ void update_res(const F32Args &A, int stage, int f, int eb,
                                           double (&x)[KM], double (&c)[KM], double (&d)[KM],
                                           bool last) {
  const int g = lane() >> 4, col = lane() & 15;
  const bool valid = eb * 16 + col < A.mesh.ne;
  const double dt = A.dt;
#pragma unroll
  for (int s = 0; s < KM; s++) x[s] = valid ? x[s] : 0.0;
  double y[KM];
  if (stage == 3) {
    const double third = 1.0 / 3.0, w = dt / 6.0;
#pragma unroll
    for (int s = 0; s < KM; s++) {
      y[s] = c[s] + (third * d[s] + w * x[s]);
      c[s] = y[s];
    }
    if (last) {
      const size_t mb = (size_t)eb * 16 * NM;
#pragma unroll
      for (int s = 0; s < KM; s++)
        if (4 * s + g < NM) A.st.c0[f][mb + (4 * s + g) * 16 + col] = y[s];
    }
  } else {
    const double h = stage == 2 ? dt : 0.5 * dt;
#pragma unroll
    for (int s = 0; s < KM; s++) {
      y[s] = c[s] + h * x[s];
      const double inc = y[s] - c[s];
      d[s] = stage == 0 ? inc : stage == 1 ? d[s] + 2.0 * inc : d[s] + inc;
    }
  }
#pragma unroll
  for (int s = 0; s < KM; s++) x[s] = y[s];
}

// traces_uv for the linear variant with the central flux, its per-lane data
// from registers: hh0 (the core depth) and the face normal of each trace row
// this lane stores (tnx, tny: row drow(rb, r, g) at 4 rb + r).
template <int NC, int NG, int NGF>
__device__ __forceinline__ void traces_uv_res(const F32Args &A, const double *O, int eb, int w,
                                              const double (&xu)[S<NC, NG, NGF>::KM],
                                              const double (&xv)[S<NC, NG, NGF>::KM], double hh0,
                                              const double (&tnx)[4 * S<NC, NG, NGF>::RBE],
                                              const double (&tny)[4 * S<NC, NG, NGF>::RBE]) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, RBE = Sh::RBE;
  const int l = lane(), g = l >> 4, col = l & 15;
  double *trL = w ? A.st.trL[1] : A.st.trL[0];
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double at[KM];
    afrag64<NM>(O + Sh::P64.tr, rb, at);
    doublex4 Tu = chain64(at, xu, zero64()), Tv = chain64(at, xv, zero64());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int kj = drow(rb, r, g);
      if (kj < NE3)
        tr_put<NE3, 2>(trL, eb, col, 1, kj,
                       hh0 * (Tu[r] * tnx[4 * rb + r] + Tv[r] * tny[4 * rb + r]));
    }
  }
}

// Stage hand-off between neighbouring blocks (sync: [1] the abort word,
// [PERSIST_LINE (1 + b)] the number of stages block b has finished, each
// count on a cache line of its own).  After a stage a wave makes its stores
// visible to the device (release fence) and sets its block's count; before
// the next it waits until every block that holds a face neighbour of its
// elements has finished as many stages: their traces of the new stage input
// are then written, and they are done reading the traces this stage
// overwrites (the traces are double-buffered, alternating per stage, and the
// neighbour relation is symmetric), so no grid-wide barrier is needed.  Then
// it takes the device's stores (acquire fence).  Every block finishes stage
// k - 1 waiting only for counts of k - 1, so the waits end; they are also
// bounded: a wave that gives up (or sees another's abort) raises the abort
// word and returns false, so every wave leaves the launch and the host
// reports it.
__device__ __forceinline__ void stage_done(unsigned *sync, int eb, unsigned k) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane() == 0)
    __hip_atomic_store(sync + PERSIST_LINE * (1 + eb), k, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool wait_nbrs(unsigned *sync, const Nbr3 &N, int eb, unsigned k) {
  const unsigned *fl = sync + PERSIST_LINE;
  const int b0 = PERSIST_LINE * (N.nb0 >= 0 ? N.nb0 >> 4 : eb),
            b1 = PERSIST_LINE * (N.nb1 >= 0 ? N.nb1 >> 4 : eb),
            b2 = PERSIST_LINE * (N.nb2 >= 0 ? N.nb2 >> 4 : eb);
  for (unsigned it = 0;; it++) {
    const unsigned f0 = __hip_atomic_load(fl + b0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                   f1 = __hip_atomic_load(fl + b1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                   f2 = __hip_atomic_load(fl + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (__ballot(f0 < k || f1 < k || f2 < k) == 0) break;
    if (it >= kSpinLimit ||
        __hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
      if (lane() == 0) __hip_atomic_store(sync + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

template <int NC, int NG, int NGF>  // (one per shape's object file)
__global__ void sync_reset_kernel(unsigned *sync, int n, unsigned abort0) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) sync[i] = i == 1 ? abort0 : 0u;
}

template <int NC, int NG, int NGF, int WG>
__global__ __launch_bounds__(64 * WG) __attribute__((amdgpu_waves_per_eu(1, 1))) void lin_persistent_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, RBE = Sh::RBE;
  constexpr Pack64 P = Sh::P64;
  constexpr int NGL = img64_size(NM, NM) / 2;  // the wall projection's image
  __shared__ floatx4 lds[P.total / 2];
  __shared__ floatx4 lgl[NGL];
  const int b0 = first_block<WG>();
  if (b0 >= nblk) return;
  stage_lds<WG>(lds, reinterpret_cast<const floatx4 *>(A.ops.pack64), P.total / 2);
  stage_lds<WG>(lgl, reinterpret_cast<const floatx4 *>(A.ops.gl64), NGL);
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int eb = b0 + wv;
  if (eb >= nblk) return;  // (no block: nobody waits for it)
  const double *O = reinterpret_cast<const double *>(lds);
  const int g = lane() >> 4, col = lane() & 15, e = eb * 16 + col;
  const bool valid = e < A.mesh.ne;
  const Nbr3 N = nbr3(A.mesh, valid ? e : 0);
  // the block's state: c = the step's start (rows past NM repeat row NM - 1,
  // as `update` reads them), x = the stage input (zero past the mesh and
  // past NM, as bload gives it), d = the running sum of the increments
  double c[3][KM], x[3][KM], d[3][KM];
  // and what the stage's epilogue reads of the mesh, which does not change:
  // the wall flag and normal, the core depth and the face normal of each
  // trace row the lane stores
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const bool wall = A.mesh.bdl[valid ? e : 0] != 0 && valid;
  const double wnx = gg[14], wny = gg[15];
  const double hh0 = (valid ? 1.0 : 0.0) * core_depth<0>(A.fld, valid ? e : 0);
  double tnx[4 * RBE], tny[4 * RBE];
#pragma unroll
  for (int rb = 0; rb < RBE; rb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int kj = drow(rb, r, g), k = (kj < NE3 ? kj : NE3 - 1) / NGF;
      tnx[4 * rb + r] = gg[5 + k];
      tny[4 * rb + r] = gg[8 + k];
    }
  {
    const size_t mb = (size_t)eb * 16 * NM;
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
      for (int s = 0; s < KM; s++) {
        const int m = 4 * s + g;
        c[f][s] = A.st.c0[f][mb + (m < NM ? m : NM - 1) * 16 + col];
        x[f][s] = valid && m < NM ? c[f][s] : 0.0;
        d[f][s] = 0.0;
      }
  }
  unsigned done = 0;  // stages finished
  for (int n = 0; n < A.steps; n++)
    for (int st = 0; st < 4; st++) {
      const bool last = n + 1 == A.steps && st == 3;
      if (done > 0 && !wait_nbrs(A.sync, N, eb, done)) return;
      {
        doublex4 G[3][Sh::RBM];
        core64<NC, NG, NGF, 0, false, false, 0, true>(A, st, O + P.dx, O + P.dy, O + P.lf, N,
                                                      eb, G, x);
        phase_fence();
        // finish64's epilogue (VAR 0) with the register update
        const int w = (st + 1) & 1;
#pragma unroll
        for (int s = 0; s < KM; s++) x[2][s] = G[2][s >> 2][s & 3];
        update_res<NM>(A, st, 2, eb, x[2], c[2], d[2], last);
        traces_h<NC, NG, NGF>(A, O, eb, w, x[2]);
        phase_fence();
#pragma unroll
        for (int s = 0; s < KM; s++) {
          x[0][s] = G[0][s >> 2][s & 3];
          x[1][s] = G[1][s >> 2][s & 3];
        }
        if (__ballot(wall))
          wall_pair64<NM, false>(reinterpret_cast<const double *>(lgl), wall, wnx, wny, x[0],
                                 x[1]);
        update_res<NM>(A, st, 0, eb, x[0], c[0], d[0], last);
        update_res<NM>(A, st, 1, eb, x[1], c[1], d[1], last);
        traces_uv_res<NC, NG, NGF>(A, O, eb, w, x[0], x[1], hh0, tnx, tny);
        // the next stage input as bload would give it
#pragma unroll
        for (int f = 0; f < 3; f++)
#pragma unroll
          for (int s = 0; s < KM; s++) x[f][s] = valid && 4 * s + g < NM ? x[f][s] : 0.0;
      }
      if (last) break;
      stage_done(A.sync, eb, ++done);
    }
}

// ---------------------------------------------------------------------------
// Nonlinear stages B + C (src/calvarnl_fs.f:115-332 and the ns/inv
// variants) and the volume half of E (src/calvarnl_fs.f:382-420): the
// velocity gradient q, the viscous flux F = mu0 q, the quadrature-point
// sources (advection, Coriolis, wind stress, bottom friction) and the mass
// flux u (h + hb - H0) in FP32 (H0 = core_depth), and their volume divergence in FP32.
// Stores that volume part (gvol, FP32, as computed) and the flux traces
// (trS, FP32); the linear core and every face term are E+G's.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR, int WG, int FORM>
__global__ __launch_bounds__(64 * WG) __attribute__((amdgpu_waves_per_eu(bc_waves(NC, FORM)))) void bc_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NE3 = Sh::NE3, KM = Sh::KM, KE = Sh::KE, KG = Sh::KG,
                RBM = Sh::RBM, RBE = Sh::RBE, RBG = Sh::RBG;
  constexpr Pack32 P = Sh::P32;
  constexpr int N32 = P.total / 4;  // 16-byte words
  __shared__ floatx4 lds[N32];
  const int b0 = first_block<WG>();
  if (b0 >= nblk) return;
  stage_lds<WG>(lds, reinterpret_cast<const floatx4 *>(A.ops.pack32), N32);
  __syncthreads();
  const int eb = wave_block(b0);
  if (eb >= nblk) return;
  const float *O = reinterpret_cast<const float *>(lds);
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const size_t mb = (size_t)e0 * NM;
  // loads in the order of their use by the gathers below: the neighbour
  // lists (the gathers' addresses need them; the wait counter is in
  // order), then the stage input, then the geometry
  const Nbr3 N = nbr3(A.mesh, valid ? e : 0);
  float bu[KM], bv[KM], bh[KM];
  bload<NM>(stage_in(A, A.stage, 0) + mb, g, col, valid, bu);
  bload<NM>(stage_in(A, A.stage, 1) + mb, g, col, valid, bv);
  bload<NM>(stage_in(A, A.stage, 2) + mb, g, col, valid, bh);
  const float mu0 = A.fld.mu0, mb0 = A.fld.mb0;
  const float *gg = A.mesh.geo + (size_t)(valid ? e : 0) * GEO;
  const float ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.f;
  const bool wall = A.mesh.bdl[valid ? e : 0] != 0 && valid;
  const float g14 = gg[14], g15 = gg[15];
  const float nwx = wall ? g14 : 0.f, nwy = wall ? g15 : 0.f;

  // B: q = weak gradient of (u, v) (src/calvarnl_fs.f:129-158); the face
  // terms use the averaged velocity traces times n/2
  floatx4 Q[4][RBM];  // qxx, qxy, qyx, qyy
  {
    float fb[4][KE];
    const float *trA = (A.stage > 0 && (A.stage & 1)) ? A.st.trA[1] : A.st.trA[0];
#pragma unroll
    for (int s = 0; s < KE; s++) {
      const int kj = 4 * s + g, kc = kj < NE3 ? kj : NE3 - 1, k = kc / NGF;
      const bool ok = valid && kj < NE3;
      const float sg = 0.5f * gg[11 + k], nx = sg * gg[5 + k], ny = sg * gg[8 + k];
      const FacePt fp = face_pt<NGF>(N, e, kc);
      const float fu = face_mean<NGF, 2>(trA, fp, 0);
      const float fv = face_mean<NGF, 2>(trA, fp, 1);
      fb[0][s] = ok ? fu * nx : 0.f;
      fb[1][s] = ok ? fu * ny : 0.f;
      fb[2][s] = ok ? fv * nx : 0.f;
      fb[3][s] = ok ? fv * ny : 0.f;
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      float adx[KM], ady[KM], al[KE];
      afrag32<NM>(O + P.dx, rb, adx);
      afrag32<NM>(O + P.dy, rb, ady);
      afrag32<NE3>(O + P.lf, rb, al);
      floatx4 Au = chain32(adx, bu, zero32()), Bu = chain32(ady, bu, zero32());
      floatx4 Av = chain32(adx, bv, zero32()), Bv = chain32(ady, bv, zero32());
      floatx4 L0 = chain32(al, fb[0], zero32()), L1 = chain32(al, fb[1], zero32());
      floatx4 L2 = chain32(al, fb[2], zero32()), L3 = chain32(al, fb[3], zero32());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        Q[0][rb][r] = di * (-by * Au[r] + ay * Bu[r] + L0[r]);
        Q[1][rb][r] = di * (bx * Au[r] - ax * Bu[r] + L1[r]);
        Q[2][rb][r] = di * (-by * Av[r] + ay * Bv[r] + L2[r]);
        Q[3][rb][r] = di * (bx * Av[r] - ax * Bv[r] + L3[r]);
      }
    }
  }
  // (the wave's vote on walls only here, at its first use: it waits for the
  // wall flags, and with them for every load issued before them)
  const bool anyw = __ballot(wall) != 0;
  if (VAR == 1 && anyw) {
    // free-slip wall projection of the tensor (src/calvarnl_fs.f:165-204):
    // only the t-n component goes through the reduced basis
    float vb[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) {
      const int rb = s >> 2, r = s & 3;
      vb[s] = (Q[0][rb][r] * nwy - Q[2][rb][r] * nwx) * nwx +
              (Q[1][rb][r] * nwy - Q[3][rb][r] * nwx) * nwy;
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      float ag[KM];
      afrag32<NM>(A.ops.gl32, rb, ag);
      floatx4 d = chain32(ag, vb, zero32());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        Q[0][rb][r] += nwx * nwy * d[r];
        Q[2][rb][r] -= nwx * nwx * d[r];
        Q[1][rb][r] += nwy * nwy * d[r];
        Q[3][rb][r] -= nwx * nwy * d[r];
      }
    }
  }
  float q0[KM], q1[KM], q2[KM], q3[KM];
  d2b(Q[0], q0);
  d2b(Q[1], q1);
  d2b(Q[2], q2);
  d2b(Q[3], q3);
  // D: traces of the viscous flux F = mu0 q (the -g eta part belongs to the
  // FP64 core), normal components
  {
  const Nrm3 nr = normals(gg);
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    float at[KM];
    afrag32<NM>(O + P.tr, rb, at);
    floatx4 T0 = chain32(at, q0, zero32()), T1 = chain32(at, q1, zero32());
    floatx4 T2 = chain32(at, q2, zero32()), T3 = chain32(at, q3, zero32());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int kj = drow(rb, r, g);
      if (kj < NE3) {
        int k = kj / NGF;
        float nx = pick3(nr.x, k), ny = pick3(nr.y, k);
        tr_put<NE3, 3>(A.st.trS, eb, col, 0, kj, mu0 * (T0[r] * nx + T1[r] * ny));
        tr_put<NE3, 3>(A.st.trS, eb, col, 1, kj, mu0 * (T2[r] * nx + T3[r] * ny));
      }
    }
  }
  }
  // E, volume part for (u, v): div F with the Jacobian folded into the
  // operands (-by Fxx + bx Fxy with Dx, ay Fxx - ax Fxy with Dy; v likewise);
  // it starts the accumulators the projection below adds the sources to
  floatx4 Pr[4][RBM];
  {
    float o0[KM], o1[KM], o2[KM], o3[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) {
      o0[s] = mu0 * (-by * q0[s] + bx * q1[s]);
      o1[s] = mu0 * (ay * q0[s] - ax * q1[s]);
      o2[s] = mu0 * (-by * q2[s] + bx * q3[s]);
      o3[s] = mu0 * (ay * q2[s] - ax * q3[s]);
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      float adx[KM], ady[KM];
      afrag32<NM>(O + P.dx, rb, adx);
      afrag32<NM>(O + P.dy, rb, ady);
      floatx4 Vu = chain32(ady, o1, chain32(adx, o0, zero32()));
      floatx4 Vv = chain32(ady, o3, chain32(adx, o2, zero32()));
#pragma unroll
      for (int r = 0; r < 4; r++) {
        Pr[0][rb][r] = di * Vu[r];
        Pr[1][rb][r] = di * Vv[r];
      }
      Pr[2][rb] = zero32();
      Pr[3][rb] = zero32();
    }
  }
  phase_fence();
  // C: values at the quadrature points, products, projection
  // (src/calvarnl_fs.f:245-296, src/nonlin.f:35-133).  A product at point
  // q = 16 qb + 4 r + g is already the B operand of k-step 4 qb + r of the
  // projection.
  const float *fc3 = A.fld.fc3 + (size_t)(valid ? e : 0) * 3;
  const float fc0 = fc3[0], fc1 = fc3[1], fc2 = fc3[2];
  // the block's quadrature-point values: one 1 KiB wave load per k-step,
  // each k-block's issued before the products that need them
  const floatx4 *qf = reinterpret_cast<const floatx4 *>(A.fld.qf) + (size_t)eb * KG * 64 + l;
  // (the last row block of points holds only point NG - 1, at k-step KG - 1)
  constexpr bool TAIL1 = NG % 16 == 1;
  static_assert(!TAIL1 || KG == 4 * (RBG - 1) + 1, "tail point layout");
#pragma unroll
  for (int qb = 0; qb < RBG; qb++) {
    floatx4 F[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (4 * qb + r < KG) F[r] = qf[(4 * qb + r) * 64];
    floatx4 Vu, Vv, Vh, Vxx, Vxy, Vyx, Vyy;
    if (TAIL1 && qb == RBG - 1) {
      // one point (the last) in this row block: lane-local dot products over
      // the lane's modes, summed over the four lane groups, in place of
      // 7 x KM MFMAs with one useful row
      float t[7] = {};
#pragma unroll
      for (int s = 0; s < KM; s++) {
        const float a = O[P.tail + 4 * s + g];
        t[0] += a * bu[s];
        t[1] += a * bv[s];
        t[2] += a * bh[s];
        t[3] += a * q0[s];
        t[4] += a * q1[s];
        t[5] += a * q2[s];
        t[6] += a * q3[s];
      }
#pragma unroll
      for (int i = 0; i < 7; i++) {
        t[i] += __shfl_xor(t[i], 16);
        t[i] += __shfl_xor(t[i], 32);
      }
      Vu = floatx4{t[0], 0.f, 0.f, 0.f};
      Vv = floatx4{t[1], 0.f, 0.f, 0.f};
      Vh = floatx4{t[2], 0.f, 0.f, 0.f};
      Vxx = floatx4{t[3], 0.f, 0.f, 0.f};
      Vxy = floatx4{t[4], 0.f, 0.f, 0.f};
      Vyx = floatx4{t[5], 0.f, 0.f, 0.f};
      Vyy = floatx4{t[6], 0.f, 0.f, 0.f};
    } else {
      float ap[KM];
      afrag32<NM>(O + P.psi, qb, ap);
      Vu = chain32(ap, bu, zero32());
      Vv = chain32(ap, bv, zero32());
      Vh = chain32(ap, bh, zero32());
      Vxx = chain32(ap, q0, zero32());
      Vxy = chain32(ap, q1, zero32());
      Vyx = chain32(ap, q2, zero32());
      Vyy = chain32(ap, q3, zero32());
    }
    float pg[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (4 * qb + r >= KG) {
        pg[0][r] = pg[1][r] = pg[2][r] = pg[3][r] = 0.f;
        continue;
      }
      // (rows past NG and elements past the mesh give 0, by selects)
      const int q = drow(qb, r, g), qc = q < NG ? q : NG - 1;
      const bool ok = valid && q < NG;
      const float f = fc0 + fc1 * O[P.yq + qc] + fc2 * O[P.xq + qc];
      // (the hardware reciprocal, within 1 ulp, in place of the correctly
      // rounded division's eight-instruction sequence)
      const float hk = Vh[r] + F[r][0], rhk = __builtin_amdgcn_rcpf(hk);
      const float u = Vu[r], v = Vv[r];
      const float p0 = -(u * Vxx[r] + v * (Vxy[r] - f)) + (F[r][1] - mb0 * u) * rhk;
      const float p1 = -(u * (Vyx[r] + f) + v * Vyy[r]) + (F[r][2] - mb0 * v) * rhk;
      // the flux H0 u is the FP64 core's (H0 = core_depth); only u (h + hb - H0) here
      const float hf = Vh[r] + F[r][3];
      pg[0][r] = ok ? p0 : 0.f;
      pg[1][r] = ok ? p1 : 0.f;
      pg[2][r] = ok ? u * hf : 0.f;
      pg[3][r] = ok ? v * hf : 0.f;
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      floatx4 aw = reinterpret_cast<const floatx4 *>(O + P.psiw)[(rb * RBG + qb) * 64 + l];
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (4 * qb + r < KG) {
#pragma unroll
          for (int f = 0; f < 4; f++) Pr[f][rb] = mfma32(aw[r], pg[f][r], Pr[f][rb]);
        }
    }
  }
  phase_fence();
  // mass flux (wall-projected), its normal traces and its volume divergence
  float uh[KM], vh[KM];
  d2b(Pr[2], uh);
  d2b(Pr[3], vh);
  if (anyw) {
    float a1[KM], a2[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) {
      if (VAR == 2) {  // src/calvarnl_ns.f:247-265
        a1[s] = wall ? uh[s] : 0.f;
        a2[s] = wall ? vh[s] : 0.f;
      } else {  // src/calvarnl_fs.f:298-315
        a1[s] = nwx * uh[s] + nwy * vh[s];
        a2[s] = 0.f;
      }
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      float ag[KM];
      afrag32<NM>(A.ops.gl32, rb, ag);
      floatx4 d1 = chain32(ag, a1, zero32());
      floatx4 d2 = zero32();
      if (VAR == 2) d2 = chain32(ag, a2, zero32());
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (4 * rb + r < KM) {
          if (VAR == 2) {
            uh[4 * rb + r] += d1[r];
            vh[4 * rb + r] += d2[r];
          } else {
            uh[4 * rb + r] += nwx * d1[r];
            vh[4 * rb + r] += nwy * d1[r];
          }
        }
    }
  }
  {
  const Nrm3 nr = normals(gg);
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    float at[KM];
    afrag32<NM>(O + P.tr, rb, at);
    floatx4 Tu = chain32(at, uh, zero32()), Tv = chain32(at, vh, zero32());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int kj = drow(rb, r, g);
      if (kj < NE3) {
        int k = kj / NGF;
        tr_put<NE3, 3>(A.st.trS, eb, col, 2, kj, Tu[r] * pick3(nr.x, k) + Tv[r] * pick3(nr.y, k));
      }
    }
  }
  }
  // the FP32 volume part of the three tendencies, stored as computed
#pragma unroll
  for (int s = 0; s < KM; s++) {
    float u = uh[s], v = vh[s];
    uh[s] = by * u - bx * v;
    vh[s] = -ay * u + ax * v;
  }
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    float adx[KM], ady[KM];
    afrag32<NM>(O + P.dx, rb, adx);
    afrag32<NM>(O + P.dy, rb, ady);
    floatx4 Vh = chain32(ady, vh, chain32(adx, uh, zero32()));
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = drow(rb, r, g);
      if (m < NM) {
        A.st.gvol[0][mb + m * 16 + col] = Pr[0][rb][r];
        A.st.gvol[1][mb + m * 16 + col] = Pr[1][rb][r];
        A.st.gvol[2][mb + m * 16 + col] = di * Vh[r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Nonlinear stages E + G (src/calvarnl_fs.f:382-544): the FP64 linear core
// (volume terms from the FP64 stage input, face terms from its traces), the
// face values of the flux traces (lifted with the core's, in FP64) and the
// B+C volume part, wall projection, RK4 stage, traces of the next stage
// input.  The core runs here rather than in B+C so that the B+C -> E+G
// hand-off is FP32 and B+C reads the stage input once.  RUS: the local
// Lax-Friedrichs (Rusanov) option is on.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR, int ROLE, bool RUS>
__device__ __forceinline__ void eg_part(const F32Args &A, const double *O64, int eb) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, RBM = Sh::RBM;
  constexpr Pack64 P64 = Sh::P64;
  const int l = lane(), g = l >> 4, col = l & 15, e0 = eb * 16, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const size_t mb = (size_t)e0 * NM;
  const Nbr3 N = nbr3(A.mesh, valid ? e : 0);
  // the FP64 gravity-wave core (volume and face terms, from the FP64 stage
  // input and traces) with the face values of the flux traces lifted along,
  // then the FP32 volume part B+C stored (gvol) added to it (read only now,
  // so that its registers are free while the core keeps its loads in flight;
  // read with the non-temporal hint, as it is read once, while bc_kernel
  // stores it without, so that where it fits in L2 it is found there)
  doublex4 G[3][RBM];
  core64<NC, NG, NGF, VAR, RUS, true, ROLE>(A, A.stage, O64 + P64.dx, O64 + P64.dy, O64 + P64.lf,
                                            N, eb, G);
#pragma unroll
  for (int f = (ROLE == 2 ? 2 : 0); f < (ROLE == 1 ? 2 : 3); f++)
#pragma unroll
    for (int rb = 0; rb < RBM; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = drow(rb, r, g), mc = m < NM ? m : NM - 1;
        const float v = __builtin_nontemporal_load(A.st.gvol[f] + mb + mc * 16 + col);
        G[f][rb][r] += (valid && m < NM) ? (double)v : 0.0;
      }
  phase_fence();
  finish64<NC, NG, NGF, VAR, ROLE>(A, A.stage, O64, eb, G);
}

// Nonlinear stages E + G.  Split (eg_split): each block of 16 elements gets
// two waves, one for the u and v equations (core, lift, wall projection,
// RK4 stage, velocity traces) and one for the eta equation (core, lift, RK4
// stage, eta traces); WG waves per workgroup then cover WG / 2 blocks.
// Otherwise one wave does the whole block.
template <int NC, int NG, int NGF, int VAR, int WG, bool RUS, int FORM>
__global__ __launch_bounds__(64 * WG) __attribute__((amdgpu_waves_per_eu(eg_waves(NC, FORM)))) void eg_kernel(F32Args A, int nblk) {
  using Sh = S<NC, NG, NGF>;
  constexpr Pack64 P64 = Sh::P64;
  constexpr int N64 = P64.total / 2;  // 16-byte words
  constexpr bool SPLIT = eg_split(NC, FORM);
  __shared__ floatx4 lds[N64];
  // The sweep runs backwards through each XCD's range of blocks: B+C went
  // forwards, so the hand-off it wrote last is still in that XCD's L2 when
  // this starts, and the next B+C, forwards again, starts where this ended
  // (C2: 1.418 -> 1.387 ms per RK4 step).
  const int b0 = last_block<SPLIT ? WG / 2 : WG>();
  if (b0 >= nblk) return;
  stage_lds<WG>(lds, reinterpret_cast<const floatx4 *>(A.ops.pack64), N64);
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int eb = b0 + (SPLIT ? wv >> 1 : wv);
  if (eb >= nblk) return;
  const double *O64 = reinterpret_cast<const double *>(lds);  // FP64 LF, TR, Dx, Dy
  if (!SPLIT) eg_part<NC, NG, NGF, VAR, 0, RUS>(A, O64, eb);
  else if (wv & 1) eg_part<NC, NG, NGF, VAR, 2, RUS>(A, O64, eb);
  else eg_part<NC, NG, NGF, VAR, 1, RUS>(A, O64, eb);
}

// Basis change, FP64.  to_psi: c = U a from the reference layout a[e*NM + m]
// into the blocked layout; from_psi: a = U^-1 c back.
// Basis change of a field between the reference's modal basis (a,
// element-major, NM per element) and the mass-orthonormal one (c, blocked
// by 16 elements: c[(e & ~15) * NM + i * 16 + (e & 15)]), c = U a with U
// upper triangular.  One thread per element; U sits in LDS and every lane
// reads it at the same address.  The element's coefficients are read
// straight from global memory: an LDS-staged variant (contiguous reads)
// measured slower, its uniform LDS reads of U raising register use to 244.
// (blockIdx.y: the field, at a + y * sa, c + y * sc)
template <int NM>
__global__ __launch_bounds__(256) void to_psi_kernel(int ne, const double *__restrict__ U,
                                                     const double *__restrict__ a,
                                                     double *__restrict__ c, size_t sa,
                                                     size_t sc) {
  a += blockIdx.y * sa;
  c += blockIdx.y * sc;
  __shared__ double Us[NM * NM];
  for (int i = threadIdx.x; i < NM * NM; i += 256) Us[i] = U[i];
  __syncthreads();
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  double x[NM];
#pragma unroll
  for (int k = 0; k < NM; k++) x[k] = a[(size_t)e * NM + k];
  double *ce = c + (size_t)(e & ~15) * NM + (e & 15);
#pragma unroll
  for (int i = 0; i < NM; i++) {
    double s = 0.0;
#pragma unroll
    for (int k = i; k < NM; k++) s += Us[i + NM * k] * x[k];
    ce[i * 16] = s;
  }
}

// a = U^-1 c by back substitution (the inverse of to_psi_kernel), 64
// elements per workgroup.  The element-major output goes through an LDS
// tile (row stride NM + 1) so that the global writes are contiguous.
template <int NM>
__global__ __launch_bounds__(PSI_WG) void from_psi_kernel(int ne, const double *__restrict__ U,
                                                          const double *__restrict__ c,
                                                          double *__restrict__ a, size_t sc,
                                                          size_t sa) {
  c += blockIdx.y * sc;
  a += blockIdx.y * sa;
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
    for (int i = 0; i < NM; i++) x[i] = cb[i * 16];
#pragma unroll
    for (int i = NM - 1; i >= 0; i--) {
      double s = x[i];
#pragma unroll
      for (int k = i + 1; k < NM; k++) s -= Us[i + NM * k] * x[k];
      x[i] = s / Us[i + NM * i];
    }
#pragma unroll
    for (int i = 0; i < NM; i++) T[threadIdx.x * (NM + 1) + i] = x[i];
  }
  __syncthreads();
  double *dst = a + (size_t)e0 * NM;
  for (int t = threadIdx.x; t < n * NM; t += PSI_WG) dst[t] = T[(t / NM) * (NM + 1) + t % NM];
}

// Element kernels of one right-hand side at a chosen workgroup size (waves
// of 16 elements per workgroup) and kernel form: the tile the autotuner
// picks.  The Lax-Friedrichs option runs form 0 only.
template <int NC, int NG, int NGF, int WG, int FORM>
int launch_elem(const F32Args &A, int phase, int nblk) {
  // a multiple of the 8 XCDs; a split E+G gives each block two waves
  const bool two = (phase == 5 && eg_split(NC, FORM)) || (phase == 3 && lin_split(FORM));
  const int grid = 8 * cdiv(cdiv(nblk, two ? WG / 2 : WG), 8);
  hipStream_t st = A.stream;
  const int v = A.variant;
  const bool lf = A.lf_alpha != 0.0;
  if (phase == 3) {
    if (lf) lin_kernel<NC, NG, NGF, WG, true, 0><<<8 * cdiv(cdiv(nblk, WG), 8), 64 * WG, 0, st>>>(A, nblk);
    else lin_kernel<NC, NG, NGF, WG, false, FORM><<<grid, 64 * WG, 0, st>>>(A, nblk);
  } else if (phase == 4) {
    if (v == 1) bc_kernel<NC, NG, NGF, 1, WG, FORM><<<grid, 64 * WG, 0, st>>>(A, nblk);
    else if (v == 2) bc_kernel<NC, NG, NGF, 2, WG, FORM><<<grid, 64 * WG, 0, st>>>(A, nblk);
    else bc_kernel<NC, NG, NGF, 3, WG, FORM><<<grid, 64 * WG, 0, st>>>(A, nblk);
  } else if (lf) {
    const int g0 = 8 * cdiv(cdiv(nblk, eg_split(NC, 0) ? WG / 2 : WG), 8);
    if (v == 1) eg_kernel<NC, NG, NGF, 1, WG, true, 0><<<g0, 64 * WG, 0, st>>>(A, nblk);
    else if (v == 2) eg_kernel<NC, NG, NGF, 2, WG, true, 0><<<g0, 64 * WG, 0, st>>>(A, nblk);
    else eg_kernel<NC, NG, NGF, 3, WG, true, 0><<<g0, 64 * WG, 0, st>>>(A, nblk);
  } else {
    if (v == 1) eg_kernel<NC, NG, NGF, 1, WG, false, FORM><<<grid, 64 * WG, 0, st>>>(A, nblk);
    else if (v == 2) eg_kernel<NC, NG, NGF, 2, WG, false, FORM><<<grid, 64 * WG, 0, st>>>(A, nblk);
    else eg_kernel<NC, NG, NGF, 3, WG, false, FORM><<<grid, 64 * WG, 0, st>>>(A, nblk);
  }
  return 0;
}

// Launches the persistent kernel when the device can hold its whole grid at
// once (its occupancy per CU times the CUs), else returns 1 and launches
// nothing.
template <int NC, int NG, int NGF, int WG>
int launch_persist(const F32Args &A, int nblk) {
  const int grid = 8 * cdiv(cdiv(nblk, WG), 8);
  const void *kern = reinterpret_cast<const void *>(&lin_persistent_kernel<NC, NG, NGF, WG>);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * WG, 0) != hipSuccess)
    return -2;
  if ((long)per_cu * cus < grid) return 1;
  sync_reset_kernel<NC, NG, NGF><<<1, 256, 0, A.stream>>>(A.sync, PERSIST_LINE * (1 + nblk),
                                                            A.abort0);
  lin_persistent_kernel<NC, NG, NGF, WG><<<grid, 64 * WG, 0, A.stream>>>(A, nblk);
  return 0;
}

template <int NC, int NG, int NGF, int WG>
int launch_form(const F32Args &A, int phase, int nblk, int form) {
  if (form == 1) return launch_elem<NC, NG, NGF, WG, 1>(A, phase, nblk);
  if (form == 2) return launch_elem<NC, NG, NGF, WG, 2>(A, phase, nblk);
  return launch_elem<NC, NG, NGF, WG, 0>(A, phase, nblk);
}

template <int NC, int NG, int NGF>
int launch(const F32Args &A, int phase, const double *in, double *out) {
  constexpr int NM = nm_of(NC);
  const int ne = A.mesh.ne, nblk = cdiv(ne, 16);
  const int grid = 8 * cdiv(cdiv(nblk, NW), 8);  // a multiple of the 8 XCDs
  hipStream_t st = A.stream;
  const int v = A.variant;
  switch (phase) {
    case 0:  // to psi (A.nf fields)
      to_psi_kernel<NM><<<dim3((ne + 255) / 256, A.nf > 1 ? A.nf : 1), 256, 0, st>>>(
          ne, A.ops.U, in, out, A.sin, A.sout);
      break;
    case 1:  // from psi (A.nf fields)
      from_psi_kernel<NM><<<dim3((ne + PSI_WG - 1) / PSI_WG, A.nf > 1 ? A.nf : 1), PSI_WG, 0, st>>>(
          ne, A.ops.U, in, out, A.sin, A.sout);
      break;
    case 2:  // traces of the first stage input
      if (v == 0) prep_kernel<NC, NG, NGF, 0><<<grid, NT, 0, st>>>(A, nblk);
      else if (v == 1) prep_kernel<NC, NG, NGF, 1><<<grid, NT, 0, st>>>(A, nblk);
      else if (v == 2) prep_kernel<NC, NG, NGF, 2><<<grid, NT, 0, st>>>(A, nblk);
      else prep_kernel<NC, NG, NGF, 3><<<grid, NT, 0, st>>>(A, nblk);
      break;
    case 3:  // linear right-hand side
    case 4:  // nonlinear B+C
    case 5: {  // nonlinear E+G
      if (phase > 3 && v == 0) return -1;
      // tile code: waves per workgroup + 16 x kernel form
      const int code = A.tile[phase == 5 ? 1 : 0], wg = code & 15, form = code >> 4;
      if (form < 0 || form >= NFORMS) return -1;
      if (wg == 2) launch_form<NC, NG, NGF, 2>(A, phase, nblk, form);
      else if (wg == 8) launch_form<NC, NG, NGF, 8>(A, phase, nblk, form);
      else launch_form<NC, NG, NGF, 4>(A, phase, nblk, form);
      break;
    }
    case 6: {  // persistent linear RK4 (A.steps steps), small meshes
      if (v != 0 || A.lf_alpha != 0.0 || !A.sync || A.steps < 1 || A.steps > PERSIST_STEPS)
        return -1;
      if (nblk > persist_blocks(NC)) return 1;  // (the per-stage launches are faster)
      // one wave per workgroup, so each block's wave has a CU to itself
      // (timed on the 512-element basin: 18.2 us per step, against 19.3 at
      // two and 25.1 at four), and one form (all three equations per wave,
      // which keeps the whole state of a block together); the tile codes of
      // the per-stage kernels do not apply
      if constexpr (persist_blocks(NC) > 0) {
        const int r = launch_persist<NC, NG, NGF, 1>(A, nblk);
        if (r) return r;  // 1: the grid does not fit at once (nothing launched)
      }
      break;
    }
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace f32

#if defined(SPOC_NC)
#define SPOC_CAT4(a, b, c, d) a##_##b##_##c##_##d
#define SPOC_NAME4(a, b, c, d) SPOC_CAT4(a, b, c, d)
int SPOC_NAME4(rhs_f32, SPOC_NC, SPOC_NG, SPOC_NGF)(const F32Args &a, int phase, const double *in,
                                                  double *out) {
  return f32::launch<SPOC_NC, SPOC_NG, SPOC_NGF>(a, phase, in, out);
}
#endif

}  // namespace spoc
