// Data of the FP32-MFMA right-hand side.
//
// This path works in the coordinates of the mass-orthonormal basis
// psi = U^{-T} phi (U the upper Cholesky factor of the reference mass matrix,
// src/shwwnl.f:153): the state is c = U a, the mass matrix is the identity,
// and every operator is a well-conditioned dense matrix formed once in FP64
// and rounded to FP32.  (In the reference's own modal basis the mass matrix
// has condition number ~2e8 at p = 6, which turns FP32 rounding of the
// test-function integrals into ~1e-4 relative errors after the solve.)
#pragma once
#include <hip/hip_runtime.h>

namespace spoc {

constexpr int GEO = 16;  // per-element geometry words (see F32Mesh)

// Operators (device, FP32, row-major with rows padded to 16 and the
// contraction length padded to 4 or 16, zeros in the padding).
struct F32Ops {
  const float *DX, *DY;  // [NMr][NMk]  U^-T agx^T U^-1, U^-T agy^T U^-1
  const float *LF;       // [NMr][NEk]  U^-T (wif * pico)^T : edge lift
  const float *PSI;      // [NGr][NMk]  pi2 U^-1 : values at quadrature points
  const float *PSIW;     // [NMr][NGr]  (W pi2 U^-1)^T : projection
  const float *TR;       // [NEr][NMk]  pico U^-1 : traces at edge points
  const float *GL;       // [NM][NM]    U Pl U^T - I (wall projector minus identity)
  const float *xq, *yq;  // quadrature points on the reference triangle
  const double *U;       // [NM*NM] upper factor, column-major (basis change)
  // low parts: A = hi + lo with hi = fl32(A), lo = fl32(A - hi), so that a
  // product costs two MFMAs and carries ~46 significant bits of the operator
  const float *DXl, *DYl, *LFl, *PSIl, *PSIWl, *TRl;
  // FP64 images of DX, DY, LF, TR (same padded layouts) and G - I, for the
  // linear gravity-wave core that runs on FP64 MFMA
  const double *DX64, *DY64, *LF64, *TR64, *GL64;
};

// Per-element mesh data of the FP32 path.
//   geo[e*GEO + 0..4]  ax bx ay by 1/det
//   geo[e*GEO + 5..7]  x-component of the face normal of local edge k
//   geo[e*GEO + 8..10] y-component of the face normal of local edge k
//   geo[e*GEO + 11..13] its(k, e) as +-1
//   geo[e*GEO + 14..15] unit wall normal normt (wall elements)
struct F32Mesh {
  int ne;
  const float *geo;
  const double *geo64;  // same 16 words in FP64
  const int *nbr;  // [ne*3] neighbour element across local edge k, -1 on a wall
  const int *nbe;  // [ne*3] that neighbour's local edge
  const int *bdl;  // [ne]
};

struct F32Fields {
  const float *hbq, *w0q, *z0q;  // [ne][NG] values at quadrature points
  const float *fc3;              // [ne][3] Coriolis f = fc0 + fc1*eta + fc2*xi
  const float *hbdq;             // [ne][NG] hb - hh0 at quadrature points (blocked)
  float mu0, mb0, g0, hh0;
  double g0d, hh0d;
};

struct F32State {
  double *c0[3], *acc[3];  // [ne][NM] psi coordinates, FP64
  float *xs[3];            // [ne][NM] FP32 shadow of the stage input
  float *trA;              // [ne][3][2][NGF] traces of the stage input
  float *trS;              // [ne][3][3][NGF] combined flux traces (BC -> EG)
  float *F[4], *src[2], *uh[2];  // [ne][NM] BC -> EG
  double *k[3];            // [ne][NM] psi tendencies (stage < 0)
  double *x64[3];          // [ne][NM] FP64 stage input (linear core)
  double *trL;             // [ne][3][2][NGF] FP64 traces: T_h, hh0 (T_u, T_v).n
  double *kL[3];           // [ne][NM] linear-core tendencies
};

struct F32Args {
  int variant;
  int stage;  // -1: tendencies only, 0..3: RK4 stage
  int hybrid; // nonlinear variants: linear gravity-wave core in FP64 (kL)
  double dt;
  hipStream_t stream;
  F32Ops ops;
  F32Mesh mesh;
  F32Fields fld;
  F32State st;
};

}  // namespace spoc
