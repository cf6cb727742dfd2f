// Data of the mixed-precision (FP32 + FP64 MFMA) right-hand side.
//
// This path works in the coordinates of the mass-orthonormal basis
// psi = U^{-T} phi (U the upper Cholesky factor of the reference mass matrix,
// src/shwwnl.f:153): the state is c = U a, the mass matrix is the identity,
// and every operator is a dense matrix formed once in FP64.  (In the
// reference's own modal basis the mass matrix has condition number ~2e8 at
// p = 6, which turns FP32 rounding of the test-function integrals into ~1e-4
// relative errors after the solve.)
//
// Register layout shared by every kernel: for a block of 16 elements, lane
// l = 16 g + c (c = element column) holds mode m = 4 s + g of element c at
// k-step s.  That is the B-operand layout of v_mfma_*_16x16x4, and it is
// also the D (result) layout of the FP64 instruction, so an FP64 product can
// feed the next product from registers.  The FP32 instruction's D layout
// holds rows 4 g + r of each 16-row block instead; the FP32 operator images
// store their rows permuted by (16 b + 4 i + j) <-> (16 b + 4 j + i), which
// makes their results come out in the shared layout too.
#pragma once
#include <hip/hip_runtime.h>

namespace spoc {

constexpr int GEO = 16;  // per-element geometry words (see F32Mesh)

__host__ __device__ constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// Operator images.  An R x K operator is the A operand of v_mfma_*_16x16x4:
// row block rb (16 rows) at k-step s (4 columns); lane l = i + 16 g holds
// A[16 rb + i][4 s + g].  FP32 images keep 4 k-steps per lane together (one
// 16-byte load): img[((rb * KQ + q) * 64 + l) * 4 + t] with s = 4 q + t,
// KQ = ceil(K / 16).  FP64 images keep 2: img[((rb * KH + h) * 64 + l) * 2 + t]
// with s = 2 h + t, KH = ceil(K / 8).  Padding is zero.
__host__ __device__ constexpr int img32_size(int R, int K) { return cdiv(R, 16) * cdiv(K, 16) * 256; }
__host__ __device__ constexpr int img64_size(int R, int K) { return cdiv(R, 16) * cdiv(K, 8) * 128; }
// Physical row of an FP32 image that holds logical row i (an involution).
__host__ __device__ constexpr int sigma_row(int i) {
  return (i & ~15) | ((i & 3) << 2) | ((i >> 2) & 3);
}

// FP32 operators, one buffer, offsets in floats.  B+C stages all of it in
// LDS.
struct Pack32 {
  int dx, dy, lf, tr, psi, psiw, xq, yq, tail, total;
};
__host__ __device__ constexpr Pack32 pack32(int nm, int ng, int ne3) {
  Pack32 p{};
  int o = 0;
  p.dx = o;   o += img32_size(nm, nm);   // U^-T agx^T U^-1
  p.dy = o;   o += img32_size(nm, nm);   // U^-T agy^T U^-1
  p.lf = o;   o += img32_size(nm, ne3);  // U^-T (wif pico)^T: edge lift
  p.tr = o;   o += img32_size(ne3, nm);  // pico U^-1: traces at the edge points
  p.psi = o;  o += img32_size(ng, nm);   // pi2 U^-1: values at the quadrature points
  p.psiw = o; o += img32_size(nm, ng);   // (W pi2 U^-1)^T: projection
  p.xq = o;   o += cdiv(ng, 16) * 16;    // quadrature points
  p.yq = o;   o += cdiv(ng, 16) * 16;
  p.tail = o; o += cdiv(nm, 16) * 16;    // row ng-1 of pi2 U^-1, modes in order (a
                                         // last row block holding one point)
  p.total = o;
  return p;
}

// FP64 operators (same meaning), offsets in doubles.  The face operators
// come first: the prep kernel stages [0, dx); E+G and the linear kernel
// stage all of it.
struct Pack64 {
  int lf, tr, dx, dy, total;
};
__host__ __device__ constexpr Pack64 pack64(int nm, int ne3) {
  Pack64 p{};
  int o = 0;
  p.lf = o; o += img64_size(nm, ne3);
  p.tr = o; o += img64_size(ne3, nm);
  p.dx = o; o += img64_size(nm, nm);
  p.dy = o; o += img64_size(nm, nm);
  p.total = o;
  return p;
}

// Transposed FP64 operators for the adjoint, offsets in doubles.
struct Pack64T {
  int dxt, dyt, lft, trt, total_lin, psi, psiwt, psit, xq, yq, total;
};
__host__ __device__ constexpr Pack64T pack64t(int nm, int ng, int ne3) {
  Pack64T p{};
  int o = 0;
  p.dxt = o; o += img64_size(nm, nm);    // Dx^T
  p.dyt = o; o += img64_size(nm, nm);    // Dy^T
  p.lft = o; o += img64_size(ne3, nm);   // LF^T: modes -> edge points
  p.trt = o; o += img64_size(nm, ne3);   // TR^T: edge points -> modes
  p.total_lin = o;                       // (what the linear adjoint uses)
  p.psi = o;  o += img64_size(ng, nm);   // PSI: modes -> quadrature points
  p.psiwt = o; o += img64_size(ng, nm);  // PSIW^T = W PSI
  p.psit = o; o += img64_size(nm, ng);   // PSI^T: quadrature points -> modes
  p.xq = o;  o += cdiv(ng, 8) * 8;       // quadrature points
  p.yq = o;  o += cdiv(ng, 8) * 8;
  p.total = o;
  return p;
}

struct F32Ops {
  const float *pack32;   // Pack32 layout
  const double *pack64;  // Pack64 layout
  const double *pack64t; // Pack64T layout
  const float *gl32;     // img32 of G - I, G = U R^T (R amm R^T)^-1 R U^T (wall projector)
  const double *gl64;    // img64 of G - I
  const double *U;       // [NM*NM] upper factor, column-major (basis change)
};

// Per-element mesh data.
//   geo[e*GEO + 0..4]   ax bx ay by 1/det
//   geo[e*GEO + 5..7]   x-component of the face normal of local edge k
//   geo[e*GEO + 8..10]  y-component of the face normal of local edge k
//   geo[e*GEO + 11..13] its(k, e) as +-1
//   geo[e*GEO + 14..15] unit wall normal normt (wall elements)
struct F32Mesh {
  int ne;
  const float *geo;
  const double *geo64;  // same 16 words in FP64
  const int *nbr;       // [ne*3] neighbour element across local edge k, -1 on a wall
  const int *nbe;       // [ne*3] that neighbour's local edge
  const int *bdl;       // [ne]
};

// Element-blocked arrays: element e = 16 b + c, row r of an array with R
// rows per element at [(b * R + r) * 16 + c].
struct F32Fields {
  // Values at the quadrature points, four per point (hb, w0, z0, hb - hcore),
  // ordered for the MFMA lane layout: quadrature k-step ks (points
  // 4 ks .. 4 ks + 3) of element block b is 64 float4 at [(b * KG + ks) * 64],
  // lane l holding point 4 ks + (l >> 4) of element 16 b + (l & 15), so that
  // one k-step is one contiguous 1 KiB wave load.  Points past NG hold
  // (1, 0, 0, 0).
  const float *qf;
  const float *fc3;              // [ne][3] Coriolis f = fc0 + fc1*eta + fc2*xi
  const double *hh0e;            // [ne] depth of the linear variant: hb(1,1) of the
                                 // element's own mesh (ensemble member)
  const double *hcore;           // [ne] depth H0 of the nonlinear variants' FP64
                                 // core: the element's constant mode hb(1,e)
  float mu0, mb0;
  double g0;
};

// Depth H0 of the FP64 core's -div(H0 u).  The linear variant's depth is the
// constant hb(1,1) (src/calvar_lin.f:65).  For the nonlinear variants any
// per-element constant keeps the split (eta + H) u = H0 u + (eta + H - H0) u
// exact (H0 u lies in the element's polynomial space; the central flux of
// H0 u averages each side's own), and the element's constant mode hb(1,e)
// leaves the FP32 remainder only the depth's variation inside the element.
template <int VAR>
__host__ __device__ __forceinline__ double core_depth(const F32Fields &f, int e) {
  return VAR == 0 ? f.hh0e[e] : f.hcore[e];
}

struct F32State {
  double *c0[3];           // [NM rows] psi coordinates: step start
  double *ys[3][3];        // [NM rows] stage inputs of RK4 stages 1, 2, 3 (Y2, Y3, Y4)
  double *k[3];            // [NM rows] psi tendencies (stage < 0)
  float *gvol[3];          // [NM rows] FP32 volume and source part of the tendency
                           // (B+C -> E+G; the FP64 core is E+G's own)
  float *trA[2];           // [2*NE3 rows] T_u, T_v of the stage input (by stage parity,
                           // like trL: the local Lax-Friedrichs option reads them in E+G)
  double *trL[2];          // [2*NE3 rows] T_eta, hh0 (T_u, T_v).n_face (wall-projected);
                           // RK4 stage s reads trL[s & 1] and writes trL[(s + 1) & 1],
                           // because one launch reads neighbours' traces and writes its own
  float *trS;              // [3*NE3 rows] flux traces (B+C -> E+G)
  // adjoint (psi coordinates, FP64): lam = adjoint of the step's end state,
  // xbar = its running sum over the stages, aw = the stage input (adjoint of
  // the stage tendency), avol = partial stage result, abar = the lifted
  // adjoints at the element's own edge points [4*NE3 rows: eta trace,
  // mass-flux trace, the two viscous-flux traces], abar2 = those of the
  // velocity gradient [2*NE3 rows]
  double *lam[3], *xbar[3], *aw[3], *avol[3];
  double *abar, *abar2;
  // nonlinear adjoint: aF (4) = adjoint of F = mu q, aUH (2) = of the raw
  // mass flux, aL (2) = wall-projected adjoint of the (u, v) tendency,
  // xb = the stage's base state (the linearisation point) and btr its
  // velocity traces [2*NE3 rows], hbbar = accumulated adjoint of hb
  double *aF[4], *aUH[2], *aL[2];
  double *xb[3];
  double *btr;
  double *hbbar;
};

// Workgroup sizes (waves of 16 elements) the element kernels can run at,
// and the number of compiled forms of each (register budget, E+G split); the
// autotuner picks one pair per kernel, coded as waves + 16 x form.
constexpr int TILE_CHOICES[3] = {2, 4, 8};
constexpr int TILE_FORMS = 3;

// Persistent linear RK4 kernel: meshes of at most persist_blocks(nc) blocks of
// 16 elements, where it measured faster than the per-stage launches (5-20%;
// every wave's release writes back its XCD's L2, a cost that grows with the
// waves per XCD: at p = 3 on 2,048 elements it is already slower, and at
// p >= 6 it never won), PERSIST_BLOCKS at most (far fewer waves than the
// chip holds, so all are resident at once), and at most PERSIST_STEPS steps per launch
// (each block's count of finished stages, up to 4 x steps, stays far below
// 2^32).  Its sync words: the abort word at [1], block b's count at
// [PERSIST_LINE * (1 + b)], each count on a 128-byte line of its own.
constexpr int PERSIST_BLOCKS = 96;
__host__ __device__ constexpr int persist_blocks(int nc) {
  return nc <= 3 ? PERSIST_BLOCKS : nc == 4 ? 64 : nc == 5 ? 48 : 0;
}
constexpr int PERSIST_STEPS = 1 << 20;
constexpr int PERSIST_LINE = 32;
constexpr int PERSIST_SYNC_WORDS = PERSIST_LINE * (1 + PERSIST_BLOCKS);

struct F32Args {
  int variant;
  int stage;  // -1: tendencies only, 0..3: RK4 stage
  int steps;  // persistent RK4 kernel: steps to run in the launch
  unsigned *sync;  // persistent RK4 kernel: [1] abort (a wait that ran out),
                   // [PERSIST_LINE (1 + b)] stages block b has finished; zeroed
                   // before the launch (abort0: the abort word's start value, 1
                   // only to test the failure path)
  unsigned abort0;
  // phases 0 and 1 (basis change): nf fields at once, field f at
  // in + f * sin, out + f * sout (nf 0 counts as 1)
  int nf;
  size_t sin, sout;
  int tile[2];  // waves per workgroup: [0] lin_kernel or bc_kernel, [1] eg_kernel
  double dt;
  // local Lax-Friedrichs (Rusanov) penalty of the gravity-wave core's face
  // flux: alpha * sqrt(g hh0) times the jumps of eta and (u, v); 0: the
  // reference's central flux
  double lf_alpha;
  hipStream_t stream;
  F32Ops ops;
  F32Mesh mesh;
  F32Fields fld;
  F32State st;
};

}  // namespace spoc
