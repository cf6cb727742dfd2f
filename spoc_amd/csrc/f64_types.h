// Device data of the plain FP64 path.
//
// Layout: element-blocked by 64 (one wavefront): element e = 64 b + c, row r
// of an array with R rows per element at [(b * R + r) * 64 + c], so that a
// wave's access to one row of its 64 elements is one contiguous 512-byte
// segment.  Arrays are padded to whole blocks.  Elements are in the
// library's internal numbering.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

namespace spoc {

constexpr int F64_WB = 64;  // elements per block

__host__ __device__ __forceinline__ size_t f64_at(int e, int r, int R) {
  return ((size_t)(e >> 6) * R + r) * F64_WB + (e & (F64_WB - 1));
}

// rows of the per-element geometry
enum { G_AX = 0, G_BX, G_AY, G_BY, G_DETAI, G_NX, G_NY, G_ROWS };
// rows of the per-element integers: faces of local edges 0..2 (0-based),
// its of local edges 0..2, wall flag
enum { I_ITM = 0, I_ITS = 3, I_BDL = 6, I_ROWS };

struct F64Mesh {
  int ne, nf, nblk, nwall;
  const int *tseg;      // (2, nf): 1-based elements, 0 = wall
  const int *ssit;      // (2, nf): local edge in each element, 1..3
  const double *norm;   // (2, nf)
  const double *geo;    // blocked, G_ROWS rows
  const int *ib;        // blocked, I_ROWS rows
  const int *wall;      // [nwall] the wall elements
};

struct F64Args {
  int variant;          // SPOC_LIN..SPOC_INV
  int stage;            // -1: tendencies into k; 0..3: RK4 stage
  double dt, dt6;       // dt6 = dt / 6, formed on the host
  hipStream_t stream;
  DevTables tab;
  F64Mesh mesh;
  double mu0, mb0, g0;
  const double *hh0e;   // [ne] hb(1,1) of the element's mesh (linear variant's depth)
  const double *hb, *w0, *z0, *fc;  // blocked NM rows
  const double *x[3];   // stage input u, v, eta (blocked NM rows)
  double *u0[3], *xo[3], *acc[3], *k[3];
  // work (blocked): q (4), F (4), sources (2), mass fluxes (2), raw
  // tendencies of wall elements (3), values at the quadrature points (9 NG
  // rows), weighted products (8 NG rows)
  double *Q[4], *F[4], *S[2], *UH[2], *G[3], *MF, *WS;
  double *face;         // face values: [f][out][j], 6 outputs
  // phases 0 and 1: nf fields at once, field f at in + f * sin, out + f * sout
  // (nf 0 counts as 1)
  int nf;
  size_t sin, sout;
  TimerHook hook;
  void *hook_user;
  void timer(int slot) const {
    if (hook) hook(hook_user, slot, stream);
  }
};

// Per-shape launchers (rhs_f64.hip, one translation unit per shape).
// phase 0: in (NM, ne) -> out blocked; 1: blocked in -> out (NM, ne);
// 2: one right-hand side (with the RK4 stage of a.stage).
#define SPOC_DECL_F64(NC_, NG_, NGF_) \
  int rhs_f64_##NC_##_##NG_##_##NGF_(const F64Args &a, int phase, const double *in, double *out);
SPOC_FOR_EACH_SHAPE(SPOC_DECL_F64)
#undef SPOC_DECL_F64

}  // namespace spoc
