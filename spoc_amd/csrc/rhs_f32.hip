// FP32-MFMA right-hand side for gfx950 (v_mfma_f32_16x16x4_f32).
//
// Coordinates: the mass-orthonormal basis psi = U^{-T} phi (see f32_types.h),
// so every step of the reference's cal_var is a dense small GEMM
//   out[row, element] = Op[row, k] * X[k, element]
// with the operator rows on the MFMA M side and 16 elements on the N side.
//
// Work split: one wavefront owns a block of 16 elements for the whole
// kernel; the element operands live in that wave's slice of LDS in
// [k][16] layout (the MFMA B fragment), operator fragments are loaded from
// global memory (L1/L2-resident, a few KB) into registers once per row block
// and reused across fields.  Per-element scalars (Jacobian terms, normals)
// multiply the D fragments lane-locally because an MFMA output column is one
// element.
//
// Edge coupling: every element stores the FP32 traces of its fields at the
// Gauss-Legendre points of its three edges; the consumer averages its own
// trace with the neighbour's (reversed along the shared edge), as the
// reference's face average does (src/calflu.f:41-62).  Both sides of a face
// read the same two stored numbers, so the face flux is single-valued.
//
// Kernels per right-hand side:  linear: lin_kernel;  nonlinear: bc_kernel
// (gradient q, quadrature-point products, mass fluxes) then eg_kernel
// (u, v, eta tendencies).  Both end in the FP64 RK4 stage update and the
// traces of the next stage input.
#include <hip/hip_runtime.h>

#include "common.h"
#include "f32_types.h"

namespace spoc {
namespace f32 {

typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int NT = 256, NW = 4;

constexpr int up(int x, int m) { return (x + m - 1) / m * m; }

template <int NC, int NG, int NGF>
struct S {
  static constexpr int NM = nm_of(NC), NE3 = 3 * NGF;
  static constexpr int NMr = up(NM, 16), NMk = up(NM, 4);
  static constexpr int NGr = up(NG, 16);
  static constexpr int NEr = up(NE3, 16), NEk = up(NE3, 4);
  static constexpr int KM = NMk / 4, KE = NEk / 4;
  static constexpr int RBM = NMr / 16, RBG = NGr / 16, RBE = NEr / 16;
};

__device__ __forceinline__ floatx4 mfma(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int lane() { return threadIdx.x & 63; }

// Order this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Operators are applied as hi + lo parts (two MFMAs per k-step) unless the
// build sets SPOC_F32_SPLIT=0: a single FP32 rounding of the operators
// shifts the discrete wave frequencies by ~6e-8 relative, which grows into
// ~1e-6 relative error after 1000 steps at p = 6.
#ifndef SPOC_F32_SPLIT
#define SPOC_F32_SPLIT 1
#endif

template <int KS>
struct Frag {
  float h[KS];
#if SPOC_F32_SPLIT
  float l[KS];
#endif
};

// A fragments of row block rb, k-steps ks0..ks0+KS-1 of a row-major operator:
// lane l holds A[16*rb + l%16][4*ks + l/16].
template <int KS>
__device__ __forceinline__ void afrag(const float *__restrict__ A, const float *__restrict__ Alo,
                                      int lda, int rb, int ks0, Frag<KS> &a) {
  int l = lane();
  size_t o = (size_t)(rb * 16 + (l & 15)) * lda + 4 * ks0 + (l >> 4);
#pragma unroll
  for (int s = 0; s < KS; s++) a.h[s] = A[o + 4 * s];
#if SPOC_F32_SPLIT
#pragma unroll
  for (int s = 0; s < KS; s++) a.l[s] = Alo[o + 4 * s];
#else
  (void)Alo;
#endif
}

// acc += A * B, B in LDS as [4*KS][16]: lane l reads B[4s + l/16][l%16].
template <int KS>
__device__ __forceinline__ floatx4 chain(const Frag<KS> &a, const float *B, floatx4 acc) {
  int l = lane();
  const float *p = B + (l >> 4) * 16 + (l & 15);
#pragma unroll
  for (int s = 0; s < KS; s++) {
    float b = p[64 * s];
#if SPOC_F32_SPLIT
    acc = mfma(a.l[s], b, acc);
#endif
    acc = mfma(a.h[s], b, acc);
  }
  return acc;
}

// acc += A * B with the B fragment already in registers (b[s] = B[4s + l/16][l%16]).
template <int KS>
__device__ __forceinline__ floatx4 chain_r(const Frag<KS> &a, const float (&b)[KS], floatx4 acc) {
#pragma unroll
  for (int s = 0; s < KS; s++) {
#if SPOC_F32_SPLIT
    acc = mfma(a.l[s], b[s], acc);
#endif
    acc = mfma(a.h[s], b[s], acc);
  }
  return acc;
}

__device__ __forceinline__ floatx4 zero4() { return floatx4{0.f, 0.f, 0.f, 0.f}; }

// FP64 MFMA (v_mfma_f64_16x16x4_f64): the A and B lane layouts are those of
// the FP32 instruction, but the four D values of lane l are rows
// l/16, 4 + l/16, 8 + l/16, 12 + l/16 of column l%16 (the FP32 instruction
// holds rows 4(l/16) .. 4(l/16)+3).
typedef double doublex4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int drow64(int rb, int g, int r) { return rb * 16 + 4 * r + g; }
__device__ __forceinline__ doublex4 mfma64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 zero4d() { return doublex4{0.0, 0.0, 0.0, 0.0}; }

template <int KS>
__device__ __forceinline__ void afrag64(const double *__restrict__ A, int lda, int rb,
                                        double (&a)[KS]) {
  int l = lane();
  const double *p = A + (size_t)(rb * 16 + (l & 15)) * lda + (l >> 4);
#pragma unroll
  for (int s = 0; s < KS; s++) a[s] = p[4 * s];
}

template <int KS>
__device__ __forceinline__ doublex4 chain64(const double (&a)[KS], const double *B, doublex4 acc) {
  int l = lane();
  const double *p = B + (l >> 4) * 16 + (l & 15);
#pragma unroll
  for (int s = 0; s < KS; s++) acc = mfma64(a[s], p[64 * s], acc);
  return acc;
}

template <int KS>
__device__ __forceinline__ doublex4 chain64_r(const double (&a)[KS], const double (&b)[KS],
                                              doublex4 acc) {
#pragma unroll
  for (int s = 0; s < KS; s++) acc = mfma64(a[s], b[s], acc);
  return acc;
}

template <int NM>
__device__ __forceinline__ double gl_apply64(const double *__restrict__ GL, const double *v, int i) {
  double d = 0.0;
  for (int j = 0; j < NM; j++) d += GL[i * NM + j] * v[j];
  return d;
}

// Wall projection of a vector pair held in LDS at stride 16 (column c),
// FP64: free slip (wallmode 1) puts the normal component on the reduced
// basis, no slip (2) both components.  Lanes < NM take part; V holds 128
// doubles of scratch.
template <int NM>
__device__ void wall_project64(const double *__restrict__ GL, double *gx, double *gy, double nx,
                               double ny, int wallmode, double *V) {
  int l = lane();
  if (wallmode == 2) {
    if (l < NM) {
      V[l] = gx[16 * l];
      V[64 + l] = gy[16 * l];
    }
    wave_sync();
    if (l < NM) {
      double d1 = gl_apply64<NM>(GL, V, l), d2 = gl_apply64<NM>(GL, V + 64, l);
      gx[16 * l] += d1;
      gy[16 * l] += d2;
    }
  } else {
    if (l < NM) V[l] = nx * gx[16 * l] + ny * gy[16 * l];
    wave_sync();
    if (l < NM) {
      double d = gl_apply64<NM>(GL, V, l);
      gx[16 * l] += nx * d;
      gy[16 * l] += ny * d;
    }
  }
  wave_sync();
}

// (G_l - I) applied to a vector held in LDS (wall elements; lanes < NM).
template <int NM>
__device__ __forceinline__ float gl_apply(const float *__restrict__ GL, const float *v, int i) {
  float d = 0.f;
  for (int j = 0; j < NM; j++) d += GL[i * NM + j] * v[j];
  return d;
}

// ---------------------------------------------------------------------------
// Traces of the stage input at the 3*NGF edge points, written per element as
// trA[e][k][2][NGF].  Nonlinear: (T_u, T_v).  Linear: (T_h, S_uv) with
// S_uv = T_u * nfx + T_v * nfy on the face normal of each edge.
// X holds u, v, h of the wave's 16 elements in [NMk][16] layout.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF>
__device__ void emit_traces(const F32Args &A, const float *X, int e0, bool lin) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NMk = Sh::NMk, KM = Sh::KM, NE3 = Sh::NE3, RBE = Sh::RBE;
  const int l = lane(), g = l >> 4, col = l & 15, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const float *geo = A.mesh.geo + (size_t)(valid ? e : 0) * GEO;
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    Frag<KM> at;
    afrag<KM>(A.ops.TR, A.ops.TRl, NMk, rb, 0, at);
    floatx4 T0, T1;
    if (lin) {
      T0 = chain<KM>(at, X + 2 * NMk * 16, zero4());  // h
      floatx4 Tu = chain<KM>(at, X, zero4()), Tv = chain<KM>(at, X + NMk * 16, zero4());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = rb * 16 + 4 * g + r, k = row / NGF;
        float nx = k < 3 ? geo[5 + k] : 0.f, ny = k < 3 ? geo[8 + k] : 0.f;
        T1[r] = Tu[r] * nx + Tv[r] * ny;
      }
    } else {
      T0 = chain<KM>(at, X, zero4());
      T1 = chain<KM>(at, X + NMk * 16, zero4());
    }
    if (valid) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = rb * 16 + 4 * g + r;
        if (row < NE3) {
          int k = row / NGF, j = row % NGF;
          float *t = A.st.trA + ((size_t)e * 3 + k) * 2 * NGF;
          t[j] = T0[r];
          t[NGF + j] = T1[r];
        }
      }
    }
  }
}

// FP64 traces of the stage input for the hybrid path.  X holds u, v, h in
// FP64, [3][NMk][16].  Writes trA (FP32 T_u, T_v, for the FP32 gradient)
// and trL (FP64 T_h, and S = hh0 (T_u nfx + T_v nfy) of the wall-projected
// mass flux, src/calvarnl_fs.f:298-321) for the linear core.  Destroys X's
// u, v columns of wall elements.
template <int NC, int NG, int NGF>
__device__ void emit_traces64(const F32Args &A, double *X, double *V, int e0, int wallmode) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NMk = Sh::NMk, KM = Sh::KM, NE3 = Sh::NE3, RBE = Sh::RBE;
  const int l = lane(), g = l >> 4, col = l & 15, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  const double *geo = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double hh0 = A.fld.hh0d;
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    double at[KM];
    afrag64<KM>(A.ops.TR64, NMk, rb, at);
    doublex4 Tu = chain64<KM>(at, X, zero4d()), Tv = chain64<KM>(at, X + NMk * 16, zero4d());
    doublex4 Th = chain64<KM>(at, X + 2 * NMk * 16, zero4d());
    if (valid) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = drow64(rb, g, r);
        if (row < NE3) {
          int k = row / NGF, j = row % NGF;
          float *ta = A.st.trA + ((size_t)e * 3 + k) * 2 * NGF;
          ta[j] = (float)Tu[r];
          ta[NGF + j] = (float)Tv[r];
          double *tl = A.st.trL + ((size_t)e * 3 + k) * 2 * NGF;
          tl[j] = Th[r];
          tl[NGF + j] = hh0 * (Tu[r] * geo[5 + k] + Tv[r] * geo[8 + k]);
        }
      }
    }
  }
  // wall elements: the mass flux is projected before its trace is taken
  unsigned long long wm = __ballot(valid && g == 0 && A.mesh.bdl[e] != 0);
  if (!wm) return;
  wave_sync();
  while (wm) {
    int c = __ffsll((long long)wm) - 1;
    wm &= wm - 1;
    const double *gw = A.mesh.geo64 + (size_t)(e0 + c) * GEO;
    wall_project64<NM>(A.ops.GL64, X + c, X + NMk * 16 + c, gw[14], gw[15], wallmode, V);
    if (l < NE3) {
      int k = l / NGF, j = l % NGF;
      double tu = 0.0, tv = 0.0;
      for (int m = 0; m < NM; m++) {
        double t = A.ops.TR64[(size_t)l * NMk + m];
        tu += t * X[16 * m + c];
        tv += t * X[NMk * 16 + 16 * m + c];
      }
      A.st.trL[((size_t)(e0 + c) * 3 + k) * 2 * NGF + NGF + j] = hh0 * (tu * gw[5 + k] + tv * gw[8 + k]);
    }
    wave_sync();
  }
}

template <int NGF>
__device__ __forceinline__ double face_avg64(const double *__restrict__ tr, const F32Mesh &M, int e,
                                             int k, int j, int f) {
  int nb = M.nbr[e * 3 + k];
  if (nb < 0) return tr[((size_t)e * 3 + 0) * 2 * NGF + f * NGF + j];
  int kn = M.nbe[e * 3 + k];
  double a = tr[((size_t)e * 3 + k) * 2 * NGF + f * NGF + j];
  double b = tr[((size_t)nb * 3 + kn) * 2 * NGF + f * NGF + (NGF - 1 - j)];
  return 0.5 * (a + b);
}

// Copy NF blocked FP32 fields [eb][NM][16] of block e0 into LDS [NMk][16].
template <int NM, int NMk>
__device__ __forceinline__ void load_block(float *dst, const float *__restrict__ src, int e0) {
  int l = lane();
  const float *s = src + (size_t)e0 * NM;
  for (int i = l; i < NMk * 16; i += 64) dst[i] = i < NM * 16 ? s[i] : 0.f;
}

// Average of own and neighbour edge traces for element e, local edge k,
// point j (local orientation), field f of NF stored per edge.  Walls use
// the element's own trace on its edge 1, like the reference's wall faces
// (src/calflu.f:31-39).
template <int NGF, int NF>
__device__ __forceinline__ float face_avg(const float *__restrict__ tr, const F32Mesh &M, int e,
                                          int k, int j, int f) {
  int nb = M.nbr[e * 3 + k];
  if (nb < 0) return tr[((size_t)e * 3 + 0) * NF * NGF + f * NGF + j];
  int kn = M.nbe[e * 3 + k];
  float a = tr[((size_t)e * 3 + k) * NF * NGF + f * NGF + j];
  float b = tr[((size_t)nb * 3 + kn) * NF * NGF + f * NGF + (NGF - 1 - j)];
  return 0.5f * (a + b);
}

// RK4 stage update of one value in FP64 (src/shwwnl.f:242-284) and the FP32
// shadow of the next stage input.  Returns that shadow value.
__device__ __forceinline__ float rk_stage(const F32Args &A, int f, size_t idx, float kf,
                                          double *x64 = nullptr) {
  double k = (double)kf;
  if (A.hybrid) k += A.st.kL[f][idx];
  int s = A.stage;
  if (s < 0) {
    A.st.k[f][idx] = k;
    return 0.f;
  }
  double *c0 = A.st.c0[f] + idx, *acc = A.st.acc[f] + idx;
  double x;
  if (s == 0) {
    *acc = k;
    x = *c0 + A.dt * k * 0.5;
  } else if (s == 1) {
    *acc = *acc + 2.0 * k;
    x = *c0 + A.dt * k * 0.5;
  } else if (s == 2) {
    *acc = *acc + 2.0 * k;
    x = *c0 + A.dt * k;
  } else {
    x = *c0 + (A.dt / 6.0) * (*acc + k);
    *c0 = x;
  }
  float xf = (float)x;
  A.st.xs[f][idx] = xf;
  if (A.hybrid) A.st.x64[f][idx] = x;
  if (x64) *x64 = x;
  return xf;
}

// Epilogue shared by the element kernels: wall projection of (gx, gy), RK
// update, traces of the next stage input.  G[f][rb] hold the 3 tendencies in
// D layout; W is 3*NMk*16 floats of this wave's LDS, V 128 floats.
template <int NC, int NG, int NGF>
__device__ void finish(const F32Args &A, floatx4 (&G)[3][S<NC, NG, NGF>::RBM], float *W,
                       float *V, int e0, int wallmode) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NMk = Sh::NMk, RBM = Sh::RBM;
  const int l = lane(), g = l >> 4, col = l & 15, e = e0 + col;
  const bool valid = e < A.mesh.ne;
  // wall elements of this block (one ballot bit per column)
  unsigned long long wm = __ballot(valid && g == 0 && A.mesh.bdl[e] != 0);
  if (wm) {
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
      for (int rb = 0; rb < RBM; rb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = rb * 16 + 4 * g + r;
          if (row < NMk) W[(f * NMk + row) * 16 + col] = G[f][rb][r];
        }
    wave_sync();
    while (wm) {
      int c = __ffsll((long long)wm) - 1;
      wm &= wm - 1;
      const float *gg = A.mesh.geo + (size_t)(e0 + c) * GEO;
      float nx = gg[14], ny = gg[15];
      float *gx = W + c, *gy = W + NMk * 16 + c;  // stride 16
      if (wallmode == 2) {  // no-slip: both components on the reduced basis
        if (l < NM) {
          V[l] = gx[l * 16];
          V[64 + l] = gy[l * 16];
        }
        wave_sync();
        if (l < NM) {
          float d1 = gl_apply<NM>(A.ops.GL, V, l), d2 = gl_apply<NM>(A.ops.GL, V + 64, l);
          gx[l * 16] += d1;
          gy[l * 16] += d2;
        }
      } else {  // free slip: normal component on the reduced basis
        if (l < NM) V[l] = nx * gx[l * 16] + ny * gy[l * 16];
        wave_sync();
        if (l < NM) {
          float d = gl_apply<NM>(A.ops.GL, V, l);
          gx[l * 16] += nx * d;
          gy[l * 16] += ny * d;
        }
      }
      wave_sync();
    }
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
      for (int rb = 0; rb < RBM; rb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = rb * 16 + 4 * g + r;
          if (row < NMk) G[f][rb][r] = W[(f * NMk + row) * 16 + col];
        }
    wave_sync();
  }
  // RK update; next stage input into W as [3][NMk][16] (FP64 in the hybrid
  // path, where W must hold 6*NMk*16 floats)
  double *W64 = reinterpret_cast<double *>(W);
#pragma unroll
  for (int f = 0; f < 3; f++)
#pragma unroll
    for (int rb = 0; rb < RBM; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = rb * 16 + 4 * g + r;
        float xf = 0.f;
        double x = 0.0;
        if (row < NM && valid) xf = rk_stage(A, f, (size_t)e0 * NM + row * 16 + col, G[f][rb][r], &x);
        if (row < NMk) {
          if (A.hybrid) W64[(f * NMk + row) * 16 + col] = x;
          else W[(f * NMk + row) * 16 + col] = xf;
        }
      }
  if (A.stage < 0) return;
  wave_sync();
  if (A.hybrid) emit_traces64<NC, NG, NGF>(A, W64, reinterpret_cast<double *>(V), e0, wallmode);
  else emit_traces<NC, NG, NGF>(A, W, e0, A.variant == 0);
}

// ---------------------------------------------------------------------------
// Linear RHS (src/calvar_lin.f:7-287) + RK4 stage.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF>
__global__ __launch_bounds__(NT) void lin_kernel(F32Args A) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NMk = Sh::NMk, NEk = Sh::NEk, KM = Sh::KM, KE = Sh::KE,
                NE3 = Sh::NE3, RBM = Sh::RBM;
  constexpr int WSZ = 16 * 3 * NMk + 128;
  __shared__ float lds[NW * WSZ];
  const int w = threadIdx.x >> 6, l = lane(), g = l >> 4, col = l & 15;
  float *X = lds + w * WSZ, *V = X + 3 * NMk * 16;
  const int ne = A.mesh.ne, e0 = (blockIdx.x * NW + w) * 16;
  if (e0 >= ne) return;  // whole wave idle: no block-level barriers follow
  const int e = e0 + col;
  const bool valid = e < ne;
  const float g0 = A.fld.g0, H = A.fld.hh0;
  // volume operands: -g h, H (by u - bx v), H (-ay u + ax v)
  for (int i = l; i < NMk * 16; i += 64) {
    int m = i >> 4, c = i & 15, ee = e0 + c;
    float xf = 0.f, xu = 0.f, xv = 0.f;
    if (m < NM && ee < ne) {
      const float *gg = A.mesh.geo + (size_t)ee * GEO;
      size_t o = (size_t)e0 * NM + i;
      float u = A.st.xs[0][o], v = A.st.xs[1][o], h = A.st.xs[2][o];
      xf = -g0 * h;
      xu = H * (gg[3] * u - gg[1] * v);
      xv = H * (-gg[2] * u + gg[0] * v);
    }
    X[i] = xf;
    X[NMk * 16 + i] = xu;
    X[2 * NMk * 16 + i] = xv;
  }
  const float *gg = A.mesh.geo + (size_t)(valid ? e : 0) * GEO;
  float ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.f;
  // edge operands as B fragments in registers: row kj = 4s + l/16 of this
  // lane's element (the factor 1/2 is the edge Jacobian of tri_f)
  float fbx[KE], fby[KE], fbh[KE];
#pragma unroll
  for (int s = 0; s < KE; s++) {
    int kj = 4 * s + g;
    fbx[s] = fby[s] = fbh[s] = 0.f;
    if (kj < NE3 && valid) {
      int k = kj / NGF, j = kj % NGF;
      float sg = gg[11 + k];
      float fh = -g0 * face_avg<NGF, 2>(A.st.trA, A.mesh, e, k, j, 0);
      float fs = face_avg<NGF, 2>(A.st.trA, A.mesh, e, k, j, 1);
      fbx[s] = 0.5f * fh * sg * gg[5 + k];
      fby[s] = 0.5f * fh * sg * gg[8 + k];
      fbh[s] = -0.5f * H * sg * fs;
    }
  }
  wave_sync();
  floatx4 G[3][RBM];
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    Frag<KM> adx, ady;
    Frag<KE> al;
    afrag<KM>(A.ops.DX, A.ops.DXl, NMk, rb, 0, adx);
    afrag<KM>(A.ops.DY, A.ops.DYl, NMk, rb, 0, ady);
    afrag<KE>(A.ops.LF, A.ops.LFl, NEk, rb, 0, al);
    floatx4 Ax = chain<KM>(adx, X, zero4()), Ay = chain<KM>(ady, X, zero4());
    floatx4 Lx = chain_r<KE>(al, fbx, zero4()), Ly = chain_r<KE>(al, fby, zero4());
    floatx4 Gh = chain<KM>(adx, X + NMk * 16, zero4());
    Gh = chain<KM>(ady, X + 2 * NMk * 16, Gh);
    Gh = chain_r<KE>(al, fbh, Gh);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      G[0][rb][r] = di * (-by * Ax[r] + ay * Ay[r] + Lx[r]);
      G[1][rb][r] = di * (bx * Ax[r] - ax * Ay[r] + Ly[r]);
      G[2][rb][r] = di * Gh[r];
    }
  }
  wave_sync();
  finish<NC, NG, NGF>(A, G, X, V, e0, 1);
}

// ---------------------------------------------------------------------------
// Nonlinear stages B + C (src/calvarnl_fs.f:115-332 and the ns/inv variants).
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void bc_kernel(F32Args A) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NMk = Sh::NMk, NEk = Sh::NEk, KM = Sh::KM, KE = Sh::KE,
                NE3 = Sh::NE3, RBM = Sh::RBM, RBG = Sh::RBG, RBE = Sh::RBE, NGr = Sh::NGr;
  constexpr int WSZ = 16 * (7 * NMk + 64) + 128;
  __shared__ float lds[NW * WSZ];
  const int w = threadIdx.x >> 6, l = lane(), g = l >> 4, col = l & 15;
  float *X = lds + w * WSZ;        // u, v, h          [3][NMk][16]
  float *Q = X + 3 * NMk * 16;     // qxx qxy qyx qyy  [4][NMk][16]
  float *P = Q + 4 * NMk * 16;     // products         [4][16][16]
  float *V = P + 4 * 256;          // wall scratch     [128]
  const int ne = A.mesh.ne, e0 = (blockIdx.x * NW + w) * 16;
  if (e0 >= ne) return;
  const int e = e0 + col;
  const bool valid = e < ne;
  const float mu0 = A.fld.mu0, g0 = A.fld.g0, mb0 = A.fld.mb0;
  for (int f = 0; f < 3; f++) load_block<NM, NMk>(X + f * NMk * 16, A.st.xs[f], e0);
  const float *gg = A.mesh.geo + (size_t)(valid ? e : 0) * GEO;
  const float ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.f;
  // velocity traces averaged over each face, times the outward normal / 2,
  // as B fragments in registers (row kj = 4s + l/16 of this lane's element)
  float fb[4][KE];
#pragma unroll
  for (int s = 0; s < KE; s++) {
    int kj = 4 * s + g;
    fb[0][s] = fb[1][s] = fb[2][s] = fb[3][s] = 0.f;
    if (kj < NE3 && valid) {
      int k = kj / NGF, j = kj % NGF;
      float sg = 0.5f * gg[11 + k], nx = sg * gg[5 + k], ny = sg * gg[8 + k];
      float fu = face_avg<NGF, 2>(A.st.trA, A.mesh, e, k, j, 0);
      float fv = face_avg<NGF, 2>(A.st.trA, A.mesh, e, k, j, 1);
      fb[0][s] = fu * nx;
      fb[1][s] = fu * ny;
      fb[2][s] = fv * nx;
      fb[3][s] = fv * ny;
    }
  }
  wave_sync();
  // B: q = weak gradient (src/calvarnl_fs.f:129-158); mass matrix = identity
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    Frag<KM> adx, ady;
    Frag<KE> al;
    afrag<KM>(A.ops.DX, A.ops.DXl, NMk, rb, 0, adx);
    afrag<KM>(A.ops.DY, A.ops.DYl, NMk, rb, 0, ady);
    afrag<KE>(A.ops.LF, A.ops.LFl, NEk, rb, 0, al);
#pragma unroll
    for (int f = 0; f < 2; f++) {
      // gx1 = -agx^T u, gy1 = -agy^T u
      floatx4 Ax = chain<KM>(adx, X + f * NMk * 16, zero4());
      floatx4 Ay = chain<KM>(ady, X + f * NMk * 16, zero4());
      floatx4 La = chain_r<KE>(al, fb[2 * f], zero4());
      floatx4 Lb = chain_r<KE>(al, fb[2 * f + 1], zero4());
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = rb * 16 + 4 * g + r;
        if (row < NMk) {
          Q[((2 * f) * NMk + row) * 16 + col] = di * (-by * Ax[r] + ay * Ay[r] + La[r]);
          Q[((2 * f + 1) * NMk + row) * 16 + col] = di * (bx * Ax[r] - ax * Ay[r] + Lb[r]);
        }
      }
    }
  }
  wave_sync();
  if (VAR == 1) {
    // free-slip wall projection of the tensor (src/calvarnl_fs.f:165-204):
    // only the t-n component goes through the reduced basis
    unsigned long long wm = __ballot(valid && g == 0 && A.mesh.bdl[e] != 0);
    while (wm) {
      int c = __ffsll((long long)wm) - 1;
      wm &= wm - 1;
      const float *gw = A.mesh.geo + (size_t)(e0 + c) * GEO;
      float nx = gw[14], ny = gw[15];
      float *qxx = Q + c, *qxy = Q + NMk * 16 + c, *qyx = Q + 2 * NMk * 16 + c,
            *qyy = Q + 3 * NMk * 16 + c;
      if (l < NM)
        V[l] = (qxx[16 * l] * ny - qyx[16 * l] * nx) * nx + (qxy[16 * l] * ny - qyy[16 * l] * nx) * ny;
      wave_sync();
      if (l < NM) {
        float d = gl_apply<NM>(A.ops.GL, V, l);
        qxx[16 * l] += nx * ny * d;
        qyx[16 * l] -= nx * nx * d;
        qxy[16 * l] += ny * ny * d;
        qyy[16 * l] -= nx * ny * d;
      }
      wave_sync();
    }
  }
  // D: F = mu0 q - g0 h, and its edge traces combined with the face normal
  {
    size_t o = (size_t)e0 * NM;
    for (int i = l; i < NM * 16; i += 64) {
      // the -g0 h part belongs to the FP64 linear core in the hybrid path
      float gh = A.hybrid ? 0.f : g0 * X[2 * NMk * 16 + i];
      A.st.F[0][o + i] = mu0 * Q[i] - gh;
      A.st.F[1][o + i] = mu0 * Q[NMk * 16 + i];
      A.st.F[2][o + i] = mu0 * Q[2 * NMk * 16 + i];
      A.st.F[3][o + i] = mu0 * Q[3 * NMk * 16 + i] - gh;
    }
  }
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    Frag<KM> at;
    afrag<KM>(A.ops.TR, A.ops.TRl, NMk, rb, 0, at);
    floatx4 T0 = chain<KM>(at, Q, zero4()), T1 = chain<KM>(at, Q + NMk * 16, zero4());
    floatx4 T2 = chain<KM>(at, Q + 2 * NMk * 16, zero4());
    floatx4 T3 = chain<KM>(at, Q + 3 * NMk * 16, zero4());
    floatx4 Th = chain<KM>(at, X + 2 * NMk * 16, zero4());
    if (valid) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = rb * 16 + 4 * g + r;
        if (row < NE3) {
          int k = row / NGF, j = row % NGF;
          float nx = gg[5 + k], ny = gg[8 + k];
          float *t = A.st.trS + ((size_t)e * 3 + k) * 3 * NGF;
          float gth = A.hybrid ? 0.f : g0 * Th[r];
          t[j] = (mu0 * T0[r] - gth) * nx + mu0 * T1[r] * ny;
          t[NGF + j] = mu0 * T2[r] * nx + (mu0 * T3[r] - gth) * ny;
        }
      }
    }
  }
  // C: values at the quadrature points, products, projection
  // (src/calvarnl_fs.f:245-296, src/nonlin.f:35-133)
  floatx4 Pr[4][RBM];
#pragma unroll
  for (int f = 0; f < 4; f++)
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) Pr[f][rb] = zero4();
  const float *fc3 = A.fld.fc3 + (size_t)(valid ? e : 0) * 3;
  const float fc0 = fc3[0], fc1 = fc3[1], fc2 = fc3[2];
  for (int qb = 0; qb < RBG; qb++) {
    Frag<KM> ap;
    afrag<KM>(A.ops.PSI, A.ops.PSIl, NMk, qb, 0, ap);
    floatx4 Vu = chain<KM>(ap, X, zero4()), Vv = chain<KM>(ap, X + NMk * 16, zero4());
    floatx4 Vh = chain<KM>(ap, X + 2 * NMk * 16, zero4());
    floatx4 Vxx = chain<KM>(ap, Q, zero4()), Vxy = chain<KM>(ap, Q + NMk * 16, zero4());
    floatx4 Vyx = chain<KM>(ap, Q + 2 * NMk * 16, zero4());
    floatx4 Vyy = chain<KM>(ap, Q + 3 * NMk * 16, zero4());
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int q = qb * 16 + 4 * g + r;
      float gx = 0.f, gy = 0.f, uh = 0.f, vh = 0.f;
      if (q < NG && valid) {
        size_t o = ((size_t)e0 * NG) + (size_t)q * 16 + col;
        float f = fc0 + fc1 * A.ops.yq[q] + fc2 * A.ops.xq[q];
        float hk = Vh[r] + A.fld.hbq[o];
        float u = Vu[r], v = Vv[r];
        gx = -(u * Vxx[r] + v * (Vxy[r] - f)) + (A.fld.w0q[o] - mb0 * u) / hk;
        gy = -(u * (Vyx[r] + f) + v * Vyy[r]) + (A.fld.z0q[o] - mb0 * v) / hk;
        // hybrid: the flux hh0 u is the FP64 core's; only u (h + hb - hh0) here
        float hf = A.hybrid ? Vh[r] + A.fld.hbdq[o] : hk;
        uh = u * hf;
        vh = v * hf;
      }
      int pr = (4 * g + r) * 16 + col;
      P[pr] = gx;
      P[256 + pr] = gy;
      P[512 + pr] = uh;
      P[768 + pr] = vh;
    }
    wave_sync();
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      Frag<4> aw;
      afrag<4>(A.ops.PSIW, A.ops.PSIWl, NGr, rb, 4 * qb, aw);
#pragma unroll
      for (int f = 0; f < 4; f++) Pr[f][rb] = chain<4>(aw, P + 256 * f, Pr[f][rb]);
    }
    wave_sync();
  }
  // outputs: sources, mass fluxes (wall-projected), combined flux traces
  float *UH = X, *VH = X + NMk * 16;
#pragma unroll
  for (int rb = 0; rb < RBM; rb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int row = rb * 16 + 4 * g + r;
      if (row < NMk) {
        UH[row * 16 + col] = Pr[2][rb][r];
        VH[row * 16 + col] = Pr[3][rb][r];
      }
      if (row < NM && valid) {
        size_t o = (size_t)e0 * NM + row * 16 + col;
        A.st.src[0][o] = Pr[0][rb][r];
        A.st.src[1][o] = Pr[1][rb][r];
      }
    }
  wave_sync();
  {
    unsigned long long wm = __ballot(valid && g == 0 && A.mesh.bdl[e] != 0);
    while (wm) {
      int c = __ffsll((long long)wm) - 1;
      wm &= wm - 1;
      const float *gw = A.mesh.geo + (size_t)(e0 + c) * GEO;
      float nx = gw[14], ny = gw[15];
      float *u = UH + c, *v = VH + c;
      if (VAR == 2) {  // src/calvarnl_ns.f:247-265
        if (l < NM) {
          V[l] = u[16 * l];
          V[64 + l] = v[16 * l];
        }
        wave_sync();
        if (l < NM) {
          float d1 = gl_apply<NM>(A.ops.GL, V, l), d2 = gl_apply<NM>(A.ops.GL, V + 64, l);
          u[16 * l] += d1;
          v[16 * l] += d2;
        }
      } else {  // src/calvarnl_fs.f:298-315
        if (l < NM) V[l] = nx * u[16 * l] + ny * v[16 * l];
        wave_sync();
        if (l < NM) {
          float d = gl_apply<NM>(A.ops.GL, V, l);
          u[16 * l] += nx * d;
          v[16 * l] += ny * d;
        }
      }
      wave_sync();
    }
  }
  {
    size_t o = (size_t)e0 * NM;
    for (int i = l; i < NM * 16; i += 64) {
      A.st.uh[0][o + i] = UH[i];
      A.st.uh[1][o + i] = VH[i];
    }
  }
#pragma unroll
  for (int rb = 0; rb < RBE; rb++) {
    Frag<KM> at;
    afrag<KM>(A.ops.TR, A.ops.TRl, NMk, rb, 0, at);
    floatx4 Tu = chain<KM>(at, UH, zero4()), Tv = chain<KM>(at, VH, zero4());
    if (valid) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = rb * 16 + 4 * g + r;
        if (row < NE3) {
          int k = row / NGF, j = row % NGF;
          A.st.trS[((size_t)e * 3 + k) * 3 * NGF + 2 * NGF + j] = Tu[r] * gg[5 + k] + Tv[r] * gg[8 + k];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Nonlinear stages E + G (src/calvarnl_fs.f:382-544) + RK4 stage.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void eg_kernel(F32Args A) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NMk = Sh::NMk, NEk = Sh::NEk, KM = Sh::KM, KE = Sh::KE,
                NE3 = Sh::NE3, RBM = Sh::RBM;
  constexpr int WSZ = 16 * 6 * NMk + 256;
  __shared__ __attribute__((aligned(16))) float lds[NW * WSZ];
  const int w = threadIdx.x >> 6, l = lane(), g = l >> 4, col = l & 15;
  float *X = lds + w * WSZ, *V = X + 6 * NMk * 16;
  const int ne = A.mesh.ne, e0 = (blockIdx.x * NW + w) * 16;
  if (e0 >= ne) return;
  const int e = e0 + col;
  const bool valid = e < ne;
  // volume operands with the Jacobian terms folded in:
  //   u: -by Fxx + bx Fxy (with agx), ay Fxx - ax Fxy (with agy); v likewise;
  //   eta: by uh - bx vh (agx), -ay uh + ax vh (agy)
  for (int i = l; i < NMk * 16; i += 64) {
    int m = i >> 4, c = i & 15, ee = e0 + c;
    float o0 = 0.f, o1 = 0.f, o2 = 0.f, o3 = 0.f, o4 = 0.f, o5 = 0.f;
    if (m < NM && ee < ne) {
      const float *gg = A.mesh.geo + (size_t)ee * GEO;
      float ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3];
      size_t o = (size_t)e0 * NM + i;
      float fxx = A.st.F[0][o], fxy = A.st.F[1][o], fyx = A.st.F[2][o], fyy = A.st.F[3][o];
      float uh = A.st.uh[0][o], vh = A.st.uh[1][o];
      o0 = -by * fxx + bx * fxy;
      o1 = ay * fxx - ax * fxy;
      o2 = -by * fyx + bx * fyy;
      o3 = ay * fyx - ax * fyy;
      o4 = by * uh - bx * vh;
      o5 = -ay * uh + ax * vh;
    }
    X[i] = o0;
    X[NMk * 16 + i] = o1;
    X[2 * NMk * 16 + i] = o2;
    X[3 * NMk * 16 + i] = o3;
    X[4 * NMk * 16 + i] = o4;
    X[5 * NMk * 16 + i] = o5;
  }
  // combined flux traces averaged over each face, times its / 2, as B
  // fragments in registers; the eta equation's lift carries -1/det
  float fb[3][KE];
  const float *gg = A.mesh.geo + (size_t)(valid ? e : 0) * GEO;
#pragma unroll
  for (int s = 0; s < KE; s++) {
    int kj = 4 * s + g;
    fb[0][s] = fb[1][s] = fb[2][s] = 0.f;
    if (kj < NE3 && valid) {
      int k = kj / NGF, j = kj % NGF;
      float sg = 0.5f * gg[11 + k];
      fb[0][s] = sg * face_avg<NGF, 3>(A.st.trS, A.mesh, e, k, j, 0);
      fb[1][s] = sg * face_avg<NGF, 3>(A.st.trS, A.mesh, e, k, j, 1);
      fb[2][s] = -sg * face_avg<NGF, 3>(A.st.trS, A.mesh, e, k, j, 2);
    }
  }
  wave_sync();
  const float di = valid ? gg[4] : 0.f;
  floatx4 G[3][RBM];
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    Frag<KM> adx, ady;
    Frag<KE> al;
    afrag<KM>(A.ops.DX, A.ops.DXl, NMk, rb, 0, adx);
    afrag<KM>(A.ops.DY, A.ops.DYl, NMk, rb, 0, ady);
    afrag<KE>(A.ops.LF, A.ops.LFl, NEk, rb, 0, al);
#pragma unroll
    for (int f = 0; f < 3; f++) {
      floatx4 acc = chain<KM>(adx, X + (2 * f) * NMk * 16, zero4());
      acc = chain<KM>(ady, X + (2 * f + 1) * NMk * 16, acc);
      acc = chain_r<KE>(al, fb[f], acc);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = rb * 16 + 4 * g + r;
        float src = 0.f;
        if (f < 2 && row < NM && valid) src = A.st.src[f][(size_t)e0 * NM + row * 16 + col];
        G[f][rb][r] = src + di * acc[r];
      }
    }
  }
  wave_sync();
  finish<NC, NG, NGF>(A, G, X, V, e0, VAR == 2 ? 2 : 1);
}

// Stage input -> FP32 shadow and its traces (before the first stage).
template <int NC, int NG, int NGF>
__global__ __launch_bounds__(NT) void prep_kernel(F32Args A) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NMk = Sh::NMk;
  constexpr int WSZ = 6 * NMk * 16 + 256;
  __shared__ __attribute__((aligned(16))) float lds[NW * WSZ];
  const int w = threadIdx.x >> 6, l = lane();
  float *X = lds + w * WSZ;
  double *X64 = reinterpret_cast<double *>(X);
  const int ne = A.mesh.ne, e0 = (blockIdx.x * NW + w) * 16;
  if (e0 >= ne) return;
  for (int f = 0; f < 3; f++)
    for (int i = l; i < NMk * 16; i += 64) {
      int m = i >> 4, c = i & 15;
      double v = 0.0;
      if (m < NM && e0 + c < ne) {
        size_t o = (size_t)e0 * NM + i;
        v = A.st.c0[f][o];
        A.st.xs[f][o] = (float)v;
        if (A.hybrid) A.st.x64[f][o] = v;
      }
      if (A.hybrid) X64[f * NMk * 16 + i] = v;
      else X[f * NMk * 16 + i] = (float)v;
    }
  wave_sync();
  if (A.hybrid)
    emit_traces64<NC, NG, NGF>(A, X64, reinterpret_cast<double *>(X + 6 * NMk * 16), e0,
                               A.variant == 2 ? 2 : 1);
  else
    emit_traces<NC, NG, NGF>(A, X, e0, A.variant == 0);
}

// ---------------------------------------------------------------------------
// Linear gravity-wave core of the nonlinear variants in FP64 on MFMA
// (hybrid path): the -g grad(eta) terms of the u, v equations and the
// -div(hh0 u) term of the eta equation, with the reference's wall
// projections (src/calvarnl_fs.f:298-321, 436-471; src/calvarnl_ns.f:247-265,
// 403-421).  In psi coordinates these are exact operators, no quadrature.
// Writes kL, which the FP32 element kernel adds before the RK update.
// ---------------------------------------------------------------------------
template <int NC, int NG, int NGF, int VAR>
__global__ __launch_bounds__(NT) void core64_kernel(F32Args A) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM, NMk = Sh::NMk, NEk = Sh::NEk, KM = Sh::KM, KE = Sh::KE,
                NE3 = Sh::NE3, RBM = Sh::RBM;
  constexpr int WSZ = 3 * NMk * 16 + 128;
  __shared__ double lds[NW * WSZ];
  const int w = threadIdx.x >> 6, l = lane(), g = l >> 4, col = l & 15;
  double *X = lds + w * WSZ, *V = X + 3 * NMk * 16;
  const int ne = A.mesh.ne, e0 = (blockIdx.x * NW + w) * 16;
  if (e0 >= ne) return;
  const int e = e0 + col;
  const bool valid = e < ne;
  const double g0 = A.fld.g0d, hh0 = A.fld.hh0d;
  const int wallmode = VAR == 2 ? 2 : 1;
  for (int f = 0; f < 3; f++)
    for (int i = l; i < NMk * 16; i += 64) {
      int m = i >> 4, c = i & 15;
      double v = 0.0;
      if (m < NM && e0 + c < ne) v = A.st.x64[f][(size_t)e0 * NM + i];
      X[f * NMk * 16 + i] = f < 2 ? hh0 * v : v;
    }
  wave_sync();
  // mass flux hh0 (u, v), wall-projected
  unsigned long long wm = __ballot(valid && g == 0 && A.mesh.bdl[e] != 0);
  {
    unsigned long long m2 = wm;
    while (m2) {
      int c = __ffsll((long long)m2) - 1;
      m2 &= m2 - 1;
      const double *gw = A.mesh.geo64 + (size_t)(e0 + c) * GEO;
      wall_project64<NM>(A.ops.GL64, X + c, X + NMk * 16 + c, gw[14], gw[15], wallmode, V);
    }
  }
  // volume operands: by uL - bx vL (agx), -ay uL + ax vL (agy), -g0 h
  for (int i = l; i < NMk * 16; i += 64) {
    int c = i & 15;
    const double *gc = A.mesh.geo64 + (size_t)(e0 + c < ne ? e0 + c : e0) * GEO;
    double uL = X[i], vL = X[NMk * 16 + i];
    X[i] = gc[3] * uL - gc[1] * vL;
    X[NMk * 16 + i] = -gc[2] * uL + gc[0] * vL;
    X[2 * NMk * 16 + i] = -g0 * X[2 * NMk * 16 + i];
  }
  wave_sync();
  const double *gg = A.mesh.geo64 + (size_t)(valid ? e : 0) * GEO;
  const double ax = gg[0], bx = gg[1], ay = gg[2], by = gg[3], di = valid ? gg[4] : 0.0;
  double fbx[KE], fby[KE], fbh[KE];
#pragma unroll
  for (int s = 0; s < KE; s++) {
    int kj = 4 * s + g;
    fbx[s] = fby[s] = fbh[s] = 0.0;
    if (kj < NE3 && valid) {
      int k = kj / NGF, j = kj % NGF;
      double sg = gg[11 + k];
      double fh = -g0 * face_avg64<NGF>(A.st.trL, A.mesh, e, k, j, 0);
      double fs = face_avg64<NGF>(A.st.trL, A.mesh, e, k, j, 1);
      fbx[s] = 0.5 * fh * sg * gg[5 + k];
      fby[s] = 0.5 * fh * sg * gg[8 + k];
      fbh[s] = -0.5 * sg * fs;
    }
  }
  doublex4 G[3][RBM];
#pragma unroll
  for (int rb = 0; rb < RBM; rb++) {
    double adx[KM], ady[KM], al[KE];
    afrag64<KM>(A.ops.DX64, NMk, rb, adx);
    afrag64<KM>(A.ops.DY64, NMk, rb, ady);
    afrag64<KE>(A.ops.LF64, NEk, rb, al);
    doublex4 Ax = chain64<KM>(adx, X + 2 * NMk * 16, zero4d());
    doublex4 Ay = chain64<KM>(ady, X + 2 * NMk * 16, zero4d());
    doublex4 Lx = chain64_r<KE>(al, fbx, zero4d()), Ly = chain64_r<KE>(al, fby, zero4d());
    doublex4 Gh = chain64<KM>(adx, X, zero4d());
    Gh = chain64<KM>(ady, X + NMk * 16, Gh);
    Gh = chain64_r<KE>(al, fbh, Gh);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      G[0][rb][r] = di * (-by * Ax[r] + ay * Ay[r] + Lx[r]);
      G[1][rb][r] = di * (bx * Ax[r] - ax * Ay[r] + Ly[r]);
      G[2][rb][r] = di * Gh[r];
    }
  }
  if (wm) {
    wave_sync();
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
      for (int rb = 0; rb < RBM; rb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = drow64(rb, g, r);
          if (row < NMk) X[(f * NMk + row) * 16 + col] = G[f][rb][r];
        }
    wave_sync();
    while (wm) {
      int c = __ffsll((long long)wm) - 1;
      wm &= wm - 1;
      const double *gw = A.mesh.geo64 + (size_t)(e0 + c) * GEO;
      wall_project64<NM>(A.ops.GL64, X + c, X + NMk * 16 + c, gw[14], gw[15], wallmode, V);
    }
#pragma unroll
    for (int f = 0; f < 2; f++)
#pragma unroll
      for (int rb = 0; rb < RBM; rb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int row = drow64(rb, g, r);
          if (row < NMk) G[f][rb][r] = X[(f * NMk + row) * 16 + col];
        }
  }
  if (!valid) return;
#pragma unroll
  for (int f = 0; f < 3; f++)
#pragma unroll
    for (int rb = 0; rb < RBM; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = drow64(rb, g, r);
        if (row < NM) A.st.kL[f][(size_t)e0 * NM + row * 16 + col] = G[f][rb][r];
      }
}

// Basis change, FP64.  to_psi: c = U a from the reference layout a[e*NM + m]
// into the blocked layout; from_psi: a = U^{-1} c by back substitution.
template <int NM>
__global__ void to_psi_kernel(int ne, const double *__restrict__ U, const double *__restrict__ a,
                              double *__restrict__ c) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * NM) return;
  int e = t / NM, i = t % NM;
  double s = 0.0;
  for (int k = i; k < NM; k++) s += U[i + NM * k] * a[(size_t)e * NM + k];
  c[(size_t)(e & ~15) * NM + i * 16 + (e & 15)] = s;
}

template <int NM>
__global__ void from_psi_kernel(int ne, const double *__restrict__ U, const double *__restrict__ c,
                                double *__restrict__ a) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double x[NM];
  const double *cb = c + (size_t)(e & ~15) * NM + (e & 15);
#pragma unroll
  for (int i = 0; i < NM; i++) x[i] = cb[i * 16];
#pragma unroll
  for (int i = NM - 1; i >= 0; i--) {
    double s = x[i];
#pragma unroll
    for (int k = i + 1; k < NM; k++) s -= U[i + NM * k] * x[k];
    x[i] = s / U[i + NM * i];
  }
#pragma unroll
  for (int i = 0; i < NM; i++) a[(size_t)e * NM + i] = x[i];
}

template <int NC, int NG, int NGF>
int launch(const F32Args &A, int phase, const double *in, double *out) {
  using Sh = S<NC, NG, NGF>;
  constexpr int NM = Sh::NM;
  const int ne = A.mesh.ne;
  const int nb = (ne + 16 * NW - 1) / (16 * NW);
  hipStream_t st = A.stream;
  switch (phase) {
    case 0:  // to psi
      to_psi_kernel<NM><<<(ne * NM + 255) / 256, 256, 0, st>>>(ne, A.ops.U, in, out);
      break;
    case 1:  // from psi
      from_psi_kernel<NM><<<(ne + 127) / 128, 128, 0, st>>>(ne, A.ops.U, in, out);
      break;
    case 2:
      prep_kernel<NC, NG, NGF><<<nb, NT, 0, st>>>(A);
      break;
    case 3:  // linear right-hand side
      lin_kernel<NC, NG, NGF><<<nb, NT, 0, st>>>(A);
      break;
    case 4:  // nonlinear B+C
      if (A.variant == 1) bc_kernel<NC, NG, NGF, 1><<<nb, NT, 0, st>>>(A);
      else if (A.variant == 2) bc_kernel<NC, NG, NGF, 2><<<nb, NT, 0, st>>>(A);
      else bc_kernel<NC, NG, NGF, 3><<<nb, NT, 0, st>>>(A);
      break;
    case 6:  // FP64 linear core (hybrid nonlinear path)
      if (A.variant == 1) core64_kernel<NC, NG, NGF, 1><<<nb, NT, 0, st>>>(A);
      else if (A.variant == 2) core64_kernel<NC, NG, NGF, 2><<<nb, NT, 0, st>>>(A);
      else core64_kernel<NC, NG, NGF, 3><<<nb, NT, 0, st>>>(A);
      break;
    case 5:  // nonlinear E+G
      if (A.variant == 1) eg_kernel<NC, NG, NGF, 1><<<nb, NT, 0, st>>>(A);
      else if (A.variant == 2) eg_kernel<NC, NG, NGF, 2><<<nb, NT, 0, st>>>(A);
      else eg_kernel<NC, NG, NGF, 3><<<nb, NT, 0, st>>>(A);
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
int SPOC_NAME4(rhs_f32, SPOC_NC, SPOC_NG, SPOC_NGF)(const F32Args &a, int phase, const double *in,
                                                  double *out) {
  return f32::launch<SPOC_NC, SPOC_NG, SPOC_NGF>(a, phase, in, out);
}
#endif

}  // namespace spoc
