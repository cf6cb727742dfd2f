// Device helpers shared by the MFMA kernels (rhs_f32.hip, adjoint.hip):
// operator fragments, products, element-blocked loads, edge traces, the
// XCD-contiguous block mapping and the masked wall projection.  Layouts are
// described in f32_types.h.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"
#include "f32_types.h"

namespace spoc {
namespace f32 {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef double doublex2 __attribute__((ext_vector_type(2)));
constexpr int NW = 4, NT = 64 * NW;  // waves (element blocks) per workgroup
constexpr int PSI_WG = 64;            // elements per workgroup of the basis changes

template <int NC, int NG, int NGF>
struct S {
  static constexpr int NM = nm_of(NC), NE3 = 3 * NGF;
  static constexpr int KM = cdiv(NM, 4), KE = cdiv(NE3, 4), KG = cdiv(NG, 4);     // k-steps
  static constexpr int RBM = cdiv(NM, 16), RBE = cdiv(NE3, 16), RBG = cdiv(NG, 16);  // row blocks
  static constexpr Pack32 P32 = pack32(NM, NG, NE3);
  static constexpr Pack64 P64 = pack64(NM, NE3);
};

__device__ __forceinline__ int lane() { return threadIdx.x & 63; }

__device__ __forceinline__ floatx4 mfma32(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma64(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ floatx4 zero32() { return floatx4{0.f, 0.f, 0.f, 0.f}; }
__device__ __forceinline__ doublex4 zero64() { return doublex4{0.0, 0.0, 0.0, 0.0}; }

// Keep the scheduler from interleaving the phases of a kernel: hoisting the
// loads of a later phase above an earlier one only raises register
// pressure (and with it lowers occupancy).
__device__ __forceinline__ void phase_fence() { __builtin_amdgcn_sched_barrier(0); }

// Mode (or row) held by register r of row block rb in lane group g.
__device__ __forceinline__ int drow(int rb, int r, int g) { return 16 * rb + 4 * r + g; }

// A fragments, k-steps 0 .. ceil(K/4)-1 of row block rb.
template <int K>
__device__ __forceinline__ void afrag32(const float *img, int rb, float (&a)[cdiv(K, 4)]) {
  constexpr int KS = cdiv(K, 4), KQ = cdiv(K, 16);
  const floatx4 *p = reinterpret_cast<const floatx4 *>(img) + (size_t)rb * KQ * 64 + lane();
#pragma unroll
  for (int q = 0; q < KQ; q++) {
    floatx4 v = p[q * 64];
#pragma unroll
    for (int t = 0; t < 4; t++)
      if (4 * q + t < KS) a[4 * q + t] = v[t];
  }
}
template <int K>
__device__ __forceinline__ void afrag64(const double *img, int rb, double (&a)[cdiv(K, 4)]) {
  constexpr int KS = cdiv(K, 4), KH = cdiv(K, 8);
  const doublex2 *p = reinterpret_cast<const doublex2 *>(img) + (size_t)rb * KH * 64 + lane();
#pragma unroll
  for (int h = 0; h < KH; h++) {
    doublex2 v = p[h * 64];
#pragma unroll
    for (int t = 0; t < 2; t++)
      if (2 * h + t < KS) a[2 * h + t] = v[t];
  }
}

template <int KS>
__device__ __forceinline__ floatx4 chain32(const float (&a)[KS], const float (&b)[KS], floatx4 acc) {
#pragma unroll
  for (int s = 0; s < KS; s++) acc = mfma32(a[s], b[s], acc);
  return acc;
}
template <int KS>
__device__ __forceinline__ doublex4 chain64(const double (&a)[KS], const double (&b)[KS],
                                            doublex4 acc) {
#pragma unroll
  for (int s = 0; s < KS; s++) acc = mfma64(a[s], b[s], acc);
  return acc;
}

// Result fragments (row blocks) -> B fragments (k-steps).
template <int KS, int RB>
__device__ __forceinline__ void d2b(const floatx4 (&D)[RB], float (&b)[KS]) {
  static_assert(KS <= 4 * RB, "k-steps exceed the rows");
#pragma unroll
  for (int s = 0; s < KS; s++) b[s] = D[s >> 2][s & 3];
}

// B fragments of an element-blocked [rows][16] array (rows >= R and
// elements past the mesh read as 0).  The loads themselves are
// unconditional (row clamped to R - 1; the blocked arrays are padded to
// whole blocks of 16 elements), so they issue back to back with no branch.
template <int R, int KS, typename T, typename V>
__device__ __forceinline__ void bload(const T *base, int g, int col, bool valid, V (&b)[KS]) {
#pragma unroll
  for (int s = 0; s < KS; s++) {
    const int m = 4 * s + g, mc = m < R ? m : R - 1;
    const T v = base[mc * 16 + col];
    b[s] = (valid && m < R) ? (V)v : V(0);
  }
}

// Offset of row kj of field 0 of element e in a blocked trace array with NF
// fields ([block][NF][NE3][16]); field f adds f * NE3 * 16.
template <int NE3, int NF>
__device__ __forceinline__ size_t tr_off(int e, int kj) {
  return ((size_t)(e >> 4) * (NF * NE3) + kj) * 16 + (e & 15);
}
template <int NE3, int NF, typename T>
__device__ __forceinline__ T tr_at(const T *tr, int e, int f, int kj) {
  return tr[tr_off<NE3, NF>(e, kj) + (size_t)f * NE3 * 16];
}
template <int NE3, int NF, typename T>
__device__ __forceinline__ void tr_put(T *tr, int eb, int col, int f, int kj, T v) {
  tr[((size_t)eb * (NF * NE3) + f * NE3 + kj) * 16 + col] = v;
}

// Average of own and neighbour traces of element e, local edge k, point j
// (local orientation), field f.  Walls use the element's own trace on its
// local edge 1, like the reference's wall faces (src/calflu.f:31-39).
template <int NGF, int NF, typename T>
__device__ __forceinline__ T face_avg(const T *__restrict__ tr, const F32Mesh &M, int e, int k,
                                      int j, int f) {
  constexpr int NE3 = 3 * NGF;
  int nb = M.nbr[e * 3 + k];
  if (nb < 0) return tr_at<NE3, NF>(tr, e, f, j);
  int kn = M.nbe[e * 3 + k];
  T a = tr_at<NE3, NF>(tr, e, f, k * NGF + j);
  T b = tr_at<NE3, NF>(tr, nb, f, kn * NGF + (NGF - 1 - j));
  return T(0.5) * (a + b);
}

// The same average without branches, for the hot kernels: the neighbours of
// the lane's element are loaded once (nbr3), each edge point resolves to the
// two (element, row) pairs it averages (face_pt), and face_mean reads them
// from a trace array.  A wall point averages the element's own trace on its
// local edge 1 with itself: 0.5 (t + t) == t exactly.
struct Nbr3 {
  int nb0, nb1, nb2;  // facing element across local edge k (negative: wall)
  int kn0, kn1, kn2;  // its local edge
};
// (each value passes through an empty asm, which keeps it a register value:
// otherwise a select between them can be folded into a load from a stack
// copy at a runtime offset, i.e. scratch memory)
__device__ __forceinline__ int in_reg(int v) {
  asm("" : "+v"(v));
  return v;
}
__device__ __forceinline__ Nbr3 nbr3(const F32Mesh &M, int e) {
  const int *n = M.nbr + e * 3, *k = M.nbe + e * 3;
  return Nbr3{n[0], n[1], n[2], k[0], k[1], k[2]};
}
struct FacePt {
  int ea, ka, eb, kb;
};
template <int NGF>
__device__ __forceinline__ FacePt face_pt(const Nbr3 &N, int e, int kj) {
  const int k = kj / NGF, j = kj - k * NGF;
  const int n0 = in_reg(N.nb0), n1 = in_reg(N.nb1), n2 = in_reg(N.nb2);
  const int m0 = in_reg(N.kn0), m1 = in_reg(N.kn1), m2 = in_reg(N.kn2);
  const int nb = k == 0 ? n0 : (k == 1 ? n1 : n2);
  const int kn = k == 0 ? m0 : (k == 1 ? m1 : m2);
  const bool wall = nb < 0;
  return {e, wall ? j : kj, wall ? e : nb, wall ? j : kn * NGF + (NGF - 1 - j)};
}
template <int NGF, int NF, typename T>
__device__ __forceinline__ T face_mean(const T *__restrict__ tr, const FacePt &p, int f) {
  constexpr int NE3 = 3 * NGF;
  return T(0.5) * (tr_at<NE3, NF>(tr, p.ea, f, p.ka) + tr_at<NE3, NF>(tr, p.eb, f, p.kb));
}

// Element block of this wave.  The workgroups on XCD x (blockIdx % 8, the
// hardware's round-robin placement) take one contiguous eighth of the
// blocks, so the neighbour traces they read were written through the same L2.
template <int W = NW>
__device__ __forceinline__ int first_block() {
  int g8 = gridDim.x >> 3;
  return ((blockIdx.x & 7) * g8 + (blockIdx.x >> 3)) * W;
}

// The same ranges, each swept from its end: workgroup j of an XCD takes the
// j-th group of blocks counted from the end of that XCD's range.
template <int W = NW>
__device__ __forceinline__ int last_block() {
  int g8 = gridDim.x >> 3;
  return ((blockIdx.x & 7) * g8 + (g8 - 1 - (blockIdx.x >> 3))) * W;
}

// Element block of wave threadIdx.x / 64 of a workgroup whose first block is
// b0, as a scalar (the compiler cannot see that it is wave-uniform): the
// addresses of the element's rows then split into a scalar base and a
// 32-bit lane offset.
__device__ __forceinline__ int wave_block(int b0) {
  return __builtin_amdgcn_readfirstlane(b0 + (int)(threadIdx.x >> 6));
}

// Copy n 16-byte words of operator images to LDS (whole workgroup), eight
// loads in flight per thread before their stores.
template <int W = NW>
__device__ __forceinline__ void stage_lds(floatx4 *dst, const floatx4 *__restrict__ src, int n) {
  constexpr int T = 64 * W, B = 8;
  for (int i0 = threadIdx.x; i0 < n; i0 += B * T) {
    floatx4 v[B];
    // unconditional loads (the index clamped to the last word), so that all
    // B are in flight together (conditional ones were issued one after the
    // other); the stores are masked
#pragma unroll
    for (int j = 0; j < B; j++) v[j] = src[min(i0 + j * T, n - 1)];
#pragma unroll
    for (int j = 0; j < B; j++)
      if (i0 + j * T < n) dst[i0 + j * T] = v[j];
  }
}

// The element's face normals (x and y components for local edges 1..3,
// geometry words 5..10), read as two 16-byte loads, and the pick of one by
// local edge with selects.
struct Nrm3 {
  float x[3], y[3];
};
__device__ __forceinline__ Nrm3 normals(const float *gg) {
  const floatx4 a = *reinterpret_cast<const floatx4 *>(gg + 4);
  const floatx4 b = *reinterpret_cast<const floatx4 *>(gg + 8);
  return Nrm3{{a[1], a[2], a[3]}, {b[0], b[1], b[2]}};
}
__device__ __forceinline__ float pick3(const float (&v)[3], int k) {
  return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

// ---------------------------------------------------------------------------
// FP64 pieces shared by the linear and the nonlinear paths.
// ---------------------------------------------------------------------------

// Wall projection of a vector pair (a, b) held as B fragments of this wave's
// 16 elements, in place: free slip (NS == false) puts the normal component
// on the reduced basis (src/calvarnl_fs.f:298-315), no slip (NS) both
// components (src/calvarnl_ns.f:247-265).  G - I is applied by MFMA to the
// columns of wall elements only (the others are masked to zero).
template <int NM, bool NS>
__device__ __forceinline__ void wall_pair64(const double *__restrict__ gl, bool wall, double nx,
                                            double ny, double (&a)[cdiv(NM, 4)],
                                            double (&b)[cdiv(NM, 4)]) {
  constexpr int KM = cdiv(NM, 4), RBM = cdiv(NM, 16);
  if (NS) {
    double ma[KM], mb[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) {
      ma[s] = wall ? a[s] : 0.0;
      mb[s] = wall ? b[s] : 0.0;
    }
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double ag[KM];
      afrag64<NM>(gl, rb, ag);
      doublex4 d1 = chain64(ag, ma, zero64()), d2 = chain64(ag, mb, zero64());
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (4 * rb + r < KM) {
          a[4 * rb + r] += d1[r];
          b[4 * rb + r] += d2[r];
        }
    }
  } else {
    double v[KM];
#pragma unroll
    for (int s = 0; s < KM; s++) v[s] = wall ? nx * a[s] + ny * b[s] : 0.0;
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) {
      double ag[KM];
      afrag64<NM>(gl, rb, ag);
      doublex4 d = chain64(ag, v, zero64());
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (4 * rb + r < KM) {
          a[4 * rb + r] += nx * d[r];
          b[4 * rb + r] += ny * d[r];
        }
    }
  }
}

}  // namespace f32
}  // namespace spoc
