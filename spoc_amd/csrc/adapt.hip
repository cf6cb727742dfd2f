// Refinement indicator of the adaptive driver on the device: checkconv and
// cal_conv (src/adapt.f:1260-1332, 1122-1192), FP64.
//
// The decisions are threshold tests on the indicator, so the indicator must
// equal the reference's bit for bit.  The reference's face loop computes,
// for face i with elements k1, k2 on local sides c1, c2, the traces
// fbdy1(j) = sum_m a(m,k1) pico(j,m,c1) and fbdy2(j) = sum_m a(m,k2)
// pico(ngf+1-j,m,c2), each summed over m in order with separate multiplies
// and adds, and |fbdy2(j) - fbdy1(j)|.  Those sums are element quantities,
// the trace T(e,k,j) of element e on its side k at point j: a first pass
// forms every T in the reference's order (this file is compiled without
// FMA contraction), a second takes, for each element, the differences of
// its faces in the face's own orientation, so both elements of a face see
// the same differences.  Every maximum uses the reference's "replace if
// smaller" test, whose
// result does not depend on the order, so maxima over a mesh are reduced
// in chunks of consecutive elements (each inside one ensemble member), then
// per member.
#include <hip/hip_runtime.h>

#include "common.h"

namespace spoc {
namespace adapt {

constexpr int CHUNK = 2048;  // elements per reduction chunk (host builds the list)
constexpr int BT = 256;      // threads per workgroup

__device__ __forceinline__ int member_of(const int *__restrict__ off, int nmem, int e) {
  int lo = 0, hi = nmem;  // off[lo] <= e < off[lo + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Largest of the workgroup's values (every thread gets it).
__device__ double block_max(double v, double *lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int s = BT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = lds[threadIdx.x + s];
      if (lds[threadIdx.x] < o) lds[threadIdx.x] = o;
    }
    __syncthreads();
  }
  const double r = lds[0];
  __syncthreads();
  return r;
}

// part[f * nch + c] = max |a_f(1, e)| over chunk c (src/adapt.f:1317-1320).
__global__ __launch_bounds__(BT) void chunk_vmax(int nm, int nch, const int *__restrict__ cst,
                                                 const double *__restrict__ a0,
                                                 const double *__restrict__ a1,
                                                 const double *__restrict__ a2,
                                                 double *__restrict__ part) {
  __shared__ double lds[BT];
  const int c = blockIdx.x, e0 = cst[c], e1 = cst[c + 1];
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const double *a = f == 0 ? a0 : (f == 1 ? a1 : a2);
    double mx = 0.0;
    for (int e = e0 + threadIdx.x; e < e1; e += BT) {
      const double v = fabs(a[(size_t)e * nm]);
      if (mx < v) mx = v;
    }
    mx = block_max(mx, lds);
    if (threadIdx.x == 0) part[(size_t)f * nch + c] = mx;
  }
}

// out[f * nmem + r] = max of part[f * nch + c] over member r's chunks
// cof[r] .. cof[r + 1] - 1, f < nf.
__global__ void member_max(int nmem, int nf, int nch, const int *__restrict__ cof,
                           const double *__restrict__ part, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nmem * nf) return;
  const int r = t % nmem, f = t / nmem;
  double mx = 0.0;
  for (int c = cof[r]; c < cof[r + 1]; c++) {
    const double v = part[(size_t)f * nch + c];
    if (mx < v) mx = v;
  }
  out[(size_t)f * nmem + r] = mx;
}

// T[(e 3 + f) NE3 + k NGF + j] = trace of field f of element e on its local
// side k at point j, summed over the modes in order (the sums fbdy1 and,
// at the reversed point, fbdy2 of src/adapt.f:1289-1302).  One thread per
// (element, side, point), EPB elements per workgroup; the elements'
// coefficients and pico(j, m, k) are staged in LDS first.
template <int NM, int NGF>
__global__ __launch_bounds__(BT) void traces(int ne, const double *__restrict__ a0,
                                             const double *__restrict__ a1,
                                             const double *__restrict__ a2,
                                             const double *__restrict__ pico,
                                             double *__restrict__ T) {
  constexpr int NE3 = 3 * NGF, EPB = BT / NE3;
  __shared__ double P[NGF * NM * 3];  // [j + NGF (m + NM k)]
  __shared__ double X[3][EPB * NM];
  const int eb = blockIdx.x * EPB, nel = min(EPB, ne - eb);
  for (int i = threadIdx.x; i < NGF * NM * 3; i += BT) P[i] = pico[i];
  for (int i = threadIdx.x; i < nel * NM; i += BT) {
    X[0][i] = a0[(size_t)eb * NM + i];
    X[1][i] = a1[(size_t)eb * NM + i];
    X[2][i] = a2[(size_t)eb * NM + i];
  }
  __syncthreads();
  const int le = threadIdx.x / NE3, kj = threadIdx.x % NE3;
  if (le >= nel) return;
  const int k = kj / NGF, j = kj - k * NGF;
  const double *p = P + j + NGF * NM * k;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int m = 0; m < NM; m++) {
    const double pm = p[NGF * m];
    s0 = s0 + X[0][le * NM + m] * pm;
    s1 = s1 + X[1][le * NM + m] * pm;
    s2 = s2 + X[2][le * NM + m] * pm;
  }
  double *t = T + (size_t)(eb + le) * 3 * NE3 + kj;
  t[0] = s0;
  t[NE3] = s1;
  t[2 * NE3] = s2;
}

// One thread per (element, local side k, point j), EPB elements per
// workgroup: |fbdy2 - fbdy1| at point j of the face on side k for the three
// fields (0 on a wall); the element's first thread takes the largest over
// its 3 NGF threads from 0 (src/adapt.f:1275-1313), divides by its
// member's vmax (0 when that is <= 1e-6; src/adapt.f:1321-1329) and, with
// decide, sets the flags (src/adapt.f:1159-1176).
template <int NM, int NGF>
__global__ __launch_bounds__(BT) void indicator(
    int ne, int nmem, const int *__restrict__ off, const int4 *__restrict__ side,
    const double *__restrict__ T, const double *__restrict__ vmax, double *__restrict__ amax,
    int decide, double fac1, double fac2, int *__restrict__ fconv, int *__restrict__ fderaff) {
  constexpr int NE3 = 3 * NGF, EPB = BT / NE3;
  __shared__ double D[EPB * NE3][3];
  const int le = threadIdx.x / NE3, kj = threadIdx.x % NE3;
  const int e = blockIdx.x * EPB + le;
  const bool live = le < EPB && e < ne;
  if (live) {
    const int k = kj / NGF, j = kj - k * NGF;
    const int4 s = side[3 * e + k];  // the face's (k1, c1, k2, c2), 0-based; k2 < 0: wall
    double dif[3] = {0.0, 0.0, 0.0};
    if (s.z >= 0) {
      const double *t1 = T + (size_t)s.x * 3 * NE3 + s.y * NGF + j;
      const double *t2 = T + (size_t)s.z * 3 * NE3 + s.w * NGF + (NGF - 1 - j);
#pragma unroll
      for (int f = 0; f < 3; f++) dif[f] = fabs(t2[f * NE3] - t1[f * NE3]);
    }
#pragma unroll
    for (int f = 0; f < 3; f++) D[threadIdx.x][f] = dif[f];
  }
  __syncthreads();
  if (!live || kj != 0) return;
  double mx[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < NE3; t++)
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const double d = D[threadIdx.x + t][f];
      if (mx[f] < d) mx[f] = d;
    }
  const int mem = member_of(off, nmem, e);
  double v[3];
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const double vm = vmax[(size_t)f * nmem + mem];
    v[f] = vm > 1.e-6 ? mx[f] / vm : 0.0;
    amax[(size_t)f * ne + e] = v[f];
  }
  if (decide) {
    fconv[e] = (v[0] > fac1 || v[1] > fac1 || v[2] > fac1) ? 1 : 0;
    fderaff[e] = (v[0] < fac2 && v[1] < fac2 && v[2] < fac2) ? 1 : 0;
  }
}

// part[c] = largest normalised indicator over chunk c (the member's supmax,
// src/adapt.f:1176-1178); part[nch + c] = 1 if some element of the chunk
// asks for coarsening (deraff), else 0.
__global__ __launch_bounds__(BT) void chunk_decide(int ne, int nch, const int *__restrict__ cst,
                                                   const double *__restrict__ amax,
                                                   const int *__restrict__ fderaff,
                                                   double *__restrict__ part) {
  __shared__ double lds[BT];
  const int c = blockIdx.x, e0 = cst[c], e1 = cst[c + 1];
  double mx = 0.0, any = 0.0;
  for (int e = e0 + threadIdx.x; e < e1; e += BT) {
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const double v = amax[(size_t)f * ne + e];
      if (mx < v) mx = v;
    }
    if (fderaff[e]) any = 1.0;
  }
  mx = block_max(mx, lds);
  any = block_max(any, lds);
  if (threadIdx.x == 0) {
    part[c] = mx;
    part[nch + c] = any;
  }
}

template <int NM, int NGF>
void launch_indicator(hipStream_t st, int ne, int nmem, const int *off, const int4 *side,
                      const double *const *a, const double *pico, double *T,
                      const double *vmax, double *amax, int decide, double fac1, double fac2,
                      int *flags) {
  constexpr int EPB = BT / (3 * NGF);
  const int nb = (ne + EPB - 1) / EPB;
  traces<NM, NGF><<<nb, BT, 0, st>>>(ne, a[0], a[1], a[2], pico, T);
  indicator<NM, NGF><<<nb, BT, 0, st>>>(ne, nmem, off, side, T, vmax, amax, decide, fac1, fac2,
                                        flags, flags + ne);
}

}  // namespace adapt

// Device work behind spoc_checkconv / spoc_cal_conv.  a[3]: the state
// (nm, ne) element-major; side[3 e + k]: the face on element e's local side
// k as (first element, its side, second element or -1, its side), 0-based
// (tseg and ssit of the face, src/adapt.f:1280-1287); off[0..nmem]:
// element ranges of the members;
// cst[0..nch]: chunk starts (each chunk inside one member), cof[0..nmem]:
// each member's first chunk.  work: 3 nch + 3 nmem + 2 nch + 2 nmem
// doubles; T: 9 ngf ne doubles (traces); amax: 3 ne doubles (out,
// normalised); flags: 2 ne ints (fconv, fderaff; with decide); with decide,
// res[0..2 nmem) = per member (supmax, then deraff as 0/1) at
// work + 3 nch + 3 nmem + 2 nch.
int adapt_launch(hipStream_t st, int nc, int ngf, int ne, int nmem, int nch,
                 const double *const *a, const int4 *side, const double *pico,
                 const int *off, const int *cst, const int *cof,
                 double *work, double *T, double *amax, bool decide, double fac1, double fac2,
                 int *flags) {
  const int nm = nm_of(nc);
  double *vpart = work, *vmax = vpart + 3 * (size_t)nch;
  double *dpart = vmax + 3 * (size_t)nmem, *res = dpart + 2 * (size_t)nch;
  adapt::chunk_vmax<<<nch, adapt::BT, 0, st>>>(nm, nch, cst, a[0], a[1], a[2], vpart);
  adapt::member_max<<<(3 * nmem + 63) / 64, 64, 0, st>>>(nmem, 3, nch, cof, vpart, vmax);
  bool found = false;
#define SPOC_ADAPT_CASE(NC_, NG_, NGF_)                                                      \
  if (!found && nc == NC_ && ngf == NGF_) {                                                \
    adapt::launch_indicator<nm_of(NC_), NGF_>(st, ne, nmem, off, side, a, pico, T, vmax,  \
                                              amax, decide ? 1 : 0, fac1, fac2, flags);    \
    found = true;                                                                          \
  }
  SPOC_FOR_EACH_SHAPE(SPOC_ADAPT_CASE)
#undef SPOC_ADAPT_CASE
  if (!found) return -1;
  if (decide) {
    adapt::chunk_decide<<<nch, adapt::BT, 0, st>>>(ne, nch, cst, amax, flags + ne, dpart);
    adapt::member_max<<<(2 * nmem + 63) / 64, 64, 0, st>>>(nmem, 2, nch, cof, dpart, res);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int adapt_chunk() { return adapt::CHUNK; }

}  // namespace spoc
