// Refinement indicator of the adaptive driver on the device: checkconv and
// cal_conv (src/adapt.f:1260-1332, 1122-1192), FP64.
//
// The decisions are threshold tests on the indicator, so the indicator must
// equal the reference's bit for bit: each edge trace is summed over the
// modes in the reference's order with separate multiplies and adds (this
// file is compiled without FMA contraction), the division by the field's
// largest mean is IEEE, and every maximum is taken with the reference's
// "replace if smaller" test, which gives the same value in any order.  So
// the work can be spread freely: one thread per (face, field) for the
// traces, one per element for the element maxima and the decisions, one
// workgroup per (member, field) for the reductions.
#include <hip/hip_runtime.h>

#include "common.h"

namespace spoc {
namespace adapt {

// Largest |trace difference| over the ngf points of face i for field f
// (src/adapt.f:1281-1313): fmax[f * nf + i]; faces without a second element
// get 0.  a[f] is field f, (nm, ne) element-major; pico[j + ngf (m + nm k)].
__global__ void face_jump(int nf, int nm, int ngf, const double *__restrict__ a0,
                          const double *__restrict__ a1, const double *__restrict__ a2,
                          const int *__restrict__ tseg, const int *__restrict__ ssit,
                          const double *__restrict__ pico, double *__restrict__ fmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, f = blockIdx.y;
  if (i >= nf) return;
  const double *a = f == 0 ? a0 : (f == 1 ? a1 : a2);
  const int k1 = tseg[2 * i] - 1, k2 = tseg[2 * i + 1] - 1;
  double mx = 0.0;
  if (k2 >= 0) {
    const double *x1 = a + (size_t)k1 * nm, *x2 = a + (size_t)k2 * nm;
    const double *p1 = pico + (size_t)ngf * nm * (ssit[2 * i] - 1);
    const double *p2 = pico + (size_t)ngf * nm * (ssit[2 * i + 1] - 1);
    for (int j = 0; j < ngf; j++) {
      double s1 = 0.0, s2 = 0.0;
      for (int m = 0; m < nm; m++) {
        s1 = s1 + x1[m] * p1[j + ngf * m];
        s2 = s2 + x2[m] * p2[(ngf - 1 - j) + ngf * m];
      }
      const double dif = fabs(s2 - s1);
      if (mx < dif) mx = dif;
    }
  }
  fmax[(size_t)f * nf + i] = mx;
}

// Raw indicator of element e (before normalisation): the largest fmax over
// its faces that have a second element, starting from 0 (src/adapt.f:1275-1277,
// 1309-1310).
__global__ void elem_jump(int ne, int nf, const int *__restrict__ itm,
                          const int *__restrict__ tseg, const double *__restrict__ fmax,
                          double *__restrict__ amax) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
#pragma unroll
  for (int f = 0; f < 3; f++) {
    double mx = 0.0;
    for (int k = 0; k < 3; k++) {
      const int fc = itm[3 * e + k] - 1;
      if (tseg[2 * fc + 1] > 0) {
        const double d = fmax[(size_t)f * nf + fc];
        if (mx < d) mx = d;
      }
    }
    amax[(size_t)f * ne + e] = mx;
  }
}

// Largest of a workgroup's values (every thread gets it).
__device__ double block_max(double v, double *lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
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

// Largest |a(1, e)| of field blockIdx.y over the elements of member
// blockIdx.x (src/adapt.f:1317-1320): vmax[f * nmem + r].
__global__ void member_vmax(int nm, const int *__restrict__ off, const double *__restrict__ a0,
                            const double *__restrict__ a1, const double *__restrict__ a2,
                            double *__restrict__ vmax) {
  __shared__ double lds[256];
  const int r = blockIdx.x, f = blockIdx.y, nmem = gridDim.x;
  const double *a = f == 0 ? a0 : (f == 1 ? a1 : a2);
  double mx = 0.0;
  for (int e = off[r] + threadIdx.x; e < off[r + 1]; e += blockDim.x) {
    const double v = fabs(a[(size_t)e * nm]);
    if (mx < v) mx = v;
  }
  mx = block_max(mx, lds);
  if (threadIdx.x == 0) vmax[(size_t)f * nmem + r] = mx;
}

// Normalised indicators and per-element decisions (src/adapt.f:1321-1329,
// 1159-1176): amax /= vmax of the element's member (0 when vmax <= 1e-6);
// fconv = some field > fac1, fderaff = every field < fac2.
__global__ void elem_decide(int ne, int nmem, const int *__restrict__ off,
                            const double *__restrict__ vmax, double *__restrict__ amax,
                            double fac1, double fac2, int *__restrict__ fconv,
                            int *__restrict__ fderaff) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int lo = 0, hi = nmem;  // member r with off[r] <= e < off[r + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  double v[3];
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const double vm = vmax[(size_t)f * nmem + lo];
    double *p = amax + (size_t)f * ne + e;
    v[f] = vm > 1.e-6 ? *p / vm : 0.0;
    *p = v[f];
  }
  fconv[e] = (v[0] > fac1 || v[1] > fac1 || v[2] > fac1) ? 1 : 0;
  fderaff[e] = (v[0] < fac2 && v[1] < fac2 && v[2] < fac2) ? 1 : 0;
}

// Per member: the largest indicator of the three fields (supmax) and
// whether some element asks for coarsening (deraff), src/adapt.f:1162-1178.
__global__ void member_decide(int ne, const int *__restrict__ off,
                              const double *__restrict__ amax, const int *__restrict__ fderaff,
                              double *__restrict__ out) {
  __shared__ double lds[256];
  const int r = blockIdx.x;
  double mx = -1.0, any = 0.0;
  for (int e = off[r] + threadIdx.x; e < off[r + 1]; e += blockDim.x) {
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
    out[2 * r] = mx;
    out[2 * r + 1] = any;
  }
}

}  // namespace adapt

// Device work behind spoc_checkconv / spoc_cal_conv.  a[3]: the state
// (nm, ne) element-major; off[0..nmem]: element ranges of the members;
// work: 3 nf + 3 nmem + 2 nmem doubles; amax: 3 ne doubles (out,
// normalised); flags: 2 ne ints (fconv, fderaff; when decide); out: 2 nmem
// doubles (supmax, deraff; when decide).
int adapt_launch(hipStream_t st, int ne, int nf, int nm, int ngf, int nmem, const double *const *a,
                 const int *tseg, const int *ssit, const int *itm, const double *pico,
                 const int *off, double *work, double *amax, bool decide, double fac1,
                 double fac2, int *flags, double *out) {
  double *fmax = work, *vmax = work + 3 * (size_t)nf;
  if (nf > 0)
    adapt::face_jump<<<dim3((nf + 127) / 128, 3), 128, 0, st>>>(nf, nm, ngf, a[0], a[1], a[2],
                                                                tseg, ssit, pico, fmax);
  adapt::elem_jump<<<(ne + 127) / 128, 128, 0, st>>>(ne, nf, itm, tseg, fmax, amax);
  adapt::member_vmax<<<dim3(nmem, 3), 256, 0, st>>>(nm, off, a[0], a[1], a[2], vmax);
  adapt::elem_decide<<<(ne + 127) / 128, 128, 0, st>>>(ne, nmem, off, vmax, amax, fac1, fac2,
                                                       flags, flags + ne);
  if (decide) adapt::member_decide<<<nmem, 256, 0, st>>>(ne, off, amax, flags + ne, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace spoc
