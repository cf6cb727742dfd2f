// Global integrals of the state (src/energy.f:3-96), FP64.
//
// Per-element terms are computed in parallel with the reference's
// expression order; the three global sums then run in the reference's
// element (and mode) order on one thread each, so that the result equals
// the reference's sequential accumulation bit for bit.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"

namespace spoc {

namespace {

// The partial integrals of spoc_energy_noalloc: ne (nm + 2) doubles, at most
// kScratchDoubles (64 MiB), held by every device the library runs on until
// the process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: only the per-device address from hipGetSymbolAddress
// (energy_scratch) is passed to kernels.
constexpr size_t kScratchDoubles = (size_t)1 << 23;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

}  // namespace

size_t energy_scratch_doubles() { return kScratchDoubles; }

// The device address of `scratch` on `device`, which must be current; looked
// up once per device.  0 on success, -1 device out of range, -2 HIP error.
int energy_scratch(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -2;
    g_scratch[device] = (double *)p;
  }
  *out = g_scratch[device];
  return 0;
}

__global__ void energy_terms(int ne, int nm, int ng, const double *__restrict__ u,
                             const double *__restrict__ v, const double *__restrict__ h,
                             const double *__restrict__ hb, const double *__restrict__ pi2,
                             const double *__restrict__ wi, const double *__restrict__ ammo1,
                             const double *__restrict__ deta, double g0, double *__restrict__ ke,
                             double *__restrict__ pe, double *__restrict__ mass) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const double *ue = u + (size_t)e * nm, *ve = v + (size_t)e * nm, *he = h + (size_t)e * nm;
  const double *hbe = hb + (size_t)e * nm;
  double d = deta[e];
  for (int j = 0; j < nm; j++) mass[(size_t)e * nm + j] = ammo1[j] * d * he[j];
  double ak = 0.0, ap = 0.0;
  for (int q = 0; q < ng; q++) {
    double mu = 0.0, mv = 0.0, mh = 0.0, mhb = 0.0;
    for (int m = 0; m < nm; m++) {
      double p = pi2[q + ng * m];
      mu = mu + ue[m] * p;
      mv = mv + ve[m] * p;
      mh = mh + he[m] * p;
      mhb = mhb + (he[m] + hbe[m]) * p;
    }
    double mk = mhb * (mu * mu + mv * mv);
    ak = ak + wi[q] * mk * pi2[q];
    double mp = mh * mh;
    ap = ap + wi[q] * mp * pi2[q];
  }
  ke[e] = ak * d * 0.5;
  pe[e] = ap * d * 0.5 * g0;
}

// One thread per (range, quantity): elements [off[r], off[r+1]) of the
// caller's numbering, in that order; order[e] (may be null: the identity)
// is where the caller's element e sits in the arrays.
__global__ void energy_sum(int nr, const int *__restrict__ off, int nm,
                           const int *__restrict__ order, const double *__restrict__ ke,
                           const double *__restrict__ pe, const double *__restrict__ mass,
                           double *__restrict__ out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nr) return;
  int r = t / 3, w = t % 3;
  size_t a = off[r], b = off[r + 1];
  double s = 0.0;
  for (size_t e = a; e < b; e++) {
    const size_t i = order ? (size_t)order[e] : e;
    if (w == 0)
      s = s + ke[i];
    else if (w == 1)
      s = s + pe[i];
    else
      for (int m = 0; m < nm; m++) s = s + mass[i * nm + m];
  }
  out[t] = s;
}

// out[3 r + (0, 1, 2)] = (ke, pe, mass) of elements [off[r], off[r+1]), r < nr.
int energy_launch(hipStream_t st, int ne, int nm, int ng, const double *u, const double *v,
                  const double *h, const double *hb, const double *pi2, const double *wi,
                  const double *ammo1, const double *deta, double g0, int nr, const int *off,
                  const int *order, double *scratch, double *out) {
  double *ke = scratch, *pe = scratch + ne, *mass = scratch + 2 * (size_t)ne;
  energy_terms<<<(ne + 127) / 128, 128, 0, st>>>(ne, nm, ng, u, v, h, hb, pi2, wi, ammo1, deta,
                                                 g0, ke, pe, mass);
  energy_sum<<<(3 * nr + 63) / 64, 64, 0, st>>>(nr, off, nm, order, ke, pe, mass, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// flag[0] |= 1 if any of the n doubles of a is not finite (NaN or Inf).
__global__ void nonfinite_kernel(size_t n, const double *__restrict__ a, int *__restrict__ flag) {
  bool bad = false;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    bad = bad || !isfinite(a[i]);
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// Sets *flag (device) to 1 if a[0..n) holds a non-finite value, else 0.
int nonfinite_launch(hipStream_t st, size_t n, const double *a, int *flag) {
  if (hipMemsetAsync(flag, 0, sizeof(int), st) != hipSuccess) return -2;
  const size_t want = (n + 255) / 256;
  const unsigned grid = (unsigned)(want < 2048 ? (want ? want : 1) : 2048);
  nonfinite_kernel<<<grid, 256, 0, st>>>(n, a, flag);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace spoc
