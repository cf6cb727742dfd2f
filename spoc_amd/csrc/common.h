// Shared definitions for the gfx950 kernels and the host C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace spoc {

constexpr int nm_of(int nc) { return (nc + 1) * (nc + 2) / 2; }
// ip(nc, i) = sum_{j<i} (nc + 1 - j): offset of the first mode with first
// index a = i (src/mat1.f:372-381).
__host__ __device__ constexpr int ip_of(int nc, int i) { return i * (nc + 1) - i * (i - 1) / 2; }

// Supported (p, triangle rule, edge rule) combinations.  The face rule is
// p+1 Gauss-Legendre points in every configuration; the triangle rule must
// have at least nm points (src/shwwnl.f:153-156 otherwise fails DPOTRF).
#define SPOC_FOR_EACH_SHAPE(X) \
  X(2, 6, 3)                   \
  X(3, 12, 4)                  \
  X(4, 25, 5)                  \
  X(5, 25, 6)                  \
  X(6, 33, 7)                  \
  X(7, 42, 8)

// Operator tables on the device, FP64, in the layouts the kernels read.
struct DevTables {
  const double *agxT;  // agxT[i + NM*j] = agx(j, i)
  const double *agyT;  // agyT[i + NM*j] = agy(j, i)
  const double *pi2;   // pi2[q + NG*m]   = pi2(q, m)
  const double *pi2T;  // pi2T[m + NM*q]  = pi2(q, m)
  const double *pico;  // pico[j + NGF*(m + NM*k)] = pico(j, m, k)
  const double *picoT; // picoT[i + NM*(j + NGF*k)] = pico(j, i, k)
  const double *U;     // U[k + NM*i] = amm(k, i) after DPOTRF (upper factor)
  const double *UL;    // UL[k + NM*i] = amml(k, i) after DPOTRF
  const double *dia1;  // 1 / U(i, i)
  const double *dia2;  // 1 / UL(i, i)
  const double *wi;    // triangle weights
  const double *wif;   // edge weights
  const double *ammo1; // ammo(:, 1), the mass integrals of the modes
};

// Mesh operands on the device (0-based where noted).
struct DevMesh {
  int ne, nf;
  const int *itm;      // (3, ne): 1-based face of local edge k
  const int *its;      // (3, ne): +-1
  const int *tseg;     // (2, nf): 1-based elements, 0 = wall
  const int *ssit;     // (2, nf): local edge in each element, 1..3
  const double *norm;  // (2, nf)
  const double *vect;  // (5, ne): ax, bx, ay, by, 1/det
  const double *normt; // (2, ne)
  const int *bdl;      // (ne)
  const double *deta;  // (ne): det as src/energy.f:37-42 computes it
};

struct Physics {
  double mu0, mb0, g0, hh0;
  // hb(1,1) of each element's own mesh (an ensemble member's first element):
  // the constant depth of the linear variant (src/calvar_lin.f:65)
  const double *hh0e;
};

}  // namespace spoc
