// Arguments shared by the host context and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace spoc {

// Optional per-launch timing hook: called with a slot index before/after
// each kernel of one right-hand side.
typedef void (*TimerHook)(void *user, int slot, hipStream_t st);

struct RhsArgs {
  int variant;            // SPOC_LIN..SPOC_INV
  hipStream_t stream;
  DevTables tab;
  DevMesh mesh;
  Physics phys;
  const double *x[3];     // stage input u, v, h  (NM, ne)
  const double *hb, *w0, *z0, *fc;
  // outputs
  double *k[3];           // tendencies when stage < 0
  double *u0[3], *xout[3], *acc[3];
  int stage;              // -1: tendencies only; 0..3 RK4 stage
  double dt;
  // work
  double *work[8];        // F(4), src(2), uh, vh
  const double *const *Fptr;  // device array holding work[0..3]
  double *face;           // nf * 6 * NGF
  TimerHook hook;
  void *hook_user;
  void timer(int slot, hipStream_t st) const {
    if (hook) hook(hook_user, slot, st);
  }
};

int rhs_f64(int nc, int ng, int ngf, const RhsArgs &a);

}  // namespace spoc

namespace spoc {
#define SPOC_DECL_F64(NC_, NG_, NGF_) int rhs_f64_##NC_##_##NG_##_##NGF_(const RhsArgs &a);
SPOC_FOR_EACH_SHAPE(SPOC_DECL_F64)
#undef SPOC_DECL_F64
}  // namespace spoc
