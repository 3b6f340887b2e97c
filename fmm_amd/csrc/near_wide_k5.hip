// Near-field pair kernels of near_kernels.h for the (multipole order m, field order F) pairs of tensor
// order m + F = 5 outside m, F <= 2, every damping kind: one multipole set per workgroup row, tensor
// built and contracted in registers (src/fmm.F90:548-621 at the orders the reference infers from dim
// and field_order).  One translation unit per group so that the build compiles them in parallel.
#include "near_kernels.h"

namespace fmmh {

bool launch_near_wide_k5(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s) {
    if (m == 5 && F == 0) { launch_damp<5, 0, true>(a, nblocks, damp, part, s); return true; }
    if (m == 4 && F == 1) { launch_damp<4, 1, true>(a, nblocks, damp, part, s); return true; }
    if (m == 3 && F == 2) { launch_damp<3, 2, true>(a, nblocks, damp, part, s); return true; }
    if (m == 2 && F == 3) { launch_damp<2, 3, true>(a, nblocks, damp, part, s); return true; }
    if (m == 1 && F == 4) { launch_damp<1, 4, true>(a, nblocks, damp, part, s); return true; }
    if (m == 0 && F == 5) { launch_damp<0, 5, true>(a, nblocks, damp, part, s); return true; }
    return false;
}

}  // namespace fmmh
