// Near-field pair kernels of near_kernels.h for the (multipole order m, field order F) pairs of tensor
// order m + F = 6 outside m, F <= 2, damping '' and THOLE: one multipole set per workgroup row, tensor
// built and contracted in registers (src/fmm.F90:548-621 at the orders the reference infers from dim
// and field_order).  One translation unit per group so that the build compiles them in parallel.
#include "near_kernels.h"

namespace fmmh {

bool launch_near_wide_k6a(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s) {
    if (damp != DAMP_NONE && damp != DAMP_THOLE) return false;
    if (m == 6 && F == 0) {
        if (damp == DAMP_NONE) launch_sets<6, 0, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<6, 0, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 5 && F == 1) {
        if (damp == DAMP_NONE) launch_sets<5, 1, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<5, 1, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 4 && F == 2) {
        if (damp == DAMP_NONE) launch_sets<4, 2, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<4, 2, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 3 && F == 3) {
        if (damp == DAMP_NONE) launch_sets<3, 3, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<3, 3, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 2 && F == 4) {
        if (damp == DAMP_NONE) launch_sets<2, 4, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<2, 4, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 1 && F == 5) {
        if (damp == DAMP_NONE) launch_sets<1, 5, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<1, 5, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 0 && F == 6) {
        if (damp == DAMP_NONE) launch_sets<0, 6, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<0, 6, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    return false;
}

}  // namespace fmmh
