// Near-field pair kernels of near_kernels.h for the (multipole order m, field order F) pairs of tensor
// order m + F = 7 outside m, F <= 2, damping '' and THOLE: one multipole set per workgroup row, tensor
// built and contracted in registers (src/fmm.F90:548-621 at the orders the reference infers from dim
// and field_order).  One translation unit per group so that the build compiles them in parallel.
// Damped pairs of tensor order 7 contract the tensor up to order 6 only, as the reference (its damped
// tensors stop at T6, src/tensors_damped_thole.F90:55-62 and the same lines in _amoeba and _erf).
#include "near_kernels.h"

namespace fmmh {

bool launch_near_wide_k7a(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s) {
    if (damp != DAMP_NONE && damp != DAMP_THOLE) return false;
    if (m == 6 && F == 1) {
        if (damp == DAMP_NONE) launch_sets<6, 1, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<6, 1, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 5 && F == 2) {
        if (damp == DAMP_NONE) launch_sets<5, 2, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<5, 2, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 4 && F == 3) {
        if (damp == DAMP_NONE) launch_sets<4, 3, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<4, 3, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 3 && F == 4) {
        if (damp == DAMP_NONE) launch_sets<3, 4, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<3, 4, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 2 && F == 5) {
        if (damp == DAMP_NONE) launch_sets<2, 5, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<2, 5, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    if (m == 1 && F == 6) {
        if (damp == DAMP_NONE) launch_sets<1, 6, DAMP_NONE, true>(a, nblocks, part, s);
        else launch_sets<1, 6, DAMP_THOLE, true>(a, nblocks, part, s);
        return true;
    }
    return false;
}

}  // namespace fmmh
