// Near-field pair sums: the FMM P2P phase (src/fmm.F90:623-747) and the
// all-pairs direct sum (src/fmm.F90:749-824) share one kernel.
//
// Work decomposition: one 64-lane wavefront per target block (a leaf of the
// octree, or a run of 64 consecutive sites for the direct sum).  Each lane
// holds one target site; sources are streamed in a wave-uniform loop so that
// their coordinates, scaled multipoles, damping factor and original index are
// scalar (SMEM) loads broadcast to all lanes for free -- no LDS traffic and no
// bank conflicts.  The tensor recursion lives entirely in VGPRs (tensors.h).
//
// Exclusions follow the reference exactly: an excluded pair is evaluated at
// displacement (1e10, 1e10, 1e10) (src/fmm.F90:714-721, 802-806), which
// leaves the -q/(sqrt(3) 1e10) potential artefact in place.
#include "common.h"
#include "tensors.h"
#include "kernels.h"

namespace fmmh {

constexpr int NEAR_EXCL_REG = 4;   // exclusions kept in registers; more are read from memory

template <int M, int F, int DAMP>
__global__ __launch_bounds__(64) void k_near(NearArgs a) {
    constexpr int K = M + F;
    constexpr int NM = ncomp(M);
    constexpr int NF = ncomp(F);
    const int b = a.blk[blockIdx.x];
    const int lane = threadIdx.x;
    const int t0 = a.tgt_start[b];
    const int tc = a.tgt_count[b];
    const int e0 = a.list_off[b];
    const int e1 = a.list_off[b + 1];
    for (int base = 0; base < tc; base += 64) {
        const bool active = base + lane < tc;
        const int k = t0 + base + (active ? lane : 0);
        const double xi = a.xs[k], yi = a.ys[k], zi = a.zs[k];
        const double dfi = DAMP != DAMP_NONE ? a.dfs[k] : 0.0;
        int ex[NEAR_EXCL_REG];
#pragma unroll
        for (int q = 0; q < NEAR_EXCL_REG; ++q) ex[q] = q < a.nexcl ? a.excl[(int64_t)q * a.lde + k] : -1;
        double acc[NF];
#pragma unroll
        for (int t = 0; t < NF; ++t) acc[t] = 0.0;
        for (int e = e0; e < e1; ++e) {
            const int sb = a.list_src[e];
            const int s0 = a.src_start[sb];
            const int s1 = s0 + a.src_count[sb];
            for (int j = s0; j < s1; ++j) {
                double dx = xi - a.xs[j];
                double dy = yi - a.ys[j];
                double dz = zi - a.zs[j];
                const int oj = a.orig[j] + 1;
                bool excluded = false;
#pragma unroll
                for (int q = 0; q < NEAR_EXCL_REG; ++q) excluded |= (ex[q] == oj);
                for (int q = NEAR_EXCL_REG; q < a.nexcl; ++q) excluded |= (a.excl[(int64_t)q * a.lde + k] == oj);
                if (excluded) { dx = 1e10; dy = 1e10; dz = 1e10; }
                double aij = 0.0;
                if constexpr (DAMP != DAMP_NONE) aij = dfi * a.dfs[j];
                double m[NM];
#pragma unroll
                for (int s = 0; s < NM; ++s) m[s] = a.ms[(int64_t)s * a.ldm + j];
                double g[K + 1];
                radial<K, DAMP>(dx * dx + dy * dy + dz * dz, aij, g);
                double T[ncomp(K)];
                md_tensor<K>(dx, dy, dz, g, T);
                contract<M, F>(m, T, acc);
            }
        }
        if (active) {
            if (a.accumulate) {
#pragma unroll
                for (int t = 0; t < NF; ++t) a.out[(int64_t)t * a.ldo + k] += acc[t];
            } else {
#pragma unroll
                for (int t = 0; t < NF; ++t) a.out[(int64_t)t * a.ldo + k] = acc[t];
            }
        }
    }
}

template <int M, int F>
static void launch_damp(const NearArgs& a, int nblocks, int damp, hipStream_t s) {
    switch (damp) {
        case DAMP_NONE: hipLaunchKernelGGL((k_near<M, F, DAMP_NONE>), dim3(nblocks), dim3(64), 0, s, a); break;
        case DAMP_THOLE: hipLaunchKernelGGL((k_near<M, F, DAMP_THOLE>), dim3(nblocks), dim3(64), 0, s, a); break;
        case DAMP_AMOEBA: hipLaunchKernelGGL((k_near<M, F, DAMP_AMOEBA>), dim3(nblocks), dim3(64), 0, s, a); break;
        default: hipLaunchKernelGGL((k_near<M, F, DAMP_ERF>), dim3(nblocks), dim3(64), 0, s, a); break;
    }
}

template <int M>
static void launch_f(const NearArgs& a, int nblocks, int F, int damp, hipStream_t s) {
    switch (F) {
        case 0: launch_damp<M, 0>(a, nblocks, damp, s); break;
        case 1: launch_damp<M, 1>(a, nblocks, damp, s); break;
        default: launch_damp<M, 2>(a, nblocks, damp, s); break;
    }
}

bool near_supported(int m, int F) { return m >= 0 && m <= NEAR_MMAX && F >= 0 && F <= NEAR_FMAX; }

void launch_near(const NearArgs& a, int nblocks, int m, int F, int damp, hipStream_t s) {
    if (nblocks <= 0) return;
    switch (m) {
        case 0: launch_f<0>(a, nblocks, F, damp, s); break;
        case 1: launch_f<1>(a, nblocks, F, damp, s); break;
        default: launch_f<2>(a, nblocks, F, damp, s); break;
    }
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
