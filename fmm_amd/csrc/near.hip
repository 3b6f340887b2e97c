// Near-field pair sums, orders <= 2: instantiations of the kernels of
// near_kernels.h for source multipole order m <= 2 and field order F <= 2
// (every set count, several sets per workgroup row), the dispatch over all
// supported orders (the rest are in near_wide*.hip), and the sum of the
// partial fields of chunked leaves.
#include "near_kernels.h"

namespace fmmh {

template <int M>
static void launch_f(const NearArgs& a, int nblocks, int F, int damp, int part, hipStream_t s) {
    switch (F) {
        case 0: launch_damp<M, 0, false>(a, nblocks, damp, part, s); break;
        case 1: launch_damp<M, 1, false>(a, nblocks, damp, part, s); break;
        default: launch_damp<M, 2, false>(a, nblocks, damp, part, s); break;
    }
}

bool near_supported(int m, int F) {
    return m >= 0 && F >= 0 && m <= NEAR_MMAX && F <= NEAR_FMAX && m + F <= NEAR_KMAX;
}

void launch_near_part(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s) {
    if (nblocks <= 0) return;
    if (!near_supported(m, F)) throw Error(FMMH_ERR_UNSUPPORTED, "near-field orders outside the compiled range");
    if (m <= 2 && F <= 2) {
        switch (m) {
            case 0: launch_f<0>(a, nblocks, F, damp, part, s); break;
            case 1: launch_f<1>(a, nblocks, F, damp, part, s); break;
            default: launch_f<2>(a, nblocks, F, damp, part, s); break;
        }
    } else if (!(launch_near_wide_k34(a, nblocks, m, F, damp, part, s) ||
                 launch_near_wide_k5(a, nblocks, m, F, damp, part, s) ||
                 launch_near_wide_k6a(a, nblocks, m, F, damp, part, s) ||
                 launch_near_wide_k6b(a, nblocks, m, F, damp, part, s) ||
                 launch_near_wide_k7a(a, nblocks, m, F, damp, part, s) ||
                 launch_near_wide_k7b(a, nblocks, m, F, damp, part, s))) {
        throw Error(FMMH_ERR_INTERNAL, "no near-field kernel for these orders");
    }
    FMMH_HIP(hipGetLastError());
}

void launch_near_excl_rt(const NearArgs& a, int nblocks, int m, int F, int damp, hipStream_t s) {
    NearArgs b = a;
    b.m_rt = m;
    b.f_rt = F;
    const dim3 g(nblocks, a.nsets);
    switch (damp) {
        case DAMP_NONE: hipLaunchKernelGGL((k_near_excl<-1, -1, DAMP_NONE, 1>), g, dim3(64), 0, s, b); break;
        case DAMP_THOLE: hipLaunchKernelGGL((k_near_excl<-1, -1, DAMP_THOLE, 1>), g, dim3(64), 0, s, b); break;
        case DAMP_AMOEBA: hipLaunchKernelGGL((k_near_excl<-1, -1, DAMP_AMOEBA, 1>), g, dim3(64), 0, s, b); break;
        default: hipLaunchKernelGGL((k_near_excl<-1, -1, DAMP_ERF, 1>), g, dim3(64), 0, s, b); break;
    }
    FMMH_HIP(hipGetLastError());
}

// One workgroup per leaf; leaves that are not chunked return at once.
__global__ void k_near_fix(NearArgs a, const int* __restrict__ leaf_node, const int* __restrict__ lk,
                           const int* __restrict__ lpart, int nf) {
    const int l = blockIdx.x;
    const int K = lk[l];
    if (K <= 1) return;
    const int node = leaf_node[l];
    const int t0 = a.tgt_start[node], tc = a.tgt_count[node];
    const int p0 = lpart[l];
    const int total = tc * nf * a.nsets;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int i = idx % tc;
        const int ut = idx / tc;                  // set * nf + component
        const double* src = a.part + (int64_t)ut * a.ldp + p0 + i;
        double v = src[0];
        for (int k = 1; k < K; ++k) v += src[(int64_t)k * tc];
        double* o = a.out + (int64_t)ut * a.ldo + t0 + i;
        if (a.accumulate) *o += v;
        else *o = v;
    }
}

void launch_near_fix(const NearArgs& a, const int* leaf_node, const int* lk, const int* lpart, int nl, int nf,
                     hipStream_t s) {
    if (nl <= 0 || !a.part) return;
    hipLaunchKernelGGL(k_near_fix, dim3(nl), dim3(256), 0, s, a, leaf_node, lk, lpart, nf);
    FMMH_HIP(hipGetLastError());
}

void launch_near(const NearArgs& a, int nblocks, int m, int F, int damp, hipStream_t s) {
    launch_near_part(a, nblocks, m, F, damp, 1, s);
    launch_near_part(a, nblocks, m, F, damp, 2, s);
}

}  // namespace fmmh
