// Class-grouped M2L (src/fmm.F90:826-882) for one multipole set at a time.
//
// Every far pair (i, j) belongs to a displacement class c = (v, level_i -
// level_j) (m2l.hip), and the pair's contribution is
//   L~_i(t) += scale_c(t) sum_s T~_c[s + t] M~_j(s),
// so all pairs of a class apply the same n_m x n_m Hankel matrix H_c.  Per
// class this is a GEMM, H_c x [M~_j1 M~_j2 ...], run here on
// v_mfma_f64_16x16x4f64: a workgroup of four wavefronts takes one tile of up
// to 32 pairs of one class, stages the class tensor and the 32 source
// expansions in LDS, and each wavefront accumulates a quarter of the 16-row
// target tiles.  A fragments are gathered from the class tensor through a
// host-built index table (packed index of s + t per fragment lane); the
// level-difference factor of the class is a per-row scale.
//
// The Hankel matrix runs over the reduced (harmonic) components only -- the
// (p + 1)^2 with z <= 1, sources folded onto them (m2l.hip, the Laplace
// identity T_{k+2x} + T_{k+2y} + T_{k+2z} = 0): 100 x 100 instead of 220 x
// 220 at p = 9.  Results are written per pair to its CSR slot, and
// k_m2l_reduce sums each target's slots in CSR order, so the sum order is
// fixed (deterministic) and independent of the tiling; the local components
// with t_z >= 2 are rebuilt after it (launch_le_unfold).
//
// Mixed precision (F32): the same tiles on v_mfma_f32_16x16x4f32 with FP32
// sources and FP32 accumulation over s; the class tensor stays FP64 in LDS
// and each A element is rounded once after its row scale; per-pair results
// are FP32 and the reduction accumulates them in FP64 (exact for any
// realistic number of pairs per target, so no compensation term is needed).
//
// MFMA operand layouts (measured on gfx950): A lane l -> A[l % 16][l / 16],
// B lane l -> B[l / 16][l % 16] for both; accumulator register q of lane l
// -> D[4 q + l / 16][l % 16] for f64_16x16x4f64 and D[4 (l / 16) + q][l % 16]
// for f32_16x16x4f32.
#include "common.h"
#include "kernels.h"

namespace fmmh {

using v4d = double __attribute__((ext_vector_type(4)));
using v4f = float __attribute__((ext_vector_type(4)));

template <bool F32> struct GemmT;
template <> struct GemmT<false> {
    using E = double;
    using Acc = v4d;
    static __device__ __forceinline__ Acc mfma(E a, E b, Acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int q, int g4) { return 4 * q + g4; }
};
template <> struct GemmT<true> {
    using E = float;
    using Acc = v4f;
    static __device__ __forceinline__ Acc mfma(E a, E b, Acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int q, int g4) { return 4 * g4 + q; }
};

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

__global__ __launch_bounds__(256) void k_m2l_tiles(const int* __restrict__ flag, const int* __restrict__ cls,
                                                   int n_pairs, int ncls, int* __restrict__ class_off,
                                                   int* __restrict__ tile_cnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_pairs && flag[k]) class_off[cls[k]] = k;   // cls = exclusive scan of flag = class id at a start
    if (k == 0) {
        class_off[ncls] = n_pairs;
        tile_cnt[ncls] = 0;
    }
}

__global__ __launch_bounds__(256) void k_m2l_tile_count(const int* __restrict__ class_off, int ncls,
                                                        int* __restrict__ tile_cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncls) tile_cnt[c] = (class_off[c + 1] - class_off[c] + M2L_GEMM_TN - 1) / M2L_GEMM_TN;
}

template <int P, bool F32>
__global__ __launch_bounds__(256) void k_m2l_gemm(M2LGemmArgs a) {
    using G = GemmT<F32>;
    using E = typename G::E;
    constexpr int NK = rdim(P);                 // reduced components: rows and k of the GEMM
    constexpr int NT = ncomp(2 * P);
    constexpr int MT = (NK + 15) / 16;          // 16-row target tiles
    constexpr int KS = (NK + 3) / 4;            // 4-deep source steps
    constexpr int MW = (MT + 3) / 4;            // target tiles per wavefront
    constexpr int TN = M2L_GEMM_TN;
    constexpr int NTL = TN / 16;
    constexpr int BLD = NK | 1;                 // odd LDS row stride of the source tile
    __shared__ double Ts[NT];
    __shared__ E Bs[TN * BLD];
    __shared__ int ecsr[TN];
    __shared__ int jsrc[TN];
    const int b = blockIdx.x;
    const int total = a.tile_off[a.ncls];
    if (b >= total) return;                     // grid is an upper bound on the tile count
    int lo = 0, hi = a.ncls;                    // class of this tile: last c with tile_off[c] <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.tile_off[mid] <= b) lo = mid;
        else hi = mid;
    }
    const int c = lo;                           // class, or (class, order) group
    const int cls = a.grp_class ? a.grp_class[c] : c;
    const int q = a.grp_order ? a.grp_order[c] : P;
    const int nq = rdim(q);                     // coefficients applied for this tile's pairs
    const int MTq = (nq + 15) / 16, KSq = (nq + 3) / 4;
    const int k0 = a.class_off[c] + (b - a.tile_off[c]) * TN;
    const int cnt = min(TN, a.class_off[c + 1] - k0);
    const int dl = a.class_dl[cls];
    const int tid = threadIdx.x;
    const int ntq = ncomp(2 * q);
    for (int x = tid; x < ntq; x += 256) Ts[x] = a.Tc[(int64_t)cls * NT + x];
    if (tid < TN) {
        const int e = tid < cnt ? a.pair_csr[k0 + tid] : -1;
        ecsr[tid] = e;
        jsrc[tid] = e >= 0 ? src_node(a.src[e]) : -1;
    }
    __syncthreads();
    const E* __restrict__ cms = F32 ? (const E*)a.CMs32 : (const E*)a.CMs;
    for (int x = tid; x < TN * NK; x += 256) {
        const int n = x / NK, s = x - n * NK;
        const int j = jsrc[n];
        Bs[n * BLD + s] = (j >= 0 && s < nq) ? cms[(int64_t)j * NK + s] : E(0);
    }
    __syncthreads();
    const int w = tid >> 6, lane = tid & 63;
    const int c16 = lane & 15, g4 = lane >> 4;
    double rs[MW];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi) {
        const int mt = w + 4 * mi;
        const int deg = mt < MT ? a.tdeg[mt * 16 + c16] : 0;
        rs[mi] = dl > 0 ? ldexp(1.0, -dl * deg) : (dl < 0 ? ldexp(1.0, -dl * (deg + 1)) : 1.0);
    }
    typename G::Acc acc[MW][NTL];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi)
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) acc[mi][nt] = typename G::Acc{0, 0, 0, 0};
    const int* __restrict__ fx = a.fidx + lane;
    // the fragment indices come from global memory (L2): KC source steps' worth are loaded together,
    // so one wait covers KC x MW loads instead of one per step
    constexpr int KC = 8;
    for (int k0 = 0; k0 < KSq; k0 += KC) {
        int fi[KC][MW];
#pragma unroll
        for (int kk = 0; kk < KC; ++kk)
#pragma unroll
            for (int mi = 0; mi < MW; ++mi) {
                const int ks = k0 + kk, mt = w + 4 * mi;
                fi[kk][mi] = (ks < KSq && mt < MTq) ? fx[(mt * KS + ks) * 64] : -1;
            }
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
            const int ks = k0 + kk;
            if (ks >= KSq) break;
            const int s = ks * 4 + g4;
            E bv[NTL];
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) bv[nt] = s < NK ? Bs[(nt * 16 + c16) * BLD + s] : E(0);
#pragma unroll
            for (int mi = 0; mi < MW; ++mi) {
                const int mt = w + 4 * mi;
                if (mt < MTq) {
                    const int f = fi[kk][mi];
                    const E av = (f >= 0 && mt * 16 + c16 < nq) ? E(Ts[f] * rs[mi]) : E(0);
#pragma unroll
                    for (int nt = 0; nt < NTL; ++nt) acc[mi][nt] = G::mfma(av, bv[nt], acc[mi][nt]);
                }
            }
        }
    }
#pragma unroll
    for (int mi = 0; mi < MW; ++mi) {
        const int mt = w + 4 * mi;
        if (mt >= MTq) continue;
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) {
            const int n = nt * 16 + c16;
            if (n >= cnt) continue;
            E* __restrict__ out = (F32 ? (E*)a.pbuf32 : (E*)a.pbuf) + (int64_t)ecsr[n] * NK;
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int t = mt * 16 + G::row(r4, g4);
                if (t < nq) out[t] = acc[mi][nt][r4];
            }
        }
    }
}

// L_i(t) = r_i^{-(|t|+1)} sum over the CSR slots of target i, in CSR order
// (FP64 accumulation for either slot type), for the reduced components t
// (slots of nk = (p + 1)^2 coefficients; written at their packed index tfull[t]
// of the ncomp(p)-long local expansion).
// A slot of a pair with a learned order q holds only its first rdim(q) coefficients.
template <class E>
__global__ __launch_bounds__(128) void k_m2l_reduce(const E* __restrict__ pbuf, const int* __restrict__ off,
                                                    const double* __restrict__ r, const int* __restrict__ tdeg,
                                                    const int* __restrict__ tfull,
                                                    const unsigned char* __restrict__ order,
                                                    double* __restrict__ LE, int nk, int nm) {
    const int i = blockIdx.x;
    const int e0 = off[i], e1 = off[i + 1];
    const double rinv = 1.0 / r[i];
    for (int t = threadIdx.x; t < nk; t += 128) {
        double sum = 0.0;
        if (order) {
            const int dt = tdeg[t];
#pragma unroll 8
            for (int e = e0; e < e1; ++e)
                if (dt <= order[e]) sum += (double)pbuf[(int64_t)e * nk + t];
        } else {
            // 16 slots loaded per batch (the adds stay in CSR order)
            int e = e0;
            for (; e + 16 <= e1; e += 16) {
                E v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = pbuf[(int64_t)(e + u) * nk + t];
#pragma unroll
                for (int u = 0; u < 16; ++u) sum += (double)v[u];
            }
            for (; e < e1; ++e) sum += (double)pbuf[(int64_t)e * nk + t];
        }
        double f = rinv;
        for (int k = 0; k < tdeg[t]; ++k) f *= rinv;
        LE[(int64_t)i * nm + tfull[t]] = sum * f;
    }
}

void launch_m2l_tiles(const int* flag, const int* cls, int n_pairs, int ncls, int* class_off, int* tile_cnt,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_m2l_tiles, dim3(nblk(std::max(n_pairs, 1), 256)), dim3(256), 0, s, flag, cls, n_pairs, ncls,
                       class_off, tile_cnt);
    if (ncls > 0)
        hipLaunchKernelGGL(k_m2l_tile_count, dim3(nblk(ncls, 256)), dim3(256), 0, s, class_off, ncls, tile_cnt);
    FMMH_HIP(hipGetLastError());
}

template <bool F32>
static void launch_gemm_p(const M2LGemmArgs& a, int grid, hipStream_t s) {
    dim3 g(grid), b(256);
    switch (a.p) {
        case 0: hipLaunchKernelGGL((k_m2l_gemm<0, F32>), g, b, 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_m2l_gemm<1, F32>), g, b, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_m2l_gemm<2, F32>), g, b, 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_m2l_gemm<3, F32>), g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_m2l_gemm<4, F32>), g, b, 0, s, a); break;
        case 5: hipLaunchKernelGGL((k_m2l_gemm<5, F32>), g, b, 0, s, a); break;
        case 6: hipLaunchKernelGGL((k_m2l_gemm<6, F32>), g, b, 0, s, a); break;
        case 7: hipLaunchKernelGGL((k_m2l_gemm<7, F32>), g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_m2l_gemm<8, F32>), g, b, 0, s, a); break;
        case 9: hipLaunchKernelGGL((k_m2l_gemm<9, F32>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_m2l_gemm<10, F32>), g, b, 0, s, a); break;
    }
}

void launch_m2l_gemm(const M2LGemmArgs& a, int grid, hipStream_t s) {
    if (grid <= 0) return;
    if (a.f32) launch_gemm_p<true>(a, grid, s);
    else launch_gemm_p<false>(a, grid, s);
    FMMH_HIP(hipGetLastError());
}

void launch_m2l_reduce(const M2LGemmArgs& a, const int* off, const double* r, double* LE, int num_nodes,
                       hipStream_t s) {
    if (num_nodes <= 0) return;
    const int nm = ncomp(a.p), nk = rdim(a.p);
    if (a.f32)
        hipLaunchKernelGGL(k_m2l_reduce<float>, dim3(num_nodes), dim3(128), 0, s, a.pbuf32, off, r, a.tdeg, a.tfull,
                           a.order, LE, nk, nm);
    else
        hipLaunchKernelGGL(k_m2l_reduce<double>, dim3(num_nodes), dim3(128), 0, s, a.pbuf, off, r, a.tdeg, a.tfull,
                           a.order, LE, nk, nm);
    FMMH_HIP(hipGetLastError());
}

// FP32 copy of one set's scaled source expansions (mixed precision).
__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

void launch_to_f32(const double* in, float* out, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_f32, dim3(nblk(n, 256)), dim3(256), 0, s, in, out, n);
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
