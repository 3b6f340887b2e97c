// Undamped Cartesian tensors T_k(v) = d^k (1/|v|), |k| <= K, for one vector
// per 64-lane workgroup, by the Hermite recursion over degree levels in LDS
// (one barrier per level):
//   A^{(n)}_{000} = g_n = (-1)^n (2n-1)!! |v|^{-(2n+1)},
//   A^{(n)}_{t+1,u,w} = t A^{(n+1)}_{t-1,u,w} + x A^{(n+1)}_{t,u,w}  (likewise in y, z).
// A holds level k's entries for n = 0..K-k at level_base[k]; T_idx with
// |idx| = k is A[level_base[k] + idx - ncomp(k - 1)].  Used for the M2L class
// tensors and the periodic shell sums.
#pragma once
#include "common.h"

namespace fmmh {

constexpr int HERMITE_KMAX = 2 * PMAX;

__host__ __device__ constexpr int hermite_aux_size(int K) { return (K + 1) * (K + 2) * (K + 3) * (K + 4) / 24; }

// zmax < K: only the entries with v <= zmax (the recursion for (t, u, v) reads entries with v' <= v, so
// they are complete; the others are left unset).
__device__ inline void hermite_aux(double vx, double vy, double vz, int K, double* A, const int* __restrict__ xyz,
                                   const int* __restrict__ level_base, int zmax = HERMITE_KMAX) {
    if ((int)threadIdx.x <= K) {
        const double R2 = vx * vx + vy * vy + vz * vz;
        const double inv = 1.0 / sqrt(R2);
        const double inv2 = inv * inv;
        double g = inv;
        for (int n = 1; n <= (int)threadIdx.x; ++n) g *= -(2.0 * n - 1.0) * inv2;
        A[threadIdx.x] = g;
    }
    __syncthreads();
    for (int k = 1; k <= K; ++k) {
        const int nd = (k + 1) * (k + 2) / 2, ndm1 = k * (k + 1) / 2, ndm2 = (k - 1) * k / 2;
        const int vmax = k < zmax ? k : zmax;
        const int ndc = (vmax + 1) * (k + 1) - vmax * (vmax + 1) / 2;   // components with v <= vmax
        const int nent = ndc * (K - k + 1);
        const int base = level_base[k], bm1 = level_base[k - 1], bm2 = k >= 2 ? level_base[k - 2] : 0;
        const int c0 = ncomp(k - 1), c1 = ncomp(k - 2), c2 = ncomp(k - 3);
        for (int x = threadIdx.x; x < nent; x += 64) {
            const int n = x / ndc;
            int l = x - n * ndc, t, u, v;
            if (vmax == k) {   // every component, in packed order
                const int pk = xyz[c0 + l];
                t = pk & 255; u = (pk >> 8) & 255; v = pk >> 16;
            } else {           // v = 0, 1, .., vmax, each with its k - v + 1 components
                v = 0;
                while (l > k - v) {
                    l -= k - v + 1;
                    ++v;
                }
                u = l;
                t = k - v - u;
                l = xyz2idx(t, u, v) - c0;
            }
            double val;
            if (t > 0) {
                val = vx * A[bm1 + (n + 1) * ndm1 + (xyz2idx(t - 1, u, v) - c1)];
                if (t > 1) val = __builtin_fma((double)(t - 1), A[bm2 + (n + 1) * ndm2 + (xyz2idx(t - 2, u, v) - c2)], val);
            } else if (u > 0) {
                val = vy * A[bm1 + (n + 1) * ndm1 + (xyz2idx(t, u - 1, v) - c1)];
                if (u > 1) val = __builtin_fma((double)(u - 1), A[bm2 + (n + 1) * ndm2 + (xyz2idx(t, u - 2, v) - c2)], val);
            } else {
                val = vz * A[bm1 + (n + 1) * ndm1 + (xyz2idx(t, u, v - 1) - c1)];
                if (v > 1) val = __builtin_fma((double)(v - 1), A[bm2 + (n + 1) * ndm2 + (xyz2idx(t, u, v - 2) - c2)], val);
            }
            A[base + n * nd + l] = val;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double hermite_T(const double* A, int idx, const int* __restrict__ xyz,
                                            const int* __restrict__ level_base) {
    const int pk = xyz[idx];
    const int k = (pk & 255) + ((pk >> 8) & 255) + (pk >> 16);
    return A[level_base[k] + (idx - ncomp(k - 1))];
}

}  // namespace fmmh
