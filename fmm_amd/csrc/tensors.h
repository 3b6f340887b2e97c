// Cartesian interaction tensors T_{tuv} = d^t/dx^t d^u/dy^u d^v/dz^v f(R)
// for the four kernels of peter-reinholdt/fmm:
//   ''      f = 1/R                              src/tensors.F90:128-139 (+ T1..T12)
//   THOLE   f = (1 - (1 + aR/2) e^{-aR}) / R     src/tensors_damped_thole.F90:65-101
//   AMOEBA  T1 = -(1 - e^{-a^3 R^3}) r / R^3, T0 = 0   src/tensors_damped_amoeba.F90:65-98
//   ERF     f = erf(sqrt(a) R) / R               src/tensors_damped_erf.F90:65-99
// The reference ships closed forms per order (and a coefficient recursion
// for high orders).  Here every kind goes through one path: the radial
// derivative sequence g_n = ((1/R) d/dR)^n f, written as
//   g_n = (-1)^n (2n-1)!! R^{-(2n+1)} lambda_n,
// with lambda_n a damping factor from a short polynomial recursion, then the
// Hermite-type recursion A^{(n)}_{t+1,u,v} = t A^{(n+1)}_{t-1,u,v} + x A^{(n+1)}_{t,u,v}
// (A^{(n)}_{000} = g_n, T = A^{(0)}), which needs no per-order code.
// The near-field form below is fully unrolled at compile time so that the
// whole recursion lives in VGPRs.
#pragma once
#include "common.h"

namespace fmmh {

constexpr int POLY_N = 8;        // damping polynomials for n = 0..7
constexpr int POLY_D = 2 * POLY_N + 2;

struct PolyTable {
    double c[POLY_N][POLY_D];
    int deg[POLY_N];
};

// lambda_n polynomials (ascending powers):
//   THOLE  P_0 = 1 + u/2,  P_{n+1} = P_n + u (P_n - P_n') / (2n+1)          (u = aR)
//   AMOEBA Q_1 = 1,        Q_{n+1} = Q_n + 3v (Q_n - Q_n') / (2n+1)         (v = a^3 R^3)
//   ERF    S_0 = 0,        S_{n+1} = S_n + w (1 - S_n' + 2 w S_n) / (2n+1)  (w = sqrt(a) R)
// obtained by differentiating g_n once more and collecting the exponential.
__host__ __device__ constexpr PolyTable make_poly(int kind) {
    PolyTable t{};
    for (int n = 0; n < POLY_N; ++n) {
        t.deg[n] = 0;
        for (int i = 0; i < POLY_D; ++i) t.c[n][i] = 0.0;
    }
    int n0 = 0;
    if (kind == DAMP_THOLE) { t.c[0][0] = 1.0; t.c[0][1] = 0.5; t.deg[0] = 1; n0 = 0; }
    if (kind == DAMP_AMOEBA) { t.c[1][0] = 1.0; t.deg[1] = 0; n0 = 1; }
    if (kind == DAMP_ERF) { t.deg[0] = 0; n0 = 0; }
    for (int n = n0; n + 1 < POLY_N; ++n) {
        double inv = 1.0 / (2.0 * n + 1.0);
        int d = t.deg[n];
        double dp[POLY_D] = {};
        for (int i = 1; i <= d; ++i) dp[i - 1] = t.c[n][i] * i;
        double q[POLY_D] = {};
        for (int i = 0; i <= d; ++i) q[i] += t.c[n][i];
        if (kind == DAMP_THOLE) {
            for (int i = 0; i <= d; ++i) q[i + 1] += t.c[n][i] * inv;
            for (int i = 0; i < d; ++i) q[i + 1] -= dp[i] * inv;
            t.deg[n + 1] = d + 1;
        } else if (kind == DAMP_AMOEBA) {
            for (int i = 0; i <= d; ++i) q[i + 1] += 3.0 * t.c[n][i] * inv;
            for (int i = 0; i < d; ++i) q[i + 1] -= 3.0 * dp[i] * inv;
            t.deg[n + 1] = d + 1;
        } else {
            q[1] += inv;
            for (int i = 0; i < d; ++i) q[i + 1] -= dp[i] * inv;
            for (int i = 0; i <= d; ++i) q[i + 2] += 2.0 * t.c[n][i] * inv;
            t.deg[n + 1] = d + 2;
        }
        for (int i = 0; i < POLY_D; ++i) t.c[n + 1][i] = q[i];
    }
    return t;
}

template <int KIND>
struct Poly {
    static constexpr PolyTable tab = make_poly(KIND);
};

template <int KIND, int N>
__device__ __forceinline__ double poly_eval(double x) {
    constexpr int D = Poly<KIND>::tab.deg[N];
    double s = Poly<KIND>::tab.c[N][D];
#pragma unroll
    for (int i = D - 1; i >= 0; --i) s = __builtin_fma(s, x, Poly<KIND>::tab.c[N][i]);
    return s;
}

// g[0..K] for one pair; R2 = |r|^2, a = damping factor product.
template <int K, int DAMP>
__device__ __forceinline__ void radial(double R2, double a, double (&g)[K + 1]) {
    const double inv = rsqrt(R2);
    const double inv2 = inv * inv;
    double b = inv;
    if constexpr (DAMP == DAMP_NONE) {
        g[0] = b;
#pragma unroll
        for (int n = 1; n <= K; ++n) { b *= -(2.0 * n - 1.0) * inv2; g[n] = b; }
    } else if constexpr (DAMP == DAMP_THOLE) {
        const double u = a * (R2 * inv);
        const double e = exp(-u);
        g[0] = b * (1.0 - e * poly_eval<DAMP_THOLE, 0>(u));
        if constexpr (K >= 1) { b *= -1.0 * inv2; g[1] = b * (1.0 - e * poly_eval<DAMP_THOLE, 1>(u)); }
        if constexpr (K >= 2) { b *= -3.0 * inv2; g[2] = b * (1.0 - e * poly_eval<DAMP_THOLE, 2>(u)); }
        if constexpr (K >= 3) { b *= -5.0 * inv2; g[3] = b * (1.0 - e * poly_eval<DAMP_THOLE, 3>(u)); }
        if constexpr (K >= 4) { b *= -7.0 * inv2; g[4] = b * (1.0 - e * poly_eval<DAMP_THOLE, 4>(u)); }
        if constexpr (K >= 5) { b *= -9.0 * inv2; g[5] = b * (1.0 - e * poly_eval<DAMP_THOLE, 5>(u)); }
        if constexpr (K >= 6) { b *= -11.0 * inv2; g[6] = b * (1.0 - e * poly_eval<DAMP_THOLE, 6>(u)); }
    } else if constexpr (DAMP == DAMP_AMOEBA) {
        const double R = R2 * inv;
        const double v = a * a * a * (R2 * R);
        const double e = exp(-v);
        g[0] = 0.0;
        if constexpr (K >= 1) { b *= -1.0 * inv2; g[1] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 1>(v)); }
        if constexpr (K >= 2) { b *= -3.0 * inv2; g[2] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 2>(v)); }
        if constexpr (K >= 3) { b *= -5.0 * inv2; g[3] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 3>(v)); }
        if constexpr (K >= 4) { b *= -7.0 * inv2; g[4] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 4>(v)); }
        if constexpr (K >= 5) { b *= -9.0 * inv2; g[5] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 5>(v)); }
        if constexpr (K >= 6) { b *= -11.0 * inv2; g[6] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 6>(v)); }
    } else {
        const double w = sqrt(a) * (R2 * inv);
        const double e = 1.1283791670955126 * exp(-w * w);   // 2/sqrt(pi) e^{-w^2}
        const double er = erf(w);
        g[0] = b * (er - e * poly_eval<DAMP_ERF, 0>(w));
        if constexpr (K >= 1) { b *= -1.0 * inv2; g[1] = b * (er - e * poly_eval<DAMP_ERF, 1>(w)); }
        if constexpr (K >= 2) { b *= -3.0 * inv2; g[2] = b * (er - e * poly_eval<DAMP_ERF, 2>(w)); }
        if constexpr (K >= 3) { b *= -5.0 * inv2; g[3] = b * (er - e * poly_eval<DAMP_ERF, 3>(w)); }
        if constexpr (K >= 4) { b *= -7.0 * inv2; g[4] = b * (er - e * poly_eval<DAMP_ERF, 4>(w)); }
        if constexpr (K >= 5) { b *= -9.0 * inv2; g[5] = b * (er - e * poly_eval<DAMP_ERF, 5>(w)); }
        if constexpr (K >= 6) { b *= -11.0 * inv2; g[6] = b * (er - e * poly_eval<DAMP_ERF, 6>(w)); }
    }
}

// Register-resident Hermite recursion; T[i] for packed index i, orders <= K.
template <int K>
__device__ __forceinline__ void md_tensor(double x, double y, double z, const double (&g)[K + 1],
                                          double (&T)[ncomp(K)]) {
    double A[K + 1][ncomp(K)];
#pragma unroll
    for (int n = 0; n <= K; ++n) A[n][0] = g[n];
#pragma unroll
    for (int k = 1; k <= K; ++k) {
#pragma unroll
        for (int t = k; t >= 0; --t) {
#pragma unroll
            for (int u = k - t; u >= 0; --u) {
                const int v = k - t - u;
                const int i = xyz2idx(t, u, v);
#pragma unroll
                for (int n = 0; n <= K - k; ++n) {
                    double val;
                    if (t > 0) {
                        val = x * A[n + 1][xyz2idx(t - 1, u, v)];
                        if (t > 1) val = __builtin_fma((double)(t - 1), A[n + 1][xyz2idx(t - 2, u, v)], val);
                    } else if (u > 0) {
                        val = y * A[n + 1][xyz2idx(t, u - 1, v)];
                        if (u > 1) val = __builtin_fma((double)(u - 1), A[n + 1][xyz2idx(t, u - 2, v)], val);
                    } else {
                        val = z * A[n + 1][xyz2idx(t, u, v - 1)];
                        if (v > 1) val = __builtin_fma((double)(v - 1), A[n + 1][xyz2idx(t, u, v - 2)], val);
                    }
                    A[n][i] = val;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < ncomp(K); ++i) T[i] = A[0][i];
}

// Exponents of packed index i (compile-time enumeration).
struct Exps { int x, y, z; };
__host__ __device__ constexpr Exps idx2xyz(int i) {
    int k = 0;
    while (ncomp(k) <= i) ++k;
    for (int x = k; x >= 0; --x)
        for (int y = k - x; y >= 0; --y)
            if (xyz2idx(x, y, k - x - y) == i) return Exps{x, y, k - x - y};
    return Exps{0, 0, 0};
}

// acc[t] += sum_s m[s] T_{s+t}, |s| <= M, |t| <= F; m already carries the
// Taylor/symmetry factor (-1)^{|s|+1} / (sx! sy! sz!) of src/fmm.F90:597-612.
template <int M, int F>
__device__ __forceinline__ void contract(const double (&m)[ncomp(M)], const double (&T)[ncomp(M + F)],
                                         double (&acc)[ncomp(F)]) {
#pragma unroll
    for (int t = 0; t < ncomp(F); ++t) {
#pragma unroll
        for (int s = 0; s < ncomp(M); ++s) {
            const Exps et = idx2xyz(t), es = idx2xyz(s);
            acc[t] = __builtin_fma(m[s], T[xyz2idx(et.x + es.x, et.y + es.y, et.z + es.z)], acc[t]);
        }
    }
}

}  // namespace fmmh
