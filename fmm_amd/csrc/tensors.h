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

// a * b + c with the constant c read from scalar registers (v_fma_f64, VOP3): the compiler
// otherwise keeps such constants in VGPRs and copies one before every accumulate-in-place
// FMA of a Horner chain.  Plain arithmetic, no memory access.
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

template <int KIND, int N>
__device__ __forceinline__ double poly_eval(double x) {
    constexpr int D = Poly<KIND>::tab.deg[N];
    if constexpr (D == 0) {
        return Poly<KIND>::tab.c[N][0];
    } else {
        const double top = Poly<KIND>::tab.c[N][D];
        double s = fma_sc(x, top, Poly<KIND>::tab.c[N][D - 1]);
#pragma unroll
        for (int i = D - 2; i >= 0; --i) s = fma_sc(s, x, Poly<KIND>::tab.c[N][i]);
        return s;
    }
}

// 2^(j/32), j = 0..31, correctly rounded (decimal arithmetic at 60 digits, then rounded to
// the nearest double); staged in LDS by the kernels that call exp_table.
static __constant__ const double kExp2By32[32] = {
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,
    0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,
    0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,
    0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,
    0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,
    0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,
    0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,
    0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0,
};

// e^x for x <= 0 by Tang's table-driven method: x = (32 m + j) ln2/32 + r, |r| <= ln2/64,
// e^x = 2^m 2^(j/32) e^r, with e^r - 1 from its degree-6 Taylor polynomial (remainder below
// 2^-58 relative) and ln2/32 split in two parts (the high part has 32 significant bits, so
// k ln2/32 is exact for |k| < 2^21).  About 1 ulp; 0 below -745 (arguments under -1000
// are clamped there).  tab: kExp2By32, usually staged in LDS.
__device__ __forceinline__ double exp_table(double x, const double* tab) {
    x = fmax(x, -1000.0);
    const double kd = __builtin_rint(x * 0x1.71547652b82fep+5);        // 32 / ln2
    double r = __builtin_fma(kd, -0x1.62e42fee00000p-6, x);             // ln2/32, high part
    r = __builtin_fma(kd, -0x1.a39ef35793c76p-38, r);                   // ln2/32, low part
    const int k = (int)kd;
    double q = fma_sc(r, 1.0 / 720.0, 1.0 / 120.0);
    q = fma_sc(q, r, 1.0 / 24.0);
    q = fma_sc(q, r, 1.0 / 6.0);
    q = __builtin_fma(q, r, 0.5);
    q = __builtin_fma(q, r, 1.0);
    q *= r;                                                              // e^r - 1
    const double t = tab[k & 31];
    return ldexp(__builtin_fma(t, q, t), k >> 5);
}

// 1/sqrt(x) for finite x > 0: the hardware estimate y refined by one Newton-Raphson step
// carried to second order, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2.  x = 0 gives NaN.
__device__ __forceinline__ double rsqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-x * y, y, 1.0);
    return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}

// g[0..K] for one pair; R2 = |r|^2, inv = 1/|r| (0 makes every g_n exactly
// 0: a pair that must not contribute), a = damping factor product.
// TAB: exponentials by exp_table on etab (kExp2By32), else the library exp.
template <int K, int DAMP, bool TAB = false>
__device__ __forceinline__ void radial_inv(double R2, double inv, double a, double (&g)[K + 1],
                                           const double* etab = nullptr) {
    static_assert(DAMP == DAMP_NONE || K <= KMAX_NEAR, "damped tensors stop at order 6");
    const double inv2 = inv * inv;
    double b = inv;
    if constexpr (DAMP == DAMP_NONE) {
        g[0] = b;
#pragma unroll
        for (int n = 1; n <= K; ++n) { b *= -(2.0 * n - 1.0) * inv2; g[n] = b; }
    } else if constexpr (DAMP == DAMP_THOLE && K <= 3) {
        // the first four polynomials differ by single leading terms: 1 + u/2, 1 + u (1 + u/2),
        // + u^3/6, + u^4/30 (the table's values), so they are built one from the next
        static_assert(Poly<DAMP_THOLE>::tab.deg[3] == 4, "Thole polynomial table");
        const double u = a * (R2 * inv);
        const double e = TAB ? exp_table(-u, etab) : exp(-u);
        const double p0 = __builtin_fma(u, 0.5, 1.0);
        g[0] = b * (1.0 - e * p0);
        if constexpr (K >= 1) {
            const double p1 = __builtin_fma(u, p0, 1.0);
            b *= -1.0 * inv2;
            g[1] = b * (1.0 - e * p1);
            if constexpr (K >= 2) {
                const double u2 = u * u;
                const double p2 = fma_sc(u2 * u, Poly<DAMP_THOLE>::tab.c[2][3], p1);
                b *= -3.0 * inv2;
                g[2] = b * (1.0 - e * p2);
                if constexpr (K >= 3) {
                    const double p3 = fma_sc(u2 * u2, Poly<DAMP_THOLE>::tab.c[3][4], p2);
                    b *= -5.0 * inv2;
                    g[3] = b * (1.0 - e * p3);
                }
            }
        }
    } else if constexpr (DAMP == DAMP_THOLE) {
        const double u = a * (R2 * inv);
        const double e = TAB ? exp_table(-u, etab) : exp(-u);
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
        const double e = TAB ? exp_table(-v, etab) : exp(-v);
        g[0] = 0.0;
        if constexpr (K >= 1) { b *= -1.0 * inv2; g[1] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 1>(v)); }
        if constexpr (K >= 2) { b *= -3.0 * inv2; g[2] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 2>(v)); }
        if constexpr (K >= 3) { b *= -5.0 * inv2; g[3] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 3>(v)); }
        if constexpr (K >= 4) { b *= -7.0 * inv2; g[4] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 4>(v)); }
        if constexpr (K >= 5) { b *= -9.0 * inv2; g[5] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 5>(v)); }
        if constexpr (K >= 6) { b *= -11.0 * inv2; g[6] = b * (1.0 - e * poly_eval<DAMP_AMOEBA, 6>(v)); }
    } else {
        const double w = sqrt(a) * (R2 * inv);
        const double e = 1.1283791670955126 * (TAB ? exp_table(-w * w, etab) : exp(-w * w));   // 2/sqrt(pi) e^{-w^2}
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

template <int K, int DAMP>
__device__ __forceinline__ void radial(double R2, double a, double (&g)[K + 1]) {
    radial_inv<K, DAMP>(R2, rsqrt(R2), a, g);
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

// Tensor order a pair of source order M and field order F needs.  The damped
// kinds stop at order 6: the reference's Tn_damp_* fill T0..T6 only and leave
// the higher components at the zero they were allocated with
// (src/tensors_damped_thole.F90:55-62, same lines in _amoeba and _erf;
// src/fmm.F90:576-577), so their terms with |s| + |t| > 6 vanish.
template <int M, int F, int DAMP>
__host__ __device__ constexpr int pair_tensor_order() {
    return (DAMP != DAMP_NONE && M + F > KMAX_NEAR) ? KMAX_NEAR : M + F;
}

// acc[t] += sum_s m[s] T_{s+t}, |s| <= M, |t| <= F, |s + t| <= KT (T holds the
// orders <= KT; KT = M + F unless the kind truncates); m already carries the
// Taylor/symmetry factor (-1)^{|s|+1} / (sx! sy! sz!) of src/fmm.F90:597-612.
template <int M, int F, int KT = M + F>
__device__ __forceinline__ void contract(const double (&m)[ncomp(M)], const double (&T)[ncomp(KT)],
                                         double (&acc)[ncomp(F)]) {
#pragma unroll
    for (int t = 0; t < ncomp(F); ++t) {
#pragma unroll
        for (int s = 0; s < ncomp(M); ++s) {
            const Exps et = idx2xyz(t), es = idx2xyz(s);
            if (et.x + et.y + et.z + es.x + es.y + es.z > KT) continue;
            acc[t] = __builtin_fma(m[s], T[xyz2idx(et.x + es.x, et.y + es.y, et.z + es.z)], acc[t]);
        }
    }
}

// Closed forms of the same contraction for the source / field orders the
// near field uses most, built directly from the radial derivatives g_n
// (d_c g_n = x_c g_{n+1}): with the source as the operator
// D = m_0 + sum_a m_a d_a + sum_{a<=b} m_ab d_a d_b,
//   D f   = m_0 g0 + (d + tr) g1 + Q g2,       d = sum_a m_a x_a, tr = m_xx + m_yy + m_zz,
//                                             Q = sum_{a<=b} m_ab x_a x_b, q_c = d_c Q
//   d_c D f = x_c A1 + m_c g1 + q_c g2,      A1 = m_0 g1 + (d + tr) g2 + Q g3
//   d_c d_e D f (m_ab = 0) = delta_ce A1 + x_c x_e B + g2 (x_c m_e + m_c x_e),  B = m_0 g2 + d g3.
// Far fewer operations than building every T component and contracting
// them; the orders without a closed form here (M = F = 2) use contract().
template <int M, int F>
constexpr bool has_closed_form() { return M <= 2 && F <= 2 && !(M == 2 && F == 2); }

template <int M, int F>
__device__ __forceinline__ void contract_closed(double x, double y, double z, const double (&g)[M + F + 1],
                                                const double (&m)[ncomp(M)], double (&acc)[ncomp(F)]) {
    static_assert(has_closed_form<M, F>(), "no closed form for these orders");
    const double m0 = m[0];
    double dt = 0.0, d = 0.0, Q = 0.0, qx = 0.0, qy = 0.0, qz = 0.0;
    if constexpr (M >= 1) {
        d = __builtin_fma(m[1], x, __builtin_fma(m[2], y, m[3] * z));
        dt = d;
    }
    if constexpr (M >= 2) {
        // packed quadrupole: 4 xx, 5 xy, 6 xz, 7 yy, 8 yz, 9 zz
        dt = d + ((m[4] + m[7]) + m[9]);
        qx = __builtin_fma(2.0 * m[4], x, __builtin_fma(m[5], y, m[6] * z));
        qy = __builtin_fma(m[5], x, __builtin_fma(2.0 * m[7], y, m[8] * z));
        qz = __builtin_fma(m[6], x, __builtin_fma(m[8], y, 2.0 * m[9] * z));
        Q = 0.5 * __builtin_fma(qx, x, __builtin_fma(qy, y, qz * z));
    }
    acc[0] = __builtin_fma(m0, g[0], acc[0]);
    if constexpr (M >= 1) acc[0] = __builtin_fma(dt, g[1], acc[0]);
    if constexpr (M >= 2) acc[0] = __builtin_fma(Q, g[2], acc[0]);
    if constexpr (F >= 1) {
        double A1 = m0 * g[1];
        if constexpr (M >= 1) A1 = __builtin_fma(dt, g[2], A1);
        if constexpr (M >= 2) A1 = __builtin_fma(Q, g[3], A1);
        acc[1] = __builtin_fma(x, A1, acc[1]);
        acc[2] = __builtin_fma(y, A1, acc[2]);
        acc[3] = __builtin_fma(z, A1, acc[3]);
        if constexpr (M >= 1) {
            acc[1] = __builtin_fma(m[1], g[1], acc[1]);
            acc[2] = __builtin_fma(m[2], g[1], acc[2]);
            acc[3] = __builtin_fma(m[3], g[1], acc[3]);
        }
        if constexpr (M >= 2) {
            acc[1] = __builtin_fma(qx, g[2], acc[1]);
            acc[2] = __builtin_fma(qy, g[2], acc[2]);
            acc[3] = __builtin_fma(qz, g[2], acc[3]);
        }
        if constexpr (F >= 2) {
            // Phi_ce = delta_ce A1 + x_c w_e + u_c x_e,  u = g2 m,  w = B r + u
            double B = m0 * g[2];
            if constexpr (M >= 1) B = __builtin_fma(d, g[3], B);
            if constexpr (M >= 1) {
                const double ux = g[2] * m[1], uy = g[2] * m[2], uz = g[2] * m[3];
                const double wx = __builtin_fma(B, x, ux), wy = __builtin_fma(B, y, uy), wz = __builtin_fma(B, z, uz);
                acc[4] = __builtin_fma(x, wx + ux, acc[4] + A1);
                acc[5] = __builtin_fma(x, wy, __builtin_fma(ux, y, acc[5]));
                acc[6] = __builtin_fma(x, wz, __builtin_fma(ux, z, acc[6]));
                acc[7] = __builtin_fma(y, wy + uy, acc[7] + A1);
                acc[8] = __builtin_fma(y, wz, __builtin_fma(uy, z, acc[8]));
                acc[9] = __builtin_fma(z, wz + uz, acc[9] + A1);
            } else {
                const double wx = B * x, wy = B * y, wz = B * z;
                acc[4] = __builtin_fma(x, wx, acc[4] + A1);
                acc[5] = __builtin_fma(x, wy, acc[5]);
                acc[6] = __builtin_fma(x, wz, acc[6]);
                acc[7] = __builtin_fma(y, wy, acc[7] + A1);
                acc[8] = __builtin_fma(y, wz, acc[8]);
                acc[9] = __builtin_fma(z, wz, acc[9] + A1);
            }
        }
    }
}

// One pair's contribution acc[t] += sum_s m[s] T_{s+t}(x, y, z) with damping
// factor product a; closed form where there is one.  The radial part is computed as
// the near-field pair loop computes it (rsqrt_pos, exp_table on etab).
// The pair's contribution from its radial derivatives gr[0..KT] (KT =
// pair_tensor_order): closed form where there is one, else the tensor and the
// truncated contraction.
template <int M, int F, int DAMP>
__device__ __forceinline__ void pair_contract(double x, double y, double z,
                                              const double (&gr)[pair_tensor_order<M, F, DAMP>() + 1],
                                              const double (&m)[ncomp(M)], double (&acc)[ncomp(F)]) {
    constexpr int KT = pair_tensor_order<M, F, DAMP>();
    if constexpr (has_closed_form<M, F>()) {
        contract_closed<M, F>(x, y, z, gr, m, acc);
    } else {
        double T[ncomp(KT)];
        md_tensor<KT>(x, y, z, gr, T);
        contract<M, F, KT>(m, T, acc);
    }
}

template <int M, int F, int DAMP>
__device__ __forceinline__ void pair_field(double x, double y, double z, double R2, double a,
                                           const double (&m)[ncomp(M)], double (&acc)[ncomp(F)],
                                           const double* etab) {
    constexpr int KT = pair_tensor_order<M, F, DAMP>();
    double gr[KT + 1];
    radial_inv<KT, DAMP, DAMP != DAMP_NONE>(R2, rsqrt_pos(R2), a, gr, etab);
    pair_contract<M, F, DAMP>(x, y, z, gr, m, acc);
}

}  // namespace fmmh
