// P2M and L2P for many multipole sets on the FP64 matrix cores
// (src/fmm.F90:437-472, 520-534 and 1046-1132).  Both are GEMMs over the
// sites of a leaf with the sets as columns:
//   P2M  CM[t][set] = sum_(j, s) coef(t, s) d_j^(t - s) M_j(s)[set]
//        (rows: the expansion coefficients t; k: site j and source
//        component s; d_j = x_j - leaf centre; coef and the monomial index
//        from the P2M shift table)
//   L2P  F_j(c)[set] = sum_t d_j^(t - c) / (t - c)! L(t)[set]
//        (rows: site j; k: the (p + 1)^2 harmonic components of t, see
//        k_l2p_basis; one accumulator per field component c, order <= 1)
// P2M: one workgroup of four wavefronts per (leaf, 16 sets): monomials of a
// chunk of sites and the multipoles are staged in LDS, each wavefront runs its
// 16-row tiles of v_mfma_f64_16x16x4f64, A gathered from the monomials
// through a table built in LDS per workgroup.  L2P: a per-site basis pass,
// then one workgroup of two wavefronts per (leaf, 16 sets) with the leaf's
// local expansion staged in LDS.  FP64 throughout, like the one-set kernels
// (expansions.hip); the sums run in another order, so a set of a many-set
// call agrees with its one-set call to rounding (the vector-ALU many-set
// kernels, used below EXP_MFMA_MIN_SETS sets, agree bit for bit).
#include "common.h"
#include "kernels.h"

namespace fmmh {

using v4d = double __attribute__((ext_vector_type(4)));

namespace {

constexpr int XM_W = 4;        // wavefronts per workgroup
constexpr int XM_CH = 16;      // sites per P2M chunk
constexpr int XM_PMAX = EXP_MFMA_PMAX;

__device__ __forceinline__ void powers(double d, int p, double* w) {
    double v = 1.0;
    for (int k = 0; k <= p; ++k) {
        w[k] = v;
        v *= d;
    }
}

// NS source components (a.nsrc); LDS: monomials of a chunk of sites, the B operand and the (t, s) table.
// NTL 16-set tiles per workgroup: each gathered A element feeds NTL MFMAs (each tile's sums are those of a
// workgroup of one tile).
template <int P, int NS, int NTL>
__global__ __launch_bounds__(64 * XM_W) void k_p2m_mfma(UpArgs a) {
    constexpr int NM = ncomp(P);
    constexpr int MT = (NM + 15) / 16;
    constexpr int MW = (MT + XM_W - 1) / XM_W;
    constexpr int nsrc = NS;
    constexpr int kmax = (XM_CH * NS + 3) / 4 * 4;
    constexpr int BW = 16 * NTL;            // sets per workgroup
    __shared__ double mono[XM_CH * NM];
    __shared__ double B[kmax * BW];
    __shared__ double tabC[NM * NS];
    __shared__ int tabE[NM * NS];
    __shared__ double pw[XM_CH][3][XM_PMAX + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int node = a.leaf_node[blockIdx.x];
    const int set0 = blockIdx.y * BW;
    const int s0 = a.start[node], c = a.count[node];
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    for (int x = tid; x < NM * nsrc; x += 64 * XM_W) {
        tabE[x] = -1;
        tabC[x] = 0.0;
    }
    __syncthreads();
    for (int t = tid; t < NM; t += 64 * XM_W)
        for (int b = a.p2m.off[t]; b < a.p2m.off[t + 1]; ++b) {
            const int sc = a.p2m.s[b];
            tabE[t * nsrc + sc] = a.p2m.e[b];
            tabC[t * nsrc + sc] = a.p2m.coef[b];
        }
    v4d acc[MW][NTL];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi)
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) acc[mi][nt] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < c; j0 += XM_CH) {
        const int cc = min(XM_CH, c - j0);
        __syncthreads();   // readers of the previous chunk are done (and the tables are in place)
        if (tid < 3 * cc) {
            const int jj = tid / 3, ax = tid - 3 * jj;
            const int j = s0 + j0 + jj;
            const double d = ax == 0 ? a.xs[j] - cx : (ax == 1 ? a.ys[j] - cy : a.zs[j] - cz);
            powers(d, P, pw[jj][ax]);
        }
        for (int x = tid; x < kmax * BW; x += 64 * XM_W) {
            const int k = x / BW, u = x - k * BW;
            const int jj = k / nsrc, sc = k - jj * nsrc;
            B[x] = (jj < cc && set0 + u < a.nsets)
                       ? a.ms_raw[((int64_t)(set0 + u) * a.dim + sc) * a.ldm + s0 + j0 + jj] : 0.0;
        }
        __syncthreads();
        for (int x = tid; x < cc * NM; x += 64 * XM_W) {
            const int jj = x / NM, e = x - jj * NM;
            const int pk = a.xyz[e];
            mono[jj * NM + e] = (pw[jj][0][pk & 255] * pw[jj][1][(pk >> 8) & 255]) * pw[jj][2][pk >> 16];
        }
        __syncthreads();
        const int nk = (cc * nsrc + 3) / 4;
        for (int ks = 0; ks < nk; ++ks) {
            const int k = 4 * ks + kq;
            const int jj = k / nsrc, sc = k - jj * nsrc;
            const bool kv = jj < cc;
            double bv[NTL];
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) bv[nt] = B[k * BW + 16 * nt + row];
#pragma unroll
            for (int mi = 0; mi < MW; ++mi) {
                const int t = (w + XM_W * mi) * 16 + row;
                double av = 0.0;
                if (t < NM && kv) {
                    const int e = tabE[t * nsrc + sc];
                    if (e >= 0) av = tabC[t * nsrc + sc] * mono[jj * NM + e];
                }
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt)
                    acc[mi][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[nt], acc[mi][nt], 0, 0, 0);
            }
        }
    }
    // accumulator register q of lane l holds D[coefficient 4 q + l / 16][set l % 16]
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt) {
        const int set = set0 + 16 * nt + row;
        if (set >= a.nsets) continue;
#pragma unroll
        for (int mi = 0; mi < MW; ++mi)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = (w + XM_W * mi) * 16 + 4 * q + kq;
                if (t < NM) a.CM[((int64_t)set * a.num_nodes + node) * NM + t] = acc[mi][nt][q];
            }
    }
}

// Harmonic form.  The local expansion of the far field is the Taylor expansion of a harmonic function,
// L(t) = d^t phi, so L(t + 2x) + L(t + 2y) + L(t + 2z) = 0 (the M2L builds its components with t_z >= 2
// from those with t_z <= 1 by this identity, see m2l.hip, and the L2L shifts preserve it).  Then
//   F_j(0) = sum_t d_j^t / t! L(t) = sum_{s : s_z <= 1} H_j(s) L(s)
// over the (p + 1)^2 reduced components s, with H_j(s) = H(s)(d_j) the fold of the monomials,
//   H(X, Y, rho)(d) = sum_{k, i} (-1)^k C(k, i) d^(X - 2i, Y - 2(k - i), rho + 2k) / (...)!,
// a homogeneous harmonic polynomial of degree n = X + Y + rho.  Its derivatives are basis polynomials
// again: d/dx H(X, Y, rho) = H(X - 1, Y, rho), d/dy H(X, Y, rho) = H(X, Y - 1, rho),
// d/dz H(X, Y, 1) = H(X, Y, 0), d/dz H(X, Y, 0) = -(H(X, Y - 2, 1) + H(X - 2, Y, 1)), which gives
//  (a) the basis by Euler's relation for homogeneous polynomials, n H = x H_x + y H_y + z H_z, from the
//      degree below (k_l2p_basis: about four FMAs per component, once per site);
//  (b) the field components as the same contraction against shifted components of L:
//      F_j(x) = sum_s H_j(s) L(s + x), F_j(y) = sum_s H_j(s) L(s + y),
//      F_j(z) = sum_{s_z = 0} H_j(s) L(s + z) - sum_{s_z = 1} H_j(s) [L(s + 2y - z) + L(s + 2x - z)]
//      (components past order p are zero).
// The L2P is then one GEMM per (leaf, 16 sets): rows the leaf's sites, k the reduced components,
// columns the sets, one accumulator per field component.  Only for local expansions this library built:
// those a caller passes in (interpolate_electron_fields) need not be harmonic and take the vector-ALU
// kernels.

// reduced component (X, Y, rho), rho = t_z <= 1: degree n = X + Y + rho, then w = Y + rho, then rho
__host__ __device__ constexpr int red_index(int X, int Y, int rho) {
    return (X + Y + rho) * (X + Y + rho) + (Y + rho == 0 ? 0 : 2 * (Y + rho) - 1 + rho);
}

// H_j(s) for every site j of a leaf, stored [s][sorted site] (a.ld_basis per component row).
template <int P>
__global__ __launch_bounds__(64) void k_l2p_basis(DownArgs a) {
    const int node = a.leaf_node[blockIdx.x];
    const int s0 = a.start[node], c = a.count[node];
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    double* out = a.basis;
    for (int jj = threadIdx.x; jj < c; jj += 64) {
        const int j = s0 + jj;
        const double x = a.xs[j] - cx, y = a.ys[j] - cy, z = a.zs[j] - cz;
        double h[rdim(P)];
        h[0] = 1.0;
        out[j] = 1.0;
#pragma unroll
        for (int n = 1; n <= P; ++n) {
            const double inv = 1.0 / (double)n;
#pragma unroll
            for (int w = 0; w <= n; ++w) {
#pragma unroll
                for (int rho = 0; rho <= 1; ++rho) {
                    if (rho > w) continue;
                    const int X = n - w, Y = w - rho;
                    double v = X >= 1 ? x * h[red_index(X - 1, Y, rho)] : 0.0;
                    if (Y >= 1) v = fma(y, h[red_index(X, Y - 1, rho)], v);
                    if (rho == 1) {
                        v = fma(z, h[red_index(X, Y, 0)], v);
                    } else if (X >= 2 || Y >= 2) {
                        double u = 0.0;
                        if (Y >= 2) u = h[red_index(X, Y - 2, 1)];
                        if (X >= 2) u += h[red_index(X - 2, Y, 1)];
                        v = fma(-z, u, v);
                    }
                    const int r = red_index(X, Y, rho);
                    h[r] = v * inv;
                    out[(int64_t)r * a.ld_basis + j] = h[r];
                }
            }
        }
    }
}

// LDS: the leaf's reduced local expansion for 16 sets (rows past the last component and the extra last row
// are zero) and, per field component and k, the two rows of it that make the B operand (and the sign).
constexpr int XL_W = 2;
template <int P, int F>
__global__ __launch_bounds__(64 * XL_W) void k_l2p_mfma(DownArgs a) {
    constexpr int NM = ncomp(P);
    constexpr int NK = rdim(P);
    constexpr int NF = ncomp(F);
    constexpr int KS = (NK + 3) / 4;
    constexpr int ZR = 4 * KS;             // the zero row
    __shared__ double L[(ZR + 1) * 16];
    __shared__ int gi[NF][4 * KS];         // row1 | row2 << 8 | negate << 16
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int node = a.leaf_node[blockIdx.x];
    const int set0 = blockIdx.y * 16;
    const int s0 = a.start[node], c = a.count[node];
    for (int x = tid; x < (ZR + 1) * 16; x += 64 * XL_W) {
        const int t = x >> 4, u = x & 15;
        double v = 0.0;
        if (t < NK && set0 + u < a.nsets) {
            int n = 0;
            while ((n + 1) * (n + 1) <= t) ++n;
            const int pos = t - n * n, wv = (pos + 1) >> 1, rho = (pos > 0 && (pos & 1) == 0) ? 1 : 0;
            v = a.LE[((int64_t)(set0 + u) * a.num_nodes + node) * NM + xyz2idx(n - wv, wv - rho, rho)];
        }
        L[x] = v;
    }
    for (int x = tid; x < NF * 4 * KS; x += 64 * XL_W) {
        const int cf = x / (4 * KS), t = x - cf * (4 * KS);
        int r1 = ZR, r2 = ZR, neg = 0;
        if (t < NK) {
            int n = 0;
            while ((n + 1) * (n + 1) <= t) ++n;
            const int pos = t - n * n, wv = (pos + 1) >> 1, rho = (pos > 0 && (pos & 1) == 0) ? 1 : 0;
            const int X = n - wv, Y = wv - rho;
            if (cf == 0) r1 = t;
            else if (n < P) {
                if (cf == 1) r1 = red_index(X + 1, Y, rho);
                else if (cf == 2) r1 = red_index(X, Y + 1, rho);
                else if (rho == 0) r1 = red_index(X, Y, 1);
                else {
                    r1 = red_index(X, Y + 2, 0);
                    r2 = red_index(X + 2, Y, 0);
                    neg = 1;
                }
            }
        }
        gi[cf][t] = r1 | (r2 << 8) | (neg << 16);
    }
    __syncthreads();
    const int set = set0 + row;
    for (int tile = w; tile * 16 < c; tile += XL_W) {
        const int jj = tile * 16 + row;            // A row: site of the leaf
        const bool okj = jj < c;
        const double* hb = a.basis + s0 + jj;
        v4d acc[NF];
#pragma unroll
        for (int cf = 0; cf < NF; ++cf) acc[cf] = v4d{0.0, 0.0, 0.0, 0.0};
        // rolled loop with the next A element loaded ahead (unrolled, the compiler keeps the B gather for
        // every k live: 256 VGPRs and one wavefront per SIMD at order 10)
        double an = (okj && kq < NK) ? hb[kq * a.ld_basis] : 0.0;
#pragma unroll 1
        for (int ks = 0; ks < KS; ++ks) {
            const int t = 4 * ks + kq;
            const double av = an;
            an = (okj && t + 4 < NK) ? hb[(int64_t)(t + 4) * a.ld_basis] : 0.0;
#pragma unroll
            for (int cf = 0; cf < NF; ++cf) {
                const int g = gi[cf][t];
                double bv = L[(g & 255) * 16 + row];
                if (cf == 3) {
                    bv += L[((g >> 8) & 255) * 16 + row];
                    if (g >> 16) bv = -bv;
                }
                acc[cf] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[cf], 0, 0, 0);
            }
        }
        // accumulator register q of lane l holds D[site 4 q + l / 16][set l % 16]
        if (set < a.nsets) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int sj = tile * 16 + 4 * q + kq;
                if (sj >= c) continue;
                const int k = s0 + sj;
                const int o = a.perm[k];
#pragma unroll
                for (int cf = 0; cf < NF; ++cf) {
                    double* fp = a.field + ((int64_t)set * NF + cf) * a.ld_field + o;
                    *fp = (*fp + a.near[((int64_t)set * NF + cf) * a.ld_near + k]) + acc[cf][q];
                }
            }
        }
    }
}

template <int P, int NTL>
void p2m_pt(const UpArgs& a, int num_leaves, hipStream_t s) {
    const dim3 g(num_leaves, (a.nsets + 16 * NTL - 1) / (16 * NTL)), b(64 * XM_W);
    switch (a.nsrc) {
        case 1: hipLaunchKernelGGL((k_p2m_mfma<P, 1, NTL>), g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_p2m_mfma<P, 4, NTL>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_p2m_mfma<P, 10, NTL>), g, b, 0, s, a); break;
    }
}

// two 16-set tiles per workgroup from 17 sets on (the gathered A element feeds both)
template <int P>
void p2m_p(const UpArgs& a, int num_leaves, hipStream_t s) {
    if (a.nsets > 16) p2m_pt<P, 2>(a, num_leaves, s);
    else p2m_pt<P, 1>(a, num_leaves, s);
}

template <int P, int F>
void l2p_pf(const DownArgs& a, int num_leaves, hipStream_t s) {
    hipLaunchKernelGGL((k_l2p_basis<P>), dim3(num_leaves), dim3(64), 0, s, a);
    hipLaunchKernelGGL((k_l2p_mfma<P, F>), dim3(num_leaves, (a.nsets + 15) / 16), dim3(64 * XL_W), 0, s, a);
}

template <int F>
void l2p_f(const DownArgs& a, int num_leaves, hipStream_t s) {
    switch (a.p) {
        case 0: l2p_pf<0, F>(a, num_leaves, s); break;
        case 1: l2p_pf<1, F>(a, num_leaves, s); break;
        case 2: l2p_pf<2, F>(a, num_leaves, s); break;
        case 3: l2p_pf<3, F>(a, num_leaves, s); break;
        case 4: l2p_pf<4, F>(a, num_leaves, s); break;
        case 5: l2p_pf<5, F>(a, num_leaves, s); break;
        case 6: l2p_pf<6, F>(a, num_leaves, s); break;
        case 7: l2p_pf<7, F>(a, num_leaves, s); break;
        case 8: l2p_pf<8, F>(a, num_leaves, s); break;
        case 9: l2p_pf<9, F>(a, num_leaves, s); break;
        default: l2p_pf<10, F>(a, num_leaves, s); break;
    }
}

}  // namespace

bool launch_p2m_mfma(const UpArgs& a, int num_leaves, hipStream_t s) {
    if (a.nsets < EXP_MFMA_MIN_SETS || a.p > XM_PMAX || (a.nsrc != 1 && a.nsrc != 4 && a.nsrc != 10)) return false;
    switch (a.p) {
        case 0: p2m_p<0>(a, num_leaves, s); break;
        case 1: p2m_p<1>(a, num_leaves, s); break;
        case 2: p2m_p<2>(a, num_leaves, s); break;
        case 3: p2m_p<3>(a, num_leaves, s); break;
        case 4: p2m_p<4>(a, num_leaves, s); break;
        case 5: p2m_p<5>(a, num_leaves, s); break;
        case 6: p2m_p<6>(a, num_leaves, s); break;
        case 7: p2m_p<7>(a, num_leaves, s); break;
        case 8: p2m_p<8>(a, num_leaves, s); break;
        case 9: p2m_p<9>(a, num_leaves, s); break;
        default: p2m_p<10>(a, num_leaves, s); break;
    }
    FMMH_HIP(hipGetLastError());
    return true;
}

bool launch_l2p_mfma(const DownArgs& a, int num_leaves, hipStream_t s) {
    if (!a.basis || a.nsets < EXP_MFMA_MIN_SETS || a.p > XM_PMAX || (a.nf != 1 && a.nf != 4)) return false;
    if (a.nf == 1) l2p_f<0>(a, num_leaves, s);
    else l2p_f<1>(a, num_leaves, s);
    FMMH_HIP(hipGetLastError());
    return true;
}

}  // namespace fmmh
