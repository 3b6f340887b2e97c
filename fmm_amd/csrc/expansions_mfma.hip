// P2M and L2P for many multipole sets on the FP64 matrix cores
// (src/fmm.F90:437-472, 520-534 and 1046-1132).  Both are GEMMs over the
// sites of a leaf with the sets as columns:
//   P2M  CM[t][set] = sum_(j, s) coef(t, s) d_j^(t - s) M_j(s)[set]
//        (rows: the expansion coefficients t; k: site j and source
//        component s; d_j = x_j - leaf centre; coef and the monomial index
//        from the P2M shift table)
//   L2P  F_j(c)[set] = sum_t d_j^(t - c) / (t - c)! L(t)[set]
//        (rows: site j and field component c, field order <= 1; k: t)
// One workgroup of four wavefronts per (leaf, 16 sets): monomials of a chunk
// of sites and the B operand (multipoles, or the leaf's local expansion) are
// staged in LDS, each wavefront runs its 16-row tiles of
// v_mfma_f64_16x16x4f64, A gathered from the monomials through a table built
// in LDS per workgroup.  FP64 throughout, like the one-set kernels
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
constexpr int XM_PMAX = 10;

__device__ __forceinline__ void powers(double d, int p, double* w) {
    double v = 1.0;
    for (int k = 0; k <= p; ++k) {
        w[k] = v;
        v *= d;
    }
}

// NS source components (a.nsrc); LDS: monomials of a chunk of sites, the B operand and the (t, s) table.
template <int P, int NS>
__global__ __launch_bounds__(64 * XM_W) void k_p2m_mfma(UpArgs a) {
    constexpr int NM = ncomp(P);
    constexpr int MT = (NM + 15) / 16;
    constexpr int MW = (MT + XM_W - 1) / XM_W;
    constexpr int nsrc = NS;
    constexpr int kmax = (XM_CH * NS + 3) / 4 * 4;
    __shared__ double mono[XM_CH * NM];
    __shared__ double B[kmax * 16];
    __shared__ double tabC[NM * NS];
    __shared__ int tabE[NM * NS];
    __shared__ double pw[XM_CH][3][XM_PMAX + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int node = a.leaf_node[blockIdx.x];
    const int set0 = blockIdx.y * 16;
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
    v4d acc[MW];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi) acc[mi] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < c; j0 += XM_CH) {
        const int cc = min(XM_CH, c - j0);
        __syncthreads();   // readers of the previous chunk are done (and the tables are in place)
        if (tid < 3 * cc) {
            const int jj = tid / 3, ax = tid - 3 * jj;
            const int j = s0 + j0 + jj;
            const double d = ax == 0 ? a.xs[j] - cx : (ax == 1 ? a.ys[j] - cy : a.zs[j] - cz);
            powers(d, P, pw[jj][ax]);
        }
        for (int x = tid; x < kmax * 16; x += 64 * XM_W) {
            const int k = x >> 4, u = x & 15;
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
            const double bv = B[k * 16 + row];
#pragma unroll
            for (int mi = 0; mi < MW; ++mi) {
                const int t = (w + XM_W * mi) * 16 + row;
                double av = 0.0;
                if (t < NM && kv) {
                    const int e = tabE[t * nsrc + sc];
                    if (e >= 0) av = tabC[t * nsrc + sc] * mono[jj * NM + e];
                }
                acc[mi] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[mi], 0, 0, 0);
            }
        }
    }
    // accumulator register q of lane l holds D[coefficient 4 q + l / 16][set l % 16]
    const int set = set0 + row;
    if (set >= a.nsets) return;
#pragma unroll
    for (int mi = 0; mi < MW; ++mi)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = (w + XM_W * mi) * 16 + 4 * q + kq;
            if (t < NM) a.CM[((int64_t)set * a.num_nodes + node) * NM + t] = acc[mi][q];
        }
}

// LDS: the leaf's local expansion as the B operand, the Taylor monomials of a chunk of 16 sites and the
// (t, c) table; the chunk's NF 16-row tiles go to the wavefronts in turn.
template <int P, int F>
__global__ __launch_bounds__(64 * XM_W) void k_l2p_mfma(DownArgs a) {
    constexpr int NM = ncomp(P);
    constexpr int NF = ncomp(F);
    constexpr int KS = (NM + 3) / 4;
    constexpr int SC = 16;                 // sites per chunk
    constexpr int ST = 16 / NF;            // sites per row tile
    __shared__ double L[4 * KS * 16];
    __shared__ double monoF[SC * NM];
    __shared__ int tabD[NM * NF];
    __shared__ double pw[SC][3][XM_PMAX + 1];
    __shared__ double ifact[XM_PMAX + 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int node = a.leaf_node[blockIdx.x];
    const int set0 = blockIdx.y * 16;
    const int s0 = a.start[node], c = a.count[node];
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    if (tid == 0) {
        double f = 1.0;
        ifact[0] = 1.0;
        for (int n = 1; n <= XM_PMAX; ++n) {
            f *= (double)n;
            ifact[n] = 1.0 / f;
        }
    }
    for (int x = tid; x < 4 * KS * 16; x += 64 * XM_W) {
        const int t = x >> 4, u = x & 15;
        L[x] = (t < NM && set0 + u < a.nsets) ? a.LE[((int64_t)(set0 + u) * a.num_nodes + node) * NM + t] : 0.0;
    }
    for (int x = tid; x < NM * NF; x += 64 * XM_W) {
        const int t = x / NF, cf = x - t * NF;
        const int pk = a.xyz[t];
        const int ex = (pk & 255) - (cf == 1), ey = ((pk >> 8) & 255) - (cf == 2), ez = (pk >> 16) - (cf == 3);
        tabD[x] = (ex < 0 || ey < 0 || ez < 0) ? -1 : xyz2idx(ex, ey, ez);
    }
    for (int j0 = 0; j0 < c; j0 += SC) {
        const int cc = min(SC, c - j0);
        __syncthreads();   // readers of the previous chunk are done (and L, tabD, ifact are in place)
        if (tid < 3 * cc) {
            const int jj = tid / 3, ax = tid - 3 * jj;
            const int j = s0 + j0 + jj;
            const double d = ax == 0 ? a.xs[j] - cx : (ax == 1 ? a.ys[j] - cy : a.zs[j] - cz);
            powers(d, P, pw[jj][ax]);
        }
        __syncthreads();
        for (int x = tid; x < cc * NM; x += 64 * XM_W) {
            const int jj = x / NM, e = x - jj * NM;
            const int pk = a.xyz[e];
            const int ex = pk & 255, ey = (pk >> 8) & 255, ez = pk >> 16;
            monoF[jj * NM + e] = ((pw[jj][0][ex] * ifact[ex]) * (pw[jj][1][ey] * ifact[ey])) * (pw[jj][2][ez] * ifact[ez]);
        }
        __syncthreads();
        for (int tile = w; tile * ST < cc; tile += XM_W) {
            // row r of the tile: site tile ST + r / NF of the chunk, field component r % NF
            const int jj = tile * ST + row / NF, cf = row - (row / NF) * NF;
            const bool rv = jj < cc;
            v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
            for (int ks = 0; ks < KS; ++ks) {
                const int t = 4 * ks + kq;
                double av = 0.0;
                if (rv && t < NM) {
                    const int e = tabD[t * NF + cf];
                    if (e >= 0) av = monoF[jj * NM + e];
                }
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, L[t * 16 + row], acc, 0, 0, 0);
            }
            // accumulator register q of lane l holds D[row 4 q + l / 16][set l % 16]
            const int set = set0 + row;
            if (set < a.nsets) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = 4 * q + kq;
                    const int sj = tile * ST + r / NF, c2 = r - (r / NF) * NF;
                    if (sj >= cc) continue;
                    const int k = s0 + j0 + sj;
                    const int o = a.perm[k];
                    double* fp = a.field + ((int64_t)set * NF + c2) * a.ld_field + o;
                    *fp = (*fp + a.near[((int64_t)set * NF + c2) * a.ld_near + k]) + acc[q];
                }
            }
        }
    }
}

template <int P>
void p2m_p(const UpArgs& a, int num_leaves, hipStream_t s) {
    const dim3 g(num_leaves, (a.nsets + 15) / 16), b(64 * XM_W);
    switch (a.nsrc) {
        case 1: hipLaunchKernelGGL((k_p2m_mfma<P, 1>), g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_p2m_mfma<P, 4>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_p2m_mfma<P, 10>), g, b, 0, s, a); break;
    }
}

template <int P, int F>
void l2p_pf(const DownArgs& a, int num_leaves, hipStream_t s) {
    hipLaunchKernelGGL((k_l2p_mfma<P, F>), dim3(num_leaves, (a.nsets + 15) / 16), dim3(64 * XM_W), 0, s, a);
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
    // to order 9 (at 10 the staged expansion and monomials take 82 KB of LDS, one workgroup per CU, and
    // the vector-ALU kernel is faster: 2.7 against 2.9 ms for 32 sets on the 200k-site water box)
    if (a.nsets < EXP_MFMA_MIN_SETS || a.p > 9 || (a.nf != 1 && a.nf != 4)) return false;
    if (a.nf == 1) l2p_f<0>(a, num_leaves, s);
    else l2p_f<1>(a, num_leaves, s);
    FMMH_HIP(hipGetLastError());
    return true;
}

}  // namespace fmmh
