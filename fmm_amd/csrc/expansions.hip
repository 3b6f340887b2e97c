// Far-field operators of the field engine, FP64:
//   P2M  src/fmm.F90:437-472, 520-534   leaf moments about the leaf centre
//   M2M  src/fmm.F90:387-435, 536-540   (reference: leaf -> every ancestor;
//        here child -> parent level by level, identical by the binomial
//        theorem, different summation order)
//   (M2L, src/fmm.F90:826-882, is in m2l.hip)
//   L2L  src/fmm.F90:981-1044           parent -> child, level by level
//   L2P  src/fmm.F90:1046-1132          evaluation at leaf sites, added to
//        the near field and scattered to the caller's site order
// Shift operators run one workgroup per node with one lane per target
// coefficient, driven by host-built (target, source, monomial, coefficient)
// tables.
#include "common.h"
#include "kernels.h"
#include "tensors.h"

namespace fmmh {

constexpr int MAX_NM = ncomp(PMAX);          // expansion coefficients per node
constexpr int UP_CHUNK = 16;                 // sites staged per P2M pass, at most
// Sites per P2M pass for nsrc source components (fewer for high multipole orders, whose staged
// components would crowd LDS); the one-set and many-sets kernels use the same passes, so they sum
// in the same order.
__host__ __device__ constexpr int up_chunk(int nsrc) { return nsrc <= 20 ? UP_CHUNK : (nsrc <= 35 ? 8 : 4); }

// monomials d^e for every packed e with |e| <= p, built from one-step parents.
__device__ __forceinline__ void build_mono(double* mono, int nm, double dx, double dy, double dz,
                                           const int* __restrict__ mono_parent, const int* __restrict__ mono_axis) {
    mono[0] = 1.0;
    for (int e = 1; e < nm; ++e) {
        const int ax = mono_axis[e];
        const double d = ax == 0 ? dx : (ax == 1 ? dy : dz);
        mono[e] = mono[mono_parent[e]] * d;
    }
}

// The same monomials for nv displacement vectors (d[v] = (dx, dy, dz)), built
// by the whole workgroup: per-axis powers first, then one product per entry
// (xyz[e] = ex | ey << 8 | ez << 16).  mono is [nv][ld], pw holds 3 (PMAX + 1)
// doubles per vector.  Ends with a barrier.
__device__ __forceinline__ void build_mono_block(double* mono, int ld, int nm, int nv, const double* d3,
                                                 double* pw, const int* __restrict__ xyz, int p) {
    for (int x = threadIdx.x; x < 3 * nv; x += blockDim.x) {
        const double d = d3[x];
        double* w = pw + x * (PMAX + 1);
        double v = 1.0;
        for (int k = 0; k <= p; ++k) { w[k] = v; v *= d; }
    }
    __syncthreads();
    for (int x = threadIdx.x; x < nv * nm; x += blockDim.x) {
        const int v = x / nm, e = x - v * nm;
        const int pk = xyz[e];
        const double* w = pw + 3 * v * (PMAX + 1);
        mono[v * ld + e] = (w[pk & 255] * w[(PMAX + 1) + ((pk >> 8) & 255)]) * w[2 * (PMAX + 1) + (pk >> 16)];
    }
    __syncthreads();
}

// LDS: the monomials of a chunk of sites, [UP_CHUNK][nm] doubles (dynamic, sized by the order).
// sum over the table terms b0..b1 of coef[b] mono[e[b]] src[s[b]], in table order; the terms' table
// entries are loaded eight at a time (one wait per batch instead of one per term).
__device__ __forceinline__ double shift_sum(const ShiftTab& tab, int b0, int b1, const double* mono,
                                            const double* src) {
    double sum = 0.0;
    int b = b0;
    for (; b + 8 <= b1; b += 8) {
        double cf[8];
        int ee[8], ss[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            cf[u] = tab.coef[b + u];
            ee[u] = tab.e[b + u];
            ss[u] = tab.s[b + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) sum = __builtin_fma(cf[u] * mono[ee[u]], src[ss[u]], sum);
    }
    for (; b < b1; ++b) sum = __builtin_fma(tab.coef[b] * mono[tab.e[b]], src[tab.s[b]], sum);
    return sum;
}

// LDS (dynamic): monomials [chunk][nm], then staged source components [chunk][nsrc].
__global__ __launch_bounds__(128) void k_p2m(UpArgs a) {
    extern __shared__ double p2m_lds[];
    __shared__ double d3[UP_CHUNK * 3];
    __shared__ double pw[UP_CHUNK * 3 * (PMAX + 1)];
    const int node = a.leaf_node[blockIdx.x];
    const int set = blockIdx.y;
    const int s0 = a.start[node], c = a.count[node];
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    const int nm = a.nm;
    const int nsrc = a.nsrc;
    const int chunk = up_chunk(nsrc);
    double* mono = p2m_lds;
    double* mch = p2m_lds + chunk * nm;
    const double* __restrict__ msr = a.ms_raw + (int64_t)set * a.dim * a.ldm;
    double* __restrict__ CM = a.CM + (int64_t)set * a.num_nodes * nm;
    double acc[(MAX_NM + 127) / 128];
#pragma unroll
    for (int q = 0; q < (MAX_NM + 127) / 128; ++q) acc[q] = 0.0;
    for (int j0 = 0; j0 < c; j0 += chunk) {
        const int cc = min(chunk, c - j0);
        if ((int)threadIdx.x < 3 * cc) {
            const int jj = threadIdx.x / 3, ax = threadIdx.x - 3 * jj;
            const int j = s0 + j0 + jj;
            d3[threadIdx.x] = ax == 0 ? a.xs[j] - cx : (ax == 1 ? a.ys[j] - cy : a.zs[j] - cz);
        }
        __syncthreads();
        build_mono_block(mono, nm, nm, cc, d3, pw, a.xyz, a.p);
        for (int x = threadIdx.x; x < cc * nsrc; x += blockDim.x) {
            const int jj = x / nsrc, sc = x % nsrc;
            mch[jj * nsrc + sc] = msr[(int64_t)sc * a.ldm + s0 + j0 + jj];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < (MAX_NM + 127) / 128; ++q) {
            const int t = threadIdx.x + 128 * q;
            if (t < nm) {
                const int b0 = a.p2m.off[t], b1 = a.p2m.off[t + 1];
                for (int b = b0; b < b1; ++b) {
                    const int s = a.p2m.s[b], e = a.p2m.e[b];
                    double sum = 0.0;
                    for (int jj = 0; jj < cc; ++jj) sum = __builtin_fma(mono[jj * nm + e], mch[jj * nsrc + s], sum);
                    acc[q] = __builtin_fma(a.p2m.coef[b], sum, acc[q]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < (MAX_NM + 127) / 128; ++q) {
        const int t = threadIdx.x + 128 * q;
        if (t < nm) CM[(int64_t)node * nm + t] = acc[q];
    }
}

// One workgroup per parent node: the children's moments are staged in LDS, every (child, coefficient)
// sum runs on its own lane (a child's terms in table order), and each coefficient adds its children's
// sums in child order -- the same order as summing child by child, so the many-sets form below (which
// does that) stays bit-identical.
// LDS: monomials, staged child moments and partial sums, 3 x [8][nm] doubles (dynamic).
__global__ __launch_bounds__(256) void k_m2m(UpArgs a, const int* __restrict__ nodes) {
    extern __shared__ double m2m_lds[];
    __shared__ double d3[24];
    __shared__ double pw[24 * (PMAX + 1)];
    __shared__ int kids[8];
    const int node = nodes[blockIdx.x];
    const int nm = a.nm;
    double* monos = m2m_lds;
    double* srcs = m2m_lds + 8 * nm;
    double* part = m2m_lds + 16 * nm;
    double* __restrict__ CM = a.CM + (int64_t)blockIdx.y * a.num_nodes * nm;
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    if (threadIdx.x < 8) kids[threadIdx.x] = a.children[node * 8 + threadIdx.x];
    if (threadIdx.x < 24) {
        const int o = threadIdx.x / 3, ax = threadIdx.x - 3 * o;
        const int ch = a.children[node * 8 + o];
        d3[threadIdx.x] = ch < 0 ? 0.0 : (ax == 0 ? a.cx[ch] - cx : (ax == 1 ? a.cy[ch] - cy : a.cz[ch] - cz));
    }
    __syncthreads();
    bool any = false;
#pragma unroll
    for (int o = 0; o < 8; ++o) any |= kids[o] >= 0;
    if (!any) return;   // leaves keep their P2M moments
    for (int x = threadIdx.x; x < 8 * nm; x += blockDim.x) {
        const int o = x / nm, t = x - o * nm;
        const int ch = kids[o];
        srcs[o * nm + t] = ch >= 0 ? CM[(int64_t)ch * nm + t] : 0.0;
    }
    build_mono_block(monos, nm, nm, 8, d3, pw, a.xyz, a.p);
    for (int x = threadIdx.x; x < 8 * nm; x += blockDim.x) {
        const int o = x / nm, t = x - o * nm;
        if (kids[o] < 0) continue;
        const double* __restrict__ mono = monos + o * nm;
        const double* __restrict__ src = srcs + o * nm;
        part[o * nm + t] = shift_sum(a.m2m, a.m2m.off[t], a.m2m.off[t + 1], mono, src);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nm; t += blockDim.x) {
        double acc = 0.0;
#pragma unroll
        for (int o = 0; o < 8; ++o)
            if (kids[o] >= 0) acc += part[o * nm + t];
        CM[(int64_t)node * nm + t] = acc;
    }
}

// ---------------------------------------------------------------- L2L / L2P
__global__ __launch_bounds__(128) void k_l2l(DownArgs a, const int* __restrict__ nodes) {
    __shared__ double mono[MAX_NM];
    __shared__ double lp[MAX_NM];
    __shared__ double d3[3];
    __shared__ double pw[3 * (PMAX + 1)];
    const int node = nodes[blockIdx.x];
    const int par = a.parent[node];
    const int nm = a.nm;
    double* __restrict__ LE = a.LE + (int64_t)blockIdx.y * a.num_nodes * nm;
    if (threadIdx.x < 3) {
        const int ax = threadIdx.x;
        d3[ax] = ax == 0 ? a.cx[node] - a.cx[par] : (ax == 1 ? a.cy[node] - a.cy[par] : a.cz[node] - a.cz[par]);
    }
    for (int s = threadIdx.x; s < nm; s += blockDim.x) lp[s] = LE[(int64_t)par * nm + s];
    __syncthreads();
    build_mono_block(mono, MAX_NM, nm, 1, d3, pw, a.xyz, a.p);
    for (int t = threadIdx.x; t < nm; t += blockDim.x)
        LE[(int64_t)node * nm + t] += shift_sum(a.l2l, a.l2l.off[t], a.l2l.off[t + 1], mono, lp);
}

// f_t = sum_e L_{t+e} d^e / (ex! ey! ez!), |t| <= F, |t + e| <= p, one lane per site, added to the near
// field and scattered to the caller's site order.  The order is a template parameter: every loop is
// unrolled, so the packed indices are constants (LDS reads with immediate offsets) and the |t + e| <= p
// tests vanish.  The operations and their order are those of k_l2p_sb (runtime order, many sets), so a
// set of a many-set call equals the one-set call bit for bit.
template <int P, int F>
__global__ __launch_bounds__(64) void k_l2p_p(DownArgs a) {
    constexpr int NF = ncomp(F);
    constexpr int NM = ncomp(P);
    __shared__ double L[NM];
    __shared__ double fact[PMAX + 1];
    const int node = a.leaf_node[blockIdx.x];
    const int s0 = a.start[node], c = a.count[node];
    for (int s = threadIdx.x; s < NM; s += 64) L[s] = a.LE[(int64_t)node * NM + s];
    if (threadIdx.x == 0) {
        double f = 1.0;
        fact[0] = 1.0;
        for (int n = 1; n <= PMAX; ++n) { f *= (double)n; fact[n] = 1.0 / f; }
    }
    __syncthreads();
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    for (int base = 0; base < c; base += 64) {
        const int jj = base + threadIdx.x;
        if (jj >= c) break;
        const int k = s0 + jj;
        const double dx = a.xs[k] - cx, dy = a.ys[k] - cy, dz = a.zs[k] - cz;
        double f[NF];
#pragma unroll
        for (int t = 0; t < NF; ++t) f[t] = 0.0;
        double px = 1.0;
#pragma unroll
        for (int ex = 0; ex <= P; ++ex) {
            double pxy = px;
#pragma unroll
            for (int ey = 0; ey + ex <= P; ++ey) {
                double m = pxy * fact[ex] * fact[ey];
#pragma unroll
                for (int ez = 0; ez + ey + ex <= P; ++ez) {
                    const double mz = m * fact[ez];
#pragma unroll
                    for (int t = 0; t < NF; ++t) {
                        const Exps et = idx2xyz(t);
                        if (et.x + et.y + et.z + ex + ey + ez <= P)
                            f[t] = __builtin_fma(mz, L[xyz2idx(et.x + ex, et.y + ey, et.z + ez)], f[t]);
                    }
                    m *= dz;
                }
                pxy *= dy;
            }
            px *= dx;
        }
        const int o = a.perm[k];
#pragma unroll
        for (int t = 0; t < NF; ++t) {
            double* fp = a.field + (int64_t)t * a.ld_field + o;
            *fp = (*fp + a.near[(int64_t)t * a.ld_near + k]) + f[t];
        }
    }
}

template <int F>
static void launch_l2p_p(const DownArgs& a, int num_leaves, hipStream_t s) {
    const dim3 g(num_leaves, 1);
    switch (a.p) {
        case 0: hipLaunchKernelGGL((k_l2p_p<0, F>), g, dim3(64), 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_l2p_p<1, F>), g, dim3(64), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_l2p_p<2, F>), g, dim3(64), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_l2p_p<3, F>), g, dim3(64), 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_l2p_p<4, F>), g, dim3(64), 0, s, a); break;
        case 5: hipLaunchKernelGGL((k_l2p_p<5, F>), g, dim3(64), 0, s, a); break;
        case 6: hipLaunchKernelGGL((k_l2p_p<6, F>), g, dim3(64), 0, s, a); break;
        case 7: hipLaunchKernelGGL((k_l2p_p<7, F>), g, dim3(64), 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_l2p_p<8, F>), g, dim3(64), 0, s, a); break;
        case 9: hipLaunchKernelGGL((k_l2p_p<9, F>), g, dim3(64), 0, s, a); break;
        default: hipLaunchKernelGGL((k_l2p_p<10, F>), g, dim3(64), 0, s, a); break;
    }
}

// ---------------------------------------------------------------- many sets
// The same four operators for SB sets per workgroup: monomials, shift tables and
// staged inputs are shared by the block of sets, and every (coefficient, set)
// output is summed in exactly the order of the one-set kernels above, so a set
// of a many-set call equals the one-set call bit for bit.
constexpr int SHIFT_SB = 8;        // sets per workgroup (P2M, M2M, L2L; L2P with field order <= 1)
constexpr int SHIFT_Q = (MAX_NM + 255) / 256;

// LDS (dynamic): staged source components [chunk][nsrc][SB].
template <int SB>
__global__ __launch_bounds__(128) void k_p2m_sb(UpArgs a) {
    __shared__ double mono[UP_CHUNK][MAX_NM];
    extern __shared__ double mch[];
    __shared__ double d3[UP_CHUNK * 3];
    __shared__ double pw[UP_CHUNK * 3 * (PMAX + 1)];
    constexpr int Q = (MAX_NM + 127) / 128;
    const int node = a.leaf_node[blockIdx.x];
    const int set0 = blockIdx.y * SB;
    const int nsb = min(SB, a.nsets - set0);
    const int s0 = a.start[node], c = a.count[node];
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    const int nm = a.nm;
    const int nsrc = a.nsrc;
    const int chunk = up_chunk(nsrc);
    double acc[Q][SB];
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int u = 0; u < SB; ++u) acc[q][u] = 0.0;
    for (int j0 = 0; j0 < c; j0 += chunk) {
        const int cc = min(chunk, c - j0);
        if ((int)threadIdx.x < 3 * cc) {
            const int jj = threadIdx.x / 3, ax = threadIdx.x - 3 * jj;
            const int j = s0 + j0 + jj;
            d3[threadIdx.x] = ax == 0 ? a.xs[j] - cx : (ax == 1 ? a.ys[j] - cy : a.zs[j] - cz);
        }
        __syncthreads();
        build_mono_block(&mono[0][0], MAX_NM, nm, cc, d3, pw, a.xyz, a.p);
        for (int x = threadIdx.x; x < SB * nsrc * cc; x += blockDim.x) {
            const int u = x / (nsrc * cc), r = x - u * nsrc * cc, sc = r / cc, jj = r - sc * cc;
            mch[(jj * nsrc + sc) * SB + u] = u < nsb ? a.ms_raw[((int64_t)(set0 + u) * a.dim + sc) * a.ldm + s0 + j0 + jj] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int t = threadIdx.x + 128 * q;
            if (t < nm) {
                const int b0 = a.p2m.off[t], b1 = a.p2m.off[t + 1];
                for (int b = b0; b < b1; ++b) {
                    const int sc = a.p2m.s[b], e = a.p2m.e[b];
                    double sum[SB];
#pragma unroll
                    for (int u = 0; u < SB; ++u) sum[u] = 0.0;
                    for (int jj = 0; jj < cc; ++jj) {
                        const double mo = mono[jj][e];
#pragma unroll
                        for (int u = 0; u < SB; ++u) sum[u] = __builtin_fma(mo, mch[(jj * nsrc + sc) * SB + u], sum[u]);
                    }
                    const double cf = a.p2m.coef[b];
#pragma unroll
                    for (int u = 0; u < SB; ++u) acc[q][u] = __builtin_fma(cf, sum[u], acc[q][u]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int t = threadIdx.x + 128 * q;
        if (t < nm)
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (u < nsb) a.CM[((int64_t)(set0 + u) * a.num_nodes + node) * nm + t] = acc[q][u];
    }
}

template <int SB>
__global__ __launch_bounds__(256) void k_m2m_sb(UpArgs a, const int* __restrict__ nodes) {
    __shared__ double monos[8][MAX_NM];
    __shared__ double chb[MAX_NM][SB];
    __shared__ double d3[24];
    __shared__ double pw[24 * (PMAX + 1)];
    const int node = nodes[blockIdx.x];
    const int set0 = blockIdx.y * SB;
    const int nsb = min(SB, a.nsets - set0);
    const int nm = a.nm;
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    double acc[SHIFT_Q][SB];
#pragma unroll
    for (int q = 0; q < SHIFT_Q; ++q)
#pragma unroll
        for (int u = 0; u < SB; ++u) acc[q][u] = 0.0;
    if (threadIdx.x < 24) {
        const int o = threadIdx.x / 3, ax = threadIdx.x - 3 * o;
        const int ch = a.children[node * 8 + o];
        d3[threadIdx.x] = ch < 0 ? 0.0 : (ax == 0 ? a.cx[ch] - cx : (ax == 1 ? a.cy[ch] - cy : a.cz[ch] - cz));
    }
    __syncthreads();
    build_mono_block(&monos[0][0], MAX_NM, nm, 8, d3, pw, a.xyz, a.p);
    bool any = false;
    for (int o = 0; o < 8; ++o) {
        const int ch = a.children[node * 8 + o];
        if (ch < 0) continue;
        any = true;
        __syncthreads();                      // readers of the previous child are done
        for (int x = threadIdx.x; x < SB * nm; x += blockDim.x) {
            const int u = x / nm, sc = x - u * nm;
            chb[sc][u] = u < nsb ? a.CM[((int64_t)(set0 + u) * a.num_nodes + ch) * nm + sc] : 0.0;
        }
        __syncthreads();
        const double* __restrict__ mono = monos[o];
#pragma unroll
        for (int q = 0; q < SHIFT_Q; ++q) {
            const int t = threadIdx.x + 256 * q;
            if (t < nm) {
                const int b0 = a.m2m.off[t], b1 = a.m2m.off[t + 1];
                double sum[SB];
#pragma unroll
                for (int u = 0; u < SB; ++u) sum[u] = 0.0;
                for (int b = b0; b < b1; ++b) {
                    const double cf = a.m2m.coef[b] * mono[a.m2m.e[b]];
                    const int sc = a.m2m.s[b];
#pragma unroll
                    for (int u = 0; u < SB; ++u) sum[u] = __builtin_fma(cf, chb[sc][u], sum[u]);
                }
#pragma unroll
                for (int u = 0; u < SB; ++u) acc[q][u] += sum[u];
            }
        }
    }
    if (!any) return;   // leaves keep their P2M moments
#pragma unroll
    for (int q = 0; q < SHIFT_Q; ++q) {
        const int t = threadIdx.x + 256 * q;
        if (t < nm)
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (u < nsb) a.CM[((int64_t)(set0 + u) * a.num_nodes + node) * nm + t] = acc[q][u];
    }
}

template <int SB>
__global__ __launch_bounds__(256) void k_l2l_sb(DownArgs a, const int* __restrict__ nodes) {
    __shared__ double mono[MAX_NM];
    __shared__ double lp[MAX_NM][SB];
    __shared__ double d3[3];
    __shared__ double pw[3 * (PMAX + 1)];
    const int node = nodes[blockIdx.x];
    const int par = a.parent[node];
    const int set0 = blockIdx.y * SB;
    const int nsb = min(SB, a.nsets - set0);
    const int nm = a.nm;
    if (threadIdx.x < 3) {
        const int ax = threadIdx.x;
        d3[ax] = ax == 0 ? a.cx[node] - a.cx[par] : (ax == 1 ? a.cy[node] - a.cy[par] : a.cz[node] - a.cz[par]);
    }
    for (int x = threadIdx.x; x < SB * nm; x += blockDim.x) {
        const int u = x / nm, sc = x - u * nm;
        lp[sc][u] = u < nsb ? a.LE[((int64_t)(set0 + u) * a.num_nodes + par) * nm + sc] : 0.0;
    }
    __syncthreads();
    build_mono_block(mono, MAX_NM, nm, 1, d3, pw, a.xyz, a.p);
    for (int t = threadIdx.x; t < nm; t += blockDim.x) {
        const int b0 = a.l2l.off[t], b1 = a.l2l.off[t + 1];
        double sum[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) sum[u] = 0.0;
        for (int b = b0; b < b1; ++b) {
            const double cf = a.l2l.coef[b] * mono[a.l2l.e[b]];
            const int sc = a.l2l.s[b];
#pragma unroll
            for (int u = 0; u < SB; ++u) sum[u] = __builtin_fma(cf, lp[sc][u], sum[u]);
        }
#pragma unroll
        for (int u = 0; u < SB; ++u)
            if (u < nsb) a.LE[((int64_t)(set0 + u) * a.num_nodes + node) * nm + t] += sum[u];
    }
}

template <int F, int SB>
__global__ __launch_bounds__(64) void k_l2p_sb(DownArgs a) {
    constexpr int NF = ncomp(F);
    __shared__ double L[SB][MAX_NM];
    __shared__ double fact[PMAX + 1];
    const int node = a.leaf_node[blockIdx.x];
    const int set0 = blockIdx.y * SB;
    const int nsb = min(SB, a.nsets - set0);
    const int s0 = a.start[node], c = a.count[node];
    const int nm = a.nm, p = a.p;
    for (int x = threadIdx.x; x < SB * nm; x += 64) {
        const int u = x / nm, sc = x - u * nm;
        L[u][sc] = u < nsb ? a.LE[((int64_t)(set0 + u) * a.num_nodes + node) * nm + sc] : 0.0;
    }
    if (threadIdx.x == 0) {
        double f = 1.0;
        fact[0] = 1.0;
        for (int n = 1; n <= PMAX; ++n) { f *= (double)n; fact[n] = 1.0 / f; }
    }
    __syncthreads();
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    for (int base = 0; base < c; base += 64) {
        const int jj = base + threadIdx.x;
        if (jj >= c) break;
        const int k = s0 + jj;
        const double dx = a.xs[k] - cx, dy = a.ys[k] - cy, dz = a.zs[k] - cz;
        double f[SB][NF];
#pragma unroll
        for (int u = 0; u < SB; ++u)
#pragma unroll
            for (int t = 0; t < NF; ++t) f[u][t] = 0.0;
        double px = 1.0;
        for (int ex = 0; ex <= p; ++ex) {
            double pxy = px;
            for (int ey = 0; ey + ex <= p; ++ey) {
                double m = pxy * fact[ex] * fact[ey];
                for (int ez = 0; ez + ey + ex <= p; ++ez) {
                    const double mz = m * fact[ez];
                    const int ne = ex + ey + ez;
#pragma unroll
                    for (int t = 0; t < NF; ++t) {
                        const Exps et = idx2xyz(t);
                        if (et.x + et.y + et.z + ne <= p) {
                            const int li = xyz2idx(et.x + ex, et.y + ey, et.z + ez);
#pragma unroll
                            for (int u = 0; u < SB; ++u) f[u][t] = __builtin_fma(mz, L[u][li], f[u][t]);
                        }
                    }
                    m *= dz;
                }
                pxy *= dy;
            }
            px *= dx;
        }
        const int o = a.perm[k];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            if (u >= nsb) break;
            const double* __restrict__ near = a.near + (int64_t)(set0 + u) * NF * a.ld_near;
            double* __restrict__ field = a.field + (int64_t)(set0 + u) * NF * a.ld_field;
#pragma unroll
            for (int t = 0; t < NF; ++t) {
                double* fp = field + (int64_t)t * a.ld_field + o;
                *fp = (*fp + near[(int64_t)t * a.ld_near + k]) + f[u][t];
            }
        }
    }
}

// L2P for field orders above 2 (any set count, one set per grid row): one thread per (site, field
// component), the terms of k_l2p_p in the same order with the orders as arguments.
__global__ __launch_bounds__(256) void k_l2p_rt(DownArgs a) {
    __shared__ double L[MAX_NM];
    __shared__ double fact[PMAX + 1];
    const int node = a.leaf_node[blockIdx.x];
    const int set = blockIdx.y;
    const int s0 = a.start[node], c = a.count[node];
    const int nm = a.nm, p = a.p, nf = a.nf;
    for (int x = threadIdx.x; x < nm; x += blockDim.x) L[x] = a.LE[((int64_t)set * a.num_nodes + node) * nm + x];
    if (threadIdx.x == 0) {
        double f = 1.0;
        fact[0] = 1.0;
        for (int n = 1; n <= PMAX; ++n) { f *= (double)n; fact[n] = 1.0 / f; }
    }
    __syncthreads();
    const double cx = a.cx[node], cy = a.cy[node], cz = a.cz[node];
    const double* __restrict__ near = a.near + (int64_t)set * nf * a.ld_near;
    double* __restrict__ field = a.field + (int64_t)set * nf * a.ld_field;
    for (int x = threadIdx.x; x < c * nf; x += blockDim.x) {
        const int jj = x / nf, t = x - jj * nf;
        const int k = s0 + jj;
        const int pk = a.xyz[t];
        const int tx = pk & 255, ty = (pk >> 8) & 255, tz = pk >> 16;
        const int rest = p - (tx + ty + tz);   // |e| <= p - |t|
        const double dx = a.xs[k] - cx, dy = a.ys[k] - cy, dz = a.zs[k] - cz;
        double f = 0.0, px = 1.0;
        for (int ex = 0; ex <= rest; ++ex) {
            double pxy = px;
            for (int ey = 0; ey + ex <= rest; ++ey) {
                double m = pxy * fact[ex] * fact[ey];
                for (int ez = 0; ez + ey + ex <= rest; ++ez) {
                    f = __builtin_fma(m * fact[ez], L[xyz2idx(tx + ex, ty + ey, tz + ez)], f);
                    m *= dz;
                }
                pxy *= dy;
            }
            px *= dx;
        }
        const int o = a.perm[k];
        double* fp = field + (int64_t)t * a.ld_field + o;
        *fp = (*fp + near[(int64_t)t * a.ld_near + k]) + f;
    }
}

void launch_p2m(const UpArgs& a, int num_leaves, hipStream_t s) {
    if (num_leaves <= 0) return;
    if (launch_p2m_mfma(a, num_leaves, s)) return;
    const size_t chunk = up_chunk(a.nsrc);
    if (a.nsets == 1)
        hipLaunchKernelGGL(k_p2m, dim3(num_leaves, 1), dim3(128), chunk * (a.nm + a.nsrc) * sizeof(double), s, a);
    else
        hipLaunchKernelGGL(k_p2m_sb<SHIFT_SB>, dim3(num_leaves, (a.nsets + SHIFT_SB - 1) / SHIFT_SB), dim3(128),
                           chunk * a.nsrc * SHIFT_SB * sizeof(double), s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_m2m(const UpArgs& a, const int* nodes, int count, hipStream_t s) {
    if (count <= 0) return;
    if (a.nsets == 1)
        hipLaunchKernelGGL(k_m2m, dim3(count, 1), dim3(256), (size_t)24 * a.nm * sizeof(double), s, a, nodes);
    else hipLaunchKernelGGL(k_m2m_sb<SHIFT_SB>, dim3(count, (a.nsets + SHIFT_SB - 1) / SHIFT_SB), dim3(256), 0, s, a, nodes);
    FMMH_HIP(hipGetLastError());
}

void launch_l2l(const DownArgs& a, const int* nodes, int count, hipStream_t s) {
    if (count <= 0) return;
    if (a.nsets == 1) hipLaunchKernelGGL(k_l2l, dim3(count, 1), dim3(128), 0, s, a, nodes);
    else hipLaunchKernelGGL(k_l2l_sb<SHIFT_SB>, dim3(count, (a.nsets + SHIFT_SB - 1) / SHIFT_SB), dim3(256), 0, s, a, nodes);
    FMMH_HIP(hipGetLastError());
}

void launch_l2p(const DownArgs& a, int num_leaves, hipStream_t s) {
    if (num_leaves <= 0) return;
    if (launch_l2p_mfma(a, num_leaves, s)) return;
    if (a.nf > ncomp(2)) {
        hipLaunchKernelGGL(k_l2p_rt, dim3(num_leaves, a.nsets), dim3(256), 0, s, a);
    } else if (a.nsets == 1) {
        switch (a.nf) {
            case 1: launch_l2p_p<0>(a, num_leaves, s); break;
            case 4: launch_l2p_p<1>(a, num_leaves, s); break;
            default: launch_l2p_p<2>(a, num_leaves, s); break;
        }
    } else {
        constexpr int SB2 = SHIFT_SB / 2;   // field gradients: 10 components per set in registers
        switch (a.nf) {
            case 1:
                hipLaunchKernelGGL((k_l2p_sb<0, SHIFT_SB>), dim3(num_leaves, (a.nsets + SHIFT_SB - 1) / SHIFT_SB),
                                   dim3(64), 0, s, a);
                break;
            case 4:
                hipLaunchKernelGGL((k_l2p_sb<1, SHIFT_SB>), dim3(num_leaves, (a.nsets + SHIFT_SB - 1) / SHIFT_SB),
                                   dim3(64), 0, s, a);
                break;
            default:
                hipLaunchKernelGGL((k_l2p_sb<2, SB2>), dim3(num_leaves, (a.nsets + SB2 - 1) / SB2), dim3(64), 0, s, a);
                break;
        }
    }
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
