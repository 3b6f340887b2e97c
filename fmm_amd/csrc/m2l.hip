// M2L, src/fmm.F90:826-882: L_i(t) += sum_{j in list(i)} sum_s c_s M_j(s) T_{s+t}(c_i - c_j),
// undamped, |s|, |t| <= p, tensor order 2p, c_s = (-1)^{|s|+1} / (sx! sy! sz!).
//
// MI355X formulation.  Node centres are the root centre plus multiples of
// the half-widths r_L = r_0 / 2^{L-1}, so for a far pair the displacement is
// u v with u = min(r_i, r_j) and v an integer vector.  With
// T_k(u v) = u^{-(|k|+1)} T_k(v):
//   * pairs are grouped into classes (v, level_i - level_j) -- a few hundred
//     to a few thousand per tree -- and one order-2p tensor is built per
//     class (Hermite recursion in LDS, once);
//   * source expansions are scaled by their own box size,
//     M~_j(s) = c_s M_j(s) r_j^{-|s|}, targets accumulate
//     L~_i(t) = sum 2^{...} sum_s T~_c[s+t] M~_j(s) and are unscaled by
//     r_i^{-(|t|+1)} at the end; level mismatches are exact powers of two.
// The per-pair work is then a pure Hankel contraction: one workgroup per
// target node, one lane per target coefficient, the class tensor staged in
// a double-buffered LDS slot (prefetched one pair ahead), the source
// expansion read with wave-uniform scalar loads, the loops fully unrolled
// for each expansion order so every index is a constant offset.
// Using the exact integer v instead of the rounded c_i - c_j changes each
// tensor by ~1e-16 relative, far inside the 1e-12 parity bar.
#include <type_traits>

#include "common.h"
#include "kernels.h"
#include "hermite.h"

namespace fmmh {

// Non-periodic keys: three a.vbits-bit components and the level difference in the low 8 bits.
// periodic lo word: three 21-bit components (image pairs can span the box)
constexpr int PBITS = 21;
constexpr long long POFF = 1ll << (PBITS - 1);

__device__ void m2l_key(const M2LClassArgs& a, int e, int* range);

// Class keys of the pairs; with a.wg_range, also the workgroup's largest |v_x|, |v_y|, |v_z| and
// |level difference| (for the dense class numbering).
__global__ __launch_bounds__(256) void k_m2l_keys(M2LClassArgs a) {
    __shared__ int red[4][4];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    int range[4] = {0, 0, 0, 0};
    if (e < a.n_pairs) m2l_key(a, e, range);
    if (!a.wg_range) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int v = range[k];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d));
        if (lane == 0) red[w][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4)
        a.wg_range[blockIdx.x * 4 + threadIdx.x] =
            max(max(red[0][threadIdx.x], red[1][threadIdx.x]), max(red[2][threadIdx.x], red[3][threadIdx.x]));
}

__device__ void m2l_key(const M2LClassArgs& a, int e, int* range) {
    const unsigned long long pk = a.pair_keys[e];
    const int i = (int)(pk >> 32);
    const int jenc = (int)(pk & 0xffffffffull);
    const int j = src_node(jenc), img = src_image(jenc);
    const double u = fmin(a.r[i], a.r[j]);
    const long long vx = llrint((a.cx[i] - a.cx[j]) / u);
    const long long vy = llrint((a.cy[i] - a.cy[j]) / u);
    const long long vz = llrint((a.cz[i] - a.cz[j]) / u);
    const int dl = a.level[i] - a.level[j];
    const long long voff = 1ll << (a.vbits - 1);
    const long long vlim = a.periodic ? POFF : voff;
    if (llabs(vx) >= vlim || llabs(vy) >= vlim || llabs(vz) >= vlim || dl < -100 || dl > 100)
        atomicOr(a.error, 1);
    range[0] = (int)min(llabs(vx), (long long)vlim);
    range[1] = (int)min(llabs(vy), (long long)vlim);
    range[2] = (int)min(llabs(vz), (long long)vlim);
    range[3] = min(abs(dl), 127);
    if (a.periodic) {
        a.class_key_in[e] = ((unsigned long long)(vx + POFF) << (2 * PBITS)) |
                            ((unsigned long long)(vy + POFF) << PBITS) | (unsigned long long)(vz + POFF);
        // image classes also depend on the box size u, i.e. on the target level
        const int lev = img == 0 ? 0 : a.level[i];
        a.class_hi_in[e] = ((unsigned long long)img << 16) | ((unsigned long long)lev << 8) |
                           (unsigned long long)(dl + 128);
    } else {
        const int b = a.vbits;
        a.class_key_in[e] = ((unsigned long long)(vx + voff) << (2 * b + 8)) |
                            ((unsigned long long)(vy + voff) << (b + 8)) |
                            ((unsigned long long)(vz + voff) << 8) | (unsigned long long)(dl + 128);
    }
    a.pair_idx[e] = e;
}

__global__ __launch_bounds__(256) void k_m2l_flags(M2LClassArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_pairs) return;
    if (a.periodic)
        a.flag[k] = (k == 0 || a.class_lo_final[k] != a.class_lo_final[k - 1] ||
                     a.class_hi_sorted[k] != a.class_hi_sorted[k - 1]) ? 1 : 0;
    else
        a.flag[k] = (k == 0 || a.class_key_sorted[k] != a.class_key_sorted[k - 1]) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_m2l_assign(M2LClassArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_pairs) return;
    // cls is the exclusive scan of the class-start flags: the class id is the
    // number of starts up to and including k, minus one
    const int c = a.cls[k] + a.flag[k] - 1;
    a.pair_class[a.pair_idx_sorted[k]] = c;
    if (a.flag[k]) {
        if (a.periodic) {
            a.class_key[c] = a.class_lo_final[k];
            a.class_key_hi[c] = a.class_hi_sorted[k];
            a.class_dl[c] = (int)(a.class_hi_sorted[k] & 255ull) - 128;
        } else {
            a.class_key[c] = a.class_key_sorted[k];
            a.class_dl[c] = (int)(a.class_key_sorted[k] & 255ull) - 128;
        }
    }
    if (k == a.n_pairs - 1) *a.num_classes = c + 1;
}


// One class tensor per workgroup iteration: T_k(v) for |k| <= 2p by the
// Hermite recursion over degree levels (one barrier per level), scaled by
// 2^{dl |k|} when the target is the deeper box (dl > 0).
// LDS (dynamic, sized by the order): the recursion's auxiliary array, hermite_aux_size(2p) doubles,
// then the packed-index exponent table (ncomp(2p) ints) and the level bases (2p + 2 ints), staged
// once per workgroup.
__global__ __launch_bounds__(64) void k_class_tensor(M2LClassArgs a) {
    extern __shared__ double ct_lds[];
    const int K = 2 * a.p;
    const int nT = ncomp(K);
    double* A = ct_lds;
    int* sxyz = reinterpret_cast<int*>(ct_lds + hermite_aux_size(K));
    int* slb = sxyz + nT;
    for (int i = threadIdx.x; i < nT; i += 64) sxyz[i] = a.xyz[i];
    for (int i = threadIdx.x; i < K + 2; i += 64) slb[i] = a.level_base[i];
    __syncthreads();
    const int ncls = *a.num_classes;
    for (int c = blockIdx.x; c < ncls; c += gridDim.x) {
        const unsigned long long key = a.class_key[c];
        double vx, vy, vz;
        const int dl = a.class_dl[c];
        if (a.periodic) {
            vx = (double)((long long)((key >> (2 * PBITS)) & ((1ull << PBITS) - 1)) - POFF);
            vy = (double)((long long)((key >> PBITS) & ((1ull << PBITS) - 1)) - POFF);
            vz = (double)((long long)(key & ((1ull << PBITS) - 1)) - POFF);
            const unsigned long long hi = a.class_key_hi[c];
            const int img = (int)(hi >> 16), lev = (int)((hi >> 8) & 255ull);
            if (img != 0) {
                // c_i - (c_j + D) = u (v - D / u), u the half-width at the deeper of the two levels
                const int deep = lev + (dl < 0 ? -dl : 0);
                const double u = ldexp(a.r[0], -(deep - 1));
                vx -= a.disp[3 * img] / u;
                vy -= a.disp[3 * img + 1] / u;
                vz -= a.disp[3 * img + 2] / u;
            }
        } else {
            const int b = a.vbits;
            const unsigned long long mask = (1ull << b) - 1;
            const long long voff = 1ll << (b - 1);
            vx = (double)((long long)((key >> (2 * b + 8)) & mask) - voff);
            vy = (double)((long long)((key >> (b + 8)) & mask) - voff);
            vz = (double)((long long)((key >> 8) & mask) - voff);
        }
        const int zmax = a.zmax;
        hermite_aux(vx, vy, vz, K, A, sxyz, slb, zmax);
        for (int idx = threadIdx.x; idx < nT; idx += 64) {
            const int pk = sxyz[idx];
            const int k = (pk & 255) + ((pk >> 8) & 255) + (pk >> 16);
            double val = (pk >> 16) <= zmax ? hermite_T(A, idx, sxyz, slb) : 0.0;
            if (dl > 0) val = ldexp(val, dl * k);
            a.Tc[(int64_t)c * nT + idx] = val;
        }
        __syncthreads();
    }
}

// M~_j(s) = c_s M_j(s) r_j^{-|s|} from the set-major CM; written set-major
// (s_out == 0) or set-fastest with s_out >= nsets sets per coefficient
// (padding sets zero, for the MFMA tiles).
__global__ __launch_bounds__(256) void k_scale_sources(const double* __restrict__ CM, const double* __restrict__ cs,
                                                       const double* __restrict__ r, double* __restrict__ CMs,
                                                       int num_nodes, int nm, int nsets, int s_out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int width = s_out > 0 ? s_out : nsets;
    if (x >= (int64_t)num_nodes * nm * width) return;
    int64_t js;
    int set;
    if (s_out > 0) { js = x / s_out; set = (int)(x - js * s_out); }
    else { set = (int)(x / ((int64_t)num_nodes * nm)); js = x - (int64_t)set * num_nodes * nm; }
    const int j = (int)(js / nm), s = (int)(js - (int64_t)j * nm);
    if (set >= nsets) { CMs[x] = 0.0; return; }
    int k = 0;
    while (ncomp(k) <= s) ++k;
    const double rinv = 1.0 / r[j];
    double f = 1.0;
    for (int q = 0; q < k; ++q) f *= rinv;
    CMs[x] = cs[s] * CM[(int64_t)set * num_nodes * nm + js] * f;
}

constexpr int M2L_LDS_SOURCE_P = 8;   // from this order on, source expansions are staged in LDS

__host__ __device__ constexpr int tet(int k) { return k * (k + 1) * (k + 2) / 6; }
__host__ __device__ constexpr int tri(int w) { return w * (w + 1) / 2; }

// Packed decomposition of a target component t for the index of s + t:
//   idx(s + t) = tet(|s| + |t|) + tri(ws + wt) + (sz + tz),  w = y + z
// (packed order within a degree runs over w, then z).
__device__ __forceinline__ int decode_target(int t) {
    int k = 0;
    while (ncomp(k) <= t) ++k;
    const int l = t - ncomp(k - 1);
    int w = 0;
    while (tri(w + 1) <= l) ++w;
    return k | (w << 8) | ((l - tri(w)) << 16);
}

template <int P>
__global__ __launch_bounds__(128, 4) void k_m2l(M2LArgs a) {
    constexpr int NM = ncomp(P);
    constexpr int NT = ncomp(2 * P);
    constexpr int NQ = (NM + 127) / 128;
    constexpr int NL = (NT + 127) / 128;
    static_assert(NQ <= 3, "at most 3 target components per lane");
    __shared__ double Tl[2][NT];
    __shared__ double Ms[P >= M2L_LDS_SOURCE_P ? NM : 1];
    const int i = blockIdx.x;
    const int set = blockIdx.y;                  // one multipole set per grid row
    const int64_t set_off = (int64_t)set * a.num_nodes * NM;   // set-major CMs and LE
    const int tid = threadIdx.x;
    const int e0 = a.off[i], e1 = a.off[i + 1];
    const int dec0 = decode_target(min(tid, NM - 1));
    const int dec1 = NQ > 1 ? decode_target(min(tid + 128, NM - 1)) : 0;
    const int dec2 = NQ > 2 ? decode_target(min(tid + 256, NM - 1)) : 0;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    double pre[NL];
    if (e0 < e1) {
        const double* T0 = a.Tc + (int64_t)a.pair_class[e0] * NT;
#pragma unroll
        for (int x = 0; x < NL; ++x)
            if (tid + 128 * x < NT) Tl[0][tid + 128 * x] = T0[tid + 128 * x];
    }
    __syncthreads();
    const int li = a.level[i];
    for (int e = e0; e < e1; ++e) {
        const int cur = (e - e0) & 1;
        const bool more = e + 1 < e1;
        if (more) {
            const double* Tn = a.Tc + (int64_t)a.pair_class[e + 1] * NT;
#pragma unroll
            for (int x = 0; x < NL; ++x) pre[x] = (tid + 128 * x < NT) ? Tn[tid + 128 * x] : 0.0;
        }
        const int j = src_node(a.src[e]);
        const int dl = li - a.level[j];
        const double* __restrict__ Mj = a.CMs + set_off + (int64_t)j * NM;
        const double* __restrict__ T = Tl[cur];
        if constexpr (P >= M2L_LDS_SOURCE_P) {
            // large orders: the source expansion goes through LDS as well, the
            // loops stay rolled (full unrolling would exhaust the SGPR file)
            __syncthreads();
            for (int x = tid; x < NM; x += 128) Ms[x] = Mj[x];
            __syncthreads();
        }
#pragma unroll 1
        for (int q = 0; q < NQ; ++q) {
            if (tid + 128 * q >= NM) break;
            const int dec = q == 0 ? dec0 : (q == 1 ? dec1 : dec2);
            const int k = dec & 255, w = (dec >> 8) & 255, tz = dec >> 16;
            double sum = 0.0;
            if constexpr (P >= M2L_LDS_SOURCE_P) {
#pragma unroll 1
                for (int n = 0; n <= P; ++n) {
                    const int tb = tet(n + k) + tz;
                    for (int ws = 0; ws <= n; ++ws) {
                        const int base = tb + tri(ws + w);
                        const int sb = tet(n) + tri(ws);
#pragma unroll 4
                        for (int sz = 0; sz <= ws; ++sz) sum = __builtin_fma(Ms[sb + sz], T[base + sz], sum);
                    }
                }
            } else {
#pragma unroll
                for (int n = 0; n <= P; ++n) {
                    const int tb = tet(n + k) + tz;
#pragma unroll
                    for (int ws = 0; ws <= n; ++ws) {
                        const int base = tb + tri(ws + w);
                        const int sb = tet(n) + tri(ws);
#pragma unroll
                        for (int sz = 0; sz <= ws; ++sz) sum = __builtin_fma(Mj[sb + sz], T[base + sz], sum);
                    }
                }
            }
            double scale = 1.0;
            if (dl > 0) scale = ldexp(1.0, -dl * k);
            else if (dl < 0) scale = ldexp(1.0, -dl * (k + 1));
            if (q == 0) acc0 = __builtin_fma(scale, sum, acc0);
            else if (q == 1) acc1 = __builtin_fma(scale, sum, acc1);
            else acc2 = __builtin_fma(scale, sum, acc2);
        }
        if (more) {
#pragma unroll
            for (int x = 0; x < NL; ++x)
                if (tid + 128 * x < NT) Tl[1 - cur][tid + 128 * x] = pre[x];
        }
        __syncthreads();
    }
    const double rinv = 1.0 / a.r[i];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int t = tid + 128 * q;
        if (t < NM) {
            const int dec = q == 0 ? dec0 : (q == 1 ? dec1 : dec2);
            double f = rinv;
            for (int k = 0; k < (dec & 255); ++k) f *= rinv;
            a.LE[set_off + (int64_t)i * NM + t] = (q == 0 ? acc0 : (q == 1 ? acc1 : acc2)) * f;
        }
    }
}

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------- dense class numbering
// Non-periodic classes without sorting the pairs.  With |v_x| <= Rx, |v_y| <= Ry, |v_z| <= Rz and
// |level difference| <= Rd over all pairs, (v, dl) has the dense index
//   (((v_x + Rx) (2Ry+1) + v_y + Ry) (2Rz+1) + v_z + Rz) (2Rd+1) + dl + Rd   in [0, N),
// lexicographic in (v_x, v_y, v_z, dl) like the packed keys.  Every pair marks its index, an exclusive
// scan of the marks numbers the classes in key order -- the numbering of the key sort -- and every pair
// reads its class back.  dims = Rx, Ry, Rz, Rd, N (N = 0: beyond CLASS_DENSE_CAP, error bit 4).  The
// marks are zero before and after (the compaction clears the ones it reads).
constexpr int DENSE_TILE = 4096;           // scan tile: 256 threads x 16 entries
constexpr int DENSE_TILES_MAX = CLASS_DENSE_CAP / DENSE_TILE;
static_assert(DENSE_TILES_MAX <= 2048, "one workgroup scans the tile sums, 2 per thread");

__global__ __launch_bounds__(1024) void k_dense_dims(const int* __restrict__ wg_range, int nwg, int* __restrict__ dims,
                                                     int* __restrict__ error) {
    __shared__ int red[16][4];
    int m[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nwg; b += 1024)
#pragma unroll
        for (int k = 0; k < 4; ++k) m[k] = max(m[k], wg_range[b * 4 + k]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) m[k] = max(m[k], __shfl_xor(m[k], d));
        if (lane == 0) red[w][k] = m[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long R[4], N = 1;
        for (int k = 0; k < 4; ++k) {
            R[k] = 0;
            for (int v = 0; v < 16; ++v) R[k] = max(R[k], (long long)red[v][k]);
            N *= 2 * R[k] + 1;
        }
        for (int k = 0; k < 4; ++k) dims[k] = (int)R[k];
        dims[4] = N <= CLASS_DENSE_CAP ? (int)N : 0;
        if (N > CLASS_DENSE_CAP) atomicOr(error, 4);
    }
}

struct DenseKey {
    int vx, vy, vz, dl;
};
__device__ __forceinline__ DenseKey dense_decode_key(unsigned long long key, int b) {
    const unsigned long long mask = (1ull << b) - 1;
    const int voff = 1 << (b - 1);
    return DenseKey{(int)((key >> (2 * b + 8)) & mask) - voff, (int)((key >> (b + 8)) & mask) - voff,
                    (int)((key >> 8) & mask) - voff, (int)(key & 255ull) - 128};
}
__device__ __forceinline__ int dense_index(const DenseKey& k, const int* dims) {
    return (((k.vx + dims[0]) * (2 * dims[1] + 1) + k.vy + dims[1]) * (2 * dims[2] + 1) + k.vz + dims[2]) *
               (2 * dims[3] + 1) + k.dl + dims[3];
}

__global__ __launch_bounds__(256) void k_dense_mark(const unsigned long long* __restrict__ key, int np, int vbits,
                                                    const int* __restrict__ dims, int* __restrict__ marks) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np || dims[4] == 0) return;
    marks[dense_index(dense_decode_key(key[e], vbits), dims)] = 1;
}

__device__ __forceinline__ int block_exscan_256(int v, int* wsum, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int run = incl - v;
    total = 0;
    for (int k = 0; k < 4; ++k) {
        if (k < w) run += wsum[k];
        total += wsum[k];
    }
    return run;
}

__global__ __launch_bounds__(256) void k_dense_tile_sum(const int* __restrict__ marks, const int* __restrict__ dims,
                                                        int* __restrict__ tsum) {
    __shared__ int wsum[4];
    const int N = dims[4];
    const int t = blockIdx.x;
    if (t * DENSE_TILE >= N) return;
    const int i0 = t * DENSE_TILE + threadIdx.x * 16;
    int v = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) v += i0 + k < N ? marks[i0 + k] : 0;
    int total;
    block_exscan_256(v, wsum, total);
    if (threadIdx.x == 0) tsum[t] = total;
}

__global__ __launch_bounds__(1024) void k_dense_tile_scan(const int* __restrict__ tsum, const int* __restrict__ dims,
                                                          int* __restrict__ toff, int* __restrict__ num_classes) {
    __shared__ int wsum[16];
    const int nt = (dims[4] + DENSE_TILE - 1) / DENSE_TILE;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int a0 = 2 * tid < nt ? tsum[2 * tid] : 0, a1 = 2 * tid + 1 < nt ? tsum[2 * tid + 1] : 0;
    int incl = a0 + a1;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int run = incl - a0 - a1, total = 0;
    for (int k = 0; k < 16; ++k) {
        if (k < w) run += wsum[k];
        total += wsum[k];
    }
    if (2 * tid < nt) toff[2 * tid] = run;
    if (2 * tid + 1 < nt) toff[2 * tid + 1] = run + a0;
    if (tid == 0) *num_classes = total;
}

__global__ __launch_bounds__(256) void k_dense_ids(const int* __restrict__ marks, const int* __restrict__ dims,
                                                   const int* __restrict__ toff, int* __restrict__ ids) {
    __shared__ int wsum[4];
    const int N = dims[4];
    const int t = blockIdx.x;
    if (t * DENSE_TILE >= N) return;
    const int i0 = t * DENSE_TILE + threadIdx.x * 16;
    int m[16], v = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        m[k] = i0 + k < N ? marks[i0 + k] : 0;
        v += m[k];
    }
    int total;
    int run = toff[t] + block_exscan_256(v, wsum, total);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (i0 + k < N) ids[i0 + k] = run;
        run += m[k];
    }
}

// Pairs' classes (the first nblk(np) workgroups), and the classes' keys and level differences from their
// dense indices, clearing the marks (the other workgroups, grid-stride over N).
__global__ __launch_bounds__(256) void k_dense_finish(M2LClassArgs a, const int* __restrict__ dims,
                                                      int* __restrict__ marks, const int* __restrict__ ids,
                                                      int pair_blocks) {
    if ((int)blockIdx.x < pair_blocks) {
        const int e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= a.n_pairs || dims[4] == 0) return;
        a.pair_class[e] = ids[dense_index(dense_decode_key(a.class_key_in[e], a.vbits), dims)];
        return;
    }
    const int N = dims[4];
    const int b = a.vbits;
    const int voff = 1 << (b - 1);
    const int nd = 2 * dims[3] + 1, nz = 2 * dims[2] + 1, ny = 2 * dims[1] + 1;
    for (int i = (blockIdx.x - pair_blocks) * blockDim.x + threadIdx.x; i < N;
         i += (gridDim.x - pair_blocks) * blockDim.x) {
        if (!marks[i]) continue;
        marks[i] = 0;
        int r = i;
        const int dl = r % nd - dims[3];
        r /= nd;
        const int vz = r % nz - dims[2];
        r /= nz;
        const int vy = r % ny - dims[1];
        const int vx = r / ny - dims[0];
        const int c = ids[i];
        a.class_key[c] = ((unsigned long long)(vx + voff) << (2 * b + 8)) |
                         ((unsigned long long)(vy + voff) << (b + 8)) | ((unsigned long long)(vz + voff) << 8) |
                         (unsigned long long)(dl + 128);
        a.class_dl[c] = dl;
    }
}

void launch_class_dense(const M2LClassArgs& a, int nwg, int* dims, int* marks, int* ids, int* tsum, int* toff,
                        hipStream_t s) {
    if (a.n_pairs <= 0) return;
    hipLaunchKernelGGL(k_dense_dims, dim3(1), dim3(1024), 0, s, a.wg_range, nwg, dims, a.error);
    hipLaunchKernelGGL(k_dense_mark, dim3(nblk(a.n_pairs, 256)), dim3(256), 0, s, a.class_key_in, a.n_pairs, a.vbits,
                       dims, marks);
    hipLaunchKernelGGL(k_dense_tile_sum, dim3(DENSE_TILES_MAX), dim3(256), 0, s, marks, dims, tsum);
    hipLaunchKernelGGL(k_dense_tile_scan, dim3(1), dim3(1024), 0, s, tsum, dims, toff, a.num_classes);
    hipLaunchKernelGGL(k_dense_ids, dim3(DENSE_TILES_MAX), dim3(256), 0, s, marks, dims, toff, ids);
    const int pb = nblk(a.n_pairs, 256);
    hipLaunchKernelGGL(k_dense_finish, dim3(pb + 256), dim3(256), 0, s, a, dims, marks, ids, pb);
    FMMH_HIP(hipGetLastError());
}

void launch_m2l_keys(const M2LClassArgs& a, hipStream_t s) {
    if (a.n_pairs <= 0) return;
    hipLaunchKernelGGL(k_m2l_keys, dim3(nblk(a.n_pairs, 256)), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_m2l_flags(const M2LClassArgs& a, hipStream_t s) {
    if (a.n_pairs <= 0) return;
    hipLaunchKernelGGL(k_m2l_flags, dim3(nblk(a.n_pairs, 256)), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_m2l_assign(const M2LClassArgs& a, hipStream_t s) {
    if (a.n_pairs <= 0) return;
    hipLaunchKernelGGL(k_m2l_assign, dim3(nblk(a.n_pairs, 256)), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_gather_u64(const unsigned long long* __restrict__ in,
                                                    const int* __restrict__ idx, unsigned long long* __restrict__ out,
                                                    int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = in[idx[k]];
}

void launch_gather_u64(const unsigned long long* in, const int* idx, unsigned long long* out, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_u64, dim3(nblk(n, 256)), dim3(256), 0, s, in, idx, out, n);
    FMMH_HIP(hipGetLastError());
}

void launch_m2l_class_tensors(const M2LClassArgs& a, int grid, hipStream_t s) {
    if (grid <= 0) return;
    const int K = 2 * a.p;
    const size_t lds = (size_t)hermite_aux_size(K) * sizeof(double) + (size_t)(ncomp(K) + K + 2) * sizeof(int);
    hipLaunchKernelGGL(k_class_tensor, dim3(grid), dim3(64), lds, s, a);
    FMMH_HIP(hipGetLastError());
}

// Harmonic (traceless) reduction of the many-sets M2L.  Away from R = 0 the
// Laplace tensor obeys T_{k+2x} + T_{k+2y} + T_{k+2z} = 0, so the contraction
// sum_s T_{s+t} M~(s) over |s| <= p equals the one over the 2n + 1 components
// of each degree n with s_z <= 1, of the folded sources
//   M'(x + 2i, y + 2(k - i), rho) = sum_{k, i} (-1)^k C(k, i) M~(x, y, rho + 2k),
// and the local-expansion components with t_z >= 2 follow from those with
// t_z <= 1 by the same identity,
//   L(tx, ty, 2k + rho) = (-1)^k sum_i C(k, i) L(tx + 2i, ty + 2(k - i), rho).
// The Hankel matrix shrinks from ncomp(p)^2 to (p + 1)^4 entries (220^2 ->
// 100^2 at p = 9); the class tensors stay whole.  Reduced components are
// ordered by degree, then w = y + z, then z, as the packed order runs.
constexpr int MAX_NM_FOLD = ncomp(10);   // expansion coefficients at the largest order of the M2L engines

// Reduced component r -> degree | w << 8 | z << 16 (the packing of decode_target).
__device__ __forceinline__ int decode_reduced(int r) {
    int k = 0;
    while (rdim(k) <= r) ++k;
    const int pos = r - k * k;   // k^2 = rdim(k - 1) components of lower degree
    const int w = (pos + 1) >> 1;
    const int z = (pos > 0 && (pos & 1) == 0) ? 1 : 0;
    return k | (w << 8) | (z << 16);
}

// Packed (full) index of a decoded component.
__device__ __forceinline__ int packed_of(int d) {
    const int k = d & 255, w = (d >> 8) & 255, z = d >> 16;
    return ncomp(k - 1) + tri(w) + z;
}

// Folded, box-scaled sources M'_j(s') in the MFMA layout [node][s' < rdim(p)][s_ld sets] (padding sets
// zero): the sum over the components s that fold onto s' of (-1)^k C(k, i) c_s M_j(s), times r_j^{-|s'|}.
// One workgroup per (node, 16 sets): the sets' expansions are loaded into LDS along the coefficients
// (coalesced) and written along the sets.
constexpr int FOLD_SETS = 16;
// C(k, i) for the folds: k <= 5 at orders <= 10 (2k <= the degree)
__constant__ double kFoldBinom[6][6] = {{1, 0, 0, 0, 0, 0},   {1, 1, 0, 0, 0, 0},    {1, 2, 1, 0, 0, 0},
                                        {1, 3, 3, 1, 0, 0},   {1, 4, 6, 4, 1, 0},    {1, 5, 10, 10, 5, 1}};

// The fold of one reduced source component s = (X, Y, rho): sum_{k, i} (-1)^k C(k, i) c M(X - 2i, Y - 2(k - i),
// rho + 2k), with c M read through cms(packed index).
template <class F>
__device__ __forceinline__ double fold_component(int X, int Y, int rho, F cms) {
    double acc = 0.0;
    for (int k = 0; 2 * k <= X + Y; ++k) {
        const double sg = (k & 1) ? -1.0 : 1.0;
        for (int i = 0; i <= k; ++i) {
            const int sx = X - 2 * i, sy = Y - 2 * (k - i);
            if (sx < 0 || sy < 0) continue;
            acc += (sg * kFoldBinom[k][i]) * cms(xyz2idx(sx, sy, rho + 2 * k));
        }
    }
    return acc;
}

// One set: one thread per (node, reduced component), output CMs[node][rdim(p)].
__global__ __launch_bounds__(256) void k_scale_sources_red1(const double* __restrict__ CM, const double* __restrict__ cs,
                                                            const double* __restrict__ r, double* __restrict__ CMs,
                                                            int num_nodes, int p) {
    const int nr = rdim(p), nm = ncomp(p);
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)num_nodes * nr) return;
    const int j = (int)(x / nr), sr = (int)(x - (int64_t)j * nr);
    const double* __restrict__ m = CM + (int64_t)j * nm;
    const int d = decode_reduced(sr);
    const int deg = d & 255, w = (d >> 8) & 255, rho = d >> 16;
    const double acc = fold_component(deg - w, w - rho, rho, [&](int sc) { return cs[sc] * m[sc]; });
    const double rinv = 1.0 / r[j];
    double f = 1.0;
    for (int q = 0; q < deg; ++q) f *= rinv;
    CMs[x] = acc * f;
}

// Several sets: the terms come from the host-built table (those fold_component visits, in its order, with
// the same products), the powers of 1 / r_j from LDS (the same sequence of products as one set's).
__global__ __launch_bounds__(256) void k_scale_sources_red(const double* __restrict__ CM, FoldTab ft,
                                                           const double* __restrict__ r, double* __restrict__ CMs,
                                                           int num_nodes, int p, int nsets, int s_ld) {
    __shared__ double cm[FOLD_SETS][MAX_NM_FOLD];
    __shared__ double pw[PMAX + 1];
    const int nr = rdim(p), nm = ncomp(p);
    const int j = blockIdx.x, set0 = blockIdx.y * FOLD_SETS;
    const int nsl = min(FOLD_SETS, s_ld - set0);   // output columns of this workgroup
    for (int x = threadIdx.x; x < FOLD_SETS * nm; x += blockDim.x) {
        const int u = x / nm, sc = x - u * nm;
        cm[u][sc] = set0 + u < nsets ? CM[((int64_t)(set0 + u) * num_nodes + j) * nm + sc] : 0.0;
    }
    if (threadIdx.x == 0) {
        const double rinv = 1.0 / r[j];
        double f = 1.0;
        for (int q = 0; q <= p; ++q) {
            pw[q] = f;
            f *= rinv;
        }
    }
    __syncthreads();
    for (int x = threadIdx.x; x < nr * nsl; x += blockDim.x) {
        const int sr = x / nsl, u = x - sr * nsl;
        double out = 0.0;
        if (set0 + u < nsets) {
            double acc = 0.0;
            for (int t = ft.off[sr]; t < ft.off[sr + 1]; ++t) acc += ft.w[t] * (ft.c[t] * cm[u][ft.sc[t]]);
            out = acc * pw[ft.deg[sr]];
        }
        CMs[((int64_t)j * nr + sr) * s_ld + set0 + u] = out;
    }
}

// Local-expansion components with t_z >= 2 from the reduced ones the M2L wrote
// (LE[row][ncomp(p)], rows = sets x nodes): UNFOLD_ROWS rows per workgroup are
// loaded into LDS, completed there and written back whole (coalesced).
constexpr int UNFOLD_ROWS = 8;
__global__ __launch_bounds__(256) void k_le_unfold(double* __restrict__ LE, int64_t rows, int p) {
    __shared__ double lr[UNFOLD_ROWS][MAX_NM_FOLD];
    const int nm = ncomp(p);
    const int64_t r0 = (int64_t)blockIdx.x * UNFOLD_ROWS;
    const int nrw = (int)min((int64_t)UNFOLD_ROWS, rows - r0);
    double* __restrict__ base = LE + r0 * nm;
    for (int x = threadIdx.x; x < nrw * nm; x += blockDim.x) lr[x / nm][x % nm] = base[x];
    __syncthreads();
    for (int x = threadIdx.x; x < nrw * nm; x += blockDim.x) {
        const int rw = x / nm, t = x - rw * nm;
        const int d = decode_target(t);
        const int deg = d & 255, w = (d >> 8) & 255, tz = d >> 16;
        if (tz < 2) {   // unchanged; stored again so that every cache line is written whole
            base[x] = lr[rw][t];
            continue;
        }
        const int tx = deg - w, ty = w - tz, k = tz >> 1, rho = tz & 1;
        double acc = 0.0;
        for (int i = 0; i <= k; ++i) acc += kFoldBinom[k][i] * lr[rw][xyz2idx(tx + 2 * i, ty + 2 * (k - i), rho)];
        base[x] = (k & 1) ? -acc : acc;
    }
}

void launch_scale_sources_red(const double* CM, const double* cs, const FoldTab& ft, const double* r, double* CMs,
                              int num_nodes, int p, int nsets, int s_ld, hipStream_t s) {
    if (num_nodes <= 0) return;
    if (nsets == 1 && s_ld == 1) {
        const int64_t tot = (int64_t)num_nodes * rdim(p);
        hipLaunchKernelGGL(k_scale_sources_red1, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, CM, cs, r, CMs,
                           num_nodes, p);
    } else {
        if (!ft.off) throw Error(FMMH_ERR_INTERNAL, "fold table missing");
        hipLaunchKernelGGL(k_scale_sources_red, dim3(num_nodes, (s_ld + FOLD_SETS - 1) / FOLD_SETS), dim3(256), 0, s,
                           CM, ft, r, CMs, num_nodes, p, nsets, s_ld);
    }
    FMMH_HIP(hipGetLastError());
}

void launch_le_unfold(double* LE, long long rows, int p, hipStream_t s) {
    if (rows <= 0 || p < 2) return;
    hipLaunchKernelGGL(k_le_unfold, dim3(nblk(rows, UNFOLD_ROWS)), dim3(256), 0, s, LE, (int64_t)rows, p);
    FMMH_HIP(hipGetLastError());
}

void launch_scale_sources(const double* CM, const double* cs, const double* r, double* CMs, int num_nodes, int p,
                          int nsets, int s_ld_out, hipStream_t s) {
    const int nm = ncomp(p);
    const int width = s_ld_out > 0 ? s_ld_out : nsets;
    hipLaunchKernelGGL(k_scale_sources, dim3(nblk((long long)num_nodes * nm * width, 256)), dim3(256), 0, s, CM, cs,
                       r, CMs, num_nodes, nm, nsets, s_ld_out);
    FMMH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- MFMA form
// For many multipole sets, the contraction of one pair is a small GEMM
//   L~_i[t][set] += sum_s H_c[t][s] M~_j[s][set],  H_c[t][s] = scale(t) T~_c[s+t]
// with the n_m x n_m Hankel matrix A gathered from the class tensor (staged
// in a double-buffered LDS slot, prefetched one pair ahead) through the
// host-built fragment index table (packed index of s + t per A-fragment
// lane, read from global memory: L2-resident) and an n_m x 16 source block B
// (16 sets).  One wavefront per (target node, group of 16 sets) keeps the
// whole n_m x 16 result in MFMA accumulators across all its pairs.
//   FP64 (v_mfma_f64_16x16x4f64): accumulators stay FP64 over every pair.
//   FP32 (mixed precision, v_mfma_f32_16x16x4f32): the source block is
//   staged per pair in LDS as FP32, each pair's 16-row tiles are accumulated
//   over s in FP32, then added to FP64 accumulators, so the sum over a
//   target's pairs is FP64.
// MFMA operand layouts (measured on gfx950): A lane l -> A[l % 16][l / 16],
// B lane l -> B[l / 16][l % 16]; accumulator register q of lane l ->
// D[4 q + l / 16][l % 16] for f64_16x16x4f64, D[4 (l / 16) + q][l % 16] for
// f32_16x16x4f32.  Learned per-pair orders (a.order) shorten both loops.
using v4d = double __attribute__((ext_vector_type(4)));
using v4f = float __attribute__((ext_vector_type(4)));
constexpr int M2L_MFMA_PMAX = 10;
constexpr int M2L_SETS_TILE = 16;

template <int P, bool F32>
constexpr int mfma_waves() { return ((rdim(P) + 15) / 16) < 4 ? ((rdim(P) + 15) / 16) : 4; }

// NTL = 16-set tiles per workgroup: an A fragment, gathered once, feeds NTL MFMAs.
// FP64: the row scale is applied to each A element and the pair accumulates
// straight into the FP64 MFMA accumulators.  FP32: the class tensor is kept
// as FP32 in LDS, the pair's tiles are accumulated over s in FP32 with
// unscaled A, and each tile is scaled by its rows' powers of two while it is
// added to the FP64 accumulators (rows beyond the pair's order are skipped
// there).
// The next pair's class tensor and, where 2 buffers fit in 64 KB (STAGE), its
// source block (n_m x 16 NTL sets) are loaded into registers with wide
// coalesced loads while the current pair is multiplied, then published to the
// other LDS buffer: the k-loop reads only LDS.
// One 16-byte chunk per lane from global memory straight into LDS: lane l of the
// wave writes bytes [16 l, 16 l + 16) after the wave-uniform LDS address lds.
__device__ __forceinline__ void dma16(const void* g, void* lds) {
    __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
}

// Target node of a workgroup: workgroup b runs on XCD b mod 8 (each XCD with its own L2), so XCD x
// takes chunks x, x + 8, x + 16, ... of M2L_XCD_CHUNK consecutive target nodes.  Nodes close in node
// order (depth-first over the Morton-sorted sites) share most of their M2L sources, which the
// workgroups an XCD runs at once then read from its L2; with plain order each XCD's L2 would see
// every eighth target and hold far fewer reused source blocks.  The grid is padded to whole rounds of
// chunks (m2l_xcd_grid); workgroups past the last node return at once.
constexpr int M2L_XCD_CHUNK = 16;
__device__ __forceinline__ int m2l_xcd_target(int b) {
    const int x = b % NUM_XCD, k = b / NUM_XCD;
    return ((k / M2L_XCD_CHUNK) * NUM_XCD + x) * M2L_XCD_CHUNK + k % M2L_XCD_CHUNK;
}
static int m2l_xcd_grid(int n) {
    const int r = NUM_XCD * M2L_XCD_CHUNK;
    return (n + r - 1) / r * r;
}

template <int P, bool F32>
constexpr int mfma_kpad() { return ((rdim(P) + 3) / 4 + 1) / 2 * 2; }   // k-steps over reduced components, even

// Source blocks are staged in LDS when the two buffers take at most 80 KB (with the class tensors
// and tables a workgroup then stays under 96 KB: p 10 FP32 with 16 sets runs one workgroup per CU).
template <int P, bool F32, int NTL>
constexpr bool mfma_stage_b() {
    return 2 * 4 * (mfma_kpad<P, F32>() + 2) * 16 * NTL * (F32 ? 4 : 8) <= 81920;
}

// The k-loop is branch-free: the A element of lane (row, kq) is read from the
// LDS class tensor at byte offset G[row part + source part] + z parts, where
// G[d][w] = sizeof(element) (tet(d) + tri(w)) sits in LDS and the row and
// source parts are pre-scaled byte offsets; rows past n_m read a valid
// (finite) element and are never stored; source rows past the pair's order
// multiply zeros (staged blocks are zero-filled there, global reads are
// selected to zero).  Two k-steps per iteration.
template <int P, bool F32, int NTL, bool STG>
__global__ __launch_bounds__(256) void k_m2l_mfma(M2LArgs a) {
    constexpr int NM = ncomp(P);            // local-expansion row length
    constexpr int NK = rdim(P);             // reduced (harmonic) components: the GEMM's rows and k
    constexpr int NT = ncomp(2 * P);
    constexpr int MT = (NK + 15) / 16;      // 16-row tiles of reduced target coefficients
    constexpr int KSP = mfma_kpad<P, F32>();  // 4-deep steps over source coefficients, padded
    constexpr int NW = mfma_waves<P, F32>();  // wavefronts per workgroup, tile mt to wave mt % NW
    constexpr int MW = (MT + NW - 1) / NW;  // row tiles per wavefront
    constexpr int NTH = 64 * NW;
    constexpr int NL = (NT + NTH - 1) / NTH;
    constexpr bool STAGE = STG && mfma_stage_b<P, F32, NTL>();
    constexpr int BW = 16 * NTL;            // sets per workgroup
    constexpr int NBL = STAGE ? (NK * BW + NTH - 1) / NTH : 1;
    using TE = typename std::conditional<F32, float, double>::type;
    using Acc = typename std::conditional<F32, v4f, v4d>::type;
    constexpr int ES = sizeof(TE);
    constexpr int BROWS = STAGE ? 4 * (KSP + 2) * BW : 1;   // + the pipeline's read-ahead
    constexpr int CH = 16 / ES;             // elements per 16-byte chunk
    constexpr int NTP = (NT + CH - 1) / CH * CH;
    constexpr int CPR = BW / CH;            // chunks per staged source row
    __shared__ __align__(16) TE Tl[2][NTP];
    __shared__ __align__(16) TE Bl[2][BROWS];   // staged source block [s][set]
    __shared__ unsigned char degl[16 * MT]; // degree of target row t (99 past the reduced components)
    const int i = m2l_xcd_target(blockIdx.x);
    if (i >= a.num_nodes) return;           // grid padding (the whole workgroup)
    const int sg = blockIdx.y * BW;         // first set of this workgroup
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63;
    const int row = lane & 15, kq = lane >> 4;
    const int e0 = a.off[i], e1 = a.off[i + 1];
    const TE* __restrict__ cms = F32 ? (const TE*)a.CMs32 : (const TE*)a.CMs;
    for (int x = tid; x < 16 * MT; x += NTH) degl[x] = x < NK ? (decode_reduced(x) & 255) : 99;
    // A fragment of row tile mt holds H[t = 16 mt + row][s = 4 ks + kq] = T~_c[s + t]: its byte offset in
    // a class-tensor buffer, per lane and k-step, from the host-built fragment table (kept in registers
    // for the whole target; outside the matrix an offset of 0 reads a finite element that is either
    // never stored (rows) or multiplies zero-filled source rows (columns))
    constexpr int KS = (NK + 3) / 4;        // k-steps of the fragment table
    int fo[MW][KSP], tdeg[MW];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi) {
        const int mt = w + NW * mi;
        const int t = mt * 16 + row;
        tdeg[mi] = t < NK ? (decode_reduced(t) & 255) : 99;
#pragma unroll
        for (int ks = 0; ks < KSP; ++ks) {
            const int f = (mt < MT && ks < KS) ? a.fidx[(mt * KS + ks) * 64 + lane] : -1;
            fo[mi][ks] = ES * (f >= 0 ? f : 0);
            asm volatile("" : "+v"(fo[mi][ks]));   // opaque: kept in a register, not re-loaded per pair
        }
    }
    v4d acc[MW][NTL];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi)
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) acc[mi][nt] = v4d{0.0, 0.0, 0.0, 0.0};
    // Pair info (class, source node, its level, order) is read one pair ahead and is complete
    // before the next pair's copies start, so no vector load waits behind them.
    // Next pair's class tensor and (STAGE) source block rows 0..n_q-1 into the other LDS
    // buffer: STAGE first zero-fills the rows past n_q that its k-steps read, then copies
    // 16-byte chunks straight from global memory into LDS (global_load_lds_dwordx4: lane l
    // of a wave fills bytes 16 l.. after the wave's LDS base), waited for with vmcnt before
    // the pair's closing barrier; otherwise the class tensor goes through registers.
    double pre[STAGE ? 1 : NL];
    const TE* __restrict__ tcp = (const TE*)a.Tcp;
    // Pair info runs two pairs ahead: when pair e starts, pair e + 1's class, source node and
    // order are complete (loaded during pair e - 1, covered by its closing wait) and its
    // copies start at once; pair e + 1's source level and pair e + 2's raw info are loaded
    // while pair e is multiplied.
    int cls_n = 0, j_n = 0, q_n = P, lv_n = 0;   // next pair (lv_n: in flight during this pair)
    int cls_r = 0, src_r = 0, q_r = P;            // the pair after it, raw
    auto load_raw = [&](int e) {
        cls_r = a.pair_class[e];
        src_r = a.src[e];
        q_r = a.order ? (int)a.order[e] : P;
    };
    auto take_raw = [&]() {
        cls_n = __builtin_amdgcn_readfirstlane(cls_r);
        j_n = __builtin_amdgcn_readfirstlane(src_node(src_r));
        q_n = __builtin_amdgcn_readfirstlane(q_r);
    };
    const int brow = lane / CPR;
    const int boff_l = brow * a.s_ld + (lane % CPR) * CH;
    auto fetch = [&](int buf) {
        if constexpr (STAGE) {
            const int nqn = rdim(q_n);
            const int kend = 4 * (((nqn + 3) / 4 + 1) & ~1);
            for (int x = nqn * BW + tid; x < kend * BW; x += NTH) Bl[buf][x] = TE(0);
            const TE* __restrict__ Tn = tcp + (int64_t)cls_n * NTP;
            for (int k = w; k * 64 * CH < NTP; k += NW)
                if (64 * k * CH + lane * CH < NTP) dma16(Tn + 64 * k * CH + lane * CH, &Tl[buf][64 * k * CH]);
            // instruction k copies rows 64 k / CPR + lane / CPR, chunk lane % CPR of each
            const TE* __restrict__ B = cms + (int64_t)j_n * NK * a.s_ld + sg;
            for (int k = w; k * 64 < nqn * CPR; k += NW)
                if (k * (64 / CPR) + brow < nqn) dma16(B + (int64_t)k * (64 / CPR) * a.s_ld + boff_l, &Bl[buf][64 * k * CH]);
        } else {
            const double* Tn = a.Tc + (int64_t)cls_n * NT;
#pragma unroll
            for (int x = 0; x < NL; ++x) pre[x] = (tid + NTH * x < NT) ? Tn[tid + NTH * x] : 0.0;
        }
    };
    auto publish = [&](int buf) {
        if constexpr (STAGE) {
            __builtin_amdgcn_s_waitcnt(0);
        } else {
#pragma unroll
            for (int x = 0; x < NL; ++x)
                if (tid + NTH * x < NT) Tl[buf][tid + NTH * x] = (TE)pre[x];
        }
    };
    const int li = a.level[i];
    if (e0 < e1) {
        load_raw(e0);
        take_raw();
        fetch(0);
        lv_n = a.level[j_n];
        if (e0 + 1 < e1) load_raw(e0 + 1);
        publish(0);
    }
    __syncthreads();
    for (int e = e0; e < e1; ++e) {
        const int cur = (e - e0) & 1;
        const bool more = e + 1 < e1;
        // this pair: source node, level offset, learned order q (only the leading rdim(q)
        // rows and columns)
        const int j = j_n;
        const int dl = li - __builtin_amdgcn_readfirstlane(lv_n);
        const int qo = q_n;
        if (more) {
            take_raw();
            fetch(1 - cur);
            lv_n = a.level[j_n];
            if (e + 2 < e1) load_raw(e + 2);
        }
        const int nq = rdim(qo);
        const int KSq = ((nq + 3) / 4 + 1) & ~1;
        // level-offset scale of row t as a power of two, 2^-(dl |t|) or 2^(dl (|t| + 1)), i.e.
        // 2^(sa |t| + sb); rows past the pair's order get 2^-2000, which flushes them to zero
        const int sa = -dl, sb = dl < 0 ? -dl : 0;
        int sh[F32 ? 1 : MW];
        if constexpr (!F32) {
#pragma unroll
            for (int mi = 0; mi < MW; ++mi) sh[mi] = tdeg[mi] <= qo ? __mul24(sa, tdeg[mi]) + sb : -2000;
        }
        const char* __restrict__ T = (const char*)&Tl[cur][0];
        const TE* __restrict__ Bc = Bl[cur];
        const int64_t boff = (int64_t)j * NK * a.s_ld + sg + row;
        Acc part[MW][NTL];
        // row tiles wholly past the pair's order are skipped (a wave-uniform test)
        bool tile_on[MW];
#pragma unroll
        for (int mi = 0; mi < MW; ++mi) tile_on[mi] = (w + NW * mi) * 16 < nq;
        // Software pipeline over step pairs (k0, k0 + 1), unrolled over the padded k-steps: the
        // tensor elements and source rows of step pair k0 + 2 are loaded while the MFMAs of k0 run
        // (two register sets in turn).  Loads past the pair's k-steps stay inside the buffers and
        // are never multiplied.
        TE av[2][2][MW], bv[2][2][NTL];
        auto load = [&](int k0, int sl) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
#pragma unroll
                for (int mi = 0; mi < MW; ++mi) av[sl][u][mi] = *(const TE*)(T + fo[mi][k0 + u]);
                const int s = (k0 + u) * 4 + kq;
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt) {
                    if constexpr (STAGE) {
                        bv[sl][u][nt] = Bc[s * BW + 16 * nt + row];
                    } else {
                        const bool sv = s < nq;
                        const TE v = cms[boff + (int64_t)(sv ? s : 0) * a.s_ld + 16 * nt];
                        bv[sl][u][nt] = sv ? v : TE(0);
                    }
                }
            }
        };
        load(0, 0);
#pragma unroll
        for (int k0 = 0; k0 < KSP; k0 += 2) {
            if (k0 >= KSq) continue;   // (a wave-uniform skip of the rest: no break, so the loop unrolls)
            const int sl = (k0 >> 1) & 1;
            if (k0 + 2 < KSP) load(k0 + 2, 1 - sl);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int mi = 0; mi < MW; ++mi) {
                    if (!tile_on[mi]) continue;
                    if constexpr (F32) {
#pragma unroll
                        for (int nt = 0; nt < NTL; ++nt)
                            part[mi][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                av[sl][u][mi], bv[sl][u][nt], (k0 == 0 && u == 0) ? Acc{0, 0, 0, 0} : part[mi][nt], 0,
                                0, 0);
                    } else {
                        const double x = ldexp(av[sl][u][mi], sh[mi]);
#pragma unroll
                        for (int nt = 0; nt < NTL; ++nt)
                            acc[mi][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, bv[sl][u][nt], acc[mi][nt], 0, 0, 0);
                    }
                }
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (F32) {
            // this pair's FP32 tiles, row-scaled, into the FP64 accumulators (the f32 layout keeps
            // row 4 kq + q in register q, the f64 one row 4 q + kq; remapped at the end)
#pragma unroll
            for (int mi = 0; mi < MW; ++mi) {
                if (!tile_on[mi]) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int deg = degl[(w + NW * mi) * 16 + 4 * kq + q];
                    const int shq = deg <= qo ? __mul24(sa, deg) + sb : -2000;
#pragma unroll
                    for (int nt = 0; nt < NTL; ++nt) acc[mi][nt][q] += ldexp((double)part[mi][nt][q], shq);
                }
            }
        }
        if (more) publish(1 - cur);
        __syncthreads();
    }
    const double rinv = 1.0 / a.r[i];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi) {
        const int mt = w + NW * mi;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = mt * 16 + (F32 ? 4 * kq + q : 4 * q + kq);
            if (mt >= MT || t >= NK) continue;
            const int dt = decode_reduced(t), deg = dt & 255, tf = packed_of(dt);
            double f = rinv;
            for (int k = 0; k < deg; ++k) f *= rinv;
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) {
                const int set = sg + 16 * nt + row;
                if (set < a.nsets) a.LE[((int64_t)set * a.num_nodes + i) * NM + tf] = acc[mi][nt][q] * f;
            }
        }
    }
}

int m2l_mfma_class_stride(int p, bool f32) {
    const int ch = f32 ? 4 : 2;
    return (ncomp(2 * p) + ch - 1) / ch * ch;
}

template <typename TE>
__global__ void k_class_pad(const double* __restrict__ Tc, TE* __restrict__ Tcp, int64_t n, int nT, int ntp) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int64_t c = x / ntp;
    const int k = (int)(x - c * ntp);
    Tcp[x] = k < nT ? (TE)Tc[c * nT + k] : TE(0);
}

void launch_m2l_class_pad(const double* Tc, void* Tcp, int ncls, int p, bool f32, hipStream_t s) {
    const int ntp = m2l_mfma_class_stride(p, f32);
    const int64_t n = (int64_t)ncls * ntp;
    if (n == 0) return;
    const int grid = (int)((n + 255) / 256);
    if (f32) hipLaunchKernelGGL(k_class_pad<float>, dim3(grid), dim3(256), 0, s, Tc, (float*)Tcp, n, ncomp(2 * p), ntp);
    else hipLaunchKernelGGL(k_class_pad<double>, dim3(grid), dim3(256), 0, s, Tc, (double*)Tcp, n, ncomp(2 * p), ntp);
    FMMH_HIP(hipGetLastError());
}

bool m2l_mfma_supported(int p, int nsets) { return p <= M2L_MFMA_PMAX && nsets >= 16; }

// Variants of the many-sets form (tile shapes): bit 0 = two 16-set tiles per
// workgroup (needs s_ld a multiple of 32), bit 1 = source blocks read from
// global memory instead of staged in LDS.  Every variant computes the same
// sums in the same order; they differ in speed only.
template <int P, bool F32>
static bool mfma_variant_ok(const M2LArgs& a, int v) {
    const bool two = v & 1, stage = !(v & 2);
    if (two && !(a.nsets > M2L_SETS_TILE && a.s_ld % (2 * M2L_SETS_TILE) == 0)) return false;
    if (stage && !(two ? mfma_stage_b<P, F32, 2>() : mfma_stage_b<P, F32, 1>())) return false;
    return true;
}

template <int P, bool F32>
static void launch_mfma_one(const M2LArgs& a, int num_nodes, int v, hipStream_t s) {
    const dim3 b(64 * mfma_waves<P, F32>());
    const int sets_per = (v & 1) ? 2 * M2L_SETS_TILE : M2L_SETS_TILE;
    const dim3 g(m2l_xcd_grid(num_nodes), (a.nsets + sets_per - 1) / sets_per);
    switch (v) {
        case 0: hipLaunchKernelGGL((k_m2l_mfma<P, F32, 1, true>), g, b, 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_m2l_mfma<P, F32, 2, true>), g, b, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_m2l_mfma<P, F32, 1, false>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_m2l_mfma<P, F32, 2, false>), g, b, 0, s, a); break;
    }
}

template <int P, bool F32>
static bool mfma_p(const M2LArgs& a, int num_nodes, int v, bool launch, hipStream_t s) {
    if (!mfma_variant_ok<P, F32>(a, v)) return false;
    if (launch) launch_mfma_one<P, F32>(a, num_nodes, v, s);
    return true;
}

template <bool F32>
static bool mfma_dispatch(const M2LArgs& a, int num_nodes, int v, bool launch, hipStream_t s) {
    switch (a.p) {
        case 0: return mfma_p<0, F32>(a, num_nodes, v, launch, s);
        case 1: return mfma_p<1, F32>(a, num_nodes, v, launch, s);
        case 2: return mfma_p<2, F32>(a, num_nodes, v, launch, s);
        case 3: return mfma_p<3, F32>(a, num_nodes, v, launch, s);
        case 4: return mfma_p<4, F32>(a, num_nodes, v, launch, s);
        case 5: return mfma_p<5, F32>(a, num_nodes, v, launch, s);
        case 6: return mfma_p<6, F32>(a, num_nodes, v, launch, s);
        case 7: return mfma_p<7, F32>(a, num_nodes, v, launch, s);
        case 8: return mfma_p<8, F32>(a, num_nodes, v, launch, s);
        case 9: return mfma_p<9, F32>(a, num_nodes, v, launch, s);
        default: return mfma_p<10, F32>(a, num_nodes, v, launch, s);
    }
}

bool m2l_mfma_variant_ok(const M2LArgs& a, int v) {
    return a.CMs32 ? mfma_dispatch<true>(a, 0, v, false, nullptr) : mfma_dispatch<false>(a, 0, v, false, nullptr);
}

// Default variant without tuning: two tiles where their staged blocks fit, else one tile, staged.
int m2l_mfma_default_variant(const M2LArgs& a) {
    if (m2l_mfma_variant_ok(a, 1)) return 1;
    if (m2l_mfma_variant_ok(a, 0)) return 0;
    return m2l_mfma_variant_ok(a, 3) ? 3 : 2;
}

void launch_m2l_mfma(const M2LArgs& a, int num_nodes, int variant, hipStream_t s) {
    if (num_nodes <= 0) return;
    const int v = m2l_mfma_variant_ok(a, variant) ? variant : m2l_mfma_default_variant(a);
    if (a.CMs32) mfma_dispatch<true>(a, num_nodes, v, true, s);
    else mfma_dispatch<false>(a, num_nodes, v, true, s);
    FMMH_HIP(hipGetLastError());
}

void launch_m2l(const M2LArgs& a, int num_nodes, hipStream_t s) {
    if (num_nodes <= 0) return;
    dim3 g(num_nodes, a.nsets), b(128);
    switch (a.p) {
        case 0: hipLaunchKernelGGL(k_m2l<0>, g, b, 0, s, a); break;
        case 1: hipLaunchKernelGGL(k_m2l<1>, g, b, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_m2l<2>, g, b, 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_m2l<3>, g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_m2l<4>, g, b, 0, s, a); break;
        case 5: hipLaunchKernelGGL(k_m2l<5>, g, b, 0, s, a); break;
        case 6: hipLaunchKernelGGL(k_m2l<6>, g, b, 0, s, a); break;
        case 7: hipLaunchKernelGGL(k_m2l<7>, g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL(k_m2l<8>, g, b, 0, s, a); break;
        case 9: hipLaunchKernelGGL(k_m2l<9>, g, b, 0, s, a); break;
        default: hipLaunchKernelGGL(k_m2l<10>, g, b, 0, s, a); break;
    }
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
