// Octree construction and dual-tree interaction lists, bit-exact with
// peter-reinholdt/fmm src/fmm.F90:137-385.
//
// The reference inserts sites one by one into a pointer octree.  Its result
// depends only on (a) the FP64 chain of node centres, root = centroid,
// child = parent +- r/2 per axis (src/fmm.F90:184-189, 229-240), and (b) the
// rule "a node is internal iff its subtree holds >= ncrit sites".  So every
// site can compute its own octant digits down to 21 levels independently
// (a 63-bit key, digit 4z + 2y + x per level, strict '>' tests as
// src/fmm.F90:197-206); after a stable sort by key,
//   * the sites of any node form a contiguous run;
//   * a run of length >= ncrit is internal, and a site's leaf level is
//     2 + max over the ncrit-windows that contain it of the common prefix
//     length of the window's first and last key;
//   * pre-order DFS ids (src/fmm.F90:286-299) are the order of
//     (first site, level) pairs, so one exclusive scan numbers all nodes.
// Leaf members are then re-sorted by site index, which is the reference's
// insertion order within a leaf.
//
// All centre and MAC arithmetic is FP64 with contraction disabled, using the
// same operation order as the reference (compiled without FMA), and the MAC
// distance reproduces flang's NORM2 (running maximum, scaled sum of squares).
#include <climits>

#include "common.h"
#include "kernels.h"

namespace fmmh {

__device__ __forceinline__ int common_digits(unsigned long long a, unsigned long long b) {
    unsigned long long x = a ^ b;
    if (x == 0ull) return KEY_LEVELS;
    return (__clzll((long long)x) - 1) / 3;
}

__device__ __forceinline__ int key_digit(unsigned long long key, int d) {   // d = 1..21
    return (int)((key >> (63 - 3 * d)) & 7ull);
}

// Centroid, one workgroup per axis: sum(coordinates, 1) / N with flang's
// left-to-right SUM, bit for bit (a tree built from another rounding of the
// centroid differs from the reference's).
//
// The sequential chain is evaluated in blocks without giving up exactness.
// While the running sum s stays in one binade [2^e, 2^{e+1}) (by magnitude),
// every partial sum is an integer multiple of the unit u = 2^{e-52}, and
// fl(s + x) = s + rint(x / u) u exactly whenever the result stays strictly
// inside the binade and x / u is not a half-integer (a tie would be broken by
// the parity of s).  So a block of CB_CHUNK addends becomes an exact int64
// prefix sum of rint(x / u); the first addend whose partial sum leaves
// (2^52, 2^53) units, or that is a tie, a non-finite value or too large,
// ends the block: the prefix up to it is exact, that one addend is added in
// FP64 as the reference does, and the next block starts after it with the
// unit of the new binade.  Where blocks end early (a running sum that keeps
// changing binade), a burst of CB_BURST addends is summed sequentially out of
// LDS, so the cost never exceeds the plain chain by much.
// Blocks are staged through LDS with coalesced loads (a thread then reads its
// CB_PER consecutive addends; rows padded by one slot, conflict-free), and
// the next block is prefetched into registers while the current one is
// scanned.
constexpr int CB_THREADS = 1024;
constexpr int CB_PER = 8;
constexpr int CB_CHUNK = CB_THREADS * CB_PER;
constexpr int CB_LDS = CB_CHUNK + CB_CHUNK / CB_PER;
constexpr int CB_MIN_PROGRESS = 64;
constexpr int CB_BURST = 4096;
static_assert(CB_BURST + 32 <= CB_LDS, "a burst and its read-ahead fit in the block buffer");

struct CentroidLds {
    double xs[CB_LDS];
    unsigned long long wtot[CB_THREADS / 64];
    int first_bad;
    double s_next;
};

// The left-to-right sum of col[pos, n) continued from s, by the block method (every thread returns the
// same value; one workgroup of CB_THREADS).
__device__ double centroid_range(const double* __restrict__ col, int pos, int n, double s, CentroidLds& L) {
#pragma clang fp contract(off)
    double* xs = L.xs;
    unsigned long long* wtot = L.wtot;
    int& first_bad = L.first_bad;
    double& s_next = L.s_next;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double pre[CB_PER];
    int pre_pos = -1;
    auto prefetch = [&](int p0) {
#pragma unroll
        for (int q = 0; q < CB_PER; ++q) {
            const int k = p0 + q * CB_THREADS + tid;
            pre[q] = k < n ? col[k] : 0.0;
        }
        pre_pos = p0;
    };
    while (pos < n) {
        if (!(fabs(s) >= 0x1p-1022) || !isfinite(s)) {   // 0, subnormal or non-finite: no binade unit
            s = s + col[pos];
            ++pos;
            continue;
        }
        const int e = ilogb(s);
        const unsigned long long S0 = (unsigned long long)(long long)ldexp(s, 52 - e);
        const int len = min(CB_CHUNK, n - pos);
        if (pre_pos != pos) prefetch(pos);
#pragma unroll
        for (int q = 0; q < CB_PER; ++q) {
            const int i = q * CB_THREADS + tid;
            xs[i + i / CB_PER] = pre[q];
        }
        __syncthreads();
        if (pos + len < n) prefetch(pos + len);   // the next block, if this one runs to its end
        const int base = tid * CB_PER;
        unsigned long long incl[CB_PER];
        unsigned badmask = 0u;
        unsigned long long run = 0ull;
#pragma unroll
        for (int q = 0; q < CB_PER; ++q) {
            const int i = base + q;
            unsigned long long inc = 0ull;
            if (i < len) {
                const double y = ldexp(xs[tid * (CB_PER + 1) + q], 52 - e);
                bool bad = !(fabs(y) < 0x1p54);
                if (!bad) {
                    const double r = rint(y);
                    bad = fabs(y - r) == 0.5;
                    inc = bad ? 0ull : (unsigned long long)(long long)r;
                }
                if (bad) badmask |= 1u << q;
            }
            run += inc;
            incl[q] = run;
        }
        // exclusive scan of the per-thread totals (unsigned: wraps only past the first invalid addend)
        unsigned long long x = run;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned long long y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (tid == 0) first_bad = len;
        if (lane == 63) wtot[wid] = x;
        __syncthreads();
        unsigned long long off = x - run;
        for (int w = 0; w < wid; ++w) off += wtot[w];
        int mine = len;
#pragma unroll
        for (int q = 0; q < CB_PER; ++q) {
            const int i = base + q;
            if (i < len && mine == len) {
                const long long Sk = (long long)(S0 + off + incl[q]);
                const long long a = Sk < 0 ? -Sk : Sk;
                if (((badmask >> q) & 1u) || !(a > (1ll << 52) && a < (1ll << 53))) mine = i;
            }
        }
        // (a wave minimum first: after a binade change every later thread has a bad addend)
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) mine = min(mine, __shfl_xor(mine, d));
        if (lane == 0 && mine < len) atomicMin(&first_bad, mine);
        __syncthreads();
        const int f = first_bad;
        // exact partial sum just before f (or at the end of the block)
        const int last = (f >= len ? len : f) - 1;
#pragma unroll
        for (int q = 0; q < CB_PER; ++q)
            if (base + q == last) s_next = ldexp((double)(long long)(S0 + off + incl[q]), e - 52);
        __syncthreads();
        if (last >= 0) s = s_next;
        if (f >= len) {
            pos += len;
        } else {
            s = s + xs[f + f / CB_PER];
            pos += f + 1;
            if (f < CB_MIN_PROGRESS && pos < n) {
                const int blen = min(CB_BURST, n - pos);
                __syncthreads();                 // every thread has read xs[f]
                for (int k = tid; k < blen; k += CB_THREADS) xs[k] = col[pos + k];
                __syncthreads();
                if (tid == 0) {
                    // the dependent adds in order, each batch of 16 addends read from LDS while the previous
                    // one is added (reads past blen stay inside xs, CB_LDS >= CB_BURST + 32, and are not added);
                    // the scheduling barriers keep a batch's reads ahead of the other batch's adds, so that no
                    // add waits for LDS (left to itself the compiler moves each read next to its add)
                    double t = s;
                    double a[16], b[16];
#pragma unroll
                    for (int q = 0; q < 16; ++q) a[q] = xs[q];
                    int k = 0;
                    for (; k + 32 <= blen; k += 32) {
#pragma unroll
                        for (int q = 0; q < 16; ++q) b[q] = xs[k + 16 + q];
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int q = 0; q < 16; ++q) t = t + a[q];
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int q = 0; q < 16; ++q) a[q] = xs[k + 32 + q];
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int q = 0; q < 16; ++q) t = t + b[q];
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    if (k + 16 <= blen) {
#pragma unroll
                        for (int q = 0; q < 16; ++q) t = t + a[q];
                        k += 16;
                    }
                    for (; k < blen; ++k) t = t + xs[k];
                    s_next = t;
                }
                __syncthreads();
                s = s_next;
                pos += blen;
            }
        }
        __syncthreads();   // xs, s_next and first_bad are rewritten by the next block
    }
    return s;
}

__global__ __launch_bounds__(CB_THREADS) void k_centroid(const double* __restrict__ X, int64_t ldx, int n,
                                                         double* __restrict__ root) {
    __shared__ CentroidLds L;
    const double s = centroid_range(X + (int64_t)blockIdx.x * ldx, 0, n, 0.0, L);
    if (threadIdx.x == 0) root[blockIdx.x] = s / (double)n;
}

// Long inputs: the same sum with most of it in parallel.  The addends are cut into chunks of
// CS_CHUNK; a first pass sums each chunk in any order (k_centroid_csum), so that the running sum at
// the start of chunk c is known approximately, P_c.  A second pass (k_centroid_spec) forms the
// chunk's approximate running sums P_k and marks its special addends: those after which P_k is in
// another binade than before, and those that would be a tie or too large in the unit u = 2^{e-52} of
// their binade e.  Between the special addends (at most CS_MAXSP) lie runs of addends that keep one
// binade; for each run it forms, exactly in int64, the partial sums J_k of rint(x / u), with their
// least and greatest values.  The chain (k_centroid_chain, one workgroup per axis) walks the chunks
// with the exact running sum s: when s lies in a run's binade, s = S0 u, and S0 + J_k stays strictly
// inside (2^52, 2^53) in magnitude for every k, every addition of the run was exact in the sense of
// the block method above and the run ends at (S0 + J_last) u; each special addend is added in FP64 as
// the reference does.  Whatever fails a check -- the first chunk, chunks with more special addends, a
// wrong guess -- is summed by the block method from s.  The result is the sequential sum bit for bit.
constexpr int CS_CHUNK = 8192;
constexpr int CS_T = 256;
constexpr int CS_PER = CS_CHUNK / CS_T;   // consecutive addends per thread in the second pass
constexpr int CS_LDS_CHUNKS = 256;
constexpr int CS_MAXSP = 4;               // special addends per chunk (see k_centroid_spec)
constexpr int CS_PARTS = CS_MAXSP + 1;
constexpr long long CS_LIM = 1ll << 54;   // |J| beyond this cannot stay inside a binade
struct CentroidPart {
    long long total, jmin, jmax;   // J over the part (jmin > jmax: empty part)
    int e;
};
struct CentroidChunk {
    CentroidPart part[CS_PARTS];   // the runs between the special addends
    double xk[CS_MAXSP];           // the special addends
    int kb[CS_MAXSP];              // their positions in the chunk, ascending
    int nsp;
    int bad;
};

__global__ __launch_bounds__(CS_T) void k_centroid_csum(const double* __restrict__ X, int64_t ldx, int n, int nch,
                                                         double* __restrict__ csum) {
    __shared__ double red[CS_T / 64];
    const double* __restrict__ col = X + (int64_t)blockIdx.y * ldx;
    const int c = blockIdx.x, tid = threadIdx.x;
    double x[CS_PER];
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {   // all loads in flight together
        const int k = c * CS_CHUNK + q * CS_T + tid;
        x[q] = k < n ? col[k] : 0.0;
    }
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) v += x[q];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) csum[blockIdx.y * nch + c] = (red[0] + red[1]) + (red[2] + red[3]);
}

__device__ __forceinline__ int binade_of(double p) {   // ilogb of a normal finite value, else a sentinel
    return (fabs(p) >= 0x1p-1022 && isfinite(p)) ? ilogb(p) : -100000;
}

// Block-wide sums and scans over CS_T threads (4 wavefronts).
template <class V>
__device__ __forceinline__ V cs_exscan(V v, V& total, V* wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    V incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const V y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    __syncthreads();
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    V off = incl - v;
    total = V(0);
    for (int k = 0; k < CS_T / 64; ++k) {
        if (k < w) off += wsum[k];
        total += wsum[k];
    }
    return off;
}
template <bool MAX>
__device__ __forceinline__ long long cs_minmax(long long v, long long* wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const long long y = __shfl_xor(v, d);
        v = MAX ? max(v, y) : min(v, y);
    }
    __syncthreads();
    if (lane == 0) wsum[w] = v;
    __syncthreads();
    long long r = wsum[0];
    for (int k = 1; k < CS_T / 64; ++k) r = MAX ? max(r, wsum[k]) : min(r, wsum[k]);
    return r;
}

// Block-wide exclusive scans and min / max of CS_PARTS values per thread (one pair of barriers each).
__device__ __forceinline__ void cs_exscan_parts(const long long (&v)[CS_PARTS], long long (&off)[CS_PARTS],
                                                long long (&tot)[CS_PARTS], long long* wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long incl[CS_PARTS];
#pragma unroll
    for (int p = 0; p < CS_PARTS; ++p) {
        incl[p] = v[p];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const long long y = __shfl_up(incl[p], d);
            if (lane >= d) incl[p] += y;
        }
    }
    __syncthreads();
    if (lane == 63)
#pragma unroll
        for (int p = 0; p < CS_PARTS; ++p) wsum[w * CS_PARTS + p] = incl[p];
    __syncthreads();
#pragma unroll
    for (int p = 0; p < CS_PARTS; ++p) {
        off[p] = incl[p] - v[p];
        tot[p] = 0;
        for (int k = 0; k < CS_T / 64; ++k) {
            if (k < w) off[p] += wsum[k * CS_PARTS + p];
            tot[p] += wsum[k * CS_PARTS + p];
        }
    }
}
__device__ __forceinline__ void cs_minmax_parts(long long (&mn)[CS_PARTS], long long (&mx)[CS_PARTS],
                                                long long* wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int p = 0; p < CS_PARTS; ++p) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            mn[p] = min(mn[p], (long long)__shfl_xor(mn[p], d));
            mx[p] = max(mx[p], (long long)__shfl_xor(mx[p], d));
        }
    }
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int p = 0; p < CS_PARTS; ++p) {
            wsum[w * 2 * CS_PARTS + p] = mn[p];
            wsum[w * 2 * CS_PARTS + CS_PARTS + p] = mx[p];
        }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < CS_PARTS; ++p) {
        mn[p] = wsum[p];
        mx[p] = wsum[CS_PARTS + p];
        for (int k = 1; k < CS_T / 64; ++k) {
            mn[p] = min(mn[p], wsum[k * 2 * CS_PARTS + p]);
            mx[p] = max(mx[p], wsum[k * 2 * CS_PARTS + CS_PARTS + p]);
        }
    }
}

__device__ __forceinline__ bool cs_unit_ok(double x, int e) {   // neither too large nor a tie in unit 2^{e-52}
    const double y = ldexp(x, 52 - e);
    return fabs(y) < 0x1p53 && fabs(y - rint(y)) != 0.5;
}

__global__ __launch_bounds__(CS_T) void k_centroid_spec(const double* __restrict__ X, int64_t ldx, int n, int nch,
                                                         const double* __restrict__ csum,
                                                         CentroidChunk* __restrict__ spec) {
    constexpr int LD = CS_PER + 1;   // padded rows: thread t reads its CS_PER addends without conflicts
    __shared__ double xs[CS_T * LD];
    __shared__ double wd[CS_T / 64];
    __shared__ int wi[CS_T / 64];
    __shared__ long long wl[CS_T / 64 * 2 * CS_PARTS];
    __shared__ int lastb[CS_T];
    __shared__ int s_kb[CS_MAXSP + 1], s_eb[CS_MAXSP + 1];
    __shared__ double s_xk[CS_MAXSP + 1];
    const double* __restrict__ col = X + (int64_t)blockIdx.y * ldx;
    const int c = blockIdx.x, tid = threadIdx.x;
    CentroidChunk* out = spec + blockIdx.y * nch + c;
    const int base = c * CS_CHUNK, len = min(CS_CHUNK, n - base);
    // the chunk in LDS (coalesced loads, all in flight together)
    double x[CS_PER];
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {
        const int k = q * CS_T + tid;
        x[q] = k < len ? col[base + k] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {
        const int k = q * CS_T + tid;
        xs[(k / CS_PER) * LD + k % CS_PER] = x[q];
    }
    // approximate running sum at the chunk start
    double ps = 0.0;
    for (int k = tid; k < c; k += CS_T) ps += csum[blockIdx.y * nch + k];
    double pst;
    cs_exscan(ps, pst, wd);
    // approximate running sums inside the chunk: this thread's addends k0 .. k0 + CS_PER - 1
    const int k0 = tid * CS_PER;
    double loc = 0.0;
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {
        x[q] = xs[tid * LD + q];
        loc += x[q];
    }
    double dummy;
    const double pre = pst + cs_exscan(loc, dummy, wd);
    const int e0 = binade_of(pst);
    // binades after each addend; a thread's first addend compares with the previous thread's last
    int bq[CS_PER];
    {
        double p = pre;
#pragma unroll
        for (int q = 0; q < CS_PER; ++q) {
            p += x[q];
            bq[q] = binade_of(p);
        }
    }
    lastb[tid] = bq[CS_PER - 1];
    __syncthreads();
    int nsp = 0, sentinel = 0;
    {
        int bprev = tid == 0 ? e0 : lastb[tid - 1];
#pragma unroll
        for (int q = 0; q < CS_PER; ++q) {
            if (k0 + q < len) {
                sentinel |= bq[q] == -100000;
                nsp += (bq[q] != bprev || !cs_unit_ok(x[q], bq[q])) ? 1 : 0;
            }
            bprev = bq[q];
        }
    }
    int all_sp;
    const int sp_off = cs_exscan(nsp, all_sp, wi);
    int anysent;
    cs_exscan(sentinel, anysent, wi);
    const bool bad0 = c == 0 || e0 == -100000 || anysent != 0 || all_sp > CS_MAXSP;
    if (bad0) {
        if (tid == 0) out->bad = 1;
        return;   // uniform across the workgroup
    }
    // the special addends' positions, values and the binades after them
    if (tid == 0) {
        s_eb[0] = e0;
        for (int i = 0; i <= CS_MAXSP; ++i) s_kb[i] = INT_MAX;
    }
    __syncthreads();
    {
        int bprev = tid == 0 ? e0 : lastb[tid - 1], i = sp_off;
#pragma unroll
        for (int q = 0; q < CS_PER; ++q) {
            if (k0 + q < len && (bq[q] != bprev || !cs_unit_ok(x[q], bq[q]))) {
                s_kb[i] = k0 + q;
                s_xk[i] = x[q];
                s_eb[i + 1] = bq[q];
                ++i;
            }
            bprev = bq[q];
        }
    }
    __syncthreads();
    // exact increments per run (the run of an addend: the number of special addends before it)
    long long T[CS_PARTS], mn[CS_PARTS], mx[CS_PARTS];
#pragma unroll
    for (int p = 0; p < CS_PARTS; ++p) {
        T[p] = 0;
        mn[p] = LLONG_MAX;
        mx[p] = LLONG_MIN;
    }
    int bad = 0;
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {
        const int k = k0 + q;
        if (k >= len) continue;
        int part = 0;
        bool special = false;
#pragma unroll
        for (int i = 0; i < CS_MAXSP; ++i) {
            part += k > s_kb[i] ? 1 : 0;
            special |= k == s_kb[i];
        }
        if (special) continue;
        const long long r = (long long)rint(ldexp(x[q], 52 - s_eb[part]));
#pragma unroll
        for (int p = 0; p < CS_PARTS; ++p) {
            if (p == part) {
                T[p] += r;
                bad |= T[p] > CS_LIM || T[p] < -CS_LIM;
                mn[p] = min(mn[p], T[p]);
                mx[p] = max(mx[p], T[p]);
            }
        }
    }
    long long off[CS_PARTS], tot[CS_PARTS];
    cs_exscan_parts(T, off, tot, wl);
#pragma unroll
    for (int p = 0; p < CS_PARTS; ++p) {
        if (mn[p] != LLONG_MAX) mn[p] += off[p];
        if (mx[p] != LLONG_MIN) mx[p] += off[p];
    }
    cs_minmax_parts(mn, mx, wl);
    int anybad;
    cs_exscan(bad, anybad, wi);
    if (tid == 0) {
        CentroidChunk r;
#pragma unroll
        for (int p = 0; p < CS_PARTS; ++p) r.part[p] = CentroidPart{tot[p], mn[p], mx[p], p <= all_sp ? s_eb[p] : 0};
#pragma unroll
        for (int i = 0; i < CS_MAXSP; ++i) {
            r.kb[i] = i < all_sp ? s_kb[i] : -1;
            r.xk[i] = i < all_sp ? s_xk[i] : 0.0;
        }
        r.nsp = all_sp;
        r.bad = anybad ? 1 : 0;
        *out = r;
    }
}

// One part in its binade, or false: s = S0 u exactly and S0 + J_k strictly inside the binade.
__device__ __forceinline__ bool centroid_part(double& s, const CentroidPart& p) {
    if (p.jmin > p.jmax) return true;   // empty part
    if (!(fabs(s) >= 0x1p-1022) || !isfinite(s) || ilogb(s) != p.e) return false;
    const long long S0 = (long long)ldexp(s, 52 - p.e);
    const long long lo = S0 + p.jmin, hi = S0 + p.jmax;
    const bool ok = s > 0.0 ? (lo > (1ll << 52) && hi < (1ll << 53)) : (hi < -(1ll << 52) && lo > -(1ll << 53));
    if (ok) s = ldexp((double)(S0 + p.total), p.e - 52);
    return ok;
}

__global__ __launch_bounds__(CB_THREADS) void k_centroid_chain(const double* __restrict__ X, int64_t ldx, int n,
                                                                int nch, const CentroidChunk* __restrict__ spec,
                                                                double* __restrict__ root) {
#pragma clang fp contract(off)
    __shared__ CentroidLds L;
    __shared__ CentroidChunk sspec[CS_LDS_CHUNKS];   // the first chunks' records, loaded up front
    const double* __restrict__ col = X + (int64_t)blockIdx.x * ldx;
    for (int c = threadIdx.x; c < min(nch, CS_LDS_CHUNKS); c += CB_THREADS) sspec[c] = spec[blockIdx.x * nch + c];
    __syncthreads();
    double s = 0.0;
    for (int c = 0; c < nch; ++c) {
        const CentroidChunk* sp = c < CS_LDS_CHUNKS ? &sspec[c] : &spec[blockIdx.x * nch + c];
        const int base = c * CS_CHUNK, end = min(n, base + CS_CHUNK);
        int from = base;   // where the block method takes over, end: not needed
        if (!sp->bad) {
            from = end;
            const int nsp = sp->nsp;
            for (int p = 0; p <= nsp; ++p) {
                if (!centroid_part(s, sp->part[p])) {
                    from = base + (p == 0 ? 0 : sp->kb[p - 1] + 1);
                    break;
                }
                if (p < nsp) s = s + sp->xk[p];
            }
        }
        if (from < end) s = centroid_range(col, from, end, s, L);
    }
    if (threadIdx.x == 0) root[blockIdx.x] = s / (double)n;
}

// max |x - c| per axis (exact, so any reduction order is fine).
__global__ __launch_bounds__(256) void k_maxabs(const double* __restrict__ X, int64_t ldx, int n,
                                                const double* __restrict__ root,
                                                unsigned long long* __restrict__ maxabs) {
    __shared__ double red[3][256];
    double m[3] = {0.0, 0.0, 0.0};
    const double c[3] = {root[0], root[1], root[2]};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) m[k] = fmax(m[k], fabs(X[(int64_t)k * ldx + i] - c[k]));
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = m[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 3) atomicMax(&maxabs[threadIdx.x], (unsigned long long)__double_as_longlong(red[threadIdx.x][0]));
}

__device__ __forceinline__ double root_halfwidth(const unsigned long long* maxabs) {
#pragma clang fp contract(off)
    double mx = __longlong_as_double((long long)maxabs[0]);
    double my = __longlong_as_double((long long)maxabs[1]);
    double mz = __longlong_as_double((long long)maxabs[2]);
    return fmax(mx, fmax(my, mz)) * 1.0001;
}

__device__ __forceinline__ double rmax_of(double r) {
#pragma clang fp contract(off)
    double t = 1.5 * r;
    return sqrt(t * r);
}

__global__ __launch_bounds__(256) void k_keys(const double* __restrict__ X, int64_t ldx, int n,
                                              double* __restrict__ root, const unsigned long long* __restrict__ maxabs,
                                              unsigned long long* __restrict__ keys, int* __restrict__ idx) {
#pragma clang fp contract(off)
    const double r0 = root_halfwidth(maxabs);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { root[3] = r0; root[4] = rmax_of(r0); }
    if (i >= n) return;
    const double x = X[i], y = X[ldx + i], z = X[2 * ldx + i];
    double cx = root[0], cy = root[1], cz = root[2], r = r0;
    unsigned long long key = 0ull;
    for (int d = 0; d < KEY_LEVELS; ++d) {
        const int o = (x > cx ? 1 : 0) | (y > cy ? 2 : 0) | (z > cz ? 4 : 0);
        key = (key << 3) | (unsigned long long)o;
        r = r * 0.5;
        cx = (o & 1) ? cx + r : cx - r;
        cy = (o & 2) ? cy + r : cy - r;
        cz = (o & 4) ? cz + r : cz - r;
    }
    keys[i] = key;
    idx[i] = i;
}

__global__ __launch_bounds__(256) void k_cpl(const unsigned long long* __restrict__ keys, int n, int ncrit,
                                             int* __restrict__ lmin, int* __restrict__ w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    lmin[k] = (k == 0) ? 1 : common_digits(keys[k - 1], keys[k]) + 2;
    if (k + ncrit - 1 < n) w[k] = common_digits(keys[k], keys[k + ncrit - 1]);
}

__global__ __launch_bounds__(256) void k_lambda(int n, int ncrit, const int* __restrict__ lmin,
                                                const int* __restrict__ w, int* __restrict__ lam,
                                                int* __restrict__ cnt, int* __restrict__ info) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j0 = max(0, k - ncrit + 1), j1 = min(k, n - ncrit);
    int W = 0;
    for (int j = j0; j <= j1; ++j) W = max(W, w[j]);
    if (W >= KEY_LEVELS) atomicOr(&info[1], 1);   // >= ncrit sites share a 21-level cell
    const int l = W + 2;
    lam[k] = l;
    cnt[k] = max(0, l - lmin[k] + 1);
}

// first index in [lo, hi) whose prefix (key >> shift) exceeds p; keys sorted.
__device__ __forceinline__ int upper_prefix(const unsigned long long* keys, int lo, int hi, int shift,
                                            unsigned long long p) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((keys[mid] >> shift) > p) hi = mid; else lo = mid + 1;
    }
    return lo;
}
// first index in [lo, hi) whose prefix is >= p.
__device__ __forceinline__ int lower_prefix(const unsigned long long* keys, int lo, int hi, int shift,
                                            unsigned long long p) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((keys[mid] >> shift) >= p) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_emit_nodes(int n, TreeDev t) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = t.cnt[k];
    if (c <= 0) return;
    const unsigned long long key = t.keys[k];
    const int l0 = t.lmin[k], l1 = t.lam[k];
    const int id0 = t.off[k];
    double cx = t.root[0], cy = t.root[1], cz = t.root[2], r = t.root[3];
    int lev = 1;
    for (int L = l0; L <= l1; ++L) {
        while (lev < L) {            // walk the centre chain down to level L
            const int o = key_digit(key, lev);
            r = r * 0.5;
            cx = (o & 1) ? cx + r : cx - r;
            cy = (o & 2) ? cy + r : cy - r;
            cz = (o & 4) ? cz + r : cz - r;
            ++lev;
        }
        const int id = id0 + (L - l0);
        t.level[id] = L;
        t.start[id] = k;
        const int q = L - 1;                    // prefix digits shared by the node's sites
        int end;
        if (q == 0) end = n;
        else {
            const int sh = 63 - 3 * q;
            end = upper_prefix(t.keys, k + 1, n, sh, key >> sh);
        }
        t.count[id] = end - k;
        t.is_leaf[id] = (L == l1) ? 1 : 0;
        t.cx[id] = cx; t.cy[id] = cy; t.cz[id] = cz;
        t.r[id] = r;
        t.rmax[id] = rmax_of(r);
        if (L == 1) {
            t.parent[id] = -1;
            t.octant[id] = -1;
        } else {
            t.octant[id] = key_digit(key, L - 1);
            if (L - 1 >= l0) t.parent[id] = id - 1;
            else {
                const int qp = L - 2;
                int ps = 0;
                if (qp > 0) {
                    const int sh = 63 - 3 * qp;
                    ps = lower_prefix(t.keys, 0, k, sh, key >> sh);
                }
                t.parent[id] = t.off[ps] + (L - 1 - t.lmin[ps]);
            }
        }
    }
}

// Child links, and the level counts and maximum level (gathered per workgroup in LDS, then one global
// atomic per level and workgroup: per-node global atomics on a few addresses serialise).
__global__ __launch_bounds__(256) void k_node_links(int num_nodes, TreeDev t) {
    __shared__ int cnt[KEY_LEVELS + 2];
    __shared__ int lmax;
    if (threadIdx.x < KEY_LEVELS + 2) cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) lmax = 0;
    __syncthreads();
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < num_nodes) {
        const int L = t.level[id];
        if (id > 0) t.children[t.parent[id] * 8 + t.octant[id]] = id;
        atomicAdd(&cnt[L], 1);
        atomicMax(&lmax, L);
    }
    __syncthreads();
    if (threadIdx.x < KEY_LEVELS + 2 && cnt[threadIdx.x]) atomicAdd(&t.info[3 + threadIdx.x], cnt[threadIdx.x]);
    if (threadIdx.x == 0) atomicMax(&t.info[2], lmax);
}

__global__ __launch_bounds__(256) void k_leaf_index(int num_nodes, const int* __restrict__ leaf_scan, TreeDev t) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= num_nodes) return;
    if (t.is_leaf[id]) {
        const int li = leaf_scan[id];
        t.leaf_id[id] = li;
        t.leaf_node[li] = id;
    } else {
        t.leaf_id[id] = -1;
    }
}

// Nodes grouped by level (any order within a level): positions from LDS counters per workgroup and one
// global atomic per level and workgroup.
__global__ __launch_bounds__(256) void k_level_scatter(int num_nodes, const int* __restrict__ level_off, TreeDev t) {
    __shared__ int cnt[KEY_LEVELS + 2];
    __shared__ int base[KEY_LEVELS + 2];
    if (threadIdx.x < KEY_LEVELS + 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    int L = 0, pos = 0;
    if (id < num_nodes) {
        L = t.level[id];
        pos = atomicAdd(&cnt[L], 1);
    }
    __syncthreads();
    if (threadIdx.x < KEY_LEVELS + 2 && cnt[threadIdx.x])
        base[threadIdx.x] = atomicAdd(&t.level_fill[threadIdx.x], cnt[threadIdx.x]);
    __syncthreads();
    if (id < num_nodes) t.level_nodes[level_off[L] + base[L] + pos] = id;
}

// Leaf members in ascending site index (the reference's insertion order).
__global__ __launch_bounds__(64) void k_leaf_sort(TreeDev t) {
    extern __shared__ int buf[];
    const int node = t.leaf_node[blockIdx.x];
    const int s0 = t.start[node], c = t.count[node];
    for (int i = threadIdx.x; i < c; i += 64) buf[i] = t.perm[s0 + i];
    __syncthreads();
    for (int i = threadIdx.x; i < c; i += 64) {
        const int v = buf[i];
        int rank = 0;
        for (int j = 0; j < c; ++j) rank += (buf[j] < v) ? 1 : 0;
        t.perm[s0 + rank] = v;
    }
}

__global__ __launch_bounds__(256) void k_gather(SiteArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const int o = a.perm[k];
    a.xs[k] = a.X[o];
    a.ys[k] = a.X[a.ldx + o];
    a.zs[k] = a.X[2 * a.ldx + o];
    // multipole column cc = set * dim + component, the same leading dimension for all sets
    for (int cc = 0; cc < a.dim * a.nsets; ++cc) {
        const double v = a.M[(int64_t)cc * a.ldm + o];
        if (a.ms_raw) a.ms_raw[(int64_t)cc * a.n + k] = v;
        a.ms_scaled[(int64_t)cc * a.n + k] = a.scale[cc % a.dim] * v;
    }
    a.dfs[k] = a.df ? a.df[o] : 0.0;
    for (int q = 0; q < a.nexcl; ++q) a.excl_s[(int64_t)q * a.n + k] = a.E[(int64_t)q * a.lde + o];
}

// flang NORM2 of a 3-vector (see the header comment).
__device__ __forceinline__ double norm2_3(double v0, double v1, double v2) {
#pragma clang fp contract(off)
    const double v[3] = {v0, v1, v2};
    double mx = 0.0, s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double a = fabs(v[k]);
        if (mx == 0.0) {
            mx = a;
        } else if (a > mx) {
            const double t = mx / a;
            const double tsq = t * t;
            s = s * tsq;
            s = s + tsq;
            mx = a;
        } else {
            const double t = a / mx;
            const double tt = t * t;
            s = s + tt;
        }
    }
    const double one_s = 1.0 + s;
    return mx * sqrt(one_s);
}

// One level of the dual traversal of src/fmm.F90:352-385 over a frontier of
// (target, source) node pairs.  Outputs are appended in arbitrary order and
// sorted afterwards; for a fixed target the reference's DFS appends sources
// in ascending pre-order id, which the sort reproduces.  Periodic runs
// traverse from one (root, root) seed per image (src/fmm.F90:337-340): the
// source id carries the image, whose displacement enters the MAC as
// c_i - (c_j + D) (src/fmm.F90:360-362).
// Workgroup-aggregated append to the three output lists: every thread asks for w[c] slots of counter
// c; one atomic per workgroup and counter (the counters sit on separate cache lines), threads get
// consecutive ranges in thread order.  Every thread of the workgroup must call it.
constexpr int TRAVERSE_THREADS = 256;
__device__ __forceinline__ void block_append3(int* const cnt[3], const int w[3], int pos[3]) {
    __shared__ int tot[3][TRAVERSE_THREADS / 64];
    __shared__ int base[3][TRAVERSE_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl[3] = {w[0], w[1], w[2]};
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int v = __shfl_up(incl[c], d, 64);
            if (lane >= d) incl[c] += v;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int c = 0; c < 3; ++c) tot[c][wid] = incl[c];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        int run = 0, b[TRAVERSE_THREADS / 64];
#pragma unroll
        for (int q = 0; q < TRAVERSE_THREADS / 64; ++q) {
            b[q] = run;
            run += tot[c][q];
        }
        const int g = run > 0 ? atomicAdd(cnt[c], run) : 0;
#pragma unroll
        for (int q = 0; q < TRAVERSE_THREADS / 64; ++q) base[c][q] = g + b[q];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) pos[c] = base[c][wid] + incl[c] - w[c];
}

__global__ __launch_bounds__(TRAVERSE_THREADS) void k_traverse(TraverseArgs a) {
#pragma clang fp contract(off)
    // after an overflowing step the count exceeds what was stored: read only
    // the stored part (the overflow flag is already set and the host retries)
    const int n = min(*a.n_in, a.cap_in);
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.n_clear) *a.n_clear = 0;
    const int stride = gridDim.x * TRAVERSE_THREADS;
    // the loop runs over whole workgroups so the workgroup collectives see every thread
    for (int e0 = blockIdx.x * TRAVERSE_THREADS; e0 < n; e0 += stride) {
        const int e = e0 + (int)threadIdx.x;
        int kind = -1;                 // 0 far (M2L), 1 leaf pair (P2P), 2 descend target, 3 descend source
        int i = 0, j = 0, img = 0, nchild = 0;
        if (e < n) {
            const int2 pr = a.in[e];
            i = pr.x; j = src_node(pr.y); img = src_image(pr.y);
            const double* D = a.disp + 3 * img;
            const double d0 = a.cx[i] - (a.cx[j] + D[0]);
            const double d1 = a.cy[i] - (a.cy[j] + D[1]);
            const double d2 = a.cz[i] - (a.cz[j] + D[2]);
            const double r = norm2_3(d0, d1, d2);
            const double lhs = r * a.theta;
            const double rhs = a.rmax[i] + a.rmax[j];
            const bool li = a.is_leaf[i] != 0, lj = a.is_leaf[j] != 0;
            if (lhs > rhs) kind = 0;
            else if (li && lj) kind = 1;
            else if (lj || ((a.rmax[i] > a.rmax[j]) && !li)) kind = 2;
            else kind = 3;
            if (kind >= 2) {
                const int pnode = kind == 2 ? i : j;
                for (int o = 0; o < 8; ++o) nchild += a.children[pnode * 8 + o] >= 0 ? 1 : 0;
            }
        }
        const int jenc = src_encode(img, j);
        const unsigned long long key = ((unsigned long long)i << 32) | (unsigned)jenc;
        int* const cnt[3] = {a.n_m2l, a.n_p2p, a.n_out};
        const int want[3] = {kind == 0 ? 1 : 0, kind == 1 ? 1 : 0, nchild};
        int got[3];
        block_append3(cnt, want, got);
        if (kind == 0) {
            if (got[0] < a.cap_m2l) a.m2l[got[0]] = key;
            else atomicOr(a.overflow, 1);
        }
        if (kind == 1) {
            if (got[1] < a.cap_p2p) a.p2p[got[1]] = key;
            else atomicOr(a.overflow, 2);
        }
        int pos = got[2];
        if (kind >= 2) {
            const int pnode = kind == 2 ? i : j;
            for (int o = 0; o < 8; ++o) {
                const int c = a.children[pnode * 8 + o];
                if (c < 0) continue;
                if (pos < a.cap_out)
                    a.out[pos] = kind == 2 ? make_int2(c, jenc) : make_int2(i, src_encode(img, c));
                else atomicOr(a.overflow, 4);
                ++pos;
            }
        }
    }
}

// CSR offsets by target from sorted (target << 32 | source) keys.
__global__ __launch_bounds__(256) void k_csr(const unsigned long long* __restrict__ keys, int nk, int nt,
                                             int* __restrict__ off, int* __restrict__ src) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (nk == 0) {
        for (int x = e; x <= nt; x += gridDim.x * blockDim.x) off[x] = 0;
        return;
    }
    if (e >= nk) return;
    const int t = (int)(keys[e] >> 32);
    const int tp = e == 0 ? -1 : (int)(keys[e - 1] >> 32);
    for (int x = tp + 1; x <= t; ++x) off[x] = e;
    src[e] = (int)(keys[e] & 0xffffffffull);
    if (e == nk - 1)
        for (int x = t + 1; x <= nt; ++x) off[x] = nk;
}

// CSR of an interaction list from its unsorted (target << 32 | source) keys without a device-wide sort:
// per-target counts, their exclusive scan (the offsets), a scatter into the targets' ranges, and a sort of
// each range by source id in LDS.  The sources of one target are distinct, so each one's position in its
// range is the number of smaller ones (its rank); the result equals the order of the sorted keys.
__global__ __launch_bounds__(256) void k_list_count(const unsigned long long* __restrict__ keys, int nk,
                                                    int* __restrict__ cnt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nk) atomicAdd(&cnt[(int)(keys[e] >> 32)], 1);
}

// cnt counts back down to 0 (ready for the next list)
__global__ __launch_bounds__(256) void k_list_scatter(const unsigned long long* __restrict__ keys, int nk,
                                                      const int* __restrict__ off, int* __restrict__ cnt,
                                                      unsigned* __restrict__ tmp) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nk) return;
    const unsigned long long k = keys[e];
    const int t = (int)(k >> 32);
    tmp[off[t] + atomicSub(&cnt[t], 1) - 1] = (unsigned)(k & 0xffffffffull);
}

constexpr int LIST_SORT_LDS = 2048;   // sources of one target sorted in LDS (longer ranges: in global memory)

__global__ __launch_bounds__(256) void k_list_sort(const unsigned* __restrict__ tmp, const int* __restrict__ off,
                                                   int* __restrict__ src, unsigned long long* __restrict__ keys_out) {
    __shared__ unsigned v[LIST_SORT_LDS];
    const int t = blockIdx.x;
    const int e0 = off[t], k = off[t + 1] - e0;
    if (k == 0) return;
    const bool lds = k <= LIST_SORT_LDS;
    if (lds)
        for (int x = threadIdx.x; x < k; x += blockDim.x) v[x] = tmp[e0 + x];
    __syncthreads();
    for (int x = threadIdx.x; x < k; x += blockDim.x) {
        const unsigned me = lds ? v[x] : tmp[e0 + x];
        int rank = 0;
        if (lds) {
            for (int y = 0; y < k; ++y) rank += v[y] < me ? 1 : 0;
        } else {
            for (int y = 0; y < k; ++y) rank += tmp[e0 + y] < me ? 1 : 0;
        }
        src[e0 + rank] = (int)me;
        if (keys_out) keys_out[e0 + rank] = ((unsigned long long)t << 32) | me;
    }
}

// Exclusive prefix sums of n ints in one workgroup of 1024 threads (each sums a contiguous run, then the
// runs' totals are scanned in LDS): one launch, deterministic, for the small arrays of the set-up.
// Thread t takes the contiguous run [t per, (t + 1) per): up to 16 values are loaded at once into
// registers (all loads in flight together), run totals are scanned with wavefront shuffles and one LDS
// step across the 16 wavefronts, and the run is written from the registers.
constexpr int EXSCAN_REG = 16;
__global__ __launch_bounds__(1024) void k_exscan_small(const int* __restrict__ in, int* __restrict__ out, int n) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + 1023) / 1024;
    const int b0 = min(n, t * per), b1 = min(n, b0 + per);
    int v[EXSCAN_REG];
    int sum = 0;
    if (per <= EXSCAN_REG) {
#pragma unroll
        for (int q = 0; q < EXSCAN_REG; ++q) v[q] = b0 + q < b1 ? in[b0 + q] : 0;
#pragma unroll
        for (int q = 0; q < EXSCAN_REG; ++q) sum += v[q];
    } else {
        for (int i = b0; i < b1; ++i) sum += in[i];
    }
    int x = sum;   // inclusive scan of the run totals in the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int run = x - sum;
    for (int k = 0; k < w; ++k) run += wsum[k];
    if (per <= EXSCAN_REG) {
#pragma unroll
        for (int q = 0; q < EXSCAN_REG; ++q) {
            if (b0 + q < b1) out[b0 + q] = run;
            run += v[q];
        }
    } else {
        for (int i = b0; i < b1; ++i) {
            const int u = in[i];
            out[i] = run;
            run += u;
        }
    }
}

static inline int blocks_for(int n, int b) { return (n + b - 1) / b; }

void exscan_i32_small(const int* in, int* out, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_exscan_small, dim3(1), dim3(1024), 0, s, in, out, n);
    FMMH_HIP(hipGetLastError());
}

void launch_list_csr(const unsigned long long* keys, int n_keys, int num_targets, int* cnt, int* off, unsigned* tmp,
                     void* scan_temp, size_t scan_temp_bytes, int* src, unsigned long long* keys_out,
                     hipStream_t s) {
    // cnt: num_targets + 1 zeros (left zero); off: num_targets + 1
    if (n_keys > 0)
        hipLaunchKernelGGL(k_list_count, dim3(blocks_for(n_keys, 256)), dim3(256), 0, s, keys, n_keys, cnt);
    if (num_targets + 1 <= EXSCAN_SMALL_MAX) exscan_i32_small(cnt, off, num_targets + 1, s);
    else exscan_i32(scan_temp, scan_temp_bytes, cnt, off, num_targets + 1, s);
    if (n_keys > 0) {
        hipLaunchKernelGGL(k_list_scatter, dim3(blocks_for(n_keys, 256)), dim3(256), 0, s, keys, n_keys, off, cnt,
                           tmp);
        hipLaunchKernelGGL(k_list_sort, dim3(num_targets), dim3(256), 0, s, tmp, off, src, keys_out);
    }
    FMMH_HIP(hipGetLastError());
}

size_t centroid_work_bytes(int n) {
    const size_t nch = (size_t)blocks_for(n, CS_CHUNK);
    return 3 * nch * (sizeof(double) + sizeof(CentroidChunk)) + 64;
}

void launch_centroid(const double* X, int64_t ldx, int n, TreeDev& t, void* work, hipStream_t s) {
    const int nch = blocks_for(n, CS_CHUNK);
    if (nch <= 2) {
        hipLaunchKernelGGL(k_centroid, dim3(3), dim3(CB_THREADS), 0, s, X, ldx, n, t.root);
    } else {
        CentroidChunk* spec = reinterpret_cast<CentroidChunk*>(work);
        double* csum = reinterpret_cast<double*>(spec + 3 * nch);
        hipLaunchKernelGGL(k_centroid_csum, dim3(nch, 3), dim3(CS_T), 0, s, X, ldx, n, nch, csum);
        hipLaunchKernelGGL(k_centroid_spec, dim3(nch, 3), dim3(CS_T), 0, s, X, ldx, n, nch, csum, spec);
        hipLaunchKernelGGL(k_centroid_chain, dim3(3), dim3(CB_THREADS), 0, s, X, ldx, n, nch, spec, t.root);
    }
    FMMH_HIP(hipMemsetAsync(t.maxabs, 0, 3 * sizeof(unsigned long long), s));
    int g = std::min(blocks_for(n, 256), 1024);
    hipLaunchKernelGGL(k_maxabs, dim3(g), dim3(256), 0, s, X, ldx, n, t.root, t.maxabs);
    FMMH_HIP(hipGetLastError());
}

void launch_keys(const double* X, int64_t ldx, int n, TreeDev& t, unsigned long long* keys_out, int* idx_out,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, X, ldx, n, t.root, t.maxabs, keys_out,
                       idx_out);
    FMMH_HIP(hipGetLastError());
}

void launch_site_levels(int n, int ncrit, TreeDev& t, hipStream_t s) {
    hipLaunchKernelGGL(k_cpl, dim3(blocks_for(n, 256)), dim3(256), 0, s, t.keys, n, ncrit, t.lmin, t.w);
    hipLaunchKernelGGL(k_lambda, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ncrit, t.lmin, t.w, t.lam, t.cnt,
                       t.info);
    FMMH_HIP(hipGetLastError());
}

void launch_emit_nodes(int n, TreeDev& t, hipStream_t s) {
    hipLaunchKernelGGL(k_emit_nodes, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, t);
    FMMH_HIP(hipGetLastError());
}

void launch_node_links(int num_nodes, TreeDev& t, hipStream_t s) {
    hipLaunchKernelGGL(k_node_links, dim3(blocks_for(num_nodes, 256)), dim3(256), 0, s, num_nodes, t);
    FMMH_HIP(hipGetLastError());
}

void launch_leaf_index(int num_nodes, const int* leaf_scan, TreeDev& t, hipStream_t s) {
    hipLaunchKernelGGL(k_leaf_index, dim3(blocks_for(num_nodes, 256)), dim3(256), 0, s, num_nodes, leaf_scan, t);
    FMMH_HIP(hipGetLastError());
}

void launch_level_scatter(int num_nodes, const int* level_off, TreeDev& t, hipStream_t s) {
    hipLaunchKernelGGL(k_level_scatter, dim3(blocks_for(num_nodes, 256)), dim3(256), 0, s, num_nodes, level_off, t);
    FMMH_HIP(hipGetLastError());
}

void launch_leaf_sort(int num_leaves, int max_leaf, TreeDev& t, hipStream_t s) {
    if (num_leaves <= 0) return;
    hipLaunchKernelGGL(k_leaf_sort, dim3(num_leaves), dim3(64), max_leaf * sizeof(int), s, t);
    FMMH_HIP(hipGetLastError());
}

void launch_gather(const SiteArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(a.n, 256)), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

__global__ void k_traverse_init(int* counters, int stride, int nimg, int2* frontier) {
    const int t = threadIdx.x;
    if (t < 6) counters[t * stride] = t == 0 ? nimg : 0;
    if (t < nimg) frontier[t] = make_int2(0, src_encode(t, 0));
}

__global__ void k_traverse_counts(const int* counters, int stride, int* out) {
    const int t = threadIdx.x;
    if (t < 6) out[t] = counters[t * stride];
}

void launch_traverse_init(int* counters, int stride, int nimg, int2* frontier, hipStream_t s) {
    hipLaunchKernelGGL(k_traverse_init, dim3(1), dim3(64), 0, s, counters, stride, nimg, frontier);
    FMMH_HIP(hipGetLastError());
}

void launch_traverse_counts(const int* counters, int stride, int* out, hipStream_t s) {
    hipLaunchKernelGGL(k_traverse_counts, dim3(1), dim3(64), 0, s, counters, stride, out);
    FMMH_HIP(hipGetLastError());
}

void launch_traverse(const TraverseArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_traverse, dim3(grid), dim3(TRAVERSE_THREADS), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_csr(const unsigned long long* keys, int n_keys, int num_targets, int* off, int* src, hipStream_t s) {
    int g = n_keys == 0 ? blocks_for(num_targets + 1, 256) : blocks_for(n_keys, 256);
    hipLaunchKernelGGL(k_csr, dim3(g), dim3(256), 0, s, keys, n_keys, num_targets, off, src);
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
