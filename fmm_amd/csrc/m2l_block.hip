// Target-block M2L (src/fmm.F90:826-882) for one multipole set at a time:
// FP64, every pair at the call's expansion order p <= M2LB_PMAX.
//
// The class-grouped GEMM applies one Hankel matrix H_c per displacement
// class c = (v, level_i - level_j) to the pairs of that class,
//   L~_i(t) += sum_s T~_c[s + t] M~_j(s)          (m2l.hip, m2l_gemm.hip).
// Here the pairs are first grouped by target block: up to M2LB_BT targets
// that are consecutive in (level, octant, node id) order -- targets with the
// same position inside their parent share most displacement classes -- and
// inside a block by class, in tiles of up to 16 pairs of one class (a target
// has at most one source per class, so the 16 columns of a tile belong to 16
// different targets).  A workgroup walks a contiguous, equal share of all
// tiles and accumulates the block's local expansions in LDS: per tile, the
// class tensor and the 16 scaled source expansions are gathered (one tile
// ahead, into registers), each wavefront runs RT 16-row target tiles of
// v_mfma_f64_16x16x4f64 over the 49 (p = 6) reduced source coefficients and
// adds its 16 x 16 result to the LDS rows of the tile's targets.  Rows and
// source coefficients are the reduced (harmonic) components -- the (p + 1)^2
// with z <= 1, sources folded onto them by the Laplace identity (m2l.hip) --
// instead of the ncomp(p) = 84 packed ones; the local components with
// t_z >= 2 are rebuilt after the finish kernel (launch_le_unfold).  Where a workgroup's
// share leaves a block, the LDS block is written to the block's partial slot
// (block + workgroup), so a block split over several workgroups has one
// partial per workgroup; k_m2lb_finish adds them in workgroup order, scales
// by r_i^{-(|t|+1)} and writes the local expansions.  Sums are in a fixed
// order for a given grid (deterministic); the per-pair slot buffer of the
// class-grouped GEMM (np x nm doubles written and read back) is gone.
//
// The plan (target order, tiles) depends only on the tree, the lists and the
// classes, and is kept with the cached classes.
#include "common.h"
#include "kernels.h"

namespace fmmh {

using v4d = double __attribute__((ext_vector_type(4)));

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------- target order
// Stable counting sort of the targets with pairs by (level, octant) in one workgroup: each wave owns a
// contiguous run of nodes and a column of bucket counters, the counters' scan in (bucket, wave) order
// gives every (bucket, wave) its first position, and equal buckets among a wave's 64 lanes are ranked by
// ballots.  Writes tperm (position -> node), tinv (node -> position or -1), tcount (pairs per position,
// 0 from ntg to nn) and ntg.
constexpr int TO_T = 1024;
constexpr int TO_W = TO_T / 64;
constexpr int TO_NB = 256;                 // 32 levels x 8 octants
constexpr int TO_BITS = 8;
static_assert(TO_NB * TO_W == 4 * TO_T, "the scan gives each thread 4 counters");

__device__ __forceinline__ int target_bucket(const M2LBlockArgs& a, int i) {
    if (a.m2l_off[i + 1] <= a.m2l_off[i]) return -1;
    return min(a.level[i], 31) * 8 + (a.octant[i] & 7);
}

__global__ __launch_bounds__(TO_T) void k_m2lb_order(M2LBlockArgs a) {
    __shared__ int cnt[TO_NB * TO_W];   // [bucket][wave]
    __shared__ int wsum[TO_W];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nn = a.nn;
    for (int i = tid; i < TO_NB * TO_W; i += TO_T) cnt[i] = 0;
    const int per = (nn + TO_W - 1) / TO_W;
    const int c0 = min(nn, w * per), c1 = min(nn, c0 + per);
    __syncthreads();
    for (int i = c0 + lane; i < c1; i += 64) {
        const int bk = target_bucket(a, i);
        if (bk >= 0) atomicAdd(&cnt[bk * TO_W + w], 1);
    }
    __syncthreads();
    int loc[4], tot = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        loc[j] = cnt[4 * tid + j];
        tot += loc[j];
    }
    int incl = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int run = incl - tot;
    for (int v = 0; v < w; ++v) run += wsum[v];
    int all = 0;
    for (int v = 0; v < TO_W; ++v) all += wsum[v];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        cnt[4 * tid + j] = run;
        run += loc[j];
    }
    __syncthreads();
    for (int r = c0; r < c1; r += 64) {
        const int i = r + lane;
        const bool inr = i < c1;
        const int bk = inr ? target_bucket(a, i) : -1;
        const bool valid = bk >= 0;
        unsigned long long same = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < TO_BITS; ++bit) {
            const bool on = ((bk >> bit) & 1) != 0;
            const unsigned long long m = __ballot(valid && on);
            same &= on ? m : ~m;
        }
        const int rank = __popcll(same & ((1ull << lane) - 1ull));
        int base = 0;
        if (valid) base = cnt[bk * TO_W + w];
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) cnt[bk * TO_W + w] = base + __popcll(same);
        __builtin_amdgcn_wave_barrier();
        if (inr) {
            if (valid) {
                const int k = base + rank;
                a.tperm[k] = i;
                a.tcount[k] = a.m2l_off[i + 1] - a.m2l_off[i];
                a.tinv[i] = k;
            } else {
                a.tinv[i] = -1;
            }
        }
    }
    for (int k = all + tid; k <= nn; k += TO_T) a.tcount[k] = 0;
    if (tid == 0) *a.ntg = all;
}

// ---------------------------------------------------------------- tiles
// One workgroup per target block: the classes of the block's pairs as bit masks over the block's targets
// (LDS), then per class its tiles (consecutive records at the block's pair offset, so no tile count is
// needed in advance): record = class, level difference, column count, 0, 16 source nodes (-1: unused
// column), 16 target bytes (the target's index in the block).  Columns are filled in target order (rank
// of the target among the class's targets), so the layout is deterministic.  Error bit 2 of flag: a
// target with two pairs of one class (the masks would lose one).  ntile[b] = the block's tile count.
constexpr int TL_T = 1024;
__global__ __launch_bounds__(TL_T) void k_m2lb_tiles(M2LBlockArgs a) {
    extern __shared__ unsigned long long tl_lds[];
    unsigned long long* mk = tl_lds;                          // [ncls] target masks
    int* coff = reinterpret_cast<int*>(tl_lds + a.ncls);      // [ncls] first pair of the class
    int* toff = coff + a.ncls;                                // [ncls] first tile of the class
    __shared__ int pb[M2LB_BT + 1];
    __shared__ int csr0[M2LB_BT];   // first CSR position of each target
    __shared__ int wsum[2][TL_T / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b = blockIdx.x;
    const int ntg = *a.ntg;
    const int k0 = b * M2LB_BT;
    if (k0 >= ntg) {
        if (tid == 0) a.ntile[b] = 0;
        return;
    }
    const int ntl = min(M2LB_BT, ntg - k0);
    if (tid <= ntl) pb[tid] = a.pbase[k0 + tid];
    if (tid < ntl) csr0[tid] = a.m2l_off[a.tperm[k0 + tid]];
    const int ncls = a.ncls;
    for (int c = tid; c < ncls; c += TL_T) mk[c] = 0ull;
    __syncthreads();
    const int base = pb[0];
    const int npb = pb[ntl] - base;
    auto locate = [&](int f, int& tl, int& e) {   // block pair f -> target index in the block, CSR position
        int lo = 0, hi = ntl;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pb[mid] - base <= f) lo = mid;
            else hi = mid;
        }
        tl = lo;
        e = csr0[lo] + (f - (pb[lo] - base));
    };
    for (int f = tid; f < npb; f += TL_T) {
        int tl, e;
        locate(f, tl, e);
        atomicOr(&mk[a.pair_class[e]], 1ull << tl);
    }
    __syncthreads();
    const int per = (ncls + TL_T - 1) / TL_T;
    const int c0 = min(ncls, tid * per), c1 = min(ncls, c0 + per);
    int np_own = 0, nt_own = 0;
    for (int c = c0; c < c1; ++c) {
        const int n = __popcll(mk[c]);
        np_own += n;
        nt_own += (n + 15) >> 4;
    }
    int ip = np_own, it = nt_own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int yp = __shfl_up(ip, d), yt = __shfl_up(it, d);
        if (lane >= d) {
            ip += yp;
            it += yt;
        }
    }
    if (lane == 63) {
        wsum[0][w] = ip;
        wsum[1][w] = it;
    }
    __syncthreads();
    int rp = ip - np_own, rt = it - nt_own, allp = 0, allt = 0;
    for (int v = 0; v < TL_T / 64; ++v) {
        if (v < w) {
            rp += wsum[0][v];
            rt += wsum[1][v];
        }
        allp += wsum[0][v];
        allt += wsum[1][v];
    }
    for (int c = c0; c < c1; ++c) {
        const int n = __popcll(mk[c]);
        coff[c] = rp;
        toff[c] = rt;
        const int dl = n > 0 ? a.class_dl[c] : 0;
        for (int q = 0; q * 16 < n; ++q) {
            int* rec = a.rec + (int64_t)(base + rt + q) * M2LB_REC;
            const int cq = min(16, n - 16 * q);
            rec[0] = c;
            rec[1] = dl;
            rec[2] = cq;
            rec[3] = 0;
            for (int col = cq; col < 16; ++col) rec[4 + col] = -1;
        }
        rp += n;
        rt += (n + 15) >> 4;
    }
    if (tid == 0) {
        a.ntile[b] = allt;
        if (allp != npb) atomicOr(a.flag, 2);
    }
    __syncthreads();
    for (int f = tid; f < npb; f += TL_T) {
        int tl, e;
        locate(f, tl, e);
        const int c = a.pair_class[e];
        const int rank = __popcll(mk[c] & ((1ull << tl) - 1ull));
        int* rec = a.rec + (int64_t)(base + toff[c] + (rank >> 4)) * M2LB_REC;
        const int col = rank & 15;
        rec[4 + col] = src_node(a.m2l_src[e]);
        reinterpret_cast<unsigned char*>(rec + 20)[col] = (unsigned char)tl;
    }
}

// ---------------------------------------------------------------- run
// Share h of the tiles: [h T / G, (h + 1) T / G).  The share holding global tile x is the largest h with
// floor(h T / G) <= x.  Workgroups are dealt to the XCDs round robin (workgroup g on XCD g mod 8), so
// share h = (g mod 8) G / 8 + g / 8 gives every XCD a contiguous eighth of the tiles: neighbouring
// blocks, whose sources and classes then stay in that XCD's L2 (G is a multiple of 8).
__device__ __forceinline__ int share_of(long long x, long long T, long long G) {
    return (int)(((x + 1) * G + T - 1) / T - 1);
}
__device__ __forceinline__ int share_index(int g, int G) { return (g % NUM_XCD) * (G / NUM_XCD) + g / NUM_XCD; }

template <int P, int RT>
struct M2LBShape {
    static constexpr int NM = rdim(P);              // reduced components (rows and source coefficients)
    static constexpr int NT = ncomp(2 * P);
    static constexpr int MT = (NM + 15) / 16;       // 16-row target tiles
    static constexpr int KS = (NM + 3) / 4;         // 4-deep source steps
    static constexpr int NW = (MT + RT - 1) / RT;   // wavefronts, RT row tiles each
    static constexpr int NTHR = 64 * NW;
    static constexpr int PB = (NM + 3) / 4;         // source coefficients per lane (4 lanes per column)
    static constexpr int PT = (NT + 63) / 64;       // class tensor entries per lane
    static constexpr int BLD = (4 * KS) | 1;        // odd LDS row stride of the source tile (zero pad to 4 KS)
};

// Position in the tile sequence: global tile t of block b, whose tiles are [tb, te) with records from rbase.
struct TileCursor {
    int t, b, tb, te, rbase;
};

__device__ __forceinline__ void cursor_seek(const M2LBRunArgs& a, TileCursor& u, int t) {
    int lo = 0, hi = a.nbm;   // last block b with toff[b] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.toff[mid] <= t) lo = mid;
        else hi = mid;
    }
    u.t = t;
    u.b = lo;
    u.tb = a.toff[lo];
    u.te = a.toff[lo + 1];
    u.rbase = a.pbase[lo * M2LB_BT];
}

__device__ __forceinline__ void cursor_next(const M2LBRunArgs& a, TileCursor& u, int t_end) {
    ++u.t;
    if (u.t >= t_end) return;
    while (u.t >= u.te) {
        ++u.b;
        u.tb = u.te;
        u.te = a.toff[u.b + 1];
        u.rbase = a.pbase[u.b * M2LB_BT];
    }
}

// Each wavefront of a workgroup owns RT row tiles of the block's local expansions and walks the
// workgroup's share of the tiles on its own: it gathers and stages every tile in its own LDS slot (the
// wavefronts of a workgroup read the same sources, mostly from the CU's cache) and never waits for the
// others -- no barriers.  Per tile: the fragments are read from LDS in groups of KG source steps, one
// group ahead of the MFMAs; the next tile is staged once the last group is read (LDS operations of a
// wavefront complete in order); the tile after it is gathered into registers; and the previous tile's
// results are added to the LDS rows while this tile's MFMAs run.  A wavefront issues MFMAs in order
// and the matrix core takes one 16x16x4 FP64 MFMA per 64 cycles, so the other work is interleaved
// with the MFMAs (scheduling groups) rather than placed between them; the padded buffers below keep
// that work free of branches.
template <int P, int RT>
__global__ __launch_bounds__((M2LBShape<P, RT>::NTHR)) void k_m2lb_run(M2LBRunArgs a) {
    using S = M2LBShape<P, RT>;
    constexpr int NM = S::NM, NT = S::NT, MT = S::MT, KS = S::KS, NW = S::NW;
    constexpr int PB = S::PB, PT = S::PT, BLD = S::BLD, BT = M2LB_BT;
    constexpr int LLD = BT + 1;        // odd row stride (4 row groups of a fragment hit other banks);
                                       // column BT takes the results of a tile's unused columns
    constexpr int NR = 16 * RT * NW;   // rows of every wavefront's row tiles (those past NM are dropped)
    constexpr int KG = 4;              // source steps per fragment group (RT + 1 LDS reads each)
    constexpr int NKG = (KS + KG - 1) / KG;
    constexpr int TZ = 64 * PT;        // zero slot of the class tensor (entries past NT take garbage)
    __shared__ double Lacc[NR * LLD];  // [row][target in block]
    __shared__ double Bsh[NW][16 * BLD];
    __shared__ double Tsh[NW][TZ + 1];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c16 = lane & 15, g4 = lane >> 4;
    const long long T = a.toff[a.nbm];
    const long long G = gridDim.x, g = share_index(blockIdx.x, gridDim.x);
    const int t0 = (int)(g * T / G), t1 = (int)((g + 1) * T / G);
    if (t0 >= t1) return;
    double* __restrict__ Bs = Bsh[w];
    double* __restrict__ Ts = Tsh[w];
    const int row0 = 16 * RT * w, row1 = min(NR, 16 * RT * (w + 1));   // this wavefront's rows
    for (int x = lane; x < (row1 - row0) * LLD; x += 64) Lacc[row0 * LLD + x] = 0.0;
    for (int x = lane; x < 16 * BLD; x += 64) Bs[x] = 0.0;
    if (lane == 0) Ts[TZ] = 0.0;
    // fragment indices of this wavefront's row tiles (TZ outside the matrix, and for row tiles past the
    // last: their MFMAs then add zeros), and |t| of the accumulator rows 4 q + l / 16
    int fi[RT][KS];
    int dq[RT][4];
#pragma unroll
    for (int mi = 0; mi < RT; ++mi) {
        const int mt = min(RT * w + mi, MT - 1);
        const bool live = RT * w + mi < MT;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int f = a.fidx[(mt * KS + ks) * 64 + lane];
            fi[mi][ks] = (live && f >= 0) ? f : TZ;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) dq[mi][q] = a.tdeg[mt * 16 + 4 * q + g4];
    }
    // gather roles: column gcol of the source tile, coefficients gsub + 4 i (entries past NM read the
    // next node's data: finite, and met only by zero fragment entries)
    const int gcol = lane >> 2, gsub = lane & 3;
    struct Rec {
        int cls, dl, src, tl;   // class and level difference; source of column gcol; target of column c16
    };
    auto load_rec = [&](const TileCursor& u, Rec& r) {
        const int* p = a.rec + (int64_t)(u.rbase + (min(u.t, t1 - 1) - u.tb)) * M2LB_REC;
        r.cls = p[0];
        r.dl = p[1];
        r.src = max(p[4 + gcol], 0);   // an unused column reads node 0 (its results go to column BT)
        r.tl = p[4 + c16] >= 0 ? (int)reinterpret_cast<const unsigned char*>(p + 20)[c16] : BT;
    };
    double bv[PB], tv[PT];
    auto gather = [&](const Rec& r) {
#pragma unroll
        for (int i = 0; i < PB; ++i) bv[i] = a.CMs[(int64_t)r.src * NM + gsub + 4 * i];
#pragma unroll
        for (int i = 0; i < PT; ++i) tv[i] = a.Tc[(int64_t)r.cls * NT + lane + 64 * i];
    };
    auto stage = [&]() {
#pragma unroll
        for (int i = 0; i < PB; ++i) Bs[gcol * BLD + gsub + 4 * i] = bv[i];
#pragma unroll
        for (int i = 0; i < PT; ++i) Ts[lane + 64 * i] = tv[i];
    };
    // accumulator sets alternate: tile t's results (set X) are added to the block's local expansions
    // while tile t + 1's MFMAs run (or at once where the share leaves the block)
    v4d accs[2][RT];
    int tls[2] = {BT, BT}, dls[2] = {0, 0};
    auto accumulate = [&](int Y) {
        // level difference of the class: rows scaled by 2^{-dl |t|} (dl > 0) or 2^{-dl (|t| + 1)}
        // (exact powers of two), then accumulator register q of lane l -> D[4 q + l / 16][l % 16] added
        // to the tile's target (all loads first, then all stores)
        const int tl = tls[Y], dl = dls[Y];
        double old[RT][4];
#pragma unroll
        for (int mi = 0; mi < RT; ++mi)
#pragma unroll
            for (int q = 0; q < 4; ++q) old[mi][q] = Lacc[((RT * w + mi) * 16 + 4 * q + g4) * LLD + tl];
#pragma unroll
        for (int mi = 0; mi < RT; ++mi)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int sh = dl > 0 ? -dl * dq[mi][q] : -dl * (dq[mi][q] + 1);   // 0 for dl = 0
                Lacc[((RT * w + mi) * 16 + 4 * q + g4) * LLD + tl] = old[mi][q] + ldexp(accs[Y][mi][q], sh);
            }
        tls[Y] = BT;
    };
    // pipeline: tile t in LDS (read as fragments), tile t + 1 gathered in registers (staged once tile t's
    // fragments are read), records of tiles t + 1 and t + 2 in two alternating sets (the loop is
    // unrolled by two, so no register copies wait on loads in flight)
    TileCursor cur, n1, n2, n3;
    cursor_seek(a, cur, t0);
    n1 = cur;
    cursor_next(a, n1, t1);
    n2 = n1;
    cursor_next(a, n2, t1);
    n3 = n2;
    cursor_next(a, n3, t1);
    Rec r0, rA, rB;
    load_rec(cur, r0);
    load_rec(n1, rA);
    load_rec(n2, rB);
    gather(r0);
    stage();
    gather(rA);
    int tl_cur = r0.tl, dl_cur = r0.dl;
    auto step = [&](int t, int X, Rec& rnx, Rec& rga) {
        // rnx: record of tile t + 1 (its target and level difference), reloaded with tile t + 3's;
        // rga: record of tile t + 2 (gathered during this step)
        const int tl_nx = rnx.tl, dl_nx = rnx.dl;
#pragma unroll
        for (int mi = 0; mi < RT; ++mi) accs[X][mi] = v4d{0.0, 0.0, 0.0, 0.0};
        double fa[2][KG][RT], fb[2][KG];
        auto load_group = [&](int j, int slot) {
#pragma unroll
            for (int k = 0; k < KG; ++k) {
                const int ks = j * KG + k;
                if (ks < KS) {
                    fb[slot][k] = Bs[c16 * BLD + 4 * ks + g4];
#pragma unroll
                    for (int mi = 0; mi < RT; ++mi) fa[slot][k][mi] = Ts[fi[mi][ks]];
                }
            }
        };
        load_group(0, 0);
#pragma unroll
        for (int j = 0; j < NKG; ++j) {
            if (j + 1 < NKG) load_group(j + 1, (j + 1) & 1);
            if (j == 0) accumulate(1 - X);   // tile t - 1
            if (j + 2 == NKG || NKG == 1) {  // every fragment of tile t is read: stage t + 1, gather t + 2
                stage();
                gather(rga);
                load_rec(n3, rnx);
            }
#pragma unroll
            for (int k = 0; k < KG; ++k) {
                if (j * KG + k < KS) {
#pragma unroll
                    for (int mi = 0; mi < RT; ++mi)
                        accs[X][mi] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[j & 1][k][mi], fb[j & 1][k],
                                                                           accs[X][mi], 0, 0, 0);
                }
            }
            // one MFMA, then up to 3 LDS and 2 other vector operations, in turn
#pragma unroll
            for (int k = 0; k < KG * RT; ++k) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x080, 3, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
            }
        }
        tls[X] = tl_cur;
        dls[X] = dl_cur;
        if (t + 1 == cur.te || t + 1 == t1) {
            // the share leaves this block: its partial slot is block + share (this wavefront's rows)
            accumulate(X);
            double* out = a.partial + (int64_t)(cur.b + g) * NM * BT;
            for (int x = lane; x < (min(row1, NM) - row0) * BT; x += 64) {
                const int row = row0 + x / BT, k = x % BT;
                out[row * BT + k] = Lacc[row * LLD + k];
                Lacc[row * LLD + k] = 0.0;
            }
        }
        tl_cur = tl_nx;
        dl_cur = dl_nx;
        cur = n1;
        n1 = n2;
        n2 = n3;
        cursor_next(a, n3, t1);
    };
    for (int t = t0; t < t1; t += 2) {
        step(t, 0, rA, rB);
        if (t + 1 < t1) step(t + 1, 1, rB, rA);
    }
}

// L_i(t) = r_i^{-(|t|+1)} x (the block's partials, in share order); one thread per (block, reduced row,
// target), written at the row's packed index of the nmf-long local expansion.
constexpr int FN_T = 256;
template <int NM>
__global__ __launch_bounds__(FN_T) void k_m2lb_finish(M2LBRunArgs a, const double* __restrict__ r,
                                                      double* __restrict__ LE, int nmf) {
    constexpr int BT = M2LB_BT;
    const int64_t x = (int64_t)blockIdx.x * FN_T + threadIdx.x;
    const int b = (int)(x / (NM * BT));
    if (b >= a.nbm) return;
    const int y = (int)(x - (int64_t)b * NM * BT), t = y / BT, k = y - t * BT;
    const int ntg = *a.ntg;
    const int kk = b * BT + k;
    if (kk >= ntg) return;
    const long long T = a.toff[a.nbm], G = a.grid;
    const int gf = share_of(a.toff[b], T, G), gl = share_of(a.toff[b + 1] - 1, T, G);
    double s = 0.0;
    for (int g = gf; g <= gl; ++g)   // (shares that are empty, T < G, wrote nothing)
        if ((g + 1) * T / G > g * T / G) s += a.partial[(int64_t)(b + g) * NM * BT + y];
    const int i = a.tperm[kk];
    const double rinv = 1.0 / r[i];
    double f = rinv;
    for (int d = 0; d < a.tdeg[t]; ++d) f *= rinv;
    LE[(int64_t)i * nmf + a.tfull[t]] = s * f;
}

// Nodes without far pairs: zero local expansions (as the per-pair reduction writes them).
__global__ __launch_bounds__(256) void k_m2lb_zero(const int* __restrict__ tinv, int nn, int nm,
                                                   double* __restrict__ LE) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)nn * nm) return;
    if (tinv[x / nm] < 0) LE[x] = 0.0;
}

void launch_m2lb_order(const M2LBlockArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_m2lb_order, dim3(1), dim3(TO_T), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_m2lb_tiles(const M2LBlockArgs& a, hipStream_t s) {
    const size_t lds = (size_t)a.ncls * (sizeof(unsigned long long) + 2 * sizeof(int));
    hipLaunchKernelGGL(k_m2lb_tiles, dim3(a.nbm + 1), dim3(TL_T), lds, s, a);
    FMMH_HIP(hipGetLastError());
}

template <int P>
static void run_p(const M2LBRunArgs& a, const double* r, double* LE, hipStream_t s) {
    constexpr int MT = (rdim(P) + 15) / 16;
    constexpr int RT = 1;
    using S = M2LBShape<P, RT>;
    hipLaunchKernelGGL((k_m2lb_run<P, RT>), dim3(a.grid), dim3(S::NTHR), 0, s, a);
    hipLaunchKernelGGL((k_m2lb_finish<S::NM>), dim3(nblk((long long)a.nbm * S::NM * M2LB_BT, FN_T)), dim3(FN_T), 0, s,
                       a, r, LE, ncomp(P));
    hipLaunchKernelGGL(k_m2lb_zero, dim3(nblk((long long)a.nn * ncomp(P), 256)), dim3(256), 0, s, a.tinv, a.nn,
                       ncomp(P), LE);
    launch_le_unfold(LE, a.nn, P, s);
}

void launch_m2lb_run(const M2LBRunArgs& a, int p, const double* r, double* LE, hipStream_t s) {
    switch (p) {
        case 0: run_p<0>(a, r, LE, s); break;
        case 1: run_p<1>(a, r, LE, s); break;
        case 2: run_p<2>(a, r, LE, s); break;
        case 3: run_p<3>(a, r, LE, s); break;
        case 4: run_p<4>(a, r, LE, s); break;
        case 5: run_p<5>(a, r, LE, s); break;
        default: run_p<6>(a, r, LE, s); break;
    }
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
