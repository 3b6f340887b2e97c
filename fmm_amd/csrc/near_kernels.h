// Near-field pair kernels (templates), shared by the translation units that
// instantiate them: near.hip (source order <= 2, field order <= 2, every set
// count) and near_wide*.hip (the other orders, one set per workgroup row).
// The FMM P2P phase (src/fmm.F90:623-747) and the all-pairs direct sum
// (src/fmm.F90:749-824) share these kernels.
//
// Work decomposition: one 64-lane wavefront per target block (a leaf of the
// octree, or a run of 64 consecutive sites for the direct sum); see the lane
// layout below.  The tensor recursion lives entirely in VGPRs (tensors.h).
//
// Exclusions follow the reference exactly: an excluded pair is evaluated at
// displacement (1e10, 1e10, 1e10) (src/fmm.F90:714-721, 802-806), which
// leaves the -q/(sqrt(3) 1e10) potential artefact in place.  Periodic image
// sources are displaced as x_i - (x_j + D) and never excluded
// (src/fmm.F90:711-721: exclusions apply to the central cell only).
// They are applied outside the pair loop: the loop evaluates every pair at
// its actual displacement except a site with itself (one integer compare),
// and k_near_excl then adds, per target site and distinct exclusion entry j
// of its near field, the pair at (1e10, 1e10, 1e10) minus the pair at its
// actual displacement (for j = i only the former); a site that does not
// exclude itself gets its self pair at displacement 0, i.e. a non-finite
// field, as in the reference.
#pragma once
#include "common.h"
#include "tensors.h"
#include "kernels.h"

namespace fmmh {


// Lane layout: a target block is taken in passes of nt <= 64 sites, each
// replicated over r = 64 / nt lane groups (group g = lane / nt, target
// lane % nt); group g takes every r-th source of the run (positions g mod r,
// whatever the tile size), so small leaves still fill the
// wavefront, and the pass size is chosen per block so that few lanes idle.
// At the end of a pass the groups' partial fields are summed in group order
// through LDS (deterministic).
// Source staging: the source blocks of a target block's list are one
// concatenated run (windows of NEAR_WIN list entries, whose exclusive prefix
// of source counts is scanned across the wavefront), cut into tiles of 64
// sources regardless of block boundaries.  Each tile is published to LDS
// from registers that were loaded while the previous tile was computed, so
// the global-load latency overlaps the pair arithmetic.
// Multipole sets: workgroup row blockIdx.y takes S consecutive sets; each pair
// tensor is built once and contracted with the S sets' multipoles, so the
// tensor recursion is shared (S = 1 for a single set).
// Tensor orders up to 3 (except ERF, whose erf needs more) with the prefetch
// registers fit 168 VGPRs (3 waves per SIMD) without spilling; the S > 1
// forms and the rest get more.
constexpr int NEAR_WIN = 64;

// Source-tile size for S sets of NM components: the largest power of two <= 64 with S NM TILE
// doubles in at most 8 KB (4 sets of quadrupoles: 16 sites, 3 waves per SIMD instead of 1.5;
// 4 sets of dipoles keep 64, measured faster than 32 there).
constexpr int near_tile(int snm) {
    int t = 64;
    while (t > 8 && snm * t * 8 > 8192) t /= 2;
    return t;
}

// T2 (many sets, closed-form contractions, field order <= 1, used for source order <= 1 at 3 wavefronts per
// SIMD: 32 undamped dipole sets 32.8 -> 30.7 ms): each lane takes two targets of one group, ti and
// ti + nh, so that every source's coordinates and S sets of multipoles read from LDS serve two pairs (the
// 4-set form is bound by those LDS reads as much as by its arithmetic); groups then have nh = ceil(nt / 2)
// lanes, and the sums run in another order than the one-target form's (a call selects it with
// NearArgs::pairs2, so that results do not depend on it otherwise).
template <int M, int F, int DAMP, int S, bool T2 = false>
__global__ __launch_bounds__(64, S == 1 ? ((M + F <= 3 && DAMP != DAMP_ERF) ? 4 : 1)
                                        : (T2 ? 3 : ((M + F <= 3 && DAMP != DAMP_ERF) ? 2 : 1))) void k_near(NearArgs a) {
    constexpr int KT = pair_tensor_order<M, F, DAMP>();
    constexpr int NM = ncomp(M);
    constexpr int NF = ncomp(F);
    static_assert(!T2 || (S > 1 && F <= 1 && has_closed_form<M, F>()), "two targets per lane: closed forms, F <= 1");
    constexpr int TPL = T2 ? 2 : 1;             // targets per lane
    // Source tile: TILE sites, sized so that the staged multipoles take at most 8 KB (S NM TILE
    // doubles), which keeps several workgroups per CU; all 64 lanes prefetch the tile's multipoles
    // cooperatively (NPM registers each, element e = lane + 64 i of the [S][NM][TILE] block).
    constexpr int TILE = near_tile(S * NM);
    constexpr int NPM = (S * NM * TILE + 63) / 64;
    __shared__ double sx[TILE], sy[TILE], sz[TILE], sdf[TILE];
    __shared__ double sm[S][NM][TILE];
    __shared__ double red[NF][64];
    __shared__ int wcum[NEAR_WIN + 1];            // window: exclusive prefix of source counts
    __shared__ int wstart[NEAR_WIN];              // window: first sorted site of each entry
    __shared__ int wimg[NEAR_WIN];                // window: periodic image id of each entry
    __shared__ double etab[DAMP != DAMP_NONE ? 32 : 1];   // 2^(j/32) for exp_table
    const int vb = a.blk[blockIdx.x];
    if (vb < 0) return;                        // an empty slot of the block order
    int b, e0, e1, pbase = -1;                 // target block, its list entries, partial sites
    if (a.vb_node) {
        b = a.vb_node[vb];
        e0 = a.vb_e0[vb];
        e1 = a.vb_e1[vb];
        pbase = a.vb_part[vb];
    } else {
        b = vb;
        e0 = a.list_off[b];
        e1 = a.list_off[b + 1];
    }
    if constexpr (DAMP != DAMP_NONE) {
        if (threadIdx.x < 32) etab[threadIdx.x] = kExp2By32[threadIdx.x];   // read after the window barriers
    }
    const int set0 = blockIdx.y * S;
    const int ns = min(S, a.nsets - set0);   // sets of this row (the rest are zero padding)
    const int lane = threadIdx.x;
    const int t0 = a.tgt_start[b];
    const int tc = a.tgt_count[b];
    const double* __restrict__ ms = a.ms + (int64_t)set0 * NM * a.ldm;
    double* __restrict__ out = a.out + (int64_t)set0 * NF * a.ldo;
    // Targets per pass: c passes of ct = ceil(tc / c) targets, each replicated r = 64 / ct
    // times, with c minimizing c (ceil(64 / r) + 2) -- the pair iterations per 64-source tile
    // plus about two iterations' worth of re-staging the tile (measured: weights 1 and 2 run
    // alike, 0 is 17 % slower) -- so that a block of, say, 40 targets runs as 2 passes of 20
    // (r = 3) instead of one pass with 24 idle lanes.
    int ct = min(tc, 64 * TPL);
    {
        int best = 1 << 30;
        for (int c = (tc + 64 * TPL - 1) / (64 * TPL); c <= 16; ++c) {
            const int t = (tc + c - 1) / c;
            const int rr = 64 / ((t + TPL - 1) / TPL);
            const int cost = c * ((64 + rr - 1) / rr + 2);
            if (cost < best) {
                best = cost;
                ct = t;
            }
        }
    }
    for (int base = 0; base < tc; base += ct) {
        const int nt = min(ct, tc - base);
        const int nh = (nt + TPL - 1) / TPL;     // lanes per group
        const int r = 64 / nh;                   // lane groups
        const int g = lane / nh;                 // this lane's group (== r for idle lanes)
        const int ti = lane - g * nh;            // this lane's target within the block (and ti + nh with T2)
        const bool active = g < r;
        const int k = t0 + base + (active ? ti : 0);
        const double xi = a.xs[k], yi = a.ys[k], zi = a.zs[k];
        const double dfi = DAMP != DAMP_NONE ? a.dfs[k] : 0.0;
        // the second target (T2): a lane without one repeats the first and its sums are not stored
        const int k2 = t0 + base + ((active && ti + nh < nt) ? ti + nh : (active ? ti : 0));
        const double xi2 = T2 ? a.xs[k2] : 0.0, yi2 = T2 ? a.ys[k2] : 0.0, zi2 = T2 ? a.zs[k2] : 0.0;
        const double dfi2 = (T2 && DAMP != DAMP_NONE) ? a.dfs[k2] : 0.0;
        double acc[S][NF], acc2[T2 ? S : 1][NF];
#pragma unroll
        for (int u = 0; u < S; ++u)
#pragma unroll
            for (int t = 0; t < NF; ++t) acc[u][t] = 0.0;
        if constexpr (T2) {
#pragma unroll
            for (int u = 0; u < S; ++u)
#pragma unroll
                for (int t = 0; t < NF; ++t) acc2[u][t] = 0.0;
        }
        for (int w0 = e0; w0 < e1; w0 += NEAR_WIN) {
            const int wn = min(NEAR_WIN, e1 - w0);
            int cnt = 0, st = 0, im = 0;
            if (lane < wn) {
                const int enc = a.list_src[w0 + lane];
                const int sb = src_node(enc);
                im = src_image(enc);
                st = a.src_start[sb];
                cnt = a.src_count[sb];
            }
            int incl = cnt;                      // inclusive scan across the wavefront
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(incl, d);
                if (lane >= d) incl += y;
            }
            __syncthreads();                     // readers of the previous window are done
            if (lane < wn) {
                wcum[lane] = incl - cnt;
                wstart[lane] = st;
                wimg[lane] = im;
            }
            if (lane == wn - 1) wcum[wn] = incl;
            __syncthreads();
            const int total = wcum[wn];
            // prefetch registers for tile position v = vbase + lane (lane < TILE) and the tile's
            // multipoles (all lanes)
            double px = 0.0, py = 0.0, pz = 0.0, pdf = 0.0;
            double pm[NPM];
            int cur = 0;                         // list entry holding the tile's first position
            auto fetch = [&](int vbase) {
                while (wcum[cur + 1] <= vbase) ++cur;
                const int v = vbase + lane;
                int pj = -1;
                if (lane < TILE && v < total) {
                    int e = cur;
                    while (wcum[e + 1] <= v) ++e;
                    const int j = wstart[e] + (v - wcum[e]);
                    const int img = wimg[e];
                    const double* D = a.disp + 3 * img;
                    pj = j;
                    px = a.xs[j] + D[0];
                    py = a.ys[j] + D[1];
                    pz = a.zs[j] + D[2];
                    if constexpr (DAMP != DAMP_NONE) pdf = a.dfs[j];
                }
#pragma unroll
                for (int i = 0; i < NPM; ++i) {
                    const int el = lane + 64 * i;           // [u][s][t] of the tile's multipole block
                    const int t = el % TILE, us = el / TILE;
                    const int jt = __shfl(pj, t);           // source of tile position t (-1: none)
                    const int u = us / NM, sc = us - u * NM;
                    pm[i] = (el < S * NM * TILE && u < ns && jt >= 0) ? ms[((int64_t)u * NM + sc) * a.ldm + jt] : 0.0;
                }
            };
            if (total > 0) fetch(0);
            for (int vbase = 0; vbase < total; vbase += TILE) {
                const int cn = min(TILE, total - vbase);
                if (lane < cn) {
                    sx[lane] = px;
                    sy[lane] = py;
                    sz[lane] = pz;
                    if constexpr (DAMP != DAMP_NONE) sdf[lane] = pdf;
                }
#pragma unroll
                for (int i = 0; i < NPM; ++i) {
                    const int el = lane + 64 * i;
                    if (el < S * NM * TILE) (&sm[0][0][0])[el] = pm[i];
                }
                __syncthreads();
                if (vbase + TILE < total) fetch(vbase + TILE);
                if (active) {
                    // group g takes the run positions v = g mod r, in increasing order: the sums do not
                    // depend on the tile size (single- and many-set forms add in the same order)
                    for (int jj = ((g - vbase) % r + r) % r; jj < cn; jj += r) {
                        const double dx = xi - sx[jj];
                        const double dy = yi - sy[jj];
                        const double dz = zi - sz[jj];
                        const double R2 = dx * dx + dy * dy + dz * dz;
                        // a pair at R = 0 (the self pair, or another site at the same position) is
                        // left to k_near_excl: 1/R = 0 makes its terms exactly 0
                        const double rinv = rsqrt_pos(R2);
                        const double inv = R2 > 0.0 ? rinv : 0.0;
                        double aij = 0.0;
                        if constexpr (DAMP != DAMP_NONE) aij = dfi * sdf[jj];
                        double gr[KT + 1];
                        radial_inv<KT, DAMP, DAMP != DAMP_NONE>(R2, inv, aij, gr, etab);
                        if constexpr (T2) {
                            const double ex = xi2 - sx[jj];
                            const double ey = yi2 - sy[jj];
                            const double ez = zi2 - sz[jj];
                            const double Q2 = ex * ex + ey * ey + ez * ez;
                            const double qinv = rsqrt_pos(Q2);
                            const double inv2 = Q2 > 0.0 ? qinv : 0.0;
                            double aij2 = 0.0;
                            if constexpr (DAMP != DAMP_NONE) aij2 = dfi2 * sdf[jj];
                            double gr2[KT + 1];
                            radial_inv<KT, DAMP, DAMP != DAMP_NONE>(Q2, inv2, aij2, gr2, etab);
#pragma unroll
                            for (int u = 0; u < S; ++u) {
                                double m[NM];
#pragma unroll
                                for (int s = 0; s < NM; ++s) m[s] = sm[u][s][jj];
                                contract_closed<M, F>(dx, dy, dz, gr, m, acc[u]);
                                contract_closed<M, F>(ex, ey, ez, gr2, m, acc2[u]);
                            }
                        } else if constexpr (has_closed_form<M, F>()) {
#pragma unroll
                            for (int u = 0; u < S; ++u) {
                                double m[NM];
#pragma unroll
                                for (int s = 0; s < NM; ++s) m[s] = sm[u][s][jj];
                                contract_closed<M, F>(dx, dy, dz, gr, m, acc[u]);
                            }
                        } else {
                            double T[ncomp(KT)];
                            md_tensor<KT>(dx, dy, dz, gr, T);
#pragma unroll
                            for (int u = 0; u < S; ++u) {
                                double m[NM];
#pragma unroll
                                for (int s = 0; s < NM; ++s) m[s] = sm[u][s][jj];
                                contract<M, F, KT>(m, T, acc[u]);
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        // sum the r partial fields of each target in group order, set by set (T2: the first targets, then
        // the second ones)
#pragma unroll
        for (int h = 0; h < TPL; ++h) {
#pragma unroll
            for (int u = 0; u < S; ++u) {
#pragma unroll
                for (int t = 0; t < NF; ++t) red[t][lane] = (T2 && h == 1) ? acc2[T2 ? u : 0][t] : acc[u][t];
                __syncthreads();
                const int tt = lane + h * nh;        // target of this sum within the pass
                if (lane < nh && tt < nt && u < ns) {
                    const int kk = t0 + base + tt;
#pragma unroll
                    for (int t = 0; t < NF; ++t) {
                        double v = red[t][lane];
                        for (int q = 1; q < r; ++q) v += red[t][lane + q * nh];
                        if (pbase >= 0) {
                            a.part[((int64_t)(set0 + u) * NF + t) * a.ldp + pbase + base + tt] = v;
                        } else {
                            double* o = out + ((int64_t)u * NF + t) * a.ldo + kk;
                            if (a.accumulate) *o += v;
                            else *o = v;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------- runtime-order pair field
// The wide orders' exclusion corrections (a few pairs per site) use one kernel per damping kind with the
// orders as arguments: the same radial derivatives and Hermite recursion as the register form
// (tensors.h), with the auxiliaries in per-lane memory.  Results agree with the register form to
// rounding.
static __constant__ const PolyTable kPolyTab[4] = {make_poly(DAMP_NONE), make_poly(DAMP_THOLE),
                                                   make_poly(DAMP_AMOEBA), make_poly(DAMP_ERF)};

template <int KIND>
__device__ __forceinline__ double poly_rt(int n, double x) {
    const PolyTable& t = kPolyTab[KIND];
    const int d = t.deg[n];
    double s = t.c[n][d];
    for (int i = d - 1; i >= 0; --i) s = __builtin_fma(s, x, t.c[n][i]);
    return s;
}

// g[0..K] as radial_inv<K, DAMP, true> computes them (K <= 6 for the damped kinds).
template <int DAMP>
__device__ void radial_rt(int K, double R2, double inv, double a, double* g, const double* etab) {
    const double inv2 = inv * inv;
    double b = inv, arg = 0.0, e = 0.0, er = 0.0;
    if constexpr (DAMP == DAMP_THOLE) {
        arg = a * (R2 * inv);
        e = exp_table(-arg, etab);
    } else if constexpr (DAMP == DAMP_AMOEBA) {
        arg = a * a * a * (R2 * (R2 * inv));
        e = exp_table(-arg, etab);
    } else if constexpr (DAMP == DAMP_ERF) {
        arg = sqrt(a) * (R2 * inv);
        e = 1.1283791670955126 * exp_table(-arg * arg, etab);   // 2/sqrt(pi) e^{-w^2}
        er = erf(arg);
    }
    for (int n = 0; n <= K; ++n) {
        if (n > 0) b *= -(2.0 * n - 1.0) * inv2;
        if constexpr (DAMP == DAMP_NONE) g[n] = b;
        else if constexpr (DAMP == DAMP_THOLE) g[n] = b * (1.0 - e * poly_rt<DAMP_THOLE>(n, arg));
        else if constexpr (DAMP == DAMP_AMOEBA) g[n] = n == 0 ? 0.0 : b * (1.0 - e * poly_rt<DAMP_AMOEBA>(n, arg));
        else g[n] = b * (er - e * poly_rt<DAMP_ERF>(n, arg));
    }
}

constexpr int NEAR_RT_AUX = (NEAR_KMAX + 1) * (NEAR_KMAX + 2) * (NEAR_KMAX + 3) * (NEAR_KMAX + 4) / 24;

// acc[t] += sum_s m[s] T_{s+t}, |s| <= M, |t| <= F, |s + t| <= KT, from g[0..KT].  Hermite auxiliaries
// A^{(n)}_k of degree |k| = L at lb[L] + n ndeg(L) + (packed index - ncomp(L - 1)), n = 0..KT - L.
__device__ void pair_contract_rt(int M, int F, int KT, double x, double y, double z, const double* g,
                                 const double* m, double* acc) {
    double A[NEAR_RT_AUX];
    int lb[NEAR_KMAX + 2];
    lb[0] = 0;
    for (int k = 0; k <= KT; ++k) lb[k + 1] = lb[k] + ndeg(k) * (KT - k + 1);
    for (int n = 0; n <= KT; ++n) A[n] = g[n];
    for (int k = 1; k <= KT; ++k) {
        int l = 0;
        for (int t = k; t >= 0; --t)
            for (int u = k - t; u >= 0; --u, ++l) {
                const int v = k - t - u;
                for (int n = 0; n <= KT - k; ++n) {
                    auto aux = [&](int L, int idx) { return A[lb[L] + (n + 1) * ndeg(L) + (idx - ncomp(L - 1))]; };
                    double val;
                    if (t > 0) {
                        val = x * aux(k - 1, xyz2idx(t - 1, u, v));
                        if (t > 1) val = __builtin_fma((double)(t - 1), aux(k - 2, xyz2idx(t - 2, u, v)), val);
                    } else if (u > 0) {
                        val = y * aux(k - 1, xyz2idx(t, u - 1, v));
                        if (u > 1) val = __builtin_fma((double)(u - 1), aux(k - 2, xyz2idx(t, u - 2, v)), val);
                    } else {
                        val = z * aux(k - 1, xyz2idx(t, u, v - 1));
                        if (v > 1) val = __builtin_fma((double)(v - 1), aux(k - 2, xyz2idx(t, u, v - 2)), val);
                    }
                    A[lb[k] + n * ndeg(k) + l] = val;
                }
            }
    }
    int it = 0;
    for (int ft = 0; ft <= F; ++ft)
        for (int tx = ft; tx >= 0; --tx)
            for (int ty = ft - tx; ty >= 0; --ty, ++it) {
                const int tz = ft - tx - ty;
                double sum = acc[it];
                int is = 0;
                for (int fs = 0; fs <= M; ++fs)
                    for (int sx = fs; sx >= 0; --sx)
                        for (int sy = fs - sx; sy >= 0; --sy, ++is) {
                            if (fs + ft > KT) continue;
                            const int sz = fs - sx - sy;
                            const int L = fs + ft;
                            const int idx = xyz2idx(sx + tx, sy + ty, sz + tz);
                            sum = __builtin_fma(m[is], A[lb[L] + (idx - ncomp(L - 1))], sum);
                        }
                acc[it] = sum;
            }
}

// ---------------------------------------------------------------- exclusions
// Exclusions of the near field (see the header comment).  One lane per
// target site of a target block; the exclusion entries of the site are
// de-duplicated (the reference tests membership), and an entry counts only
// if its site lies in a central-cell source block of the block's list.
// The pair loop gives 0 for every pair at R = 0; here a site gets the
// reference's non-finite self pair when it does not exclude itself or when
// another site sits at exactly its position without being excluded (the
// reference evaluates that pair at R = 0 too).  Such sites share a leaf of the
// octree, so tree calls (a.coinc_list == 0) look for them in the target block
// only; the direct sum looks through the block's whole source list.
// M < 0: runtime orders a.m_rt, a.f_rt (S = 1).
constexpr int NEAR_EXCL_STAGE = 512;
// ALLROWS (S > 1): one workgroup takes every set row of its block.  The coincidence count and the list
// search of each exclusion entry (chains of dependent loads) run once; the entries' additions are made
// row by row by the same statements as with one row per workgroup, so each set's sums stay bit for bit
// those of its single-set call.  Calls with more than EXCL_ROWS_MAX exclusion entries per site run one
// workgroup per set row.
constexpr int EXCL_ROWS_MAX = 64;

template <int M, int F, int DAMP, int S, bool ALLROWS = false>
__global__ __launch_bounds__(64) void k_near_excl(NearArgs a) {
    constexpr bool RT = M < 0;
    static_assert(!RT || S == 1, "runtime orders take one set per row");
    constexpr int NM = RT ? ncomp(NEAR_MMAX) : ncomp(M < 0 ? 0 : M);
    constexpr int NF = RT ? ncomp(NEAR_FMAX) : ncomp(F < 0 ? 0 : F);
    const int nm = RT ? ncomp(a.m_rt) : NM;
    const int nf = RT ? ncomp(a.f_rt) : NF;
    __shared__ double etab[DAMP != DAMP_NONE ? 32 : 1];   // 2^(j/32) for exp_table
    const int vb = a.blk[blockIdx.x];
    if (vb < 0) return;
    int b = vb;
    if (a.vb_node) {
        b = a.vb_node[vb];
        if (a.vb_e0[vb] != a.list_off[b]) return;   // a leaf's corrections run once, with its first chunk
    }
    if constexpr (DAMP != DAMP_NONE) {
        if (threadIdx.x < 32) etab[threadIdx.x] = kExp2By32[threadIdx.x];
        __syncthreads();
    }
    static_assert(!ALLROWS || (S > 1 && !RT), "all rows in one workgroup: sets per row, fixed orders");
    int set0 = ALLROWS ? 0 : (int)blockIdx.y * S;
    int ns = min(S, a.nsets - set0);
    const int t0 = a.tgt_start[b], tc = a.tgt_count[b];
    const int e0 = a.list_off[b], e1 = a.list_off[b + 1];
    const double* __restrict__ ms = a.ms + (int64_t)set0 * nm * a.ldm;
    double* __restrict__ out = a.out + (int64_t)set0 * nf * a.ldo;
    // the central-cell source blocks of the list, as sorted-site ranges [first, end), staged in LDS
    // (NEAR_EXCL_STAGE of them; a longer list is searched in global memory)
    __shared__ int rfirst[NEAR_EXCL_STAGE], rend[NEAR_EXCL_STAGE];
    const bool staged = e1 - e0 <= NEAR_EXCL_STAGE;
    if (staged) {
        for (int x = e0 + (int)threadIdx.x; x < e1; x += 64) {
            const int enc = a.list_src[x];
            int f = 0, en = 0;                    // image entries never match
            if (src_image(enc) == 0) {
                const int sb = src_node(enc);
                f = a.src_start[sb];
                en = f + a.src_count[sb];
            }
            rfirst[x - e0] = f;
            rend[x - e0] = en;
        }
    }
    __syncthreads();
    __shared__ double cxs[64], cys[64], czs[64];   // the block's sites, staged for the position compare
    __shared__ int cos_[64];
    for (int base = 0; base < tc; base += 64) {
        const int jt = base + (int)threadIdx.x;
        const bool act = jt < tc;
        const int k = t0 + (act ? jt : 0);
        const int myo = a.orig[k] + 1;
        const double xi = a.xs[k], yi = a.ys[k], zi = a.zs[k];
        // other sites at exactly this position (the pair loop gave them 0)
        int coinc = 0;
        if (a.coinc_list) {
            for (int x = e0; x < e1 && act; ++x) {
                const int enc = a.list_src[x];
                if (src_image(enc) != 0) continue;
                const int sb = src_node(enc), st = a.src_start[sb], en = st + a.src_count[sb];
                for (int j = st; j < en; ++j)
                    coinc += (a.xs[j] == xi && a.ys[j] == yi && a.zs[j] == zi && a.orig[j] + 1 != myo) ? 1 : 0;
            }
        } else {
            for (int cb = 0; cb < tc; cb += 64) {
                __syncthreads();                   // readers of the previous chunk are done
                const int j = t0 + cb + (int)threadIdx.x;
                if (cb + (int)threadIdx.x < tc) {
                    cxs[threadIdx.x] = a.xs[j];
                    cys[threadIdx.x] = a.ys[j];
                    czs[threadIdx.x] = a.zs[j];
                    cos_[threadIdx.x] = a.orig[j] + 1;
                }
                __syncthreads();
                const int cn = min(64, tc - cb);
                for (int q = 0; q < cn; ++q)
                    coinc += (cxs[q] == xi && cys[q] == yi && czs[q] == zi && cos_[q] != myo) ? 1 : 0;
            }
        }
        if (!act) continue;                        // after the barriers of this pass
        const double dfi = DAMP != DAMP_NONE ? a.dfs[k] : 0.0;
        const int coinc_site = coinc;
        unsigned long long nearmask = 0ull;        // ALLROWS: entries found in the list by row 0
        const int nrow = ALLROWS ? (a.nsets + S - 1) / S : 1;
        for (int row = 0; row < nrow; ++row) {
            if constexpr (ALLROWS) {
                set0 = row * S;
                ns = min(S, a.nsets - set0);
                ms = a.ms + (int64_t)set0 * nm * a.ldm;
                out = a.out + (int64_t)set0 * nf * a.ldo;
            }
            coinc = coinc_site;
            double acc[S][NF];
#pragma unroll
            for (int u = 0; u < S; ++u)
#pragma unroll
                for (int t = 0; t < NF; ++t) acc[u][t] = 0.0;
            auto add = [&](int j, double dx, double dy, double dz, double sign) {
                double aij = 0.0;
                if constexpr (DAMP != DAMP_NONE) aij = dfi * a.dfs[j];
                const double R2 = dx * dx + dy * dy + dz * dz;
                if constexpr (RT) {
                    if (ns < 1) return;
                    const int KT = (DAMP != DAMP_NONE && a.m_rt + a.f_rt > KMAX_NEAR) ? KMAX_NEAR : a.m_rt + a.f_rt;
                    double m[NM], g[NEAR_KMAX + 1];
                    for (int s = 0; s < nm; ++s) m[s] = sign * ms[(int64_t)s * a.ldm + j];
                    radial_rt<DAMP>(KT, R2, rsqrt_pos(R2), aij, g, etab);
                    pair_contract_rt(a.m_rt, a.f_rt, KT, dx, dy, dz, g, m, acc[0]);
                } else {
#pragma unroll
                    for (int u = 0; u < S; ++u) {
                        if (u >= ns) break;
                        double m[NM];
#pragma unroll
                        for (int s = 0; s < NM; ++s) m[s] = sign * ms[((int64_t)u * NM + s) * a.ldm + j];
                        pair_field<M, F, DAMP>(dx, dy, dz, R2, aij, m, acc[u], etab);
                    }
                }
            };
            bool self = false;
            for (int q = 0; q < a.nexcl; ++q) {
                const int e = a.excl[(int64_t)q * a.lde + k];
                if (e <= 0) continue;
                if (e > a.nsites) {                    // not a site: reported to the caller, entry skipped
                    if (a.bad) *a.bad = 1;
                    continue;
                }
                bool dup = false;
                for (int q2 = 0; q2 < q; ++q2) dup |= a.excl[(int64_t)q2 * a.lde + k] == e;
                if (dup) continue;
                if (e == myo) {
                    self = true;
                    add(k, 1e10, 1e10, 1e10, 1.0);
                    continue;
                }
                const int j = a.inv ? a.inv[e - 1] : e - 1;   // sorted position of the excluded site
                bool near = false;
                if (ALLROWS && row > 0) {
                    near = (nearmask >> q) & 1ull;
                } else {
                    if (staged) {
                        for (int x = 0; x < e1 - e0 && !near; ++x) near = j >= rfirst[x] && j < rend[x];
                    } else {
                        for (int x = e0; x < e1 && !near; ++x) {
                            const int enc = a.list_src[x];
                            if (src_image(enc) != 0) continue;
                            const int sb = src_node(enc);
                            const int st = a.src_start[sb];
                            near = j >= st && j < st + a.src_count[sb];
                        }
                    }
                    if (ALLROWS && near) nearmask |= 1ull << q;
                }
                if (!near) continue;
                add(j, 1e10, 1e10, 1e10, 1.0);
                const double dx = xi - a.xs[j], dy = yi - a.ys[j], dz = zi - a.zs[j];
                if (dx == 0.0 && dy == 0.0 && dz == 0.0) --coinc;   // excluded: no R = 0 pair to remove
                else add(j, dx, dy, dz, -1.0);
            }
            // the reference's pair at R = 0: non-finite
            if (!self || coinc > 0) add(k, 0.0, 0.0, 0.0, 1.0);
#pragma unroll
            for (int u = 0; u < S; ++u) {
                if (u >= ns) break;
                for (int t = 0; t < nf; ++t) out[((int64_t)u * nf + t) * a.ldo + k] += acc[u][t];
            }
        }   // rows
    }
}

// Sets per workgroup row: 4 where the tensor is cheap in registers, else 2.
template <int M, int F, int DAMP>
constexpr int near_sets() { return (M + F <= 3 && DAMP != DAMP_ERF) ? 4 : 2; }

// part 1: the pair kernel; part 2: the exclusion correction (after it on the stream).  ROWS_OF_ONE: one
// set per workgroup row whatever the set count (the wide orders, whose registers hold one set).
template <int M, int F, int DAMP, bool ROWS_OF_ONE>
static void launch_sets(const NearArgs& a, int nblocks, int part, hipStream_t s) {
    if (ROWS_OF_ONE || a.nsets == 1) {
        const dim3 g(nblocks, a.nsets);
        if (part == 1) hipLaunchKernelGGL((k_near<M, F, DAMP, 1>), g, dim3(64), 0, s, a);
        else if constexpr (ROWS_OF_ONE) launch_near_excl_rt(a, nblocks, M, F, DAMP, s);
        else hipLaunchKernelGGL((k_near_excl<M, F, DAMP, 1>), g, dim3(64), 0, s, a);
    }
    if constexpr (!ROWS_OF_ONE) {
        if (a.nsets > 1) {
            constexpr int S = near_sets<M, F, DAMP>();
            // two targets per lane: charges and dipoles (quadrupole sources spill at 3 waves per SIMD and ran
            // slower, 38.0 -> 41.4 ms on 32 AMOEBA sets)
            constexpr bool T2_OK = S == 4 && M <= 1 && F <= 1 && has_closed_form<M, F>();
            const dim3 g(nblocks, (a.nsets + S - 1) / S);
            if (part == 1) {
                if (T2_OK && a.pairs2) hipLaunchKernelGGL((k_near<M, F, DAMP, S, T2_OK>), g, dim3(64), 0, s, a);
                else hipLaunchKernelGGL((k_near<M, F, DAMP, S>), g, dim3(64), 0, s, a);
            } else if (a.nexcl <= EXCL_ROWS_MAX) {
                hipLaunchKernelGGL((k_near_excl<M, F, DAMP, S, true>), dim3(nblocks), dim3(64), 0, s, a);
            } else {
                hipLaunchKernelGGL((k_near_excl<M, F, DAMP, S>), g, dim3(64), 0, s, a);
            }
        }
    }
}

template <int M, int F, bool ROWS_OF_ONE>
static void launch_damp(const NearArgs& a, int nblocks, int damp, int part, hipStream_t s) {
    switch (damp) {
        case DAMP_NONE: launch_sets<M, F, DAMP_NONE, ROWS_OF_ONE>(a, nblocks, part, s); break;
        case DAMP_THOLE: launch_sets<M, F, DAMP_THOLE, ROWS_OF_ONE>(a, nblocks, part, s); break;
        case DAMP_AMOEBA: launch_sets<M, F, DAMP_AMOEBA, ROWS_OF_ONE>(a, nblocks, part, s); break;
        default: launch_sets<M, F, DAMP_ERF, ROWS_OF_ONE>(a, nblocks, part, s); break;
    }
}

}  // namespace fmmh
