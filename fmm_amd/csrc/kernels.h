// Kernel argument blocks and launch wrappers shared by the translation units
// of the field engine.  Everything here is device memory.
#pragma once
#include "common.h"

namespace fmmh {

// Orders compiled into the near-field kernels: source multipole order m <= 6 (dim <= 84), field order
// F <= 6, m + F <= 7.  m, F <= 2 have their own tuned forms (near.hip); the rest are near_wide*.hip.
constexpr int NEAR_MMAX = 6;
constexpr int NEAR_FMAX = 6;
constexpr int NEAR_KMAX = 7;

struct NearArgs {
    const int* blk;            // workgroup -> target block row (a leaf's node id, or a chunk id)
    const int* tgt_start;      // per target block: first sorted site
    const int* tgt_count;      //   and number of sites
    const int* list_off;       // CSR over target blocks -> source blocks
    const int* list_src;
    const int* src_start;      // per source block: first sorted site
    const int* src_count;
    const double* xs;          // sorted site coordinates (SoA)
    const double* ys;
    const double* zs;
    const double* ms;          // sorted multipoles scaled by (-1)^{|s|+1}/(sx!sy!sz!), component-major
    int64_t ldm;
    const double* dfs;         // sorted damping factors
    const int* orig;           // sorted -> original (0-based) site index
    const int* excl;           // sorted exclusion rows (1-based original indices), component-major
    int64_t lde;
    int nexcl;
    double* out;               // per sorted site, ncomp(F) components
    int64_t ldo;
    int accumulate;            // 0: out = result, 1: out += result
    int nsets;                 // multipole sets: ms is [set][component][ldm], out is [set][component][ldo]
    const double* disp;        // per image id: source displacement (x, y, z); image 0 = (0, 0, 0)
    const int* inv;            // original (0-based) site index -> sorted position (null: identity)
    // Chunked target blocks (null vb_node: blk holds target block ids, each with its whole list).
    // blk then holds chunk ids; chunk v covers list entries [vb_e0[v], vb_e1[v]) of block vb_node[v]
    // and writes its partial field to part at sites vb_part[v] + i (vb_part -1: the block's only
    // chunk, which writes out directly).
    const int* vb_node = nullptr;
    const int* vb_e0 = nullptr;
    const int* vb_e1 = nullptr;
    const int* vb_part = nullptr;
    double* part = nullptr;    // [set][component][ldp]
    int64_t ldp = 0;
    int nsites = 0;            // sites of the call: exclusion entries above it are not sites
    int* bad = nullptr;        // set to 1 when an exclusion entry is above nsites (null: not reported)
    int coinc_list = 0;        // sites at one position: 0 in the target block (octree leaves), 1 in its list
    int m_rt = 0, f_rt = 0;    // source and field orders (the runtime-order exclusion kernel)
    int pairs2 = 0;            // many sets: two targets per lane where the orders allow (another summation order)
};

bool near_supported(int m, int F);
void launch_near(const NearArgs& a, int nblocks, int m, int F, int damp, hipStream_t s);
// Exclusion correction with the orders as arguments (one set per workgroup row), for the orders outside
// m, F <= 2.
void launch_near_excl_rt(const NearArgs& a, int nblocks, int m, int F, int damp, hipStream_t s);
// The orders outside m, F <= 2, by tensor order m + F (3 and 4, 5, 6, 7; a: no damping and THOLE, b: AMOEBA
// and ERF); false if (m, F, damp) is not theirs.
bool launch_near_wide_k34(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s);
bool launch_near_wide_k5(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s);
bool launch_near_wide_k6a(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s);
bool launch_near_wide_k6b(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s);
bool launch_near_wide_k7a(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s);
bool launch_near_wide_k7b(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s);
// part 1: the pair kernel only; part 2: the exclusion correction only
void launch_near_part(const NearArgs& a, int nblocks, int m, int F, int damp, int part, hipStream_t s);
// Sum of the chunks' partial fields of every chunked leaf, in chunk order, into out (after part 1).
// lk[l]: chunks of leaf l (1: not chunked), lpart[l]: its first partial site.
void launch_near_fix(const NearArgs& a, const int* leaf_node, const int* lk, const int* lpart, int nl, int nf,
                     hipStream_t s);

// ---------------------------------------------------------------- tree
struct TreeDev {
    // per site (sorted order after the build)
    unsigned long long* keys;  // 63-bit octant-digit keys, sorted
    int* perm;                 // sorted position -> original site
    int* lmin;                 // shallowest level whose node starts at this site
    int* lam;                  // leaf level of the site
    int* cnt;                  // nodes that start at this site
    int* off;                  // exclusive scan of cnt = pre-order id of the first such node
    int* w;                    // window common-prefix lengths
    // per node (pre-order id)
    int* level;                // root = 1
    int* start;
    int* count;
    int* parent;
    int* is_leaf;
    int* octant;
    int* children;             // 8 per node, -1 = empty
    double* cx;
    double* cy;
    double* cz;
    double* r;
    double* rmax;
    int* leaf_id;              // node -> leaf index or -1
    int* leaf_node;            // leaf index -> node
    int* level_nodes;          // nodes grouped by level
    // scalars
    double* root;              // cx, cy, cz, r0, rmax0
    unsigned long long* maxabs;  // 3 bit patterns of max |x - c|
    int* info;                 // [0] num_nodes, [1] error flag, [2] max level, [3..3+KEY_LEVELS+1] level counts
    int* level_fill;
};

// work: centroid_work_bytes(n) bytes of device memory.
size_t centroid_work_bytes(int n);
void launch_centroid(const double* X, int64_t ldx, int n, TreeDev& t, void* work, hipStream_t s);
void launch_keys(const double* X, int64_t ldx, int n, TreeDev& t, unsigned long long* keys_out, int* idx_out,
                 hipStream_t s);
void launch_site_levels(int n, int ncrit, TreeDev& t, hipStream_t s);
void launch_emit_nodes(int n, TreeDev& t, hipStream_t s);
void launch_node_links(int num_nodes, TreeDev& t, hipStream_t s);
void launch_leaf_index(int num_nodes, const int* leaf_scan, TreeDev& t, hipStream_t s);
void launch_level_scatter(int num_nodes, const int* level_off, TreeDev& t, hipStream_t s);
void launch_leaf_sort(int num_leaves, int max_leaf, TreeDev& t, hipStream_t s);

// sorted site arrays
struct SiteArgs {
    const double* X; int64_t ldx;
    const double* M; int64_t ldm; int dim;
    const int* E; int64_t lde; int nexcl;
    const double* df;
    const int* perm;
    int n;
    double* xs; double* ys; double* zs;
    double* ms_raw; double* ms_scaled;   // component-major, leading dimension n (ms_raw null: not written)
    double* dfs;
    int* excl_s;                         // component-major, leading dimension n
    const double* scale;                 // per multipole component factor (device, dim entries)
    int nsets;                           // M is [set][component][ldm]; ms_* are [set][component][n]
};
void launch_gather(const SiteArgs& a, hipStream_t s);

// ---------------------------------------------------------------- lists
struct TraverseArgs {
    const int2* in; const int* n_in; int cap_in;   // n_in may exceed cap_in after an overflow
    int2* out; int* n_out; int cap_out;
    unsigned long long* m2l; int* n_m2l; int cap_m2l;
    unsigned long long* p2p; int* n_p2p; int cap_p2p;
    int* overflow;
    const double* cx; const double* cy; const double* cz; const double* rmax;
    const int* is_leaf; const int* children;
    double theta;
    const double* disp;        // per image id: displacement of the source cell; frontier sources carry the id
    int* n_clear;              // a frontier counter this step neither reads nor writes, zeroed for the next
};
void launch_traverse(const TraverseArgs& a, int grid, hipStream_t s);
// Traversal set-up: counters (6 of them, stride `stride` ints: 3 frontier counters, M2L, P2P, overflow) =
// {nimg, 0, 0, 0, 0, 0} and the frontier seeded with one (root, root) pair per periodic image.
void launch_traverse_init(int* counters, int stride, int nimg, int2* frontier, hipStream_t s);
// Copies the 6 counters into out[0..5] (contiguous, for one readback).
void launch_traverse_counts(const int* counters, int stride, int* out, hipStream_t s);
void launch_csr(const unsigned long long* keys, int n_keys, int num_targets, int* off, int* src, hipStream_t s);
// CSR (offsets, sources ascending per target, and the keys in that order when keys_out is not null) of an
// unsorted list of (target << 32 | source) keys.  cnt: num_targets + 1 ints that are zero on entry and on
// return; tmp: n_keys; scan_temp: exscan_i32_temp(num_targets + 1) bytes.
void launch_list_csr(const unsigned long long* keys, int n_keys, int num_targets, int* cnt, int* off, unsigned* tmp,
                     void* scan_temp, size_t scan_temp_bytes, int* src, unsigned long long* keys_out,
                     hipStream_t s);

// ---------------------------------------------------------------- expansions
// Shift tables: entries grouped by target component t (CSR by t); each entry
// carries the source component s, the exponent triple of the monomial, and
// the combinatorial coefficient.
struct ShiftTab {
    const int* off;      // ncomp(order_t) + 1
    const int* s;        // source packed index
    const int* e;        // packed index of the monomial exponent triple
    const double* coef;
};

struct UpArgs {
    const double* xs; const double* ys; const double* zs;
    const double* ms_raw; int64_t ldm;
    const int* leaf_node; const int* start; const int* count;
    const double* cx; const double* cy; const double* cz;
    const int* children;
    double* CM; int nm; int p;
    int nsrc;                 // multipole components entering P2M (|s| <= min(m, p))
    ShiftTab p2m; ShiftTab m2m;
    const int* mono_parent;   // packed index -> packed index with one exponent less
    const int* mono_axis;     //   and which axis was removed
    int num_nodes;            // CM is [set][node][nm]; ms_raw is [set][component][ldm]
    int nsets;                // one grid row per set
    int dim;                  // multipole components per site and set
    const int* xyz;           // packed index -> ex | ey << 8 | ez << 16
};
void launch_p2m(const UpArgs& a, int num_leaves, hipStream_t s);
// P2M and L2P (field order <= 1) of EXP_MFMA_MIN_SETS or more sets on the FP64 matrix cores
// (expansions_mfma.hip); false (nothing launched) where they do not apply.
constexpr int EXP_MFMA_MIN_SETS = 16;
constexpr int EXP_MFMA_PMAX = 10;
bool launch_p2m_mfma(const UpArgs& a, int num_leaves, hipStream_t s);
void launch_m2m(const UpArgs& a, const int* nodes, int count, hipStream_t s);

// M2L by displacement class.  For a far pair (i, j) let u = min(r_i, r_j):
// c_i - c_j = u v with v an integer vector (node centres are the root centre
// plus multiples of the finest half-width involved), so
//   T_k(c_i - c_j) = u^{-(|k|+1)} T_k(v).
// Pairs are grouped by (v, level_i - level_j); one tensor per class, and the
// per-pair work is the Hankel contraction L~_i(t) += sum_s T~_c[s+t] M~_j(s)
// on box-size-scaled expansions M~_j(s) = c_s M_j(s) r_j^{-|s|}.
struct M2LClassArgs {
    const unsigned long long* pair_keys;   // sorted (target << 32 | source), CSR order
    int n_pairs;
    const double* cx; const double* cy; const double* cz; const double* r;
    const int* level;
    unsigned long long* class_key_in;      // per pair: packed (v, level difference)
    int vbits;                              // bits per component of v in non-periodic keys (<= 18)
    int* wg_range;                          // per key workgroup: max |v_x|, |v_y|, |v_z|, |dl| (or null)
    int* pair_idx;                          // per pair: its CSR position
    unsigned long long* class_key_sorted;
    int* pair_idx_sorted;
    int* flag;                              // first of a class in sorted order
    int* cls;                               // exclusive scan of flag
    int* pair_class;                        // CSR position -> class id
    unsigned long long* class_key;          // class id -> key
    int* num_classes;                       // device scalar
    int* error;                             // device scalar: displacement out of range
    double* Tc;                             // class tensors, ncomp(2p) each
    int p;
    int zmax;                               // components with z > zmax are set to 0 (2: all the harmonic
                                            // M2L engines read; 2p: every component)
    const int* xyz;                         // packed index -> t | u << 8 | v << 16
    const int* level_base;                  // aux slot base per degree (2p + 2 entries)
    int* class_dl;                          // class id -> level_i - level_j
    // Periodic runs: a class is (v, level difference, image, and for image
    // classes the target level), keyed by two words sorted as (hi, lo);
    // the image displacement D enters the class tensor as T(v - D / u).
    int periodic;
    unsigned long long* class_hi_in;        // per pair: image | level | level difference
    unsigned long long* class_hi_sorted;    // hi words in final sorted order
    unsigned long long* class_lo_final;     // lo words in final sorted order
    unsigned long long* class_key_hi;       // class id -> hi word
    const double* disp;                     // per image id: displacement
};
void launch_m2l_keys(const M2LClassArgs& a, hipStream_t s);
// Dense class numbering of non-periodic pairs (after launch_m2l_keys with wg_range over nwg workgroups):
// pair_class, class_key, class_dl and num_classes without the key sort; error bit 4 when the displacement
// range exceeds CLASS_DENSE_CAP indices (then nothing is written).  dims: 5 ints; marks: CLASS_DENSE_CAP
// ints, zero on entry and on return; ids: CLASS_DENSE_CAP ints; tsum, toff: CLASS_DENSE_CAP / 4096 ints.
constexpr int CLASS_DENSE_CAP = 1 << 23;
void launch_class_dense(const M2LClassArgs& a, int nwg, int* dims, int* marks, int* ids, int* tsum, int* toff,
                        hipStream_t s);
void launch_m2l_flags(const M2LClassArgs& a, hipStream_t s);        // after the key sort
void launch_m2l_assign(const M2LClassArgs& a, hipStream_t s);       // after the flag scan
void launch_m2l_class_tensors(const M2LClassArgs& a, int grid, hipStream_t s);
void launch_gather_u64(const unsigned long long* in, const int* idx, unsigned long long* out, int n, hipStream_t s);

// Expansions of S multipole sets are stored set-major, CM[(set * nodes + node)
// * nm + s] and likewise LE.  The scaled sources CMs that M2L reads are
// set-major for k_m2l (s_ld = 0) and set-fastest, CMs[(node * nm + s) * s_ld +
// set] with s_ld padded to 16 or 32, for the MFMA form.
struct M2LArgs {
    const int* off; const int* src;
    const int* pair_class;
    const double* Tc;
    const double* CMs;
    double* LE;
    const int* level;
    const double* r;
    int nm; int p;
    int nsets;
    int s_ld;
    int num_nodes;
    const unsigned char* order;          // per CSR position: learned expansion order (null: p; MFMA form only)
    const int* fidx;                     // MFMA form: A-fragment index table (see M2LGemmArgs)
    const float* CMs32;                  // MFMA form, mixed precision: FP32 copy of CMs (null: FP64)
    const void* Tcp;                     // MFMA form: class tensors in its precision, stride m2l_mfma_class_stride
};
// Padded class-tensor copy for the MFMA form (16-byte aligned rows of FP32 or FP64).
int m2l_mfma_class_stride(int p, bool f32);
void launch_m2l_class_pad(const double* Tc, void* Tcp, int ncls, int p, bool f32, hipStream_t s);
void launch_scale_sources(const double* CM, const double* cs, const double* r, double* CMs, int num_nodes, int p,
                          int nsets, int s_ld_out, hipStream_t s);
// Fold of c M onto the reduced components, term by term (built on the host in the order the device fold
// visits them): reduced component sr takes terms [off[sr], off[sr + 1]), each w (c M)(sc) with w the signed
// binomial and c the Taylor factor of sc; deg[sr] its degree.
struct FoldTab {
    const int* off; const int* sc; const int* deg;
    const double* w; const double* c;
};
// Many-sets M2L in harmonic (traceless) form: folded sources [node][(p + 1)^2][s_ld] and the
// local-expansion components with t_z >= 2 rebuilt from the others after it (rows = sets x nodes).
// One set (nsets = s_ld = 1) folds with cs; several sets with the term table ft.
void launch_scale_sources_red(const double* CM, const double* cs, const FoldTab& ft, const double* r, double* CMs,
                              int num_nodes, int p, int nsets, int s_ld, hipStream_t s);
void launch_le_unfold(double* LE, long long rows, int p, hipStream_t s);
void launch_m2l(const M2LArgs& a, int num_nodes, hipStream_t s);
// MFMA form for many sets: one wavefront per (target, 16 sets) accumulates
// its n_m x 16 local expansions over all its pairs (FP64 MFMA, or FP32 MFMA
// per pair with FP64 accumulation when CMs32 is set).  CMs is set-fastest with
// s_ld >= nsets, a multiple of 16.  Supported from 16 sets on, orders <= 10.
bool m2l_mfma_supported(int p, int nsets);
// Tile-shape variants 0..3 (see m2l.hip); an invalid variant falls back to the default.
constexpr int M2L_MFMA_VARIANTS = 4;
bool m2l_mfma_variant_ok(const M2LArgs& a, int variant);
int m2l_mfma_default_variant(const M2LArgs& a);
void launch_m2l_mfma(const M2LArgs& a, int num_nodes, int variant, hipStream_t s);

// Class-grouped M2L: all pairs of one displacement class share one Hankel
// matrix H_c[t][s] = scale_c(t) T~_c[s + t], so per class the M2L is the GEMM
// H_c (n_m x n_m) x [M~_j ...] (n_m x pairs) on the FP64 matrix cores, in
// tiles of 32 pairs.  Each pair's n_m results go to its CSR slot of pbuf;
// launch_m2l_reduce then sums every target's slots in CSR order (fixed order,
// deterministic) and applies r_i^{-(|t|+1)}.
struct M2LGemmArgs {
    const int* class_off;                // ncls + 1: first class-sorted pair of each class
    const int* tile_off;                 // ncls + 1: first tile of each class, tile_off[ncls] = tiles
    int ncls;
    const int* pair_csr;                 // class-sorted pair -> CSR position
    const int* src;                      // CSR position -> source node
    const int* class_dl;                 // class -> level_i - level_j
    const double* Tc;                    // class tensors, ncomp(2p) each
    const double* CMs;                   // folded scaled source expansions of one set, [node][(p + 1)^2]
    const int* fidx;                     // [MT][KS][64]: packed index of s + t per A-fragment lane (s, t reduced
                                         //   components), -1 outside
    const int* tdeg;                     // [16 MT]: |t| of each reduced target row, 0 past (p + 1)^2
    double* pbuf;                        // [CSR position][(p + 1)^2] per-pair partial local coefficients
    int p;
    int f32;                             // mixed precision: FP32 MFMA on CMs32, slots in pbuf32
    const float* CMs32;
    float* pbuf32;
    // Per-pair expansion orders (learned order selection): tiles are then
    // runs of one (class, order) group; grp_class / grp_order give each
    // group's class and order, and only the leading rdim(order) rows and
    // columns of the Hankel matrix are applied.  order[CSR position] tells the
    // reduction how many coefficients a slot holds.  All null: every pair at p.
    const int* grp_class;
    const int* grp_order;
    const unsigned char* order;
    const int* tfull;                    // [(p + 1)^2]: packed index of each reduced component
};
// Target-block M2L (one set, FP64, every pair at order p <= M2LB_PMAX): pairs grouped by blocks of
// M2LB_BT targets in (level, octant, id) order and by class inside a block, in tiles of up to 16 pairs;
// local expansions accumulate per block in LDS, a partial per (block, workgroup) where a workgroup's
// share of the tiles leaves the block.
constexpr int M2LB_PMAX = 6;
constexpr int M2LB_BT = 64;              // targets per block (one bit each in a class mask)
constexpr int M2LB_REC = 24;             // ints per tile record: class, dl, count, 0, 16 sources, 16 bytes
constexpr int M2LB_MAX_CLASSES = 4000;   // class masks and offsets in 64 KB of LDS
constexpr int M2LB_READ_SLACK = 64;      // doubles readable past the scaled sources and class tensors
struct M2LBlockArgs {
    const int* m2l_off;                  // CSR of the far pairs
    const int* m2l_src;
    const int* pair_class;               // CSR position -> class
    const int* class_dl;                 // class -> level_i - level_j
    const int* level;                    // per node
    const int* octant;
    int nn, ncls;
    int nbm;                             // block capacity, ceil(nn / M2LB_BT)
    int* tperm;                          // [nn]: block order position -> target node
    int* tinv;                           // [nn]: node -> position, -1 without far pairs
    int* tcount;                         // [nn + 1]: pairs per position, 0 past ntg
    int* pbase;                          // [nn + 1]: exclusive scan of tcount (exscan after the order kernel)
    int* ntile;                          // [nbm + 1]: tiles per block
    int* toff;                           // [nbm + 1]: exclusive scan of ntile (exscan after the tiles kernel)
    int* rec;                            // [n_pairs][M2LB_REC]: tile records at the block's pair offset
    int* ntg;                            // device scalar: targets with far pairs
    int* flag;                           // bit 2: a target with two pairs of one class
};
void launch_m2lb_order(const M2LBlockArgs& a, hipStream_t s);
void launch_m2lb_tiles(const M2LBlockArgs& a, hipStream_t s);
struct M2LBRunArgs {
    const int* toff; const int* pbase; const int* rec; int nbm;
    const int* ntg; const int* tperm; const int* tinv; int nn;
    const double* CMs;                   // folded scaled source expansions of one set, [node][(p + 1)^2]
                                         //   (+ slack)
    const double* Tc;                    // class tensors (+ M2LB_READ_SLACK doubles)
    const int* fidx; const int* tdeg;    // the class-grouped GEMM's fragment tables (reduced components)
    const int* tfull;                    //   and the packed index of each reduced component
    double* partial;                     // [nbm + grid][(p + 1)^2][M2LB_BT]
    int grid;                            // workgroups of the run kernel
};
// Run kernel, per-block sums (scaled by r^{-(|t|+1)}) into LE, and zeros for nodes without far pairs.
void launch_m2lb_run(const M2LBRunArgs& a, int p, const double* r, double* LE, hipStream_t s);

constexpr int M2L_GEMM_PMAX = 10;
constexpr int M2L_GEMM_TN = 32;          // pairs per tile
void launch_m2l_tiles(const int* flag, const int* cls, int n_pairs, int ncls, int* class_off, int* tile_cnt,
                      hipStream_t s);
void launch_m2l_gemm(const M2LGemmArgs& a, int grid, hipStream_t s);
void launch_m2l_reduce(const M2LGemmArgs& a, const int* off, const double* r, double* LE, int num_nodes,
                       hipStream_t s);

// ---------------------------------------------------------------- learned order selection
constexpr int SEL_IN = 4;                // MLP inputs
constexpr int SEL_H = 64;                // units per hidden layer
constexpr int SEL_NPARAM = SEL_H * SEL_IN + SEL_H + SEL_H * SEL_H + SEL_H + SEL_H + 1;
constexpr int SEL_NS = 32;               // target sites per training sample

struct SelSampleArgs {
    const unsigned long long* pair_keys;  // CSR-sorted (target << 32 | source)
    const int* sample_pair;               // per sample: CSR position of its pair
    int nsamples;
    const double* cx; const double* cy; const double* cz; const double* r;
    const int* start; const int* count;
    const double* xs; const double* ys; const double* zs;
    const double* ms; int64_t ldm;        // sorted Taylor-scaled multipoles, component-major
    const double* CM; int nm; int cap;    // cell multipoles at order cap (one set)
    const double* cs;                     // Taylor factors of the expansion components
    const int* xyz; const int* level_base;
    const double* node_norm;              // A_j per node
    float* feat;                          // [nsamples][SEL_IN] (x3 left 0)
    double* err;                          // [nsamples][cap + 1] truncation error per order
    double* fref;                         // [nsamples] RMS exact field at the sampled sites
};
void launch_sel_sample(const SelSampleArgs& a, int m, hipStream_t s);
void launch_sel_node_norm(const double* CM, const double* r, const int* xyz, int num_nodes, int nm, int nsets,
                          double* out, hipStream_t s);

struct MlpArgs {
    const float* X; const float* Y; int ns;   // raw features [ns][SEL_IN], targets [ns]
    const float* norm;                        // [2 SEL_IN]: feature mean, 1 / std
    float* W;                                 // [SEL_NPARAM]
    float* part;                              // [nblk][SEL_NPARAM + 1] workgroup partial gradients (+ loss)
    float* grad;                              // [SEL_NPARAM + 1] minibatch gradient (+ loss)
    float* adam_m; float* adam_v;
    int* step;                                // device step counter (null: 0)
    float* loss_log;                          // per step: minibatch loss
    int batch; int nblk;                      // samples per step, workgroups per step (64 samples each)
    int use_perm;                             // 0: samples 0..batch-1; 1: permuted running position
    unsigned long long perm_a, perm_c;
    float lr, beta1, beta2, eps, asym;
    int lr_half;                              // steps per halving of the step size
};
void launch_mlp_grad(const MlpArgs& a, hipStream_t s);        // part + grad
void launch_mlp_train_step(const MlpArgs& a, hipStream_t s);  // grad, Adam, step + 1
void launch_sel_predict(const float* W, const float* norm, const float* X, int ns, float* Y, hipStream_t s);

struct SelOrderArgs {
    const unsigned long long* pair_keys; int n_pairs;
    const double* cx; const double* cy; const double* cz; const double* r;
    const double* node_norm;
    const double* tau;                        // device scalar: per-pair field tolerance
    const float* W; const float* norm;
    int p; int cap_train; int force_p;
    float margin, rel_lo, rel_hi;             // training range of x2 - x3
    unsigned char* order;                     // per CSR position
    unsigned long long* hist;                 // [PMAX + 1] pairs per order
};
void launch_sel_orders(const SelOrderArgs& a, hipStream_t s);
void launch_sel_groups(const int* pair_csr, const int* pair_class, const unsigned char* order, int n,
                       unsigned long long* keys, int* vals, hipStream_t s);
void launch_sel_group_flags(const unsigned long long* keys, int n, int* flag, hipStream_t s);
void launch_sel_group_info(const unsigned long long* keys, const int* flag, const int* gid, int n, int* grp_class,
                           int* grp_order, int* ngroups, hipStream_t s);
// Per-pair tolerance tau = eps_kappa * (RMS near field) / sqrt(mean M2L pairs on a site's path), with the
// smallest non-zero near-field RMS over the nsets sets (near is [set][component][ldn]).  acc holds
// 4 + nsets doubles: [0] the sum of squares used, [1] the path sum, [2] tau, [4 + s] set s's sum of squares.
void launch_sel_budget(const double* near, int64_t ldn, int n, int nf, int nsets, const int* leaf_node,
                       const int* count, const int* parent, const int* m2l_off, int num_leaves, double* acc,
                       double eps_kappa, hipStream_t s);

struct SpotArgs {
    int nspots; int num_leaves; int n;
    const int* leaf_node; const int* start; const int* count;
    const int* p2p_off; const int* p2p_src;
    const double* xs; const double* ys; const double* zs;
    const double* ms; int64_t ldm;            // Taylor-scaled multipoles [set][component][ldm]
    double* out;                              // [chunk][set][nspots][4] far-field direct sum, per source chunk
    int* site;                                // [nspots] sorted site index
    int nsets;
    int chunk; int nchunks;                   // sources per chunk (a multiple of spot_tile(m)), chunk count
};
// Source multipole orders the spot checks (and so the learned per-pair orders) cover; calls with higher
// source orders keep expansion_order for every pair.
constexpr int SPOT_MMAX = 2;
int spot_tile(int m);                  // sources per staged tile of launch_spot_far
int spot_row_blocks(int nspots);       // its workgroups along the spot sites
void launch_spot_far(const SpotArgs& a, int m, hipStream_t s);
void launch_to_f32(const double* in, float* out, int64_t n, hipStream_t s);

// M2M / L2L on FP32 MFMA for one level (shift_mfma.hip): columns are (node of the level, set) pairs.
struct LevelStarts {
    int off[KEY_LEVELS + 3];       // first slot of each level in a per-level node array
};
struct ShiftMfmaArgs {
    const int* nodes;              // the level's internal nodes
    const int* count;              //   their number (device)
    int max_count;                 //   and an upper bound (grid size)
    const int* children;           // [node][8] by octant, -1 where absent
    const double* r;               // node half-widths
    const double* src; double* dst;   // [set][node][nm]: M2M children -> parents, L2L parents -> children
    int64_t set_stride;            // num_nodes * nm
    int nsets;
    const float* A;                // 8 shift matrices (one per octant) in MFMA fragment order
    const int* xyz;                // packed index -> ex | ey << 8 | ez << 16
    // L2L of harmonic local expansions on their (p + 1)^2 components with t_z <= 1 (A: the reduced
    // matrices): reduced index -> packed index; null: the full shift
    const int* tfull;
};
int shift_mfma_frag_floats(int p, bool reduced);
void launch_shift_mfma(const ShiftMfmaArgs& a, int p, bool l2l, hipStream_t s);
void launch_internal_nodes(int num_nodes, const int* level, const int* children, const LevelStarts& ls, int* out,
                           int* cnt, hipStream_t s);

struct DownArgs {
    const double* cx; const double* cy; const double* cz;
    const int* parent;
    double* LE; int nm; int p;
    ShiftTab l2l; ShiftTab l2p;
    const int* mono_parent; const int* mono_axis;
    // L2P
    const int* leaf_node; const int* start; const int* count;
    const double* xs; const double* ys; const double* zs;
    const int* perm;
    const double* near; int64_t ld_near;   // P2P result per sorted site
    double* field; int64_t ld_field;       // accumulated at original indices
    int nf;
    int num_nodes;                         // LE is [set][node][nm]; near and field are [set][component][ld]
    int nsets;                             // one grid row per set
    const int* xyz;                        // packed index -> ex | ey << 8 | ez << 16
    // many-set L2P on the matrix cores: (p + 1)^2 x ld_basis doubles for the per-site harmonic basis.  Set
    // only where LE holds harmonic local expansions (built by this library's M2L and L2L); null selects the
    // vector-ALU kernels
    double* basis;
    int64_t ld_basis;
};
void launch_l2l(const DownArgs& a, const int* nodes, int count, hipStream_t s);
void launch_l2p(const DownArgs& a, int num_leaves, hipStream_t s);
bool launch_l2p_mfma(const DownArgs& a, int num_leaves, hipStream_t s);

// ---------------------------------------------------------------- periodic boundaries
// Lattice sum beyond the 27 cells (src/fmm.F90:884-944): at shell level l
// the 26 x 27 cells of size 3^{l-1} boxes around the central 3^l supercell
// act on the root's local expansion with the level-l supercell multipole
// (charge removed), l = 1..PSUM_LEVELS until the contribution's norm drops
// below 1e-8.  The cell tensors depend only on the box and p, so their sums
// per level are formed once (k_periodic_shells) and the level loop runs in
// one workgroup per set.  Then the dipole correction (src/fmm.F90:946-979).
constexpr int PSUM_LEVELS = 50;
struct PeriodicArgs {
    const double* box;        // 3 x 3 column-major (device)
    double* shells;           // [PSUM_LEVELS][26][ncomp(2p)] per-neighbour shell tensor sums
    int p;
    const int* xyz; const int* level_base;
    const double* CM; double* LE; int num_nodes; int nsets;   // [set][node][nm], root = node 0
    const double* cs;         // Taylor factors c_s of the expansion components
    ShiftTab m2m; const int* mono_parent; const int* mono_axis;
    const double* X; int64_t ldx; int n;                       // caller-order coordinates
    double* field; int64_t ldf; int nf;                         // [set][component][ldf]
    double fourpi, twopi, threeV;
    int* levels_used;         // per set: shell levels summed
};
void launch_periodic_shells(const PeriodicArgs& a, hipStream_t s);
void launch_periodic_sum(const PeriodicArgs& a, hipStream_t s);
void launch_dipole_correction(const PeriodicArgs& a, hipStream_t s);

// ---------------------------------------------------------------- QM/MM embedding
// classify_box (src/fmm.F90:1134-1169) against a set of target points: a
// node is too close if r theta <= rmax for any target (flang NORM2); the
// recursion descends close internal nodes only, so a node is visited iff
// every ancestor is close.  type: 0 not used, 1 far (box multipole), 2 close
// leaf (its sites join the near field).
void launch_classify(const double* targets, int64_t ldt, int ntargets, double theta, const TreeDev& t, int num_nodes,
                     int* close, int* type, hipStream_t s);
// Per-node counts for the compaction: near[node] = count if type 2, far[node] = 1 if type 1.
void launch_qmmm_counts(const int* type, const int* count, int num_nodes, int* near_cnt, int* far_cnt, hipStream_t s);
// Writes 1-based site indices of close leaves (reference leaf order) and the
// far boxes' pre-order ids (0-based) in node order.
void launch_qmmm_compact(const int* type, const TreeDev& t, int num_nodes, const int* near_off, const int* far_off,
                         int* near_sites, int* far_nodes, hipStream_t s);

// ---------------------------------------------------------------- primitives (rocPRIM)
size_t sort_pairs_u64_temp(int n);
void sort_pairs_u64(void* temp, size_t temp_bytes, const unsigned long long* kin, unsigned long long* kout,
                    const int* vin, int* vout, int n, int end_bit, hipStream_t s);
size_t sort_keys_u64_temp(int n);
void sort_keys_u64(void* temp, size_t temp_bytes, const unsigned long long* kin, unsigned long long* kout, int n,
                   int end_bit, hipStream_t s);
size_t exscan_i32_temp(int n);
void exscan_i32(void* temp, size_t temp_bytes, const int* in, int* out, int n, hipStream_t s);
// The same scan in one workgroup (one launch; for n up to EXSCAN_SMALL_MAX).
constexpr int EXSCAN_SMALL_MAX = 1 << 16;
void exscan_i32_small(const int* in, int* out, int n, hipStream_t s);

}  // namespace fmmh
