/*
 * fmm_hip.h -- C ABI of the MI355X (gfx950) field engine.
 *
 * Drop-in for the field engine of peter-reinholdt/fmm: the Fortran module
 * procedures `field_direct` (src/fmm.F90:1338-1352) and `field_fmm`
 * (src/fmm.F90:1354-1400).  A Fortran caller uses the module `fmm` of
 * fmm_amd/fortran/fmm_hip.F90, which keeps those signatures and forwards
 * here through iso_c_binding; Python callers use fmm_amd.octree, which keeps
 * the f2py `octree` signatures (CMakeLists.txt:21-39).
 *
 * Conventions shared by every entry point
 *   - Arrays are column-major, as in Fortran.  coordinates is n x 3,
 *     multipoles n x dim (dim = (m+1)(m+2)(m+3)/6 packed q,x,y,z,xx,xy,...),
 *     exclusions n x nexcl (1-based site indices, 0 = padding), field
 *     n x (F+1)(F+2)(F+3)/6.  ld_* is the leading dimension (>= n) of each.
 *   - Orders: source multipole order m <= 6 (dim <= 84) and field order
 *     F <= 6 with m + F <= 7 (src/fmm.F90:567-577 takes any; other orders
 *     return FMMH_ERR_UNSUPPORTED).  The damped kinds contract tensors up to
 *     order 6 only, as the reference, whose damped tensors stop at T6
 *     (src/tensors_damped_thole.F90:55-62).
 *   - field is ACCUMULATED into (field += result), like the reference, whose
 *     intent(out) dummy is never zeroed (src/fmm.F90:731, 813, 1113).
 *   - Exclusions are honoured the reference's way: an excluded pair is
 *     evaluated at displacement (1e10, 1e10, 1e10) (src/fmm.F90:714-721,
 *     802-806).  Every site must list itself or its field is not finite,
 *     exactly as in the reference.
 *   - Damping (damp_type) applies to near-field pairs only; far-field
 *     translations are undamped (src/fmm.F90:872-873).
 *   - comm is accepted and ignored: one rank, one GPU.
 *   - Calls are synchronous.  A context owns one HIP stream (default
 *     priority) and pooled device memory; ctx = NULL uses a per-process
 *     default context on the current device.  Calls on one context must not
 *     overlap.
 *   - Return value: FMMH_OK or an error code; fmmh_last_error() gives the
 *     message of the calling thread's last failure.
 */
#ifndef FMM_HIP_H
#define FMM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fmmh_status {
    FMMH_OK = 0,
    FMMH_ERR_ARG = 1,          /* invalid sizes, leading dimensions, exclusion index out of range */
    FMMH_ERR_DAMP = 2,         /* "Unknown damp type: " -- src/fmm.F90:594, 1372-1377 */
    FMMH_ERR_UNSUPPORTED = 3,  /* orders outside the compiled range, displacements beyond the class keys */
    FMMH_ERR_HIP = 4,          /* HIP runtime failure (allocation, launch) */
    FMMH_ERR_TREE = 5,         /* >= ncrit sites share one cell 21 levels down */
    FMMH_ERR_INTERNAL = 6
};

enum fmmh_damp {
    FMMH_DAMP_NONE = 0,        /* ''       */
    FMMH_DAMP_THOLE = 1,       /* 'THOLE'  */
    FMMH_DAMP_AMOEBA = 2,      /* 'AMOEBA' */
    FMMH_DAMP_ERF = 3          /* 'ERF'    */
};

typedef struct fmmh_context fmmh_context;

/* Context on HIP device `device` (-1: current device). */
int fmmh_context_create(fmmh_context** ctx, int32_t device);
int fmmh_context_destroy(fmmh_context* ctx);

/* Message of the calling thread's last failed call ("" if none). */
const char* fmmh_last_error(void);

/* damp_type as the reference compares it: trim(damp_type) against 'THOLE',
 * 'AMOEBA', 'ERF', '' (case-sensitive, trailing blanks ignored;
 * src/fmm.F90:585-595).  s need not be NUL-terminated.  Returns
 * FMMH_ERR_DAMP with "Unknown damp type: <s>" for anything else. */
int fmmh_parse_damp_type(const char* s, int64_t len, int32_t* code);

/* Replaces field_direct (src/fmm.F90:1338-1352; all-pairs sum of
 * particle_fields_direct, src/fmm.F90:749-824).  Host arrays.
 * damping_factors needs n_damping >= n when damp_type != FMMH_DAMP_NONE. */
int fmmh_field_direct(fmmh_context* ctx, int32_t comm,
                      const double* coordinates, int64_t ld_coordinates,
                      const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                      const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                      int32_t field_order, double* field, int64_t ld_field,
                      int32_t damp_type, const double* damping_factors, int64_t n_damping);

/* Replaces field_fmm (src/fmm.F90:1354-1400).  Host arrays.  n <= ncrit
 * falls back to the direct sum (src/fmm.F90:1381-1383).
 * is_periodic != 0: periodic boundaries as the reference implements them --
 * 27 image cells displaced by box_matrix (3 x 3, column-major, host memory in
 * every variant) through the interaction lists (exclusions in the central
 * cell only), the lattice sum beyond them on the root expansion and the
 * dipole correction (src/fmm.F90:884-979).  Needs n > ncrit ("Not enough
 * coordinates for periodic FMM, please decrease ncrit") and
 * expansion_order >= 2.  box_matrix may be NULL when is_periodic == 0. */
int fmmh_field_fmm(fmmh_context* ctx, int32_t comm,
                   const double* coordinates, int64_t ld_coordinates,
                   const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                   const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                   double theta, int32_t ncrit, int32_t expansion_order,
                   int32_t field_order, double* field, int64_t ld_field,
                   int32_t damp_type, const double* damping_factors, int64_t n_damping,
                   int32_t is_periodic, const double* box_matrix);

/* Device-resident forms of the two calls above: identical arguments, but
 * every array pointer except box_matrix is device memory of the context's device.  No host
 * copies; an exclusion entry above n is found by the exclusion pass, skipped, and reported
 * as FMMH_ERR_ARG after the call (the field is then not meaningful).  Used when the inputs
 * are already in HBM (benchmarks, GPU-side callers). */
int fmmh_field_direct_dev(fmmh_context* ctx, int32_t comm,
                          const double* coordinates, int64_t ld_coordinates,
                          const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                          const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                          int32_t field_order, double* field, int64_t ld_field,
                          int32_t damp_type, const double* damping_factors, int64_t n_damping);

int fmmh_field_fmm_dev(fmmh_context* ctx, int32_t comm,
                       const double* coordinates, int64_t ld_coordinates,
                       const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                       const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                       double theta, int32_t ncrit, int32_t expansion_order,
                       int32_t field_order, double* field, int64_t ld_field,
                       int32_t damp_type, const double* damping_factors, int64_t n_damping,
                       int32_t is_periodic, const double* box_matrix);

/* fmmh_field_direct_dev without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Same arguments and results as fmmh_field_direct_dev (bit for bit), one
 * multipole set; ctx must be a context made by fmmh_context_create (NULL is
 * refused).  The working arrays live in a fixed buffer of
 * FMMH_NOALLOC_SCRATCH_BYTES per device, reserved when the library is loaded
 * on a device and held until the process ends; calls whose working arrays do
 * not fit return FMMH_ERR_UNSUPPORTED (about 128 bytes per site for
 * quadrupoles and 3 exclusions, plus 8 per further multipole component and 4
 * per further exclusion: ~260,000 quadrupole sites).  Every call on one
 * device uses the same buffer, whatever the context: calls that overlap on
 * one device race.  Kernels of source / field orders above 2 use per-lane
 * private memory, which the HIP runtime reserves on their first launch. */
#define FMMH_NOALLOC_SCRATCH_BYTES (32ll << 20)
int fmmh_field_direct_dev_noalloc(fmmh_context* ctx, int32_t comm,
                                  const double* coordinates, int64_t ld_coordinates,
                                  const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                                  const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                                  int32_t field_order, double* field, int64_t ld_field,
                                  int32_t damp_type, const double* damping_factors, int64_t n_damping);

/* Many multipole sets on one set of sites: nsets independent field_direct /
 * field_fmm calls (same coordinates, exclusions, damping and parameters) in
 * one call that builds the tree and interaction lists once.  Replaces a loop
 * of nsets calls of src/fmm.F90:1338-1352 / 1354-1400 (e.g. the induced-
 * dipole iterations and response vectors of a polarizable embedding, which
 * call field_fmm repeatedly on fixed coordinates).
 * multipoles is n x dim x nsets and field n x nf x nsets, column-major with
 * the one leading dimension ld_multipoles / ld_field (set s starts at
 * s * dim * ld_multipoles / s * nf * ld_field).  Each set's field equals the
 * single-set call's to rounding (the M2L of >= 16 sets runs on the FP64
 * matrix cores, in another summation order). */
int fmmh_field_direct_sets(fmmh_context* ctx, int32_t comm,
                           const double* coordinates, int64_t ld_coordinates,
                           const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                           const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                           int32_t field_order, double* field, int64_t ld_field,
                           int32_t damp_type, const double* damping_factors, int64_t n_damping);

int fmmh_field_direct_sets_dev(fmmh_context* ctx, int32_t comm,
                               const double* coordinates, int64_t ld_coordinates,
                               const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                               int32_t nsets, const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                               int32_t field_order, double* field, int64_t ld_field,
                               int32_t damp_type, const double* damping_factors, int64_t n_damping);

int fmmh_field_fmm_sets(fmmh_context* ctx, int32_t comm,
                        const double* coordinates, int64_t ld_coordinates,
                        const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                        const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                        double theta, int32_t ncrit, int32_t expansion_order,
                        int32_t field_order, double* field, int64_t ld_field,
                        int32_t damp_type, const double* damping_factors, int64_t n_damping,
                        int32_t is_periodic, const double* box_matrix);

int fmmh_field_fmm_sets_dev(fmmh_context* ctx, int32_t comm,
                            const double* coordinates, int64_t ld_coordinates,
                            const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                            const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                            double theta, int32_t ncrit, int32_t expansion_order,
                            int32_t field_order, double* field, int64_t ld_field,
                            int32_t damp_type, const double* damping_factors, int64_t n_damping,
                            int32_t is_periodic, const double* box_matrix);

/* ---- QM/MM embedding (src/fmm.F90:1134-1336) ---------------------------------
 * The reference returns allocatable arrays; here each query is two calls on
 * one context: the first builds the octree of `coordinates` (n x 3), runs
 * classify_box against `target_coordinates` (ntargets x 3; a node is too
 * close if norm2(target - centre) * theta <= rmax for any target, and close
 * internal nodes are descended, src/fmm.F90:1134-1169) and returns
 * sizes[2] = {num_near_field, num_far_field}; the export call then fills
 * caller arrays of those sizes.  All arrays are host memory, column-major.
 *   idx_near_field  num_near_field sites (1-based) of the close leaves, in
 *                   node order and the reference's leaf member order
 *   box_coordinates num_far_field x 3 centres of the far boxes, node order
 *   box_indices     num_far_field 1-based node indices (pre-order), the
 *                   numbering interpolate_electron_fields expects
 *   box_multipoles  num_far_field x (p+1)(p+2)(p+3)/6 x ndens cell multipoles
 *                   of the far boxes (P2M + M2M about the box centres) */

/* Replaces get_boxes (src/fmm.F90:1231-1291). */
int fmmh_get_boxes(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates, int32_t n,
                   const double* target_coordinates, int64_t ld_targets, int32_t ntargets, double theta,
                   int32_t ncrit, int32_t expansion_order, int64_t* sizes);
int fmmh_get_boxes_export(fmmh_context* ctx, int32_t* idx_near_field, double* box_coordinates,
                          int32_t* box_indices);

/* Replaces get_multipoles (src/fmm.F90:1171-1229); multipoles is
 * n x dim x ndens (one leading dimension), dim = (m+1)(m+2)(m+3)/6, m <= 6. */
int fmmh_get_multipoles(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                        int32_t n, const double* multipoles, int64_t ld_multipoles, int32_t dim, int32_t ndens,
                        const double* target_coordinates, int64_t ld_targets, int32_t ntargets, double theta,
                        int32_t ncrit, int32_t expansion_order, int64_t* sizes);
int fmmh_get_multipoles_export(fmmh_context* ctx, int32_t* idx_near_field, double* box_coordinates,
                               double* box_multipoles);

/* Replaces interpolate_electron_fields (src/fmm.F90:1293-1336): local
 * expansions (nbox x (p+1)(p+2)(p+3)/6 x ndens) placed on the boxes
 * box_indices (1-based, as get_boxes returns them; a later row for the same
 * box wins), every other box zero, then L2L down the octree of `coordinates`
 * and L2P into field (n x nf x ndens, accumulated, nf = ncomp(field_order)). */
int fmmh_interpolate_electron_fields(fmmh_context* ctx, int32_t comm, const double* coordinates,
                                     int64_t ld_coordinates, int32_t n, const double* local_expansion,
                                     int64_t ld_local_expansion, int32_t nbox, int32_t ndens,
                                     const int32_t* box_indices, double theta, int32_t ncrit,
                                     int32_t expansion_order, int32_t field_order, double* field,
                                     int64_t ld_field);

/* ---- diagnostics (no counterpart in the reference) -------------------------
 * Spot-check direct sum: the field_direct result (src/fmm.F90:749-824) at the
 * ntargets sites targets[] (0-based, device memory like every other array),
 * accumulated into field (ntargets x nf, leading dimension ld_field).  Costs
 * ntargets x n pairs; used to state FMM errors at sizes where the full direct
 * sum is too costly. */
int fmmh_field_direct_targets_dev(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates,
                                  const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                                  const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                                  const int32_t* targets, int32_t ntargets, int32_t field_order, double* field,
                                  int64_t ld_field, int32_t damp_type, const double* damping_factors,
                                  int64_t n_damping);

/* Octree and interaction lists of field_fmm for the given sites (host
 * coordinates), exported for bit-exact comparison with the reference's
 * tree (src/fmm.F90:137-385).  fmmh_tree_build keeps the result in the
 * context and returns sizes in info[5] = {num_nodes, num_leaves, max_depth,
 * n_m2l, n_p2p}; fmmh_tree_export then fills caller arrays (node ids are
 * pre-order, 0-based; centers num_nodes x 3 column-major; parent -1 for the
 * root; nleaf = sites in the subtree; members/mem_off list each leaf's sites
 * (0-based) in the reference's order, offsets num_nodes + 1 with empty ranges
 * for internal nodes; m2l/p2p lists as CSR by target node, ascending source). */
int fmmh_tree_build(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates, int32_t n,
                    int32_t ncrit, double theta, int64_t* info);
int fmmh_tree_export(fmmh_context* ctx, double* centers, double* r, double* rmax, int32_t* depth,
                     int32_t* parent, int32_t* is_leaf, int32_t* nleaf, int32_t* mem_off, int32_t* members,
                     int32_t* m2l_off, int32_t* m2l_src, int32_t* p2p_off, int32_t* p2p_src);

/* Per-phase device times (ms) of the context's last call, measured with HIP
 * events on the context stream when timing is enabled.  names receives a
 * ';'-separated list of phase names (capacity name_cap bytes).  Returns the
 * number of phases written. */
enum fmmh_option {
    FMMH_OPT_TIMING = 1,          /* value != 0: record per-phase HIP events */
    FMMH_OPT_LIST_CAPACITY = 2,   /* scale of the initial interaction-list buffers (default 1; they grow on overflow) */
    FMMH_OPT_M2L_MFMA = 3,        /* M2L engine.  0: per-target contraction on the vector ALUs;
                                     1 (default): per-target FP64 MFMA over sets from 6 sets on (orders
                                     <= 8), else the class-grouped FP64 MFMA GEMM (orders <= 10) --
                                     accumulated per block of 64 targets on the chip for FP64 at
                                     orders <= 6 without learned orders, else through per-pair slots;
                                     2: per-target MFMA for any set count; 3: class-grouped always;
                                     4: class-grouped with per-pair slots always */
    FMMH_OPT_REUSE_TREE = 4,      /* value != 0: keep the last field_fmm octree, interaction lists and M2L
                                     class tensors, and reuse them while the coordinates (compared bitwise
                                     on the device), n, ncrit and theta are unchanged -- e.g. the induced-
                                     dipole iterations of a polarizable embedding, for which the reference
                                     rebuilds every call (src/fmm.F90:1387-1398).  Results are identical
                                     to a rebuild.  0 (default) drops the cache. */
    FMMH_OPT_PRECISION = 5,       /* 0 (default): FP64 throughout.  1: mixed -- the M2M, M2L and L2L
                                     translations run on FP32 MFMA with FP32 accumulation and FP64 sums
                                     into the expansions; P2M, L2P, the near field (with its damped
                                     tensors) and the field stay FP64.  Results then differ from the
                                     reference at the FP32 level. */
    /* Learned per-pair expansion orders (no counterpart in the reference, whose M2L applies
     * expansion_order to every pair).  With a trained selector on the context
     * (fmmh_selector_train / fmmh_selector_set) and ORDER_TOL > 0, field_fmm picks an order
     * q <= expansion_order for every M2L pair so that the far field of every multipole set keeps
     * a relative RMS error of about ORDER_TOL; without a trained selector, for periodic calls and
     * for source multipole orders above 2, every pair keeps expansion_order. */
    FMMH_OPT_ORDER_TOL = 6,       /* requested relative RMS field error of the far field (0 = off, default) */
    FMMH_OPT_ORDER_MARGIN = 7,    /* added to the predicted order before rounding up (default 0.25) */
    FMMH_OPT_SPOT_CHECKS = 8,     /* > 0: after the call, compare the far field at this many sites (evenly
                                     spaced in the octree's site order) against direct sums with the
                                     reference's far semantics (undamped, no exclusions), set by set; if
                                     the relative RMS error of any set exceeds ORDER_TOL, redo M2L..L2P
                                     at expansion_order (fallback).  0 (default): no check. */
    FMMH_OPT_ORDER_SAFETY = 9,    /* factor on the per-pair tolerance (default 1) */
    FMMH_OPT_ORDER_FORCE_FULL = 10, /* != 0: run the selection machinery but keep every pair at
                                     expansion_order (ablation) */
    FMMH_OPT_AUTOTUNE = 11,       /* 1 (default): the first call of each M2L shape for many sets (order,
                                     precision, sets per pass, learned orders) times every tile-shape
                                     variant of that kernel on its own pairs and keeps the fastest for the
                                     context's later calls (results do not depend on the choice).  0: fixed
                                     defaults.  -1: forget the choices made so far and keep tuning on. */
    FMMH_OPT_GRAPH = 12,          /* != 0: tree-reuse calls (FMMH_OPT_REUSE_TREE) without timing, learned
                                     orders or periodic images run their device work as a HIP graph: the
                                     first call of a shape runs eagerly, the next one captures and launches
                                     the graph, later ones replay it.  A shape is every pointer, size and
                                     option of the call; any reallocation of the context's buffers or
                                     rebuild of its cached tree data starts over.  Results are identical
                                     to eager calls.  0 (default): eager. */
    FMMH_OPT_NEAR_CHUNK = 13      /* near-field work split: a leaf whose pair work exceeds a cap (in
                                     source-tile iterations of one wavefront) runs as several workgroups
                                     over consecutive parts of its source list, whose partial fields are
                                     then summed in a fixed order.  0 (default): cap = total work / 16384,
                                     at least 256.  > 0: that cap (tests, tuning). */
};
int fmmh_set_option(fmmh_context* ctx, int32_t key, double value);
int fmmh_phase_times(fmmh_context* ctx, double* ms, int32_t cap, char* names, int32_t name_cap);

/* Copy `bytes` from the start of the context's pooled device buffer `name`
 * (an internal name such as "e_cm", "e_le", "m_tc") to host memory, for
 * white-box tests of intermediate results.  Fails if the buffer is unknown
 * or smaller than `bytes`. */
int fmmh_debug_buffer(fmmh_context* ctx, const char* name, void* host, int64_t bytes);

/* Counters of the last field_fmm call: {num_nodes, num_leaves, max_depth,
 * n_m2l, n_p2p, p2p_site_pairs, m2l_classes, list_retries, graph (0: eager,
 * 1: captured and launched, 2: replayed; FMMH_OPT_GRAPH), graph replays so far,
 * near-field workgroups (leaves plus extra chunks; FMMH_OPT_NEAR_CHUNK),
 * host waits on the device during the call (stream or event synchronisations
 * before the result is complete; the readback of a host-array call is one more)}. */
int fmmh_stats(fmmh_context* ctx, int64_t* out, int32_t cap);

/* ---- learned per-pair expansion orders (no counterpart in the reference) ------
 * A multilayer perceptron (4 inputs, 2 hidden layers of 64 ReLU units, FP32) maps
 * an M2L pair (i, j) to the expansion order its contribution needs:
 *   x0 = r_i / |R|, x1 = r_j / |R|   (box half-widths over the centre distance R)
 *   x2 = log10(A_j / |R|^2)          (A_j = 2-norm of M_j(s) r_j^{-|s|}, |s| <= p, max over sets)
 *   x3 = log10(tau)                  (the pair's absolute field tolerance)
 * and the call uses q = ceil(y + margin) clamped to [0, expansion_order].
 *
 * fmmh_selector_train builds the octree of a sample system (host arrays,
 * column-major n x 3 coordinates and n x dim multipoles, dim 1/4/10), runs the
 * upward pass at order cap, samples nsamples M2L pairs, measures each pair's
 * truncation error at every order 0..cap against the direct sum over its source
 * box at up to 32 sites of its target box (on the GPU), turns the error curves
 * into examples, and trains the network on the GPU (Adam, asymmetric squared
 * loss that charges under-predicted orders 4x) for `epochs` passes.  info[8] =
 * {examples, samples, first-step loss, final loss, sampling ms, training ms,
 * steps, graph replays}. */
int fmmh_selector_train(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates,
                        const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, double theta,
                        int32_t ncrit, int32_t cap, int32_t nsamples, int32_t epochs, uint64_t seed, double* info);
/* Samples of the last training run: pairs[ns][2] (target, source node, 0-based
 * pre-order ids of the octree of the training system), features[ns][4] (x3 = 0),
 * errors[ns][cap + 1], field_rms[ns] (RMS exact field at the sampled sites).
 * Any pointer may be NULL. */
int fmmh_selector_samples(fmmh_context* ctx, int32_t* pairs, float* features, double* errors, double* field_rms);
/* Per-step minibatch loss of the last training run; returns the step count. */
int fmmh_selector_loss(fmmh_context* ctx, float* loss, int32_t cap);
int32_t fmmh_selector_nparam(void);
/* Trained weights (fmmh_selector_nparam() floats: W1[64][4], b1[64], W2[64][64],
 * b2[64], W3[64], b3), feature normalisation norm[8] (mean[4], 1/std[4]) and
 * meta[4] = {cap, lowest x2 - x3, highest x2 - x3 of the training set, margin};
 * pairs with x2 - x3 above that range keep expansion_order. */
int fmmh_selector_get(fmmh_context* ctx, float* weights, float* norm, double* meta);
int fmmh_selector_set(fmmh_context* ctx, const float* weights, const float* norm, const double* meta);
/* Loss and gradient of the training objective for one batch of ns examples
 * (features[ns][4], targets[ns]) at the given weights: grad_loss[nparam] =
 * gradient, grad_loss[nparam] = mean loss.  Same kernels as training. */
int fmmh_selector_grad(fmmh_context* ctx, const float* weights, const float* norm, const float* features,
                       const float* targets, int32_t ns, float asym, float* grad_loss);
/* Network outputs y for ns raw feature rows (the inference forward pass). */
int fmmh_selector_predict(fmmh_context* ctx, const float* weights, const float* norm, const float* features,
                          int32_t ns, float* out);
/* Selection figures of the last field_fmm call: out[0] spot-check relative
 * error, the largest over the multipole sets (-1: not checked), [1] 1 if it fell
 * back to expansion_order, [2] per-pair tolerance tau, [3] M2L flops at the chosen
 * orders (2 (q + 1)^4 per pair and set: the harmonic form's Hankel matrix), [4] M2L
 * flops at expansion_order, [5] mean
 * order, [6] 1 if selection ran, [7] RMS near field used for tau (the smallest
 * non-zero one over the sets), [8 + q] pairs at order q (q = 0..10), [19] the
 * context's tolerance scale for later calls (1; after a fallback whose full-order
 * result meets the tolerance, multiplied by about tol / error, at most 1/2 and at
 * least 1/8, floor 1/256; reset to 1 by setting FMMH_OPT_ORDER_TOL or
 * FMMH_OPT_ORDER_SAFETY and by training or loading weights), [20] the spot-check
 * error of the full-order result after a fallback (-1: no fallback). */
int fmmh_selection_stats(fmmh_context* ctx, double* out, int32_t cap);

/* Tile-shape choices of the autotuner (FMMH_OPT_AUTOTUNE): up to cap (key, variant) pairs, key = order << 32 |
 * fp32 << 24 | sets << 8 | learned-orders, variant bit 0 = 32 sets per workgroup, bit 1 = source blocks from
 * global memory instead of LDS; last_ms[4] = the times of the last tuning (-1: variant not valid).  Returns the
 * number of choices held. */
int fmmh_tuning(fmmh_context* ctx, int64_t* keys, int32_t* variants, double* last_ms, int32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* FMM_HIP_H */
