// Host runtime and C ABI of the field engine (include/fmm_hip.h).
//
// Pipeline of one field_fmm call (src/fmm.F90:1354-1400), all on the
// context's stream:
//   tree     centroid -> keys -> sort -> levels -> scan -> nodes -> links
//            -> leaves -> level lists -> leaf member order      (2 host syncs)
//   lists    level-synchronous dual traversal, sorted into CSR  (1 host sync)
//   M2L      displacement classes, class tensors, Hankel contraction (1 sync)
//   gather   sorted SoA site arrays, Taylor-scaled multipoles
//   P2M, M2M (level by level), P2P, M2L, L2L (level by level), L2P + scatter
// Device buffers are pooled per context and only grow.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fmm_hip.h"
#include "common.h"
#include "kernels.h"

using namespace fmmh;

namespace {

thread_local std::string g_last_error;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct Tables {
    int p = -1, m = -1;
    int* ibuf = nullptr;
    double* dbuf = nullptr;
    float* fbuf = nullptr;
    const float* m2m_frag = nullptr;  // mixed precision: M2M / L2L matrices per octant, MFMA fragment order
    const float* l2l_frag = nullptr;
    const float* l2l_frag_red = nullptr;   // L2L of harmonic expansions on the (p + 1)^2 reduced components
    ShiftTab p2m{}, m2m{}, l2l{};
    const int* mono_parent = nullptr;
    const int* mono_axis = nullptr;
    const int* xyz = nullptr;
    const int* level_base = nullptr;
    const double* cs = nullptr;       // Taylor factors for the p expansion
    const double* scale = nullptr;    // Taylor factors for the dim source components
    int nsrc = 0;
    const int* gemm_fidx = nullptr;   // class-grouped / block M2L: A-fragment index table over the reduced
    const int* gemm_tdeg = nullptr;   //   (harmonic) components, |t| per reduced target row,
    const int* gemm_tfull = nullptr;  //   and the packed index of each reduced component
    FoldTab fold{};                   // many-sets source fold, term by term (null off: p beyond the table)
};

struct TreeOut {
    TreeDev t{};
    int num_nodes = 0, num_leaves = 0, max_depth = 0;
    std::vector<int> level_off;   // KEY_LEVELS + 3 entries; level L nodes at [off[L], off[L+1])
};

struct ListsOut {
    int* m2l_off = nullptr; int* m2l_src = nullptr; int n_m2l = 0;
    unsigned long long* m2l_keys = nullptr;   // sorted (target << 32 | source), CSR order
    int* p2p_off = nullptr; int* p2p_src = nullptr; int n_p2p = 0;
};


}  // namespace

struct fmmh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<std::string, DevBuf> bufs;
    int* h_pinned = nullptr;          // small pinned readback area
    std::map<long long, Tables> tables;
    bool timing = false;
    std::vector<std::pair<std::string, hipEvent_t>> marks;
    std::vector<hipEvent_t> event_pool;
    size_t event_used = 0;
    std::vector<std::pair<std::string, double>> phase_ms;
    int64_t stats[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int64_t host_waits = 0;           // host waits on the device during the current call
    int64_t stats_classes = 0;
    double list_cap_scale = 1.0;      // initial interaction-list capacity per node (tests force retries)
    double list_need[3] = {0, 0, 0};  // frontier capacity that sufficed, M2L and P2P entries (x 1.25), per node
    int m2l_mfma = 1;                 // FMMH_OPT_M2L_MFMA: 0 VALU, 1 auto, 2 per-target MFMA, 3 class-grouped,
                                      // 4 class-grouped with per-pair slots
    int precision = 0;                // FMMH_OPT_PRECISION: 0 FP64, 1 mixed (FP32 MFMA translations)
    // Tree / list cache (FMMH_OPT_REUSE_TREE): the last field_fmm tree, lists
    // and M2L classes, reused while coordinates (bitwise), n, ncrit and theta
    // stay the same.  The device data live in the named pool buffers, which
    // only build_tree / build_lists / the M2L class pass write.
    bool reuse_tree = false;
    bool cache_valid = false;
    int cache_n = 0, cache_ncrit = 0;
    double cache_theta = 0.0;
    bool cache_periodic = false;
    bool shells_valid = false;        // periodic shell tensor sums in "p_shells" for (shells_p, shells_box)
    int shells_p = -1;
    double shells_box[9] = {};
    double cache_box[9] = {};
    TreeOut cache_tr;
    ListsOut cache_L;
    int cache_class_p = -1;           // expansion order of the cached class tensors (-1: none)
    int class_zmax = 0;               // their components with z > class_zmax are not computed
    M2LClassArgs class_args{};        // the arguments they were computed with
    bool cache_work_valid = false;    // near-field block order of the cached tree ("s_work_blk", "s_inv")
    int near_nv = 0, near_grid = 0, near_np = 0;   // near-field chunks, workgroup slots, partial sites
    long long near_chunk_cap = 0;     // FMMH_OPT_NEAR_CHUNK (0: automatic)
    int cache_tiles_p = -1;           // class-grouped GEMM tile tables of the cached classes (-1: none)
    bool cache_block_plan = false;    // target-block M2L plan of the cached classes
    bool classes_grouped = false;     // pairs sorted by class ("m_idx_sorted", "m_flag", "m_cls") for the classes
    int class_vbits = 0;              // bits per displacement component of the non-periodic class keys
    bool dense_marks_ready = false;   // "m_dense_marks" holds zeros
    int num_cu = 0;                   // compute units of the device
    // Graph replay (FMMH_OPT_GRAPH): a tree-reuse call runs as a captured hipGraph of its device work
    // once a previous call of the same shape ran eagerly (tuning done, buffers allocated).  The shape key
    // holds every pointer, size and option the captured work depends on, the pool generation (bumped by
    // any buffer reallocation, so a replay never sees a freed address) and the cache generation (bumped
    // whenever the tree, the M2L classes, their tile tables or the near-field block order are rebuilt,
    // so a replay never skips a rebuild its captured work relied on).
    int graph_opt = 0;
    std::string graph_key, graph_pending;
    hipGraphExec_t graph_exec = nullptr;
    bool capturing = false;
    int64_t pool_gen = 0;
    int64_t cache_gen = 0;
    int64_t graph_replays = 0;
    int64_t cache_classes = 0;
    unsigned long long cache_pairs = 0;
    // last get_boxes / get_multipoles classification, kept for the export call
    int qm_num_near = -1, qm_num_far = 0, qm_nm = 0, qm_ndens = 0, qm_num_nodes = 0;
    bool qm_multipoles = false;
    int cache_retries = 0;
    int list_retries = 0;
    // Learned per-pair expansion orders: the trained MLP (device buffers
    // "sel_W", "sel_norm"), its training range, and the per-call options.
    bool sel_trained = false;
    int sel_cap = 0;                  // expansion order the selector was trained at
    float sel_rel_lo = 0.f, sel_rel_hi = 0.f;   // range of x2 - x3 seen in training
    float sel_margin = 0.25f;         // FMMH_OPT_ORDER_MARGIN
    double order_tol = 0.0;           // FMMH_OPT_ORDER_TOL: requested relative RMS field error (0: off)
    double sel_kappa = 1.0;           // FMMH_OPT_ORDER_SAFETY
    double sel_adapt = 1.0;           // tolerance scale: tightened after each fallback, reset by the options above
    double sel_full_err = -1.0;       // last call: spot-check error of the full-order result after a fallback
    int spot_checks = 0;              // FMMH_OPT_SPOT_CHECKS
    int sel_force_full = 0;           // FMMH_OPT_ORDER_FORCE_FULL: selection machinery on, every pair at p
    // Tile-shape autotuner (FMMH_OPT_AUTOTUNE): the many-sets M2L variant chosen per
    // (order, precision, sets in the chunk, learned orders on), timed on the first call that needs it.
    int autotune = 1;
    std::map<long long, int> tuned;
    std::vector<double> tune_ms;      // last tuning: milliseconds per variant (-1: not valid)
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    // last call with selection: spot-check error, fallback flag, tolerance, order histogram
    double sel_stats[8] = {-1, 0, 0, 0, 0, 0, 0, 0};
    unsigned long long sel_hist[PMAX + 1] = {};
    // last training run
    std::vector<float> train_loss;
    int train_nsamples = 0, train_cap = 0;
    std::vector<int> train_pairs;     // [nsamples][2] target, source node (pre-order ids)
    std::vector<float> train_feat;    // [nsamples][SEL_IN]
    std::vector<double> train_err;    // [nsamples][cap + 1]
    std::vector<double> train_fref;   // [nsamples]
    // last tree (for export)
    int num_nodes = 0, num_leaves = 0, max_depth = 0, n_m2l = 0, n_p2p = 0, n_sites = 0;
    std::vector<int> level_off;       // host copy, size KEY_LEVELS + 3

    template <class T>
    T* buf(const std::string& name, size_t count) {
        DevBuf& b = bufs[name];
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (bytes > b.cap) {
            if (b.p) FMMH_HIP(hipFree(b.p));
            b.p = nullptr;
            size_t nb = std::max(bytes, b.cap + b.cap / 2);
            FMMH_HIP(hipMalloc(&b.p, nb));
            b.cap = nb;
            ++pool_gen;
        }
        return (T*)b.p;
    }

    void mark(const char* name) {
        if (!timing) return;
        if (event_used == event_pool.size()) {
            hipEvent_t e;
            FMMH_HIP(hipEventCreate(&e));
            event_pool.push_back(e);
        }
        hipEvent_t e = event_pool[event_used++];
        FMMH_HIP(hipEventRecord(e, stream));
        marks.emplace_back(name, e);
    }
    void begin_timing() {
        marks.clear();
        event_used = 0;
        phase_ms.clear();
        mark("start");
    }
    void end_timing() {
        if (!timing) return;
        FMMH_HIP(hipStreamSynchronize(stream));
        for (size_t i = 1; i < marks.size(); ++i) {
            float ms = 0.f;
            FMMH_HIP(hipEventElapsedTime(&ms, marks[i - 1].second, marks[i].second));
            phase_ms.emplace_back(marks[i].first, (double)ms);
        }
    }
    void sync() {
        ++host_waits;
        FMMH_HIP(hipStreamSynchronize(stream));
    }
    // An event the host waits on for a readback while later work stays queued behind it.
    hipEvent_t rb_event = nullptr;
    hipEvent_t readback_event() {
        if (!rb_event) FMMH_HIP(hipEventCreateWithFlags(&rb_event, hipEventDisableTiming));
        return rb_event;
    }
    void graph_exec_replace(hipGraph_t g) {
        if (graph_exec) {
            FMMH_HIP(hipGraphExecDestroy(graph_exec));
            graph_exec = nullptr;
        }
        const hipError_t e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
        FMMH_HIP(hipGraphDestroy(g));
        FMMH_HIP(e);
    }
};

// Stream capture of a region: begins on construction when asked; end() returns the graph; if the
// region throws, the destructor ends the capture and drops it, so the stream is usable again.
struct GraphCapture {
    hipStream_t s;
    bool active;
    bool* flag;   // true while capturing (no host synchronisation, no autotuning)
    GraphCapture(hipStream_t st, bool on, bool* f) : s(st), active(on), flag(f) {
        if (active) {
            FMMH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            *flag = true;
        }
    }
    hipGraph_t end() {
        hipGraph_t g = nullptr;
        active = false;
        *flag = false;
        FMMH_HIP(hipStreamEndCapture(s, &g));
        return g;
    }
    ~GraphCapture() {
        if (active) {
            *flag = false;
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        }
    }
};

namespace {

std::mutex g_default_mutex;
fmmh_context* g_default = nullptr;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        f();
        return FMMH_OK;
    } catch (const fmmh::Error& e) {
        return fail(e.code, e.what());
    } catch (const std::exception& e) {
        return fail(FMMH_ERR_INTERNAL, e.what());
    }
}

// ---------------------------------------------------------------- fixed working memory
// fmmh_field_direct_dev_noalloc keeps its working arrays in this array, one copy per device that the
// library runs on, held until the process ends.  Host code never uses the name `scratch` as an address
// (on the host it is a stand-in, not device memory): the device address of each device's copy is taken
// once, with that device current, when the first context on the device is created, and kept in
// g_scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)FMMH_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The calling thread's current device must be `device`.
void scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw Error(FMMH_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    FMMH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(FMMH_ERR_HIP, "hipGetSymbolAddress returned no address for the fixed working memory");
    g_scratch_ptr[device] = ptr;
}

void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    void* ptr = device >= 0 && device < MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
    if (!ptr) throw Error(FMMH_ERR_INTERNAL, "fixed working memory of the device was not set up");
    return ptr;
}

fmmh_context* make_context(int device) {
    fmmh_context* c = new fmmh_context;
    if (device < 0) FMMH_HIP(hipGetDevice(&device));
    c->device = device;
    FMMH_HIP(hipSetDevice(device));
    FMMH_HIP(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device));
    scratch_setup(device);
    FMMH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    FMMH_HIP(hipHostMalloc((void**)&c->h_pinned, 4096, hipHostMallocDefault));
    return c;
}

fmmh_context* resolve(fmmh_context* ctx) {
    if (ctx) {
        FMMH_HIP(hipSetDevice(ctx->device));
        return ctx;
    }
    std::lock_guard<std::mutex> lk(g_default_mutex);
    if (!g_default) g_default = make_context(-1);
    FMMH_HIP(hipSetDevice(g_default->device));
    return g_default;
}

int order_from_dim(int dim) { return (int)std::lround(std::pow((double)dim * 6.0, 1.0 / 3.0)) - 2; }

double factorial(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

struct XYZ { int x, y, z; };
std::vector<XYZ> packed_components(int K) {
    std::vector<XYZ> v;
    for (int k = 0; k <= K; ++k)
        for (int x = k; x >= 0; --x)
            for (int y = k - x; y >= 0; --y) v.push_back({x, y, k - x - y});
    return v;
}

double taylor_factor(const XYZ& s) {
    int k = s.x + s.y + s.z;
    return ((k + 1) % 2 ? -1.0 : 1.0) / (factorial(s.x) * factorial(s.y) * factorial(s.z));
}

// Host-built operator tables for expansion order p and source multipole order m.
Tables& get_tables(fmmh_context* c, int p, int m) {
    long long key = (long long)p * 64 + m;
    auto it = c->tables.find(key);
    if (it != c->tables.end()) return it->second;
    const int nm = ncomp(p);
    const int K = 2 * p;
    // the packed-index table xyz also serves the field components of L2P (orders <= NEAR_FMAX)
    const int KX = std::max(K, NEAR_FMAX);
    std::vector<XYZ> comp = packed_components(std::max(KX, m));
    auto idx = [](int x, int y, int z) { return xyz2idx(x, y, z); };
    std::vector<int> ints;
    std::vector<double> dbl;
    struct Seg { size_t i0, d0; };
    auto shift_table = [&](bool up, int tmax_order, int smax_order, std::vector<int>& off, std::vector<int>& s,
                           std::vector<int>& e, std::vector<double>& coef) {
        int nt = ncomp(tmax_order);
        off.assign(nt + 1, 0);
        for (int t = 0; t < nt; ++t) {
            off[t] = (int)s.size();
            const XYZ& T = comp[t];
            for (int si = 0; si < ncomp(smax_order); ++si) {
                const XYZ& S = comp[si];
                if (up) {
                    if (S.x > T.x || S.y > T.y || S.z > T.z) continue;
                    s.push_back(si);
                    e.push_back(idx(T.x - S.x, T.y - S.y, T.z - S.z));
                    auto bin = [](int n, int k) { return factorial(n) / (factorial(k) * factorial(n - k)); };
                    coef.push_back(bin(T.x, S.x) * bin(T.y, S.y) * bin(T.z, S.z));
                } else {
                    if (S.x < T.x || S.y < T.y || S.z < T.z) continue;
                    s.push_back(si);
                    e.push_back(idx(S.x - T.x, S.y - T.y, S.z - T.z));
                    coef.push_back(1.0 / (factorial(S.x - T.x) * factorial(S.y - T.y) * factorial(S.z - T.z)));
                }
            }
        }
        off[nt] = (int)s.size();
    };
    std::vector<int> p2m_off, p2m_s, p2m_e, m2m_off, m2m_s, m2m_e, l2l_off, l2l_s, l2l_e;
    std::vector<double> p2m_c, m2m_c, l2l_c;
    int msrc = std::min(m, p);
    shift_table(true, p, msrc, p2m_off, p2m_s, p2m_e, p2m_c);
    shift_table(true, p, p, m2m_off, m2m_s, m2m_e, m2m_c);
    shift_table(false, p, p, l2l_off, l2l_s, l2l_e, l2l_c);
    std::vector<int> mono_parent(nm), mono_axis(nm);
    for (int e = 0; e < nm; ++e) {
        const XYZ& E = comp[e];
        if (e == 0) { mono_parent[e] = 0; mono_axis[e] = 0; }
        else if (E.x > 0) { mono_parent[e] = idx(E.x - 1, E.y, E.z); mono_axis[e] = 0; }
        else if (E.y > 0) { mono_parent[e] = idx(E.x, E.y - 1, E.z); mono_axis[e] = 1; }
        else { mono_parent[e] = idx(E.x, E.y, E.z - 1); mono_axis[e] = 2; }
    }
    std::vector<int> level_base(K + 2, 0);
    for (int k = 0; k <= K; ++k) level_base[k + 1] = level_base[k] + ndeg(k) * (K - k + 1);
    std::vector<int> xyz(ncomp(KX));
    for (int i = 0; i < ncomp(KX); ++i) xyz[i] = comp[i].x | (comp[i].y << 8) | (comp[i].z << 16);
    std::vector<double> cs(nm), scale(ncomp(m));
    for (int s = 0; s < nm; ++s) cs[s] = taylor_factor(comp[s]);
    for (int s = 0; s < ncomp(m); ++s) scale[s] = taylor_factor(comp[s]);
    // Class-grouped / block M2L fragment tables over the reduced (harmonic) components -- those with
    // z <= 1, by degree, then w = y + z, then z (m2l.hip): lane l of the A fragment of row tile mt and
    // source step ks holds H[t][s], t = 16 mt + l % 16, s = 4 ks + l / 16 (reduced indices).
    std::vector<int> gemm_fidx, gemm_tdeg, gemm_tfull, fold_off, fold_sc, fold_deg;
    std::vector<double> fold_w, fold_c;
    if (p <= M2L_GEMM_PMAX) {
        const int nk = (p + 1) * (p + 1);
        std::vector<XYZ> red;
        for (int n = 0; n <= p; ++n) {
            red.push_back(XYZ{n, 0, 0});
            for (int w = 1; w <= n; ++w) {
                red.push_back(XYZ{n - w, w, 0});
                red.push_back(XYZ{n - w, w - 1, 1});
            }
        }
        const int MT = (nk + 15) / 16, KS = (nk + 3) / 4;
        gemm_fidx.assign((size_t)MT * KS * 64, -1);
        gemm_tdeg.assign((size_t)MT * 16, 0);
        gemm_tfull.assign((size_t)nk, 0);
        for (int mt = 0; mt < MT; ++mt)
            for (int ks = 0; ks < KS; ++ks)
                for (int l = 0; l < 64; ++l) {
                    const int t = 16 * mt + l % 16, s = 4 * ks + l / 16;
                    if (t < nk && s < nk)
                        gemm_fidx[((size_t)mt * KS + ks) * 64 + l] =
                            idx(red[t].x + red[s].x, red[t].y + red[s].y, red[t].z + red[s].z);
                }
        for (int t = 0; t < nk; ++t) {
            gemm_tdeg[t] = red[t].x + red[t].y + red[t].z;
            gemm_tfull[t] = idx(red[t].x, red[t].y, red[t].z);
        }
        // the fold of c M onto reduced component (X, Y, rho): terms (-1)^k C(k, i) c M(X - 2i, Y - 2(k - i),
        // rho + 2k), k from 0, i from 0, those with both indices >= 0 (the order of the device fold)
        fold_off.assign(1, 0);
        for (int t = 0; t < nk; ++t) {
            const int X = red[t].x, Y = red[t].y, rho = red[t].z;
            for (int k = 0; 2 * k <= X + Y; ++k) {
                double binom = 1.0;
                for (int i = 0; i <= k; ++i) {
                    if (i > 0) binom = binom * (k - i + 1) / i;
                    const int sx = X - 2 * i, sy = Y - 2 * (k - i);
                    if (sx < 0 || sy < 0) continue;
                    const int sc = idx(sx, sy, rho + 2 * k);
                    fold_sc.push_back(sc);
                    fold_w.push_back(((k & 1) ? -1.0 : 1.0) * binom);
                    fold_c.push_back(cs[sc]);
                }
            }
            fold_off.push_back((int)fold_sc.size());
            fold_deg.push_back(X + Y + rho);
        }
    }

    // pack into one int and one double buffer
    std::vector<std::pair<std::vector<int>*, const int**>> iparts;
    Tables t;
    t.p = p; t.m = m; t.nsrc = ncomp(msrc);
    std::vector<std::vector<int>*> ilist = {&p2m_off, &p2m_s, &p2m_e, &m2m_off, &m2m_s, &m2m_e, &l2l_off, &l2l_s,
                                            &l2l_e, &mono_parent, &mono_axis, &xyz, &level_base, &gemm_fidx,
                                            &gemm_tdeg, &gemm_tfull, &fold_off, &fold_sc, &fold_deg};
    std::vector<std::vector<double>*> dlist = {&p2m_c, &m2m_c, &l2l_c, &cs, &scale, &fold_w, &fold_c};
    std::vector<size_t> ioff, doff;
    size_t ni = 0, nd = 0;
    for (auto* v : ilist) { ioff.push_back(ni); ni += v->size(); }
    for (auto* v : dlist) { doff.push_back(nd); nd += v->size(); }
    std::vector<int> hi(ni);
    std::vector<double> hd(nd);
    for (size_t k = 0; k < ilist.size(); ++k) std::copy(ilist[k]->begin(), ilist[k]->end(), hi.begin() + ioff[k]);
    for (size_t k = 0; k < dlist.size(); ++k) std::copy(dlist[k]->begin(), dlist[k]->end(), hd.begin() + doff[k]);
    FMMH_HIP(hipMalloc((void**)&t.ibuf, std::max<size_t>(ni, 1) * sizeof(int)));
    FMMH_HIP(hipMalloc((void**)&t.dbuf, std::max<size_t>(nd, 1) * sizeof(double)));
    FMMH_HIP(hipMemcpy(t.ibuf, hi.data(), ni * sizeof(int), hipMemcpyHostToDevice));
    FMMH_HIP(hipMemcpy(t.dbuf, hd.data(), nd * sizeof(double), hipMemcpyHostToDevice));
    const int* I = t.ibuf;
    const double* D = t.dbuf;
    t.p2m = ShiftTab{I + ioff[0], I + ioff[1], I + ioff[2], D + doff[0]};
    t.m2m = ShiftTab{I + ioff[3], I + ioff[4], I + ioff[5], D + doff[1]};
    t.l2l = ShiftTab{I + ioff[6], I + ioff[7], I + ioff[8], D + doff[2]};
    t.mono_parent = I + ioff[9];
    t.mono_axis = I + ioff[10];
    t.xyz = I + ioff[11];
    t.level_base = I + ioff[12];
    if (p <= M2L_GEMM_PMAX) {
        t.gemm_fidx = I + ioff[13];
        t.gemm_tdeg = I + ioff[14];
        t.gemm_tfull = I + ioff[15];
        t.fold = FoldTab{I + ioff[16], I + ioff[17], I + ioff[18], D + doff[5], D + doff[6]};
    }
    t.cs = D + doff[3];
    t.scale = D + doff[4];
    // Box-scaled M2M / L2L matrices per octant for the FP32 MFMA shifts (shift_mfma.hip):
    // M2M S[t][s] = c(t, s) u^(t-s) 2^-|t|, L2L S[t][s] = c(t, s) u^(s-t) 2^-|s|, u = (+-1, +-1, +-1) by
    // octant bits; lane l of the fragment (mt, ks) holds S[16 mt + l % 16][4 ks + l / 16].
    if (p >= 1) {
        const int KS = (nm + 3) / 4;
        const size_t nfrag = (size_t)shift_mfma_frag_floats(p, false);
        std::vector<float> hf(2 * 8 * nfrag, 0.f);
        auto fill = [&](float* dst, const std::vector<int>& off, const std::vector<int>& ss,
                        const std::vector<int>& ee, const std::vector<double>& cc, bool l2l, int o) {
            const int ux = (o & 1) ? 1 : -1, uy = (o & 2) ? 1 : -1, uz = (o & 4) ? 1 : -1;
            for (int tt = 0; tt < nm; ++tt)
                for (int b = off[tt]; b < off[tt + 1]; ++b) {
                    const int sv = ss[b];
                    const XYZ& E = comp[ee[b]];
                    const XYZ& T = comp[l2l ? sv : tt];
                    const double sign = ((E.x & 1) ? ux : 1) * ((E.y & 1) ? uy : 1) * ((E.z & 1) ? uz : 1);
                    const double v = cc[b] * sign * std::ldexp(1.0, -(T.x + T.y + T.z));
                    const int mt = tt / 16, ks = sv / 4, l = tt % 16 + 16 * (sv % 4);
                    dst[((size_t)mt * KS + ks) * 64 + l] = (float)v;
                }
        };
        for (int o = 0; o < 8; ++o) {
            fill(hf.data() + (size_t)o * nfrag, m2m_off, m2m_s, m2m_e, m2m_c, false, o);
            fill(hf.data() + (size_t)(8 + o) * nfrag, l2l_off, l2l_s, l2l_e, l2l_c, true, o);
        }
        // Reduced L2L for harmonic local expansions: R S U per octant, with U the unfold (component
        // (x, y, 2k + rho) = (-1)^k sum_i C(k, i) reduced (x + 2i, y + 2(k - i), rho)) and R the
        // restriction to the reduced rows; reduced order: by degree n, then w = y + z, then z.
        const bool red_ok = p <= M2L_GEMM_PMAX;
        const size_t nfr = red_ok ? (size_t)shift_mfma_frag_floats(p, true) : 0;
        if (red_ok) {
            const int nk = (p + 1) * (p + 1), KSr = (nk + 3) / 4;
            auto ridx = [](int X, int Y, int rho) {
                const int n = X + Y + rho, w = Y + rho;
                return n * n + (w == 0 ? 0 : 2 * w - 1 + rho);
            };
            std::vector<int> rfull(nk);
            for (int u = 0; u < nm; ++u)
                if (comp[u].z <= 1) rfull[ridx(comp[u].x, comp[u].y, comp[u].z)] = u;
            std::vector<double> U((size_t)nm * nk, 0.0);
            for (int u = 0; u < nm; ++u) {
                const int k = comp[u].z / 2, rho = comp[u].z % 2;
                double binom = 1.0;
                for (int i = 0; i <= k; ++i) {
                    if (i > 0) binom = binom * (k - i + 1) / i;
                    U[(size_t)u * nk + ridx(comp[u].x + 2 * i, comp[u].y + 2 * (k - i), rho)] += ((k & 1) ? -1.0 : 1.0) * binom;
                }
            }
            hf.resize(hf.size() + 8 * nfr, 0.f);
            std::vector<double> S((size_t)nm * nm);
            for (int o = 0; o < 8; ++o) {
                const int ux = (o & 1) ? 1 : -1, uy = (o & 2) ? 1 : -1, uz = (o & 4) ? 1 : -1;
                std::fill(S.begin(), S.end(), 0.0);
                for (int tt = 0; tt < nm; ++tt)
                    for (int b = l2l_off[tt]; b < l2l_off[tt + 1]; ++b) {
                        const int sv = l2l_s[b];
                        const XYZ& E = comp[l2l_e[b]];
                        const XYZ& T = comp[sv];
                        const double sign = ((E.x & 1) ? ux : 1) * ((E.y & 1) ? uy : 1) * ((E.z & 1) ? uz : 1);
                        S[(size_t)tt * nm + sv] = l2l_c[b] * sign * std::ldexp(1.0, -(T.x + T.y + T.z));
                    }
                float* dst = hf.data() + 16 * nfrag + (size_t)o * nfr;
                for (int tr = 0; tr < nk; ++tr) {
                    const double* Srow = S.data() + (size_t)rfull[tr] * nm;
                    for (int sr = 0; sr < nk; ++sr) {
                        double a = 0.0;
                        for (int u = 0; u < nm; ++u) a += Srow[u] * U[(size_t)u * nk + sr];
                        dst[((size_t)(tr / 16) * KSr + sr / 4) * 64 + tr % 16 + 16 * (sr % 4)] = (float)a;
                    }
                }
            }
        }
        FMMH_HIP(hipMalloc((void**)&t.fbuf, hf.size() * sizeof(float)));
        FMMH_HIP(hipMemcpy(t.fbuf, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice));
        t.m2m_frag = t.fbuf;
        t.l2l_frag = t.fbuf + 8 * nfrag;
        if (red_ok) t.l2l_frag_red = t.fbuf + 16 * nfrag;
    }
    return c->tables.emplace(key, t).first->second;
}

__global__ void k_iota(int* a, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// Target blocks of 64 consecutive sites from `first` on (n targets), each
// with the one source block 0.
__global__ void k_direct_blocks(int* start, int* count, int* list_off, int* list_src, int nb, int n, int first) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) {
        start[b] = first + 64 * b;
        count[b] = min(64, n - 64 * b);
        list_off[b] = b;
        list_src[b] = 0;
    }
    if (b == 0) list_off[nb] = nb;
}

// Site pairs of the near field: sum over leaves of (sites) x (sources of its list), from the per-leaf
// source totals; one atomic per workgroup.
__global__ __launch_bounds__(256) void k_pair_count(const long long* __restrict__ tot, const int* __restrict__ leaf_node,
                                                    const int* __restrict__ count, int nl,
                                                    unsigned long long* __restrict__ total) {
    __shared__ unsigned long long wsum[4];
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = l < nl ? (unsigned long long)tot[l] * (unsigned long long)count[leaf_node[l]] : 0ull;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (t) atomicAdd(total, t);
    }
}

// Order of the near-field workgroups.  Workgroup b runs on XCD b % 8 (the
// dispatcher deals workgroups round robin), and each XCD has its own L2: the
// leaves are cut into 8 contiguous Morton ranges of about equal pair work,
// range x goes to the slots b = 8 r + x, so an XCD's L2 holds one region's
// sources.  A leaf whose work exceeds a cap (the total over NEAR_CHUNK_DIV,
// at least NEAR_CHUNK_MIN source-tile iterations) is cut into chunks of
// consecutive list entries with about equal source counts; each chunk is its
// own workgroup and writes a partial field, which k_near_fix sums in chunk
// order (deterministic).  Within a range the heaviest chunks come first (a
// stable sort on the range and the chunk's work in 1/64 of the cap),
// because the hardware starts workgroups in index order and long workgroups
// should not finish last on an idle GPU.
constexpr int NEAR_CHUNK_DIV = 16384;
constexpr long long NEAR_CHUNK_MIN = 256;
constexpr int NEAR_CHUNK_MAX = 64;   // chunks per leaf

// Work of one leaf in source-tile iterations of k_near's wavefront: the same pass sizing as the kernel
// (c passes of ct targets, r = 64 / ct lane groups), plus a fixed cost per pass.
__device__ long long leaf_work(int tc, long long total) {
    int ct = min(tc, 64), best = 1 << 30;
    for (int c = (tc + 63) / 64; c <= 16; ++c) {
        const int t = (tc + c - 1) / c;
        const int rr = 64 / t;
        const int cost = c * ((64 + rr - 1) / rr + 2);
        if (cost < best) {
            best = cost;
            ct = t;
        }
    }
    long long w = 0;
    for (int b = 0; b < tc; b += ct) {
        const int r = 64 / min(ct, tc - b);
        w += (total + r - 1) / r + 8;
    }
    return w;
}

// One wavefront per leaf: the source count of every list entry (ecnt, kept for the chunk split) and
// the leaf's total and work.
__global__ __launch_bounds__(64) void k_leaf_work(const int* __restrict__ off, const int* __restrict__ src,
                                                  const int* __restrict__ leaf_node, const int* __restrict__ count,
                                                  int nl, int* __restrict__ ecnt, long long* __restrict__ tot,
                                                  long long* __restrict__ work) {
    const int l = blockIdx.x;
    if (l >= nl) return;
    const int node = leaf_node[l];
    const int e0 = off[node], e1 = off[node + 1];
    long long t = 0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += 64) {
        const int c = count[src_node(src[e])];
        ecnt[e] = c;
        t += c;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d);
    if (threadIdx.x == 0) {
        tot[l] = t;
        work[l] = leaf_work(count[node], t);
    }
}

// Exclusive prefix sums of n int64 values in one workgroup of 1024 threads (pre[n] = total); each
// thread sums a contiguous run, so the integer result does not depend on scheduling.
__global__ __launch_bounds__(1024) void k_scan_i64(const long long* __restrict__ in, long long* __restrict__ pre,
                                                   int n) {
    __shared__ long long wtot[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + 1023) / 1024;
    const int a = min(n, t * per), b = min(n, a + per);
    long long s = 0;
    for (int i = a; i < b; ++i) s += in[i];
    long long incl = s;   // inclusive scan over the wavefront, then over the 16 wavefront totals
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const long long y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    if (t == 0) {
        long long run = 0;
        for (int q = 0; q < 16; ++q) {
            const long long v = wtot[q];
            wtot[q] = run;
            run += v;
        }
        pre[n] = run;
    }
    __syncthreads();
    long long run = wtot[w] + incl - s;
    for (int i = a; i < b; ++i) {
        pre[i] = run;
        run += in[i];
    }
}

__device__ void leaf_chunk(const long long* __restrict__ work, const long long* __restrict__ pre,
                           const int* __restrict__ off, const int* __restrict__ leaf_node,
                           const int* __restrict__ count, int nl, long long cap_set, int l, int* __restrict__ lk,
                           int* __restrict__ lx, int* __restrict__ lps, int* hist) {
    const long long total = pre[nl];
    const long long cap = cap_set > 0 ? cap_set : max(total / NEAR_CHUNK_DIV, NEAR_CHUNK_MIN);
    const int node = leaf_node[l];
    const int entries = off[node + 1] - off[node];
    const long long w = work[l];
    int K = (int)min((w + cap - 1) / cap, (long long)NEAR_CHUNK_MAX);
    K = max(1, min(K, entries));
    // range by the work midpoint of the leaf: monotone in l, about equal work per range
    const int x = total > 0 ? (int)min((long long)(NUM_XCD - 1), ((pre[l] + w / 2) * NUM_XCD) / total) : 0;
    lk[l] = K;
    lx[l] = x;
    lps[l] = K > 1 ? K * count[node] : 0;
    atomicAdd(&hist[x], K);
}

// Per leaf: chunk count K (1 = whole list), XCD range, partial sites (K tc when chunked; with one
// extra slot for the totals at index nl so that exclusive scans over nl + 1 entries give them).
__global__ void k_leaf_chunks(const long long* __restrict__ work, const long long* __restrict__ pre,
                              const int* __restrict__ off, const int* __restrict__ leaf_node,
                              const int* __restrict__ count, int nl, long long cap_set, int* __restrict__ lk,
                              int* __restrict__ lx, int* __restrict__ lps, int* __restrict__ range_cnt) {
    __shared__ int hist[NUM_XCD];
    if (threadIdx.x < NUM_XCD) hist[threadIdx.x] = 0;
    __syncthreads();
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nl) leaf_chunk(work, pre, off, leaf_node, count, nl, cap_set, l, lk, lx, lps, hist);
    if (l == nl) {
        lk[nl] = 0;
        lps[nl] = 0;
    }
    __syncthreads();
    if (threadIdx.x < NUM_XCD && hist[threadIdx.x]) atomicAdd(&range_cnt[threadIdx.x], hist[threadIdx.x]);
}


// Per leaf (one wavefront): its chunks (consecutive list entries, boundaries at about equal cumulative
// source counts, at least one entry each) and their order buckets: range, then decreasing work in
// 1/64 of the chunk cap (VB_WORK_BITS).  The split runs on lane 0 over the entry counts staged in LDS.
constexpr int VB_STAGE = 512;
constexpr int VB_WORK_BITS = 6;
constexpr int VB_BUCKETS = NUM_XCD << VB_WORK_BITS;
__global__ __launch_bounds__(64) void k_leaf_vblocks(const long long* __restrict__ work,
                                                     const long long* __restrict__ tot,
                                                     const long long* __restrict__ pre, long long cap_set,
                                                     const int* __restrict__ off, const int* __restrict__ ecnt,
                                                     const int* __restrict__ leaf_node,
                                                     const int* __restrict__ count, int nl,
                                                     const int* __restrict__ lk, const int* __restrict__ lx,
                                                     const int* __restrict__ lvb, const int* __restrict__ lpart,
                                                     int* __restrict__ vb_node, int* __restrict__ vb_e0,
                                                     int* __restrict__ vb_e1, int* __restrict__ vb_part,
                                                     int* __restrict__ bucket) {
    __shared__ int sc[VB_STAGE];
    const int l = blockIdx.x;
    if (l >= nl) return;
    const int node = leaf_node[l];
    const int e0 = off[node], e1 = off[node + 1];
    const int K = lk[l];
    const bool staged = e1 - e0 <= VB_STAGE;
    if (K > 1 && staged)
        for (int e = e0 + (int)threadIdx.x; e < e1; e += 64) sc[e - e0] = ecnt[e];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int v0 = lvb[l];
    const int x = lx[l];
    const long long cap = cap_set > 0 ? cap_set : max(pre[nl] / NEAR_CHUNK_DIV, NEAR_CHUNK_MIN);
    const long long total = tot[l];
    const long long wl = work[l];
    int e = e0;
    long long cum = 0;
    for (int k = 0; k < K; ++k) {
        const int a = e;
        long long csrc = 0;
        if (k == K - 1) {
            e = e1;
            csrc = total - cum;
        } else {
            // close the chunk once the cumulative count reaches (k + 1) total / K, leaving at least
            // one entry for each later chunk
            const long long goal = ((long long)(k + 1) * total) / K;
            do {
                const long long c = staged ? sc[e - e0] : ecnt[e];
                cum += c;
                csrc += c;
                ++e;
            } while (cum < goal && e < e1 - (K - 1 - k));
        }
        const int v = v0 + k;
        vb_node[v] = node;
        vb_e0[v] = a;
        vb_e1[v] = e;
        vb_part[v] = K > 1 ? lpart[l] + k * count[node] : -1;
        const long long cw = total > 0 ? (wl * csrc) / total : wl / K;
        const int wq = (int)min((cw << VB_WORK_BITS) / cap, (long long)((1 << VB_WORK_BITS) - 1));
        bucket[v] = (x << VB_WORK_BITS) | ((1 << VB_WORK_BITS) - 1 - wq);
    }
}

struct RangeStarts {
    int first[NUM_XCD + 1];
};

// Workgroup slots of the chunks, in one workgroup of 1024 threads: a stable counting sort of the chunks
// by bucket (range, then decreasing work; ties in chunk id order), the chunk at sorted position k of range
// x going to slot (k - first[x]) NUM_XCD + x, unused slots -1.  Each wave owns a contiguous run of chunks
// and its own column of bucket counters; the counters' scan in (bucket, wave) order gives every
// (bucket, wave) its first position, and a wave places 64 chunks at a time, ranking equal buckets among
// its lanes by ballots.  Chunk loads are issued VB_ORDER_UNROLL rounds at a time.
constexpr int VB_ORDER_T = 1024;
constexpr int VB_ORDER_W = VB_ORDER_T / 64;
constexpr int VB_ORDER_UNROLL = 4;
constexpr int VB_BUCKET_BITS = 3 + VB_WORK_BITS;
static_assert(VB_BUCKETS * VB_ORDER_W == 8 * VB_ORDER_T, "the scan gives each thread 8 counters");
static_assert(NUM_XCD == 8, "3 range bits");
__global__ __launch_bounds__(VB_ORDER_T) void k_vb_order(const int* __restrict__ bucket, int nv, RangeStarts rs,
                                                         int* __restrict__ slots, int nslots) {
    __shared__ int cnt[VB_BUCKETS * VB_ORDER_W];   // [bucket][wave]
    __shared__ int wsum[VB_ORDER_W];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int i = tid; i < nslots; i += VB_ORDER_T) slots[i] = -1;
    for (int i = tid; i < VB_BUCKETS * VB_ORDER_W; i += VB_ORDER_T) cnt[i] = 0;
    const int per = (nv + VB_ORDER_W - 1) / VB_ORDER_W;
    const int c0 = min(nv, w * per), c1 = min(nv, c0 + per);
    __syncthreads();
    for (int r = c0; r < c1; r += 64 * VB_ORDER_UNROLL) {
        int b[VB_ORDER_UNROLL];
#pragma unroll
        for (int u = 0; u < VB_ORDER_UNROLL; ++u) {
            const int i = r + u * 64 + lane;
            b[u] = i < c1 ? bucket[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < VB_ORDER_UNROLL; ++u)
            if (b[u] >= 0) atomicAdd(&cnt[b[u] * VB_ORDER_W + w], 1);
    }
    __syncthreads();
    // exclusive scan of the counters in (bucket, wave) order: 8 consecutive counters per thread
    int loc[8], tot = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        loc[j] = cnt[8 * tid + j];
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
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        cnt[8 * tid + j] = run;
        run += loc[j];
    }
    __threadfence();   // the -1 fill is visible before any slot is written below
    __syncthreads();
    for (int r = c0; r < c1; r += 64 * VB_ORDER_UNROLL) {
        int b[VB_ORDER_UNROLL];
#pragma unroll
        for (int u = 0; u < VB_ORDER_UNROLL; ++u) {
            const int i = r + u * 64 + lane;
            b[u] = i < c1 ? bucket[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < VB_ORDER_UNROLL; ++u) {
            const int i = r + u * 64 + lane;
            const bool valid = b[u] >= 0;
            unsigned long long same = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < VB_BUCKET_BITS; ++bit) {
                const bool on = ((b[u] >> bit) & 1) != 0;
                const unsigned long long m = __ballot(valid && on);
                same &= on ? m : ~m;
            }
            // lanes of this round with the same bucket: the lowest one advances the counter
            const int rank = __popcll(same & ((1ull << lane) - 1ull));
            int base = 0;
            if (valid) base = cnt[b[u] * VB_ORDER_W + w];
            __builtin_amdgcn_wave_barrier();
            if (valid && rank == 0) cnt[b[u] * VB_ORDER_W + w] = base + __popcll(same);
            __builtin_amdgcn_wave_barrier();
            if (valid) {
                const int k = base + rank;
                const int x = b[u] >> VB_WORK_BITS;
                slots[(k - rs.first[x]) * NUM_XCD + x] = i;
            }
        }
    }
}

__global__ void k_invert_perm(const int* __restrict__ perm, int* __restrict__ inv, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) inv[perm[k]] = k;
}

__global__ void k_fill_i32(int* a, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Exclusive scan: one workgroup for small arrays, the device-wide scan otherwise (temp sized for n).
void exscan_auto(fmmh_context* c, const char* temp_name, const int* in, int* out, int n, hipStream_t s) {
    if (n <= EXSCAN_SMALL_MAX) {
        exscan_i32_small(in, out, n, s);
        return;
    }
    const size_t tb = exscan_i32_temp(n);
    exscan_i32(c->buf<char>(temp_name, tb), tb, in, out, n, s);
}

// Device flag of exclusion entries above n (only the device-array entry points can pass them; the host
// entry points check the entries before copying).  Cleared by a kernel, so that a captured call clears it
// on every replay; read back with the call's last synchronisation.
constexpr int PIN_BAD_EXCL = 960;   // int slot of the pinned readback area
int* bad_excl_flag(fmmh_context* c) {
    int* f = c->buf<int>("x_bad_excl", 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, f, 0, 1);
    return f;
}
void bad_excl_readback(fmmh_context* c) {
    FMMH_HIP(hipMemcpyAsync(c->h_pinned + PIN_BAD_EXCL, c->buf<int>("x_bad_excl", 1), sizeof(int),
                            hipMemcpyDeviceToHost, c->stream));
}
void bad_excl_check(fmmh_context* c) {   // after the stream synchronisation that follows the readback
    const int f = c->h_pinned[PIN_BAD_EXCL];
    if (f & 2) throw Error(FMMH_ERR_INTERNAL, "a target with two far pairs of one displacement class");
    if (f) throw Error(FMMH_ERR_ARG, "exclusion index outside 0..n");
}

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// Periodic images (src/fmm.F90:156-175): image id 0 is the central cell, ids
// 1..26 the neighbours in the reference's (ai, aj, ak) loop order; the
// displacement is matmul(box_matrix, aijk) summed column by column.
struct Periodic {
    bool on = false;
    int nimg = 1;
    double box[9] = {};                  // column-major 3 x 3
    double disp[NUM_IMAGES * 3] = {};
};

Periodic make_periodic(bool on, const double* box) {
    Periodic P;
    if (!on) return P;
    P.on = true;
    P.nimg = NUM_IMAGES;
    std::memcpy(P.box, box, sizeof(P.box));
    int id = 1;
    for (int ai = -1; ai <= 1; ++ai)
        for (int aj = -1; aj <= 1; ++aj)
            for (int ak = -1; ak <= 1; ++ak) {
                if (ai == 0 && aj == 0 && ak == 0) continue;
                for (int r = 0; r < 3; ++r)
                    P.disp[3 * id + r] = (box[r] * ai + box[r + 3] * aj) + box[r + 6] * ak;
                ++id;
            }
    return P;
}

// Device copy of the image displacements (one slot per image id).
const double* upload_disp(fmmh_context* c, const Periodic& P) {
    double* d = c->buf<double>("p_disp", NUM_IMAGES * 3);
    double* h = reinterpret_cast<double*>(c->h_pinned + 512);
    std::memcpy(h, P.disp, sizeof(P.disp));
    FMMH_HIP(hipMemcpyAsync(d, h, sizeof(P.disp), hipMemcpyHostToDevice, c->stream));
    return d;
}
constexpr int TRAVERSE_GRID = 1024;           // workgroups per traversal step (grid-stride over the frontier)
constexpr double M2L_GEMM_MAX_BYTES = 48e9;   // per-pair slot buffer of the class-grouped M2L
constexpr int M2L_MFMA_MIN_SETS = 6;
// Mixed precision: a level's M2M runs on FP32 MFMA when it has at least this many (node, set) columns;
// the top levels, with a handful of nodes, are bound by one workgroup's latency (eight octants in turn)
// and run the FP64 kernel instead (faster there, and exact).  L2L, one octant per workgroup, always runs
// on FP32 MFMA.
constexpr long long SHIFT_MFMA_MIN_COLS = 4096;          // auto engine: per-target MFMA from this many sets on

struct Call {
    const double* X; int64_t ldx;
    const double* M; int64_t ldm;
    int n, dim;
    const int* E; int64_t lde; int nexcl;
    int F; double* field; int64_t ldf;
    int damp; const double* df;
    int nsets;   // multipoles n x dim x nsets and field n x nf x nsets, one leading dimension each
};

void check_common(const Call& c, int64_t n_damping) {
    if (c.n < 1) throw Error(FMMH_ERR_ARG, "n must be >= 1");
    if (c.ldx < c.n || c.ldm < c.n || c.ldf < c.n || (c.nexcl > 0 && c.lde < c.n))
        throw Error(FMMH_ERR_ARG, "leading dimension smaller than n");
    if (c.nexcl < 0) throw Error(FMMH_ERR_ARG, "nexcl must be >= 0");
    if (c.nsets < 1) throw Error(FMMH_ERR_ARG, "nsets must be >= 1");
    if (c.damp < 0 || c.damp > 3) throw Error(FMMH_ERR_DAMP, "Unknown damp type code");
    if (c.damp != FMMH_DAMP_NONE && (n_damping < c.n || c.df == nullptr))
        throw Error(FMMH_ERR_ARG, "damping_factors must hold n values when damping is on");
    int m = order_from_dim(c.dim);
    if (m < 0 || ncomp(m) != c.dim) throw Error(FMMH_ERR_ARG, "dim is not (m+1)(m+2)(m+3)/6 for any order m");
    if (!near_supported(m, c.F))
        throw Error(FMMH_ERR_UNSUPPORTED, "multipole order " + std::to_string(m) + " / field order " +
                                              std::to_string(c.F) + " outside the compiled range (<= " +
                                              std::to_string(NEAR_MMAX) + " / <= " + std::to_string(NEAR_FMAX) +
                                              ", sum <= " + std::to_string(NEAR_KMAX) + ")");
}

void check_exclusions_host(const Call& c) {
    for (int q = 0; q < c.nexcl; ++q)
        for (int i = 0; i < c.n; ++i) {
            int v = c.E[(int64_t)q * c.lde + i];
            if (v < 0 || v > c.n)
                throw Error(FMMH_ERR_ARG, "exclusions(" + std::to_string(i + 1) + "," + std::to_string(q + 1) +
                                              ") = " + std::to_string(v) + " is outside 0..n");
        }
}

// ---------------------------------------------------------------- direct
void run_direct(fmmh_context* c, const Call& k) {
    hipStream_t s = c->stream;
    const int n = k.n;
    const int m = order_from_dim(k.dim);
    Tables& tb = get_tables(c, 0, m);
    const int nb = nblk(n, 64);
    int* start = c->buf<int>("d_start", nb);
    int* count = c->buf<int>("d_count", nb);
    int* loff = c->buf<int>("d_loff", nb + 1);
    int* lsrc = c->buf<int>("d_lsrc", nb);
    int* s_start = c->buf<int>("d_sstart", 1);
    int* s_count = c->buf<int>("d_scount", 1);
    int* iota = c->buf<int>("iota", std::max(n, nb));
    double* ms = c->buf<double>("d_ms", (size_t)n * k.dim * k.nsets);
    double* ms_raw = c->buf<double>("d_ms_raw", (size_t)n * k.dim * k.nsets);
    double* xs = c->buf<double>("d_xs", n);
    double* ys = c->buf<double>("d_ys", n);
    double* zs = c->buf<double>("d_zs", n);
    double* dfs = c->buf<double>("d_dfs", n);
    int* excl = c->buf<int>("d_excl", (size_t)n * std::max(k.nexcl, 1));
    hipLaunchKernelGGL(k_direct_blocks, dim3(nblk(nb, 256)), dim3(256), 0, s, start, count, loff, lsrc, nb, n, 0);
    hipLaunchKernelGGL(k_iota, dim3(nblk(std::max(n, nb), 256)), dim3(256), 0, s, iota, std::max(n, nb));
    int zero_n[2] = {0, n};
    FMMH_HIP(hipMemcpyAsync(s_start, &zero_n[0], sizeof(int), hipMemcpyHostToDevice, s));
    FMMH_HIP(hipMemcpyAsync(s_count, &zero_n[1], sizeof(int), hipMemcpyHostToDevice, s));
    SiteArgs g{k.X, k.ldx, k.M, k.ldm, k.dim, k.E, k.lde, k.nexcl, k.damp ? k.df : nullptr, iota, n,
               xs, ys, zs, ms_raw, ms, dfs, excl, tb.scale, k.nsets};
    launch_gather(g, s);
    c->mark("gather");
    NearArgs a{iota, start, count, loff, lsrc, s_start, s_count, xs, ys, zs, ms, n, dfs, iota, excl, n, k.nexcl,
               k.field, k.ldf, 1, k.nsets, upload_disp(c, Periodic{})};
    a.nsites = n;
    a.bad = bad_excl_flag(c);
    a.coinc_list = 1;
    launch_near(a, nb, m, k.F, k.damp, s);
    c->mark("direct");
    bad_excl_readback(c);
    FMMH_HIP(hipStreamSynchronize(s));
    bad_excl_check(c);
}

// ---------------------------------------------------------------- direct sum without allocation
// run_direct with every working array carved out of the fixed per-device memory (`base`, as returned by
// scratch_ptr): no allocation, no host-to-device copy.  Layout (each part 256-byte aligned): xs, ys, zs,
// dfs (n doubles each), Taylor-scaled multipoles (n x dim doubles; the direct sum needs no raw copy),
// exclusions (n x nexcl ints), site order (max(n, blocks) ints),
// block start / count / list offsets / list sources, source block start and count, exclusion flag,
// image displacement, Taylor factors.
struct NoallocLayout {
    size_t xs, ys, zs, dfs, ms, excl, iota, start, count, loff, lsrc, sstart, scount, bad, disp, scale, total;
};

NoallocLayout noalloc_layout(int n, int dim, int nexcl) {
    const size_t nb = (size_t)nblk(n, 64);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    };
    NoallocLayout L{};
    L.xs = take((size_t)n * 8);
    L.ys = take((size_t)n * 8);
    L.zs = take((size_t)n * 8);
    L.dfs = take((size_t)n * 8);
    L.ms = take((size_t)n * dim * 8);
    L.excl = take((size_t)n * std::max(nexcl, 1) * 4);
    L.iota = take(std::max((size_t)n, nb) * 4);
    L.start = take(nb * 4);
    L.count = take(nb * 4);
    L.loff = take((nb + 1) * 4);
    L.lsrc = take(nb * 4);
    L.sstart = take(4);
    L.scount = take(4);
    L.bad = take(4);
    L.disp = take(3 * 8);
    L.scale = take((size_t)dim * 8);
    L.total = off;
    return L;
}

// The one source block [0, n), the central-cell image displacement, the exclusion flag and the Taylor
// factors (-1)^{|s|+1} / (sx! sy! sz!) of the dim source components (src/fmm.F90:597-612), on the device.
__global__ void k_noalloc_setup(int n, int dim, int* sstart, int* scount, int* bad, double* disp, double* scale) {
    const int t = threadIdx.x;
    if (t == 0) {
        *sstart = 0;
        *scount = n;
        *bad = 0;
    }
    if (t < 3) disp[t] = 0.0;
    for (int s = t; s < dim; s += blockDim.x) {
        int k = 0;
        while (ncomp(k) <= s) ++k;
        int sx = 0, sy = 0, sz = 0;
        for (int x = k; x >= 0; --x)
            for (int y = k - x; y >= 0; --y)
                if (xyz2idx(x, y, k - x - y) == s) { sx = x; sy = y; sz = k - x - y; }
        double f = 1.0;
        for (int i = 2; i <= sx; ++i) f *= i;
        for (int i = 2; i <= sy; ++i) f *= i;
        for (int i = 2; i <= sz; ++i) f *= i;
        scale[s] = ((k + 1) % 2 ? -1.0 : 1.0) / f;
    }
}

void run_direct_noalloc(fmmh_context* c, const Call& k, void* base) {
    hipStream_t s = c->stream;
    const int n = k.n;
    const int m = order_from_dim(k.dim);
    const int nb = nblk(n, 64);
    const NoallocLayout L = noalloc_layout(n, k.dim, k.nexcl);
    if (L.total > NOALLOC_SCRATCH_BYTES)
        throw Error(FMMH_ERR_UNSUPPORTED, "field_direct without allocation: the working arrays of " +
                                              std::to_string(n) + " sites (" + std::to_string(L.total) +
                                              " bytes) exceed the fixed " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                              " bytes");
    unsigned char* b = static_cast<unsigned char*>(base);
    auto D = [&](size_t off) { return reinterpret_cast<double*>(b + off); };
    auto I = [&](size_t off) { return reinterpret_cast<int*>(b + off); };
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(64), 0, s, n, k.dim, I(L.sstart), I(L.scount), I(L.bad),
                       D(L.disp), D(L.scale));
    hipLaunchKernelGGL(k_direct_blocks, dim3(nblk(nb, 256)), dim3(256), 0, s, I(L.start), I(L.count), I(L.loff),
                       I(L.lsrc), nb, n, 0);
    hipLaunchKernelGGL(k_iota, dim3(nblk(std::max(n, nb), 256)), dim3(256), 0, s, I(L.iota), std::max(n, nb));
    SiteArgs g{k.X, k.ldx, k.M, k.ldm, k.dim, k.E, k.lde, k.nexcl, k.damp ? k.df : nullptr, I(L.iota), n,
               D(L.xs), D(L.ys), D(L.zs), nullptr, D(L.ms), D(L.dfs), I(L.excl), D(L.scale), 1};
    launch_gather(g, s);
    NearArgs a{I(L.iota), I(L.start), I(L.count), I(L.loff), I(L.lsrc), I(L.sstart), I(L.scount), D(L.xs), D(L.ys),
               D(L.zs), D(L.ms), n, D(L.dfs), I(L.iota), I(L.excl), n, k.nexcl, k.field, k.ldf, 1, 1, D(L.disp)};
    a.nsites = n;
    a.bad = I(L.bad);
    a.coinc_list = 1;
    launch_near(a, nb, m, k.F, k.damp, s);
    FMMH_HIP(hipMemcpyAsync(c->h_pinned + PIN_BAD_EXCL, I(L.bad), sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipStreamSynchronize(s));
    bad_excl_check(c);
}

// Direct sum at a subset of target sites (spot checks at sizes where the
// full O(N^2) sum is too costly).  The targets are appended as copies after
// the n sources in the gathered site arrays, so the near kernel sees target
// blocks [n, n + nt) against the one source block [0, n).
__global__ void k_target_perm(int* perm, const int* targets, int n, int nt, int* bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[i] = i;
    else if (i < n + nt) {
        int t = targets[i - n];
        if (t < 0 || t >= n) { *bad = 1; t = 0; }
        perm[i] = t;
    }
}

__global__ void k_add_rows(const double* __restrict__ src, int64_t lds, int first, double* __restrict__ dst,
                           int64_t ldd, int nt, int nf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    for (int t = 0; t < nf; ++t) dst[t * ldd + i] += src[t * lds + first + i];
}

void run_direct_targets(fmmh_context* c, const Call& k, const int* targets, int nt) {
    hipStream_t s = c->stream;
    const int n = k.n, ntot = n + nt;
    const int m = order_from_dim(k.dim);
    const int nf = ncomp(k.F);
    Tables& tb = get_tables(c, 0, m);
    const int nb = nblk(nt, 64);
    int* start = c->buf<int>("d_start", nb);
    int* count = c->buf<int>("d_count", nb);
    int* loff = c->buf<int>("d_loff", nb + 1);
    int* lsrc = c->buf<int>("d_lsrc", nb);
    int* s_start = c->buf<int>("d_sstart", 1);
    int* s_count = c->buf<int>("d_scount", 1);
    int* iota = c->buf<int>("iota", std::max(ntot, nb));
    int* perm = c->buf<int>("d_tperm", ntot);
    int* bad = c->buf<int>("d_tbad", 1);
    double* ms = c->buf<double>("d_ms", (size_t)ntot * k.dim);
    double* ms_raw = c->buf<double>("d_ms_raw", (size_t)ntot * k.dim);
    double* xs = c->buf<double>("d_xs", ntot);
    double* ys = c->buf<double>("d_ys", ntot);
    double* zs = c->buf<double>("d_zs", ntot);
    double* dfs = c->buf<double>("d_dfs", ntot);
    int* excl = c->buf<int>("d_excl", (size_t)ntot * std::max(k.nexcl, 1));
    double* out = c->buf<double>("d_tout", (size_t)ntot * nf);
    FMMH_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_target_perm, dim3(nblk(ntot, 256)), dim3(256), 0, s, perm, targets, n, nt, bad);
    FMMH_HIP(hipMemcpyAsync(c->h_pinned + 410, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    c->sync();
    if (c->h_pinned[410]) throw Error(FMMH_ERR_ARG, "target index outside 0..n-1");
    hipLaunchKernelGGL(k_direct_blocks, dim3(nblk(nb, 256)), dim3(256), 0, s, start, count, loff, lsrc, nb, nt, n);
    hipLaunchKernelGGL(k_iota, dim3(nblk(std::max(ntot, nb), 256)), dim3(256), 0, s, iota, std::max(ntot, nb));
    int zero_n[2] = {0, n};
    std::memcpy(c->h_pinned + 420, zero_n, sizeof(zero_n));
    FMMH_HIP(hipMemcpyAsync(s_start, c->h_pinned + 420, sizeof(int), hipMemcpyHostToDevice, s));
    FMMH_HIP(hipMemcpyAsync(s_count, c->h_pinned + 421, sizeof(int), hipMemcpyHostToDevice, s));
    SiteArgs g{k.X, k.ldx, k.M, k.ldm, k.dim, k.E, k.lde, k.nexcl, k.damp ? k.df : nullptr, perm, ntot,
               xs, ys, zs, ms_raw, ms, dfs, excl, tb.scale, 1};
    launch_gather(g, s);
    NearArgs a{iota, start, count, loff, lsrc, s_start, s_count, xs, ys, zs, ms, ntot, dfs, perm, excl, ntot,
               k.nexcl, out, ntot, 0, 1, upload_disp(c, Periodic{})};
    a.nsites = n;
    a.bad = bad_excl_flag(c);
    a.coinc_list = 1;
    launch_near(a, nb, m, k.F, k.damp, s);
    hipLaunchKernelGGL(k_add_rows, dim3(nblk(nt, 256)), dim3(256), 0, s, out, (int64_t)ntot, n, k.field, k.ldf, nt,
                       nf);
    FMMH_HIP(hipGetLastError());
    c->mark("direct_targets");
    bad_excl_readback(c);
    FMMH_HIP(hipStreamSynchronize(s));
    bad_excl_check(c);
}

// ---------------------------------------------------------------- tree
void build_tree(fmmh_context* c, const double* X, int64_t ldx, int n, int ncrit, TreeOut& out) {
    hipStream_t s = c->stream;
    TreeDev& t = out.t;
    unsigned long long* keys_in = c->buf<unsigned long long>("t_keys_in", n);
    int* idx_in = c->buf<int>("t_idx_in", n);
    t.keys = c->buf<unsigned long long>("t_keys", n);
    t.perm = c->buf<int>("t_perm", n);
    t.lmin = c->buf<int>("t_lmin", n);
    t.lam = c->buf<int>("t_lam", n);
    t.cnt = c->buf<int>("t_cnt", n);
    t.off = c->buf<int>("t_off", n);
    t.w = c->buf<int>("t_w", n);
    t.root = c->buf<double>("t_root", 8);
    t.maxabs = c->buf<unsigned long long>("t_maxabs", 3);
    t.info = c->buf<int>("t_info", 64);
    t.level_fill = c->buf<int>("t_level_fill", KEY_LEVELS + 3);
    FMMH_HIP(hipMemsetAsync(t.info, 0, 64 * sizeof(int), s));
    launch_centroid(X, ldx, n, t, c->buf<char>("t_centroid_work", centroid_work_bytes(n)), s);
    c->mark("t_centroid");
    launch_keys(X, ldx, n, t, keys_in, idx_in, s);
    c->mark("t_keys");
    size_t tb = std::max(sort_pairs_u64_temp(n), exscan_i32_temp(n));
    void* temp = c->buf<char>("t_temp", tb);
    sort_pairs_u64(temp, tb, keys_in, t.keys, idx_in, t.perm, n, 63, s);
    c->mark("t_sort");
    launch_site_levels(n, ncrit, t, s);
    exscan_i32(temp, tb, t.cnt, t.off, n, s);
    int* hp = c->h_pinned;
    FMMH_HIP(hipMemcpyAsync(hp, t.off + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipMemcpyAsync(hp + 1, t.cnt + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipMemcpyAsync(hp + 2, t.info + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    c->sync();
    c->mark("t_levels");
    if (hp[2]) throw Error(FMMH_ERR_TREE, "octree deeper than 21 levels: >= ncrit sites share one cell");
    const int nn = hp[0] + hp[1];
    out.num_nodes = nn;
    t.level = c->buf<int>("n_level", nn);
    t.start = c->buf<int>("n_start", nn);
    t.count = c->buf<int>("n_count", nn);
    t.parent = c->buf<int>("n_parent", nn);
    t.is_leaf = c->buf<int>("n_is_leaf", nn);
    t.octant = c->buf<int>("n_octant", nn);
    t.children = c->buf<int>("n_children", (size_t)nn * 8);
    t.cx = c->buf<double>("n_cx", nn);
    t.cy = c->buf<double>("n_cy", nn);
    t.cz = c->buf<double>("n_cz", nn);
    t.r = c->buf<double>("n_r", nn);
    t.rmax = c->buf<double>("n_rmax", nn);
    t.leaf_id = c->buf<int>("n_leaf_id", nn);
    t.leaf_node = c->buf<int>("n_leaf_node", nn);
    t.level_nodes = c->buf<int>("n_level_nodes", nn);
    int* leaf_scan = c->buf<int>("n_leaf_scan", nn);
    FMMH_HIP(hipMemsetAsync(t.children, 0xff, (size_t)nn * 8 * sizeof(int), s));
    launch_emit_nodes(n, t, s);
    launch_node_links(nn, t, s);
    c->mark("t_nodes");
    exscan_auto(c, "t_temp2", t.is_leaf, leaf_scan, nn, s);
    FMMH_HIP(hipMemcpyAsync(hp, t.info, (3 + KEY_LEVELS + 2) * sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipMemcpyAsync(hp + 40, leaf_scan + (nn - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipMemcpyAsync(hp + 41, t.is_leaf + (nn - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    c->sync();
    out.max_depth = hp[2];
    out.num_leaves = hp[40] + hp[41];
    out.level_off.assign(KEY_LEVELS + 3, 0);
    for (int L = 1; L <= KEY_LEVELS + 1; ++L) out.level_off[L + 1] = out.level_off[L] + hp[3 + L];
    out.level_off[0] = 0;
    int* level_off_d = c->buf<int>("t_level_off", KEY_LEVELS + 3);
    std::memcpy(hp + 64, out.level_off.data(), (KEY_LEVELS + 3) * sizeof(int));
    FMMH_HIP(hipMemcpyAsync(level_off_d, hp + 64, (KEY_LEVELS + 3) * sizeof(int), hipMemcpyHostToDevice, s));
    FMMH_HIP(hipMemsetAsync(t.level_fill, 0, (KEY_LEVELS + 3) * sizeof(int), s));
    launch_leaf_index(nn, leaf_scan, t, s);
    launch_level_scatter(nn, level_off_d, t, s);
    launch_leaf_sort(out.num_leaves, ncrit, t, s);
    // (no wait here: the level offsets' pinned slots hp[64..] are not reused before the call's last wait)
}

// overlap (may be empty): device work that does not depend on the lists, queued once behind the first
// traversal's readback, so that the GPU runs it while the host waits for the list sizes.
template <class Overlap>
void build_lists(fmmh_context* c, TreeOut& tr, double theta, ListsOut& L, int nimg, const double* disp,
                 Overlap&& overlap) {
    hipStream_t s = c->stream;
    const int nn = tr.num_nodes;
    const double sc = c->list_cap_scale;
    c->list_retries = 0;
    // initial capacities: entries per node that fit the usual trees, or, when larger, what the context's
    // previous tree needed per node (repeated calls at small theta then do not overflow and retry)
    int cap_f = std::max(64, (int)std::min(1e9, std::max(160.0, c->list_need[0]) * nn * sc));
    int cap_m = std::max(64, (int)std::min(1e9, std::max(160.0, c->list_need[1]) * nn * sc));
    int cap_p = std::max(64, (int)std::min(1e9, std::max(64.0, c->list_need[2]) * nn * sc));
    int* hp = c->h_pinned;
    for (int attempt = 0; attempt < 64; ++attempt) {
        int2* fa = c->buf<int2>("l_front_a", cap_f);
        int2* fb = c->buf<int2>("l_front_b", cap_f);
        unsigned long long* m2l = c->buf<unsigned long long>("l_m2l", cap_m);
        unsigned long long* p2p = c->buf<unsigned long long>("l_p2p", cap_p);
        // three frontier counters (rotating: step k reads f[k % 3], appends to f[(k + 1) % 3] and zeroes
        // f[(k + 2) % 3], which neither it nor the previous step's readers touch), M2L, P2P and overflow,
        // each on its own 128-byte line (the traversal's atomics on them do not contend for one line)
        constexpr int CL = 32;
        int* counters = c->buf<int>("l_counters", 6 * CL);
        int* counts_out = c->buf<int>("l_counts", 8);
        launch_traverse_init(counters, CL, nimg, fa, s);   // one (root, root) seed per image
        // Every level of the traversal deepens one side of each surviving
        // pair, so level_i + level_j grows by one per step: 2 max_depth steps
        // always drain the frontier.  The frontier size stays on the device
        // (the kernel reads it), so no host round trip per step.
        const int steps = 2 * tr.max_depth;
        for (int it = 0; it < steps; ++it) {
            int2* fin = it % 2 == 0 ? fa : fb;
            int2* fout = it % 2 == 0 ? fb : fa;
            TraverseArgs a{fin, counters + (it % 3) * CL, cap_f, fout, counters + ((it + 1) % 3) * CL, cap_f, m2l,
                           counters + 3 * CL, cap_m, p2p, counters + 4 * CL, cap_p, counters + 5 * CL, tr.t.cx,
                           tr.t.cy, tr.t.cz, tr.t.rmax, tr.t.is_leaf, tr.t.children, theta, disp,
                           counters + ((it + 2) % 3) * CL};
            launch_traverse(a, TRAVERSE_GRID, s);
        }
        launch_traverse_counts(counters, CL, counts_out, s);
        FMMH_HIP(hipMemcpyAsync(hp, counts_out, 6 * sizeof(int), hipMemcpyDeviceToHost, s));
        if (attempt == 0) {
            c->mark("l_traverse");
            FMMH_HIP(hipEventRecord(c->readback_event(), s));
            overlap();
            ++c->host_waits;
            FMMH_HIP(hipEventSynchronize(c->readback_event()));
        } else {
            c->sync();
        }
        if (hp[5]) {
            ++c->list_retries;
            if (hp[5] & 1) cap_m = (int)std::min(1e9, std::max(2.0 * cap_m, 1.5 * hp[3]));
            if (hp[5] & 2) cap_p = (int)std::min(1e9, std::max(2.0 * cap_p, 1.5 * hp[4]));
            if (hp[5] & 4) cap_f *= 2;
            continue;
        }
        if (hp[steps % 3] != 0) throw Error(FMMH_ERR_INTERNAL, "dual traversal did not drain");
        if (sc == 1.0) {   // entries per node that sufficed (lists with a margin; not under a forced capacity)
            c->list_need[0] = (double)cap_f / std::max(nn, 1);
            c->list_need[1] = 1.25 * hp[3] / std::max(nn, 1);
            c->list_need[2] = 1.25 * hp[4] / std::max(nn, 1);
        }
        L.n_m2l = hp[3];
        L.n_p2p = hp[4];
        // CSR by target, sources ascending (the reference's append order, src/fmm.F90:352-385)
        const size_t tb = exscan_i32_temp(nn + 1);
        void* temp = c->buf<char>("l_temp", tb);
        L.m2l_keys = c->buf<unsigned long long>("l_m2l_sorted", L.n_m2l);
        L.m2l_off = c->buf<int>("l_m2l_off", nn + 1);
        L.m2l_src = c->buf<int>("l_m2l_src", L.n_m2l);
        L.p2p_off = c->buf<int>("l_p2p_off", nn + 1);
        L.p2p_src = c->buf<int>("l_p2p_src", L.n_p2p);
        int* cnt = c->buf<int>("l_target_cnt", nn + 1);
        unsigned* tmp = c->buf<unsigned>("l_list_tmp", std::max(L.n_m2l, L.n_p2p));
        hipLaunchKernelGGL(k_fill_i32, dim3(nblk(nn + 1, 256)), dim3(256), 0, s, cnt, 0, nn + 1);
        launch_list_csr(m2l, L.n_m2l, nn, cnt, L.m2l_off, tmp, temp, tb, L.m2l_src, L.m2l_keys, s);
        launch_list_csr(p2p, L.n_p2p, nn, cnt, L.p2p_off, tmp, temp, tb, L.p2p_src, nullptr, s);
        return;
    }
    throw Error(FMMH_ERR_INTERNAL, "interaction list capacity did not converge");
}

// ---------------------------------------------------------------- M2L
// Displacement classes (sort + unique), one tensor per class, then the
// target-centric Hankel contraction of every set.  One host sync to size the
// class table.  Many sets go through the MFMA form in chunks of up to 32 sets.
// M2L displacement classes of the pair list (sort + unique), in two halves around one host round trip
// for the class count: m2l_classes_begin launches the key, sort, flag, scan and assign kernels and the
// count's readback (false when there are no pairs); after a stream sync m2l_classes_end sizes and fills
// the class tensors.  The count is left in cache_classes and stats_classes.
void m2l_sort_classes(fmmh_context* c, M2LClassArgs& ca, bool assign);

bool m2l_classes_begin(fmmh_context* c, TreeOut& tr, ListsOut& L, Tables& tb, int p, const Periodic& P,
                       const double* disp, M2LClassArgs& ca) {
    hipStream_t s = c->stream;
    const int np = L.n_m2l;
    if (np <= 0) return false;
    int* pair_class = c->buf<int>("m_pair_class", np);
    int* dev_scalars = c->buf<int>("m_scalars", 4);   // num_classes, error
    FMMH_HIP(hipMemsetAsync(dev_scalars, 0, 4 * sizeof(int), s));
    ca = M2LClassArgs{};
    ca.pair_keys = L.m2l_keys;
    ca.n_pairs = np;
    ca.cx = tr.t.cx; ca.cy = tr.t.cy; ca.cz = tr.t.cz; ca.r = tr.t.r; ca.level = tr.t.level;
    ca.class_key_in = c->buf<unsigned long long>("m_key_in", np);
    ca.pair_idx = c->buf<int>("m_idx_in", np);
    ca.class_key_sorted = c->buf<unsigned long long>("m_key_sorted", np);
    ca.pair_idx_sorted = c->buf<int>("m_idx_sorted", np);
    ca.flag = c->buf<int>("m_flag", np);
    ca.cls = c->buf<int>("m_cls", np);
    ca.pair_class = pair_class;
    ca.class_key = c->buf<unsigned long long>("m_class_key", np);
    ca.num_classes = dev_scalars;
    ca.error = dev_scalars + 1;
    ca.p = p;
    ca.xyz = tb.xyz;
    ca.level_base = tb.level_base;
    ca.class_dl = c->buf<int>("m_class_dl", np);
    ca.periodic = P.on ? 1 : 0;
    ca.disp = disp;
    // Far-pair displacements are below 2 r_root = 2^{max_depth} u in each component (u the smaller
    // half-width, root level 1), so max_depth + 2 bits hold v with a bit to spare; the sort then runs over
    // 3 vbits + 8 key bits instead of 62.
    ca.vbits = std::min(18, std::max(4, tr.max_depth + 2));
    if (P.on) {
        ca.class_hi_in = c->buf<unsigned long long>("m_hi_in", np);
        ca.class_hi_sorted = c->buf<unsigned long long>("m_hi_sorted", np);
        ca.class_lo_final = c->buf<unsigned long long>("m_lo_final", np);
        ca.class_key_hi = c->buf<unsigned long long>("m_class_key_hi", np);
    }
    c->cache_tiles_p = -1;
    c->cache_block_plan = false;
    c->class_vbits = ca.vbits;
    ++c->cache_gen;
    if (!P.on) {
        // dense numbering: no sort of the pairs (the grouping by class is made when an engine needs it)
        const int nwg = nblk(np, 256);
        ca.wg_range = c->buf<int>("m_wg_range", 4 * (size_t)nwg);
        launch_m2l_keys(ca, s);
        int* marks = c->buf<int>("m_dense_marks", CLASS_DENSE_CAP);
        if (!c->dense_marks_ready) {
            FMMH_HIP(hipMemsetAsync(marks, 0, sizeof(int) * (size_t)CLASS_DENSE_CAP, s));
            c->dense_marks_ready = true;
        }
        launch_class_dense(ca, nwg, c->buf<int>("m_dense_dims", 8), marks, c->buf<int>("m_dense_ids", CLASS_DENSE_CAP),
                           c->buf<int>("m_dense_tsum", CLASS_DENSE_CAP / 4096),
                           c->buf<int>("m_dense_toff", CLASS_DENSE_CAP / 4096), s);
        c->classes_grouped = false;
    } else {
        launch_m2l_keys(ca, s);
        m2l_sort_classes(c, ca, true);
    }
    FMMH_HIP(hipMemcpyAsync(c->h_pinned + 300, dev_scalars, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    return true;
}

// The pairs grouped by class by sorting the class keys: sorted order (pair_idx_sorted), class-start flags
// and their scan; with assign, also the class ids, keys and level differences and the count.
void m2l_sort_classes(fmmh_context* c, M2LClassArgs& ca, bool assign) {
    hipStream_t s = c->stream;
    const int np = ca.n_pairs;
    size_t tbytes = std::max(sort_pairs_u64_temp(np), exscan_i32_temp(np));
    void* temp = c->buf<char>("m_temp", tbytes);
    if (!ca.periodic) {
        sort_pairs_u64(temp, tbytes, ca.class_key_in, ca.class_key_sorted, ca.pair_idx, ca.pair_idx_sorted, np,
                       3 * ca.vbits + 8, s);
    } else {
        // (hi, lo) lexicographic order by two stable radix sorts: lo first, then hi
        int* idx1 = c->buf<int>("m_idx1", np);
        unsigned long long* hi1 = c->buf<unsigned long long>("m_hi1", np);
        sort_pairs_u64(temp, tbytes, ca.class_key_in, ca.class_key_sorted, ca.pair_idx, idx1, np, 63, s);
        launch_gather_u64(ca.class_hi_in, idx1, hi1, np, s);
        sort_pairs_u64(temp, tbytes, hi1, ca.class_hi_sorted, idx1, ca.pair_idx_sorted, np, 21, s);
        launch_gather_u64(ca.class_key_in, ca.pair_idx_sorted, ca.class_lo_final, np, s);
    }
    launch_m2l_flags(ca, s);
    exscan_i32(temp, tbytes, ca.flag, ca.cls, np, s);
    if (assign) launch_m2l_assign(ca, s);
    c->classes_grouped = true;
}

// The class-grouped GEMM's grouping of the pairs, for classes numbered without it (the key sort of the
// packed non-periodic keys, which the dense numbering follows).
void m2l_group_pairs(fmmh_context* c, int np) {
    if (c->classes_grouped) return;
    M2LClassArgs ca{};
    ca.n_pairs = np;
    ca.class_key_in = c->buf<unsigned long long>("m_key_in", np);
    ca.pair_idx = c->buf<int>("m_idx_in", np);
    ca.class_key_sorted = c->buf<unsigned long long>("m_key_sorted", np);
    ca.pair_idx_sorted = c->buf<int>("m_idx_sorted", np);
    ca.flag = c->buf<int>("m_flag", np);
    ca.cls = c->buf<int>("m_cls", np);
    ca.vbits = c->class_vbits;
    m2l_sort_classes(c, ca, false);
}

void m2l_classes_end(fmmh_context* c, int p, bool pending, M2LClassArgs& ca) {
    int ncls = 0;
    if (pending) {
        if (c->h_pinned[301] & 1) throw Error(FMMH_ERR_UNSUPPORTED, "M2L displacement outside the class key range");
        if (c->h_pinned[301] & 4) {   // displacement range too wide for the dense numbering: sort the keys
            m2l_sort_classes(c, ca, true);
            FMMH_HIP(hipMemcpyAsync(c->h_pinned + 300, ca.num_classes, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            c->sync();
        }
        ncls = c->h_pinned[300];
        ca.Tc = c->buf<double>("m_tc", (size_t)ncls * ncomp(2 * p) + M2LB_READ_SLACK);
        // the harmonic engines read the components with z <= 2 only; the VALU form (FMMH_OPT_M2L_MFMA = 0)
        // reads them all (run_m2l completes reduced tensors before it runs)
        ca.zmax = c->m2l_mfma != 0 ? 2 : 2 * p;
        launch_m2l_class_tensors(ca, std::min(ncls, 4096), c->stream);
        c->class_zmax = ca.zmax;
        c->class_args = ca;
    }
    c->stats_classes = ncls;
    c->cache_classes = ncls;
}

// unfold = false: the harmonic engines leave the local-expansion components with t_z >= 2 unset (for a
// downward pass that reads only the reduced components: the reduced L2L and the many-set L2P); the
// target-block form (one set) always completes them.
void run_m2l(fmmh_context* c, TreeOut& tr, ListsOut& L, Tables& tb, int p, int nsets, const double* CM, double* LE,
             bool classes_cached, const Periodic& P, const double* disp, const unsigned char* order = nullptr,
             bool unfold = true) {
    hipStream_t s = c->stream;
    const int nn = tr.num_nodes, np = L.n_m2l;
    const int nm = ncomp(p), nT = ncomp(2 * p);
    int* pair_class = c->buf<int>("m_pair_class", np);
    if (!classes_cached) {
        M2LClassArgs ca{};
        const bool pending = m2l_classes_begin(c, tr, L, tb, p, P, disp, ca);
        if (pending) c->sync();
        m2l_classes_end(c, p, pending, ca);
    }
    const int ncls = (int)c->cache_classes;
    double* Tc = np > 0 ? c->buf<double>("m_tc", (size_t)ncls * nT + M2LB_READ_SLACK) : nullptr;
    c->stats_classes = ncls;
    c->mark("m2l_classes");
    // auto: the per-target MFMA form from M2L_MFMA_MIN_SETS sets on (a 16-set tile, padded, costs about
    // as much as 6 passes of the class-grouped GEMM at p 10), or whenever the GEMM's slot buffer does not fit
    const bool gemm_fits = p <= M2L_GEMM_PMAX && np > 0 && (double)np * rdim(p) * sizeof(double) <= M2L_GEMM_MAX_BYTES;
    const bool mfma = c->m2l_mfma == 2 ? m2l_mfma_supported(p, 16)
                                       : (c->m2l_mfma == 1 && (m2l_mfma_supported(p, nsets) ||
                                                               (m2l_mfma_supported(p, 16) &&
                                                                (nsets >= M2L_MFMA_MIN_SETS || !gemm_fits))));
    if (order && c->m2l_mfma == 0)
        throw Error(FMMH_ERR_UNSUPPORTED, "learned expansion orders need an MFMA M2L engine (FMMH_OPT_M2L_MFMA != 0)");
    const size_t node_stride = (size_t)nn * nm;
    // target-block form of the class-grouped GEMM: one FP64 set at a time, every pair at order p
    const bool block_form = !mfma && (c->m2l_mfma == 1 || c->m2l_mfma == 3) && !order && c->precision != 1 &&
                            p <= M2LB_PMAX && np > 0 && ncls <= M2LB_MAX_CLASSES;
    if (block_form) {
        const int nbm = nblk(nn, M2LB_BT);
        M2LBlockArgs ba{};
        ba.m2l_off = L.m2l_off; ba.m2l_src = L.m2l_src; ba.pair_class = pair_class;
        ba.class_dl = c->buf<int>("m_class_dl", np);
        ba.level = tr.t.level; ba.octant = tr.t.octant;
        ba.nn = nn; ba.ncls = ncls; ba.nbm = nbm;
        ba.tperm = c->buf<int>("b_tperm", nn);
        ba.tinv = c->buf<int>("b_tinv", nn);
        ba.tcount = c->buf<int>("b_tcount", nn + 1);
        ba.pbase = c->buf<int>("b_pbase", nn + 1);
        ba.ntile = c->buf<int>("b_ntile", nbm + 1);
        ba.toff = c->buf<int>("b_toff", nbm + 1);
        ba.rec = c->buf<int>("b_rec", (size_t)np * M2LB_REC);
        ba.ntg = c->buf<int>("b_ntg", 1);
        ba.flag = c->buf<int>("x_bad_excl", 1);   // cleared by the near-field launch of this call
        if (!(classes_cached && c->cache_block_plan)) {
            launch_m2lb_order(ba, s);
            exscan_auto(c, "b_scan_temp", ba.tcount, ba.pbase, nn + 1, s);
            launch_m2lb_tiles(ba, s);
            exscan_auto(c, "b_scan_temp", ba.ntile, ba.toff, nbm + 1, s);
            c->cache_block_plan = true;
            ++c->cache_gen;
        }
        // two workgroups per compute unit (LDS: the block's local expansions plus the staged tile)
        const int grid = std::max(1, 2 * c->num_cu);
        const size_t red_stride = (size_t)nn * rdim(p);
        double* CMs = c->buf<double>("m_cms", red_stride * nsets + M2LB_READ_SLACK);
        for (int set = 0; set < nsets; ++set)
            launch_scale_sources_red(CM + set * node_stride, tb.cs, tb.fold, tr.t.r, CMs + set * red_stride, nn, p, 1, 1,
                                     s);
        M2LBRunArgs ra{ba.toff, ba.pbase, ba.rec, nbm, ba.ntg, ba.tperm, ba.tinv, nn, nullptr, Tc, tb.gemm_fidx,
                       tb.gemm_tdeg, tb.gemm_tfull,
                       c->buf<double>("b_partial", (size_t)(nbm + grid) * rdim(p) * M2LB_BT), grid};
        for (int set = 0; set < nsets; ++set) {
            ra.CMs = CMs + set * red_stride;
            launch_m2lb_run(ra, p, tr.t.r, LE + set * node_stride, s);
        }
        return;
    }
    // class-grouped GEMM: one set at a time, per-pair slots of np x nm doubles
    const bool gemm = !mfma && (c->m2l_mfma == 1 || c->m2l_mfma == 3 || c->m2l_mfma == 4) && gemm_fits;
    if (gemm) {
        m2l_group_pairs(c, np);
        const int* flag = c->buf<int>("m_flag", np);
        const int* cls = c->buf<int>("m_cls", np);
        const int* pair_csr = c->buf<int>("m_idx_sorted", np);
        const int* class_dl = c->buf<int>("m_class_dl", np);
        const int* grp_class = nullptr;
        const int* grp_order = nullptr;
        int ngrp = ncls;
        if (order) {
            // tiles of one (class, order) group: re-sort the class-sorted pairs by (class, order)
            unsigned long long* gk_in = c->buf<unsigned long long>("g_keys_in", np);
            unsigned long long* gk = c->buf<unsigned long long>("g_keys", np);
            int* gv_in = c->buf<int>("g_vals_in", np);
            int* gv = c->buf<int>("g_vals", np);
            int* gflag = c->buf<int>("g_flag", np);
            int* gid = c->buf<int>("g_id", np);
            int* gcls = c->buf<int>("g_class", np);
            int* gord = c->buf<int>("g_order", np);
            int* gcount = c->buf<int>("g_count", 1);
            launch_sel_groups(pair_csr, pair_class, order, np, gk_in, gv_in, s);
            int bits = 4;
            while ((1ll << (bits - 4)) < (long long)ncls) ++bits;
            size_t gb = std::max(sort_pairs_u64_temp(np), exscan_i32_temp(np));
            void* gtemp = c->buf<char>("g_temp", gb);
            sort_pairs_u64(gtemp, gb, gk_in, gk, gv_in, gv, np, bits, s);
            launch_sel_group_flags(gk, np, gflag, s);
            exscan_i32(gtemp, gb, gflag, gid, np, s);
            launch_sel_group_info(gk, gflag, gid, np, gcls, gord, gcount, s);
            FMMH_HIP(hipMemcpyAsync(c->h_pinned + 302, gcount, sizeof(int), hipMemcpyDeviceToHost, s));
            c->sync();
            ngrp = c->h_pinned[302];
            flag = gflag;
            cls = gid;
            pair_csr = gv;
            grp_class = gcls;
            grp_order = gord;
        }
        int* class_off = c->buf<int>("m_class_off", ngrp + 1);
        int* tile_cnt = c->buf<int>("m_tile_cnt", ngrp + 1);
        int* tile_off = c->buf<int>("m_tile_off", ngrp + 1);
        // tile tables of the (cached) classes without learned orders are kept with them
        if (!(classes_cached && !order && c->cache_tiles_p == p)) {
            launch_m2l_tiles(flag, cls, np, ngrp, class_off, tile_cnt, s);
            exscan_auto(c, "m_temp_tiles", tile_cnt, tile_off, ngrp + 1, s);
            c->cache_tiles_p = order ? -1 : p;
            ++c->cache_gen;
        }
        const int grid = ngrp + (np + M2L_GEMM_TN - 1) / M2L_GEMM_TN;   // >= tiles
        const bool f32 = c->precision == 1;
        const int nk = rdim(p);
        const size_t red_stride = (size_t)nn * nk;
        double* pbuf = f32 ? nullptr : c->buf<double>("m_pbuf", (size_t)np * nk);
        float* pbuf32 = f32 ? c->buf<float>("m_pbuf32", (size_t)np * nk) : nullptr;
        float* cms32 = f32 ? c->buf<float>("m_cms32", red_stride) : nullptr;
        double* CMs = c->buf<double>("m_cms", red_stride * nsets);
        for (int set = 0; set < nsets; ++set)
            launch_scale_sources_red(CM + set * node_stride, tb.cs, tb.fold, tr.t.r, CMs + set * red_stride, nn, p, 1, 1,
                                     s);
        for (int set = 0; set < nsets; ++set) {
            if (f32) launch_to_f32(CMs + set * red_stride, cms32, (int64_t)red_stride, s);
            M2LGemmArgs ga{class_off, tile_off, ngrp, pair_csr, L.m2l_src, class_dl, Tc, CMs + set * red_stride,
                           tb.gemm_fidx, tb.gemm_tdeg, pbuf, p, f32 ? 1 : 0, cms32, pbuf32, grp_class, grp_order,
                           order, tb.gemm_tfull};
            launch_m2l_gemm(ga, grid, s);
            launch_m2l_reduce(ga, L.m2l_off, tr.t.r, LE + set * node_stride, nn, s);
            if (unfold) launch_le_unfold(LE + set * node_stride, nn, p, s);
        }
        return;
    }
    // MFMA form: up to 64 sets per pass (one wavefront per target and 16 sets)
    const int chunk = mfma ? 64 : nsets;
    auto pad_sets = [](int ns) { return ns <= 16 ? 16 : ((ns + 31) / 32) * 32; };
    const int s_ld_max = mfma ? pad_sets(std::min(chunk, nsets)) : nsets;
    double* CMs = c->buf<double>("m_cms", node_stride * s_ld_max);
    const bool f32 = mfma && c->precision == 1;
    float* CMs32 = f32 ? c->buf<float>("m_cms32", node_stride * s_ld_max) : nullptr;
    void* Tcp = nullptr;
    if (mfma && ncls > 0) {
        const size_t tb_bytes = (size_t)ncls * m2l_mfma_class_stride(p, f32) * (f32 ? sizeof(float) : sizeof(double));
        Tcp = c->buf<char>("m_tcp", tb_bytes);
        launch_m2l_class_pad(Tc, Tcp, ncls, p, f32, s);
    }
    for (int c0 = 0; c0 < nsets; c0 += chunk) {
        const int ns = std::min(chunk, nsets - c0);
        const int s_ld = mfma ? pad_sets(ns) : 0;
        if (mfma) launch_scale_sources_red(CM + c0 * node_stride, tb.cs, tb.fold, tr.t.r, CMs, nn, p, ns, s_ld, s);
        else launch_scale_sources(CM + c0 * node_stride, tb.cs, tr.t.r, CMs, nn, p, ns, 0, s);
        if (f32) launch_to_f32(CMs, CMs32, (int64_t)nn * (p + 1) * (p + 1) * s_ld, s);
        M2LArgs ma{L.m2l_off, L.m2l_src, pair_class, Tc, CMs, LE + c0 * node_stride, tr.t.level, tr.t.r,
                   nm, p, ns, s_ld, nn, order, tb.gemm_fidx, CMs32, Tcp};
        if (!mfma) {
            if (c->class_zmax < 2 * p && ncls > 0) {   // the VALU form reads every tensor component
                M2LClassArgs fa = c->class_args;
                fa.zmax = 2 * p;
                fa.Tc = Tc;
                launch_m2l_class_tensors(fa, std::min(ncls, 4096), s);
                c->class_zmax = 2 * p;
            }
            launch_m2l(ma, nn, s);
            continue;
        }
        const long long key = ((long long)p << 32) | ((long long)(f32 ? 1 : 0) << 24) | ((long long)ns << 8) |
                              (order ? 1 : 0);
        auto it = c->tuned.find(key);
        if (it != c->tuned.end() || !c->autotune || c->capturing) {
            launch_m2l_mfma(ma, nn, it != c->tuned.end() ? it->second : m2l_mfma_default_variant(ma), s);
            if (unfold) launch_le_unfold(LE + c0 * node_stride, (long long)ns * nn, p, s);
            continue;
        }
        // first call of this shape: time every valid variant on this call's pairs; each one writes
        // the same local expansions (same sums in the same order), so the last run's output stands
        if (!c->tune_ev[0]) {
            FMMH_HIP(hipEventCreate(&c->tune_ev[0]));
            FMMH_HIP(hipEventCreate(&c->tune_ev[1]));
        }
        int best = m2l_mfma_default_variant(ma);
        float best_ms = 1e30f;
        c->tune_ms.assign(M2L_MFMA_VARIANTS, -1.0);
        for (int v = 0; v < M2L_MFMA_VARIANTS; ++v) {
            if (!m2l_mfma_variant_ok(ma, v)) continue;
            FMMH_HIP(hipEventRecord(c->tune_ev[0], s));
            launch_m2l_mfma(ma, nn, v, s);
            FMMH_HIP(hipEventRecord(c->tune_ev[1], s));
            FMMH_HIP(hipEventSynchronize(c->tune_ev[1]));
            float ms = 0.f;
            FMMH_HIP(hipEventElapsedTime(&ms, c->tune_ev[0], c->tune_ev[1]));
            c->tune_ms[v] = ms;
            if (ms < best_ms) { best_ms = ms; best = v; }
        }
        c->tuned[key] = best;
        if (unfold) launch_le_unfold(LE + c0 * node_stride, (long long)ns * nn, p, s);
    }
}

// ---------------------------------------------------------------- periodic
// Lattice sum into the root local expansions and the dipole correction into
// the field (src/fmm.F90:1392-1395), after M2L and before L2L.  The shell
// tensor sums depend only on the box and p and are kept between calls.
void run_periodic(fmmh_context* c, TreeOut& tr, Tables& tb, const Periodic& P, const Call& k, int p, const double* CM,
                  double* LE) {
    hipStream_t s = c->stream;
    const int nT = ncomp(2 * p);
    double* boxd = c->buf<double>("p_box", 9);
    double* hbox = reinterpret_cast<double*>(c->h_pinned + 700);
    std::memcpy(hbox, P.box, sizeof(P.box));
    FMMH_HIP(hipMemcpyAsync(boxd, hbox, sizeof(P.box), hipMemcpyHostToDevice, s));
    double* shells = c->buf<double>("p_shells", (size_t)PSUM_LEVELS * 26 * nT);
    // volume exactly as src/fmm.F90:961-963 forms it (including its h(1,2) in the last cofactor)
    auto H = [&](int r, int col) { return P.box[(r - 1) + 3 * (col - 1)]; };
    const double volume = H(1, 1) * (H(2, 2) * H(3, 3) - H(2, 3) * H(3, 2)) -
                          H(1, 2) * (H(2, 1) * H(3, 3) - H(2, 3) * H(3, 1)) +
                          H(1, 3) * (H(1, 2) * H(3, 2) - H(2, 2) * H(3, 1));
    const double pi = 3.14159265359;   // the reference's literal
    PeriodicArgs pa{boxd, shells, p, tb.xyz, tb.level_base, CM, LE, tr.num_nodes, k.nsets, tb.cs, tb.m2m,
                    tb.mono_parent, tb.mono_axis, k.X, k.ldx, k.n, k.field, k.ldf, ncomp(k.F), 4.0 * pi, 2.0 * pi,
                    3.0 * volume, c->buf<int>("p_levels", k.nsets)};
    if (!(c->shells_valid && c->shells_p == p && std::memcmp(c->shells_box, P.box, sizeof(P.box)) == 0)) {
        c->shells_valid = false;
        launch_periodic_shells(pa, s);
        c->shells_valid = true;
        c->shells_p = p;
        std::memcpy(c->shells_box, P.box, sizeof(P.box));
    }
    launch_periodic_sum(pa, s);
    launch_dipole_correction(pa, s);
    c->mark("periodic");
}

// ---------------------------------------------------------------- field_fmm
__global__ void k_coords_differ(const double* __restrict__ X, int64_t ldx, const double* __restrict__ C, int n,
                                int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool d = false;
    for (int q = 0; q < 3; ++q)
        d |= __double_as_longlong(X[q * ldx + i]) != __double_as_longlong(C[(int64_t)q * n + i]);
    if (d) *flag = 1;
}

// True when the cached tree was built from bitwise the same coordinates with
// the same n, ncrit and theta (one device compare, one host sync).
bool cached_tree_matches(fmmh_context* c, const Call& k, double theta, int ncrit, const Periodic& P) {
    if (!c->cache_valid || c->cache_n != k.n || c->cache_ncrit != ncrit || c->cache_theta != theta) return false;
    if (c->cache_periodic != P.on || (P.on && std::memcmp(c->cache_box, P.box, sizeof(P.box)) != 0)) return false;
    hipStream_t s = c->stream;
    int* flag = c->buf<int>("c_flag", 1);
    FMMH_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_coords_differ, dim3(nblk(k.n, 256)), dim3(256), 0, s, k.X, k.ldx,
                       c->buf<double>("c_X", (size_t)k.n * 3), k.n, flag);
    FMMH_HIP(hipGetLastError());
    FMMH_HIP(hipMemcpyAsync(c->h_pinned + 400, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    c->sync();
    return c->h_pinned[400] == 0;
}

// ---------------------------------------------------------------- learned orders
// Per-pair expansion orders for this call (CSR positions): source norms,
// the per-pair tolerance from the near-field RMS and the list sizes along
// each site's path (on the device), then the MLP per pair.
const unsigned char* select_orders(fmmh_context* c, TreeOut& tr, ListsOut& L, Tables& tb, int p, int nsets,
                                   const double* CM, const double* near, int n, int nf) {
    hipStream_t s = c->stream;
    const int nn = tr.num_nodes, np = L.n_m2l, nm = ncomp(p);
    double* node_norm = c->buf<double>("sel_node_norm", nn);
    launch_sel_node_norm(CM, tr.t.r, tb.xyz, nn, nm, nsets, node_norm, s);
    double* acc = c->buf<double>("sel_acc", 4 + (size_t)nsets);
    FMMH_HIP(hipMemsetAsync(acc, 0, (4 + (size_t)nsets) * sizeof(double), s));
    launch_sel_budget(near, n, n, nf, nsets, tr.t.leaf_node, tr.t.count, tr.t.parent, L.m2l_off, tr.num_leaves,
                      acc, c->order_tol * c->sel_kappa * c->sel_adapt, s);
    unsigned char* order = c->buf<unsigned char>("sel_order", np);
    unsigned long long* hist = c->buf<unsigned long long>("sel_hist", PMAX + 1);
    FMMH_HIP(hipMemsetAsync(hist, 0, (PMAX + 1) * sizeof(unsigned long long), s));
    SelOrderArgs oa{L.m2l_keys, np, tr.t.cx, tr.t.cy, tr.t.cz, tr.t.r, node_norm, acc + 2,
                    c->buf<float>("sel_W", SEL_NPARAM), c->buf<float>("sel_norm", 2 * SEL_IN), p, c->sel_cap,
                    c->sel_force_full, c->sel_margin, c->sel_rel_lo, c->sel_rel_hi, order, hist};
    launch_sel_orders(oa, s);
    c->mark("select");
    return order;
}

// Relative RMS error (field block; the potential for field order 0) of the
// far field at spot sites against the direct sum with the reference's far
// semantics.  FMM far field at a site = field - field before L2P - near; the
// direct far field is the sum of the source chunks' parts
// dir[chunk][set][spot][4], in chunk order.  k_spot_compare: workgroup
// (set, b) sums the squared error and field over entries x = 4 spot + c in
// [SPOT_CMP_X b, SPOT_CMP_X (b + 1)) into part[set][b][2]; k_spot_worst adds
// each set's parts in block order and returns the largest relative error
// (a non-finite one wins).
constexpr int SPOT_CMP_X = 1024;
__global__ __launch_bounds__(256) void k_spot_compare(const double* __restrict__ dir, const int* __restrict__ site,
                                                      const int* __restrict__ perm, const double* __restrict__ near,
                                                      int64_t ldn, const double* __restrict__ field, int64_t ldf,
                                                      const double* __restrict__ before, int64_t ldb, int nspots,
                                                      int nf, int nsets, int nchunks, double* __restrict__ part) {
    __shared__ double red[2][256];
    const int set = blockIdx.x;
    const int c0 = nf >= 4 ? 1 : 0, c1 = nf >= 4 ? 4 : 1;
    const double* __restrict__ nr = near + (int64_t)set * nf * ldn;
    const double* __restrict__ fl = field + (int64_t)set * nf * ldf;
    const double* __restrict__ bf = before + (int64_t)set * nf * ldb;
    double e2 = 0.0, f2 = 0.0;
    const int x1 = min(4 * nspots, SPOT_CMP_X * ((int)blockIdx.y + 1));
    for (int x = SPOT_CMP_X * blockIdx.y + threadIdx.x; x < x1; x += 256) {   // chunk loads are coalesced
        const int sp = x >> 2, c = x & 3;
        if (c < c0 || c >= c1) continue;
        const int k = site[sp], o = perm[k];
        const double nv = nr[(int64_t)c * ldn + k];
        const double ffar = fl[(int64_t)c * ldf + o] - bf[(int64_t)c * ldb + o] - nv;
        double d = 0.0;
        for (int ch = 0; ch < nchunks; ++ch) d += dir[((int64_t)ch * nsets + set) * nspots * 4 + x];
        e2 += (ffar - d) * (ffar - d);
        f2 += (nv + d) * (nv + d);
    }
    red[0][threadIdx.x] = e2;
    red[1][threadIdx.x] = f2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) part[((int64_t)set * gridDim.y + blockIdx.y) * 2 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_spot_worst(const double* __restrict__ part, int nsets, int nb, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double worst = 0.0;
    for (int set = 0; set < nsets; ++set) {
        double e2 = 0.0, f2 = 0.0;
        for (int b = 0; b < nb; ++b) {
            e2 += part[((int64_t)set * nb + b) * 2];
            f2 += part[((int64_t)set * nb + b) * 2 + 1];
        }
        const double e = f2 > 0.0 ? sqrt(e2 / f2) : 0.0;   // a set without field counts 0
        worst = (e > worst || e != e) ? e : worst;
    }
    out[0] = worst;
}

// The direct far field at the spot sites depends on the sources and lists only: spot_direct launches
// it once, spot_compare compares it with the FMM's far field as often as the far field is redone (the
// full-order fallback reuses it).
struct SpotPlan {
    int nspots = 0, nchunks = 0, ncb = 0;
    double* out = nullptr;
    int* site = nullptr;
    double* res = nullptr;
};

SpotPlan spot_direct(fmmh_context* c, TreeOut& tr, ListsOut& L, const double* xs, const double* ys,
                     const double* zs, const double* ms, int n, int m, int nsets, int nspots) {
    SpotPlan sp;
    sp.nspots = nspots = std::min(nspots, n);
    // source chunks: about four workgroups per compute unit in all
    const int tile = spot_tile(m);
    const int sets_blocks = (nsets + 31) / 32;
    const long long others = (long long)spot_row_blocks(nspots) * sets_blocks;
    const int want = (int)std::max(1ll, (4ll * c->num_cu + others - 1) / others);
    const int chunk = std::max(tile, (int)(((long long)n + (long long)want * tile - 1) / ((long long)want * tile)) * tile);
    sp.nchunks = (n + chunk - 1) / chunk;
    sp.out = c->buf<double>("spot_out", (size_t)sp.nchunks * nspots * 4 * nsets);
    sp.site = c->buf<int>("spot_site", nspots);
    sp.ncb = (4 * nspots + SPOT_CMP_X - 1) / SPOT_CMP_X;
    sp.res = c->buf<double>("spot_res", 1 + (size_t)nsets * sp.ncb * 2);
    SpotArgs sa{nspots, tr.num_leaves, n, tr.t.leaf_node, tr.t.start, tr.t.count, L.p2p_off, L.p2p_src, xs, ys, zs,
                ms, n, sp.out, sp.site, nsets, chunk, sp.nchunks};
    launch_spot_far(sa, m, c->stream);
    return sp;
}

double spot_compare(fmmh_context* c, TreeOut& tr, const SpotPlan& sp, int n, int nf, const double* near,
                    const Call& k, const double* before) {
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_spot_compare, dim3(k.nsets, sp.ncb), dim3(256), 0, s, sp.out, sp.site, tr.t.perm, near,
                       (int64_t)n, k.field, k.ldf, before, (int64_t)n, sp.nspots, nf, k.nsets, sp.nchunks,
                       sp.res + 1);
    hipLaunchKernelGGL(k_spot_worst, dim3(1), dim3(64), 0, s, sp.res + 1, k.nsets, sp.ncb, sp.res);
    FMMH_HIP(hipGetLastError());
    double* h = reinterpret_cast<double*>(c->h_pinned + 800);
    FMMH_HIP(hipMemcpyAsync(h, sp.res, sizeof(double), hipMemcpyDeviceToHost, s));
    c->sync();
    c->mark("spot_check");
    return h[0];
}

// Near-field workgroup order (leaves, and chunks of heavy leaves, in the XCD-aware order), in two
// halves around one host round trip: near_order_begin launches the per-leaf work, its prefix sums, the
// chunk counts and the readback of the sizes; after a stream sync near_order_end lays out the chunks,
// sorts them and fills the workgroup slots ("s_work_blk", near_grid of them).
void near_order_begin(fmmh_context* c, TreeOut& tr, ListsOut& L) {
    hipStream_t s = c->stream;
    const int nl = tr.num_leaves;
    long long* work = c->buf<long long>("s_work", nl);
    long long* pre = c->buf<long long>("s_work_pre", nl + 1);
    int* lk = c->buf<int>("s_leaf_k", nl + 1);
    int* lx = c->buf<int>("s_leaf_x", nl);
    int* lps = c->buf<int>("s_leaf_ps", nl + 1);
    int* lvb = c->buf<int>("s_leaf_vb", nl + 1);
    int* lpart = c->buf<int>("s_leaf_part", nl + 1);
    int* range_cnt = c->buf<int>("s_range_cnt", NUM_XCD);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, range_cnt, 0, NUM_XCD);
    hipLaunchKernelGGL(k_leaf_work, dim3(nl), dim3(64), 0, s, L.p2p_off, L.p2p_src, tr.t.leaf_node, tr.t.count,
                       nl, c->buf<int>("s_entry_cnt", std::max(L.n_p2p, 1)), c->buf<long long>("s_leaf_tot", nl),
                       work);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, s, work, pre, nl);
    hipLaunchKernelGGL(k_leaf_chunks, dim3(nblk(nl + 1, 256)), dim3(256), 0, s, work, pre, L.p2p_off,
                       tr.t.leaf_node, tr.t.count, nl, c->near_chunk_cap, lk, lx, lps, range_cnt);
    exscan_auto(c, "s_work_temp", lk, lvb, nl + 1, s);
    exscan_auto(c, "s_work_temp", lps, lpart, nl + 1, s);
    int* hp = c->h_pinned + 900;
    FMMH_HIP(hipMemcpyAsync(hp, range_cnt, NUM_XCD * sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipMemcpyAsync(hp + NUM_XCD, lvb + nl, sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipMemcpyAsync(hp + NUM_XCD + 1, lpart + nl, sizeof(int), hipMemcpyDeviceToHost, s));
}

void near_order_end(fmmh_context* c, TreeOut& tr, ListsOut& L) {
    hipStream_t s = c->stream;
    const int nl = tr.num_leaves;
    long long* work = c->buf<long long>("s_work", nl);
    long long* pre = c->buf<long long>("s_work_pre", nl + 1);
    int* lk = c->buf<int>("s_leaf_k", nl + 1);
    int* lx = c->buf<int>("s_leaf_x", nl);
    int* lvb = c->buf<int>("s_leaf_vb", nl + 1);
    int* lpart = c->buf<int>("s_leaf_part", nl + 1);
    int* hp = c->h_pinned + 900;
    const int nv = hp[NUM_XCD];
    if (nv <= 0 || nv > 0xffffff) throw Error(FMMH_ERR_UNSUPPORTED, "near-field workgroup count out of range");
    RangeStarts rs;
    int per = 1;
    rs.first[0] = 0;
    for (int x = 0; x < NUM_XCD; ++x) {
        rs.first[x + 1] = rs.first[x] + hp[x];
        per = std::max(per, hp[x]);
    }
    c->near_nv = nv;
    c->near_grid = per * NUM_XCD;
    c->near_np = hp[NUM_XCD + 1];
    int* vb_node = c->buf<int>("s_vb_node", nv);
    int* vb_e0 = c->buf<int>("s_vb_e0", nv);
    int* vb_e1 = c->buf<int>("s_vb_e1", nv);
    int* vb_part = c->buf<int>("s_vb_part", nv);
    int* vbucket = c->buf<int>("s_vb_bucket", nv);
    hipLaunchKernelGGL(k_leaf_vblocks, dim3(nl), dim3(64), 0, s, work, c->buf<long long>("s_leaf_tot", nl), pre,
                       c->near_chunk_cap, L.p2p_off, c->buf<int>("s_entry_cnt", std::max(L.n_p2p, 1)),
                       tr.t.leaf_node, tr.t.count, nl, lk, lx, lvb, lpart, vb_node, vb_e0, vb_e1, vb_part, vbucket);
    int* slots = c->buf<int>("s_work_blk", (size_t)c->near_grid);
    hipLaunchKernelGGL(k_vb_order, dim3(1), dim3(VB_ORDER_T), 0, s, vbucket, nv, rs, slots, c->near_grid);
}

void run_fmm(fmmh_context* c, const Call& k, double theta, int ncrit, int p, const Periodic& P) {
    c->host_waits = 0;
    hipStream_t s = c->stream;
    const double* disp = upload_disp(c, P);
    const int n = k.n;
    const int m = order_from_dim(k.dim);
    const int nf = ncomp(k.F);
    const int nm = ncomp(p);
    TreeOut tr;
    ListsOut L;
    const int S = k.nsets;
    Tables& tb = get_tables(c, p, m);
    // Sorted site arrays, expansions and the upward pass (gather, P2M, M2M): they need the tree but not the
    // lists, so a call that builds its tree runs them while the host waits for the list sizes.
    double *xs = nullptr, *ys = nullptr, *zs = nullptr, *ms_raw = nullptr, *ms = nullptr, *dfs = nullptr;
    double *CM = nullptr, *LE = nullptr;
    int* excl = nullptr;
    int* inodes = nullptr;
    int* icnt = nullptr;
    bool shift32 = false;
    UpArgs up{};
    bool upward_done = false;
    auto upward_pass = [&]() {
        const int nn = tr.num_nodes, nl = tr.num_leaves;
        xs = c->buf<double>("s_xs", n);
        ys = c->buf<double>("s_ys", n);
        zs = c->buf<double>("s_zs", n);
        ms_raw = c->buf<double>("s_ms_raw", (size_t)n * k.dim * S);
        ms = c->buf<double>("s_ms", (size_t)n * k.dim * S);
        dfs = c->buf<double>("s_dfs", n);
        excl = c->buf<int>("s_excl", (size_t)n * std::max(k.nexcl, 1));
        SiteArgs g{k.X, k.ldx, k.M, k.ldm, k.dim, k.E, k.lde, k.nexcl, k.damp ? k.df : nullptr, tr.t.perm, n,
                   xs, ys, zs, ms_raw, ms, dfs, excl, tb.scale, S};
        launch_gather(g, s);
        c->mark("gather");
        CM = c->buf<double>("e_cm", (size_t)nn * nm * S);
        LE = c->buf<double>("e_le", (size_t)nn * nm * S);
        up = UpArgs{xs, ys, zs, ms_raw, n, tr.t.leaf_node, tr.t.start, tr.t.count, tr.t.cx, tr.t.cy, tr.t.cz,
                    tr.t.children, CM, nm, p, tb.nsrc, tb.p2m, tb.m2m, tb.mono_parent, tb.mono_axis, nn, S, k.dim,
                    tb.xyz};
        launch_p2m(up, nl, s);
        // mixed precision: M2M and L2L on FP32 MFMA (box-scaled matrices per octant), FP64 accumulation,
        // over per-level lists of internal nodes
        shift32 = c->precision == 1 && p >= 1 && tb.m2m_frag;
        if (shift32) {
            inodes = c->buf<int>("s_internal", nn);
            icnt = c->buf<int>("s_internal_cnt", KEY_LEVELS + 2);
            LevelStarts ls{};
            for (int L = 0; L < KEY_LEVELS + 3; ++L) ls.off[L] = tr.level_off[L];
            launch_internal_nodes(nn, tr.t.level, tr.t.children, ls, inodes, icnt, s);
        }
        for (int lev = tr.max_depth - 1; lev >= 1; --lev) {
            int o0 = tr.level_off[lev], o1 = tr.level_off[lev + 1];
            if (shift32 && (long long)(o1 - o0) * S >= SHIFT_MFMA_MIN_COLS) {
                ShiftMfmaArgs sa{inodes + o0, icnt + lev, o1 - o0, tr.t.children, tr.t.r, CM, CM, (int64_t)nn * nm,
                                 S, tb.m2m_frag, tb.xyz};
                launch_shift_mfma(sa, p, false, s);
            } else {
                launch_m2m(up, tr.t.level_nodes + o0, o1 - o0, s);
            }
        }
        c->mark("upward");
        upward_done = true;
    };
    const bool reuse = c->reuse_tree && cached_tree_matches(c, k, theta, ncrit, P);
    if (reuse) {
        tr = c->cache_tr;
        L = c->cache_L;
        c->list_retries = c->cache_retries;
        c->mark("tree_reused");
    } else {
        c->cache_valid = false;
        c->cache_work_valid = false;
        c->cache_tiles_p = -1;
        c->cache_block_plan = false;
        ++c->cache_gen;
        build_tree(c, k.X, k.ldx, n, ncrit, tr);
        c->mark("tree");
        build_lists(c, tr, theta, L, P.nimg, disp, upward_pass);
        c->mark("lists");
    }
    const int nn = tr.num_nodes, nl = tr.num_leaves;
    c->num_nodes = nn; c->num_leaves = nl; c->max_depth = tr.max_depth;
    c->n_m2l = L.n_m2l; c->n_p2p = L.n_p2p; c->n_sites = n;
    c->level_off = tr.level_off;
    // learned per-pair orders (off: every pair at p, as the reference)
    const bool select = c->order_tol > 0.0 && c->sel_trained && !P.on && L.n_m2l > 0 && m <= SPOT_MMAX;
    // graph replay: a tree-reuse call without timing events, selection or periodic images
    bool replay = false, capture = false;
    std::string gkey;
    if (c->graph_opt && reuse && !c->timing && !select && !P.on) {
        char kb[512];
        std::snprintf(kb, sizeof kb, "%p %lld %p %lld %d %d %d %p %lld %d %d %p %lld %d %p | %.17g %d %d | %d %d %d %zu | %lld %lld",
                      (const void*)k.X, (long long)k.ldx, (const void*)k.M, (long long)k.ldm, k.n, k.dim, k.nsets,
                      (const void*)k.E, (long long)k.lde, k.nexcl, k.F, (const void*)k.field, (long long)k.ldf,
                      k.damp, (const void*)k.df, theta, ncrit, p, c->precision, c->m2l_mfma, c->autotune,
                      c->tuned.size(), (long long)c->pool_gen, (long long)c->cache_gen);
        gkey = kb;
        if (c->graph_exec && c->graph_key == gkey) replay = true;
        else if (c->graph_pending == gkey) capture = true;
        else c->graph_pending = gkey;   // this call runs eagerly; the next one of this shape is captured
    }
    c->sel_stats[0] = -1.0;
    c->sel_stats[1] = 0.0;
    c->sel_full_err = -1.0;
    if (replay) {
        FMMH_HIP(hipGraphLaunch(c->graph_exec, s));
        ++c->graph_replays;
        c->stats_classes = c->cache_classes;   // a replay runs with the cached classes
    } else {
    GraphCapture cap(s, capture, &c->capturing);
    // tree-dependent preparation, kept with a cached tree: the near-field workgroup order and the M2L
    // displacement classes, sized through one host round trip
    const bool work_cached = reuse && c->cache_work_valid;
    const bool classes_cached = reuse && c->cache_class_p == p;
    if (!work_cached) near_order_begin(c, tr, L);
    M2LClassArgs ca{};
    const bool cls_pending = !classes_cached && m2l_classes_begin(c, tr, L, tb, p, P, disp, ca);
    if (!work_cached || cls_pending) c->sync();
    if (!work_cached) {
        near_order_end(c, tr, L);
        c->cache_work_valid = c->reuse_tree;
        ++c->cache_gen;
    }
    if (!classes_cached) m2l_classes_end(c, p, cls_pending, ca);
    c->mark("prep");
    if (!upward_done) upward_pass();
    double* near = c->buf<double>("e_near", (size_t)n * nf * S);
    int* inv = c->buf<int>("s_inv", n);
    // the sorted-site inverse permutation depends on the tree only
    if (!work_cached) hipLaunchKernelGGL(k_invert_perm, dim3(nblk(n, 256)), dim3(256), 0, s, tr.t.perm, inv, n);
    const int nv = c->near_nv, ngrid = c->near_grid;
    const int np_sites = c->near_np;
    NearArgs na{c->buf<int>("s_work_blk", (size_t)ngrid), tr.t.start, tr.t.count, L.p2p_off, L.p2p_src, tr.t.start,
                tr.t.count, xs, ys, zs, ms, n, dfs, tr.t.perm, excl, n, k.nexcl, near, n, 0, S, disp, inv,
                c->buf<int>("s_vb_node", nv), c->buf<int>("s_vb_e0", nv), c->buf<int>("s_vb_e1", nv),
                c->buf<int>("s_vb_part", nv),
                np_sites > 0 ? c->buf<double>("s_near_part", (size_t)np_sites * nf * S) : nullptr, np_sites};
    na.nsites = n;
    na.bad = bad_excl_flag(c);
    na.coinc_list = 0;   // sites at one position share a leaf
    na.pairs2 = S >= EXP_MFMA_MIN_SETS ? 1 : 0;   // many sets: two targets per lane (16 sets and more)
    c->mark("p2p_order");
    launch_near_part(na, ngrid, m, k.F, k.damp, 1, s);
    if (np_sites > 0)
        launch_near_fix(na, tr.t.leaf_node, c->buf<int>("s_leaf_k", nl + 1), c->buf<int>("s_leaf_part", nl + 1),
                        nl, nf, s);
    c->mark("p2p");
    launch_near_part(na, ngrid, m, k.F, k.damp, 2, s);
    c->mark("p2p_excl");
    const unsigned char* order = select ? select_orders(c, tr, L, tb, p, S, CM, near, n, nf) : nullptr;
    // many sets in mixed precision: the downward pass runs on the reduced (harmonic) components -- the FP32
    // L2L shifts with the reduced matrices and the matrix-core L2P read only those -- so the M2L leaves the
    // components with t_z >= 2 unset
    const bool reduced_down = shift32 && S >= EXP_MFMA_MIN_SETS && p <= EXP_MFMA_PMAX && (nf == 1 || nf == 4) &&
                              !P.on && tb.l2l_frag_red && tb.gemm_tfull;
    run_m2l(c, tr, L, tb, p, S, CM, LE, true, P, disp, order, !reduced_down);
    c->mark("m2l");
    if (P.on) run_periodic(c, tr, tb, P, k, p, CM, LE);
    DownArgs da{tr.t.cx, tr.t.cy, tr.t.cz, tr.t.parent, LE, nm, p, tb.l2l, tb.l2l, tb.mono_parent, tb.mono_axis,
                tr.t.leaf_node, tr.t.start, tr.t.count, xs, ys, zs, tr.t.perm, near, n, k.field, k.ldf, nf, nn, S, tb.xyz};
    if (S >= EXP_MFMA_MIN_SETS && p <= EXP_MFMA_PMAX && (nf == 1 || nf == 4)) {
        da.basis = c->buf<double>("l2p_basis", (size_t)rdim(p) * n);   // the expansions here are harmonic
        da.ld_basis = n;
    }
    auto downward = [&]() {
        for (int lev = 2; lev <= tr.max_depth; ++lev) {
            if (shift32) {   // the parents of this level's nodes, on FP32 MFMA (one octant per workgroup)
                int o0 = tr.level_off[lev - 1], o1 = tr.level_off[lev];
                ShiftMfmaArgs sa{inodes + o0, icnt + lev - 1, o1 - o0, tr.t.children, tr.t.r, LE, LE,
                                 (int64_t)nn * nm, S, reduced_down ? tb.l2l_frag_red : tb.l2l_frag, tb.xyz,
                                 reduced_down ? tb.gemm_tfull : nullptr};
                launch_shift_mfma(sa, p, true, s);
            } else {
                int o0 = tr.level_off[lev], o1 = tr.level_off[lev + 1];
                launch_l2l(da, tr.t.level_nodes + o0, o1 - o0, s);
            }
        }
        c->mark("l2l");
        launch_l2p(da, nl, s);
        c->mark("l2p");
    };
    const bool check = select && c->spot_checks > 0;
    double* before = nullptr;
    if (check) {
        // the field before L2P, to isolate the far field at the spot sites and to restore on fallback
        before = c->buf<double>("spot_before", (size_t)n * nf * S);
        FMMH_HIP(hipMemcpy2DAsync(before, n * sizeof(double), k.field, k.ldf * sizeof(double), n * sizeof(double),
                                  (size_t)nf * S, hipMemcpyDeviceToDevice, s));
    }
    downward();
    if (check) {
        const SpotPlan spot = spot_direct(c, tr, L, xs, ys, zs, ms, n, m, S, c->spot_checks);
        const double err = spot_compare(c, tr, spot, n, nf, near, k, before);
        c->sel_stats[0] = err;
        if (!(err <= c->order_tol)) {
            // the learned orders missed the tolerance: redo the far field at expansion_order
            c->sel_stats[1] = 1.0;
            FMMH_HIP(hipMemcpy2DAsync(k.field, k.ldf * sizeof(double), before, n * sizeof(double),
                                      n * sizeof(double), (size_t)nf * S, hipMemcpyDeviceToDevice, s));
            c->cache_classes = c->stats_classes;   // the class tensors of this call are in place
            run_m2l(c, tr, L, tb, p, S, CM, LE, true, P, disp, nullptr, !reduced_down);
            c->mark("m2l_fallback");
            downward();
            // the full-order result at the same sites; where it meets the tolerance, later calls get a
            // per-pair tolerance tighter by about the factor the learned orders missed (where it does
            // not, the tolerance is out of reach at this expansion_order and tightening cannot help)
            const double err_full = spot_compare(c, tr, spot, n, nf, near, k, before);
            c->sel_full_err = err_full;
            if (err_full <= c->order_tol) {
                const double f = err > 0.0 && std::isfinite(err) ? 0.8 * c->order_tol / err : 0.5;
                c->sel_adapt = std::max(c->sel_adapt * std::min(0.5, std::max(0.125, f)), 1.0 / 256.0);
            }
        }
    }
    if (capture) {
        c->graph_exec_replace(cap.end());
        c->graph_key = gkey;
        c->graph_pending.clear();
        FMMH_HIP(hipGraphLaunch(c->graph_exec, s));
    }
    }   // device body (not replayed)
    bad_excl_readback(c);
    if (select) {
        unsigned long long* hh = reinterpret_cast<unsigned long long*>(c->h_pinned + 850);
        FMMH_HIP(hipMemcpyAsync(hh, c->buf<unsigned long long>("sel_hist", PMAX + 1),
                                (PMAX + 1) * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        double* ht = reinterpret_cast<double*>(c->h_pinned + 880);
        FMMH_HIP(hipMemcpyAsync(ht, c->buf<double>("sel_acc", 4), 4 * sizeof(double), hipMemcpyDeviceToHost, s));
        c->sync();
        double fl = 0.0, cnt = 0.0, qs = 0.0;
        for (int q = 0; q <= PMAX; ++q) {
            c->sel_hist[q] = hh[q];
            fl += (double)hh[q] * 2.0 * rdim(q) * rdim(q) * S;
            cnt += (double)hh[q];
            qs += (double)hh[q] * q;
        }
        c->sel_stats[2] = ht[2];
        c->sel_stats[3] = c->sel_stats[1] > 0.0 ? (double)L.n_m2l * 2.0 * rdim(p) * rdim(p) * S : fl;
        c->sel_stats[4] = (double)L.n_m2l * 2.0 * rdim(p) * rdim(p) * S;
        c->sel_stats[5] = cnt > 0 ? qs / cnt : 0.0;
        c->sel_stats[6] = 1.0;
        c->sel_stats[7] = std::sqrt(ht[0] / n);
    } else {
        for (int q = 0; q <= PMAX; ++q) c->sel_hist[q] = 0;
        c->sel_stats[2] = 0.0;
        c->sel_stats[3] = c->sel_stats[4] = (double)L.n_m2l * 2.0 * rdim(p) * rdim(p) * S;
        c->sel_stats[5] = p;
        c->sel_stats[6] = 0.0;
        c->sel_stats[7] = 0.0;
    }
    unsigned long long pairs = 0;
    if (reuse) {
        pairs = c->cache_pairs;
        c->sync();
    } else {
        unsigned long long* pc = c->buf<unsigned long long>("e_paircount", 1);
        FMMH_HIP(hipMemsetAsync(pc, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_pair_count, dim3(nblk(nl, 256)), dim3(256), 0, s,
                           c->buf<long long>("s_leaf_tot", nl), tr.t.leaf_node, tr.t.count, nl, pc);
        FMMH_HIP(hipMemcpyAsync(c->h_pinned + 200, pc, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        if (c->reuse_tree) {
            // keep a copy of the coordinates the cached tree was built from
            double* cx = c->buf<double>("c_X", (size_t)n * 3);
            FMMH_HIP(hipMemcpy2DAsync(cx, n * sizeof(double), k.X, k.ldx * sizeof(double), n * sizeof(double), 3,
                                      hipMemcpyDeviceToDevice, s));
        }
        c->sync();   // the call's completion; the pair count and the coordinate copy ride on it
        std::memcpy(&pairs, c->h_pinned + 200, sizeof(pairs));
    }
    if (c->reuse_tree) {
        if (!reuse) {
            c->cache_tr = tr;
            c->cache_L = L;
            c->cache_n = n; c->cache_ncrit = ncrit; c->cache_theta = theta;
            c->cache_periodic = P.on;
            std::memcpy(c->cache_box, P.box, sizeof(P.box));
            c->cache_pairs = pairs;
            c->cache_retries = c->list_retries;
            c->cache_valid = true;
        }
        c->cache_class_p = p;
        c->cache_classes = c->stats_classes;
    }
    bad_excl_check(c);   // read back before the synchronisations above
    c->stats[0] = nn; c->stats[1] = nl; c->stats[2] = tr.max_depth;
    c->stats[3] = L.n_m2l; c->stats[4] = L.n_p2p; c->stats[5] = (int64_t)pairs; c->stats[6] = c->stats_classes;
    c->stats[7] = c->list_retries;
    c->stats[8] = replay ? 2 : (capture ? 1 : 0);
    c->stats[9] = c->graph_replays;
    c->stats[10] = c->near_nv;
    c->stats[11] = c->host_waits;   // up to here; the field readback of a host-array call adds one
}

// ---------------------------------------------------------------- host <-> device staging
struct Staged {
    Call k;
    int64_t nf;
};

Staged stage_inputs(fmmh_context* c, const Call& h, int64_t n_damping) {
    hipStream_t s = c->stream;
    const int n = h.n;
    Staged st;
    st.k = h;
    st.nf = ncomp(h.F);
    double* X = c->buf<double>("h_X", (size_t)n * 3);
    double* M = c->buf<double>("h_M", (size_t)n * h.dim * h.nsets);
    int* E = c->buf<int>("h_E", (size_t)n * std::max(h.nexcl, 1));
    double* F = c->buf<double>("h_F", (size_t)n * st.nf * h.nsets);
    double* df = c->buf<double>("h_df", n);
    FMMH_HIP(hipMemcpy2DAsync(X, n * sizeof(double), h.X, h.ldx * sizeof(double), n * sizeof(double), 3,
                              hipMemcpyHostToDevice, s));
    FMMH_HIP(hipMemcpy2DAsync(M, n * sizeof(double), h.M, h.ldm * sizeof(double), n * sizeof(double),
                              (size_t)h.dim * h.nsets, hipMemcpyHostToDevice, s));
    if (h.nexcl > 0)
        FMMH_HIP(hipMemcpy2DAsync(E, n * sizeof(int), h.E, h.lde * sizeof(int), n * sizeof(int), h.nexcl,
                                  hipMemcpyHostToDevice, s));
    FMMH_HIP(hipMemcpy2DAsync(F, n * sizeof(double), h.field, h.ldf * sizeof(double), n * sizeof(double),
                              (size_t)st.nf * h.nsets, hipMemcpyHostToDevice, s));
    if (h.damp != FMMH_DAMP_NONE)
        FMMH_HIP(hipMemcpyAsync(df, h.df, n * sizeof(double), hipMemcpyHostToDevice, s));
    (void)n_damping;
    st.k.X = X; st.k.ldx = n;
    st.k.M = M; st.k.ldm = n;
    st.k.E = E; st.k.lde = n;
    st.k.field = F; st.k.ldf = n;
    st.k.df = h.damp != FMMH_DAMP_NONE ? df : nullptr;
    return st;
}

void unstage_field(fmmh_context* c, const Staged& st, double* field, int64_t ldf) {
    FMMH_HIP(hipMemcpy2DAsync(field, ldf * sizeof(double), st.k.field, st.k.n * sizeof(double),
                              st.k.n * sizeof(double), (size_t)st.nf * st.k.nsets, hipMemcpyDeviceToHost,
                              c->stream));
    c->sync();
}

Call make_call(const double* X, int64_t ldx, const double* M, int64_t ldm, int n, int dim, const int* E, int64_t lde,
               int nexcl, int F, double* field, int64_t ldf, int damp, const double* df, int nsets = 1) {
    return Call{X, ldx, M, ldm, n, dim, E, lde, nexcl, F, field, ldf, damp, df, nsets};
}

void check_fmm_args(double theta, int ncrit, int p, int is_periodic, const double* box_matrix) {
    if (is_periodic && box_matrix == nullptr) throw Error(FMMH_ERR_ARG, "periodic FMM needs box_matrix");
    if (is_periodic && p < 2)
        throw Error(FMMH_ERR_UNSUPPORTED, "periodic FMM needs expansion_order >= 2 (the dipole correction reads "
                                          "the quadrupole trace of the root expansion)");
    if (!(theta > 0.0)) throw Error(FMMH_ERR_ARG, "theta must be > 0");
    if (ncrit < 2 || ncrit > 4096) throw Error(FMMH_ERR_ARG, "ncrit must be in 2..4096");
    if (p < 0 || p > PMAX) throw Error(FMMH_ERR_UNSUPPORTED, "expansion_order must be in 0.." + std::to_string(PMAX));
}

}  // namespace

// ---------------------------------------------------------------- QM/MM embedding
// Octree of the sites, classification against the target points, and the
// compaction of near-field sites / far boxes (src/fmm.F90:1171-1291).
void qm_classify(fmmh_context* c, const double* X, int n, const double* T, int nt, double theta, int ncrit,
                 TreeOut& tr) {
    hipStream_t s = c->stream;
    c->cache_valid = false;
    build_tree(c, X, n, n, ncrit, tr);
    const int nn = tr.num_nodes;
    int* close = c->buf<int>("q_close", nn);
    int* type = c->buf<int>("q_type", nn);
    launch_classify(T, nt, nt, theta, tr.t, nn, close, type, s);
    int* near_cnt = c->buf<int>("q_near_cnt", nn + 1);
    int* far_cnt = c->buf<int>("q_far_cnt", nn + 1);
    int* near_off = c->buf<int>("q_near_off", nn + 1);
    int* far_off = c->buf<int>("q_far_off", nn + 1);
    launch_qmmm_counts(type, tr.t.count, nn, near_cnt, far_cnt, s);
    size_t tb = exscan_i32_temp(nn + 1);
    void* temp = c->buf<char>("q_temp", tb);
    exscan_i32(temp, tb, near_cnt, near_off, nn + 1, s);
    exscan_i32(temp, tb, far_cnt, far_off, nn + 1, s);
    FMMH_HIP(hipMemcpyAsync(c->h_pinned + 440, near_off + nn, sizeof(int), hipMemcpyDeviceToHost, s));
    FMMH_HIP(hipMemcpyAsync(c->h_pinned + 441, far_off + nn, sizeof(int), hipMemcpyDeviceToHost, s));
    c->sync();
    c->qm_num_near = c->h_pinned[440];
    c->qm_num_far = c->h_pinned[441];
    c->qm_num_nodes = nn;
    int* near_sites = c->buf<int>("q_near", std::max(c->qm_num_near, 1));
    int* far_nodes = c->buf<int>("q_far", std::max(c->qm_num_far, 1));
    launch_qmmm_compact(type, tr.t, nn, near_off, far_off, near_sites, far_nodes, s);
    c->sync();
}

const double* stage_host(fmmh_context* c, const char* name, const double* h, int64_t ld, int n, int cols) {
    double* d = c->buf<double>(name, (size_t)n * std::max(cols, 1));
    if (cols > 0)
        FMMH_HIP(hipMemcpy2DAsync(d, n * sizeof(double), h, ld * sizeof(double), n * sizeof(double), cols,
                                  hipMemcpyHostToDevice, c->stream));
    return d;
}

void check_qm(int n, int64_t ldx, int nt, int64_t ldt, double theta, int ncrit, int p) {
    if (n < 1 || ldx < n) throw Error(FMMH_ERR_ARG, "bad n / leading dimension of coordinates");
    if (nt < 0 || (nt > 0 && ldt < nt)) throw Error(FMMH_ERR_ARG, "bad target count / leading dimension");
    check_fmm_args(theta, ncrit, p, 0, nullptr);
}

// ================================================================ C ABI
extern "C" {

const char* fmmh_last_error(void) { return g_last_error.c_str(); }

int fmmh_context_create(fmmh_context** ctx, int32_t device) {
    return guarded([&] { *ctx = make_context(device); });
}

int fmmh_context_destroy(fmmh_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        FMMH_HIP(hipSetDevice(ctx->device));
        FMMH_HIP(hipStreamSynchronize(ctx->stream));
        for (auto& kv : ctx->bufs)
            if (kv.second.p) FMMH_HIP(hipFree(kv.second.p));
        for (auto& kv : ctx->tables) {
            FMMH_HIP(hipFree(kv.second.ibuf));
            FMMH_HIP(hipFree(kv.second.dbuf));
            if (kv.second.fbuf) FMMH_HIP(hipFree(kv.second.fbuf));
        }
        for (hipEvent_t e : ctx->event_pool) FMMH_HIP(hipEventDestroy(e));
        if (ctx->rb_event) FMMH_HIP(hipEventDestroy(ctx->rb_event));
        if (ctx->graph_exec) FMMH_HIP(hipGraphExecDestroy(ctx->graph_exec));
        FMMH_HIP(hipHostFree(ctx->h_pinned));
        FMMH_HIP(hipStreamDestroy(ctx->stream));
        delete ctx;
    });
}

int fmmh_parse_damp_type(const char* s, int64_t len, int32_t* code) {
    return guarded([&] {
        int64_t l = len;
        while (l > 0 && (s[l - 1] == ' ' || s[l - 1] == '\0')) --l;
        std::string t(s, (size_t)l);
        if (t.empty()) *code = FMMH_DAMP_NONE;
        else if (t == "THOLE") *code = FMMH_DAMP_THOLE;
        else if (t == "AMOEBA") *code = FMMH_DAMP_AMOEBA;
        else if (t == "ERF") *code = FMMH_DAMP_ERF;
        else throw Error(FMMH_ERR_DAMP, "Unknown damp type: " + std::string(s, (size_t)len));
    });
}

static int field_direct_impl(fmmh_context* ctx, bool host, const double* coordinates, int64_t ld_coordinates,
                             const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                             const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, int32_t field_order,
                             double* field, int64_t ld_field, int32_t damp_type, const double* damping_factors,
                             int64_t n_damping) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        Call h = make_call(coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, exclusions, ld_exclusions,
                           nexcl, field_order, field, ld_field, damp_type, damping_factors, nsets);
        check_common(h, n_damping);
        if (host) check_exclusions_host(h);
        c->begin_timing();
        if (host) {
            Staged st = stage_inputs(c, h, n_damping);
            c->mark("h2d");
            run_direct(c, st.k);
            unstage_field(c, st, field, ld_field);
            c->mark("d2h");
        } else {
            run_direct(c, h);
        }
        c->end_timing();
    });
}

static int field_fmm_impl(fmmh_context* ctx, bool host, const double* coordinates, int64_t ld_coordinates,
                          const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                          const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, double theta,
                          int32_t ncrit, int32_t expansion_order, int32_t field_order, double* field,
                          int64_t ld_field, int32_t damp_type, const double* damping_factors, int64_t n_damping,
                          int32_t is_periodic, const double* box_matrix) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        Call h = make_call(coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, exclusions, ld_exclusions,
                           nexcl, field_order, field, ld_field, damp_type, damping_factors, nsets);
        // src/fmm.F90:1372-1382: damp_type first, then the periodic checks
        if (damp_type < 0 || damp_type > 3) throw Error(FMMH_ERR_DAMP, "Unknown damp type code");
        if (is_periodic && n <= ncrit)
            throw Error(FMMH_ERR_UNSUPPORTED, "Not enough coordinates for periodic FMM, please decrease ncrit");
        check_fmm_args(theta, ncrit, expansion_order, is_periodic, box_matrix);
        check_common(h, n_damping);
        if (host) check_exclusions_host(h);
        c->begin_timing();
        Call k = h;
        Staged st;
        if (host) {
            st = stage_inputs(c, h, n_damping);
            k = st.k;
            c->mark("h2d");
        }
        if (n <= ncrit) run_direct(c, k);
        else run_fmm(c, k, theta, ncrit, expansion_order, make_periodic(is_periodic != 0, box_matrix));
        if (host) {
            unstage_field(c, st, field, ld_field);
            c->mark("d2h");
        }
        c->end_timing();
    });
}

int fmmh_field_direct(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                      const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                      const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, int32_t field_order,
                      double* field, int64_t ld_field, int32_t damp_type, const double* damping_factors,
                      int64_t n_damping) {
    (void)comm;
    return field_direct_impl(ctx, true, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, 1, exclusions,
                             ld_exclusions, nexcl, field_order, field, ld_field, damp_type, damping_factors,
                             n_damping);
}

int fmmh_field_direct_dev(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                          const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                          const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, int32_t field_order,
                          double* field, int64_t ld_field, int32_t damp_type, const double* damping_factors,
                          int64_t n_damping) {
    (void)comm;
    return field_direct_impl(ctx, false, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, 1,
                             exclusions, ld_exclusions, nexcl, field_order, field, ld_field, damp_type,
                             damping_factors, n_damping);
}

int fmmh_field_direct_dev_noalloc(fmmh_context* ctx, int32_t comm, const double* coordinates,
                                  int64_t ld_coordinates, const double* multipoles, int64_t ld_multipoles, int32_t n,
                                  int32_t dim, const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                                  int32_t field_order, double* field, int64_t ld_field, int32_t damp_type,
                                  const double* damping_factors, int64_t n_damping) {
    (void)comm;
    return guarded([&] {
        if (!ctx) throw Error(FMMH_ERR_ARG, "field_direct without allocation needs a context (fmmh_context_create)");
        FMMH_HIP(hipSetDevice(ctx->device));
        Call h = make_call(coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, exclusions, ld_exclusions,
                           nexcl, field_order, field, ld_field, damp_type, damping_factors, 1);
        check_common(h, n_damping);
        run_direct_noalloc(ctx, h, scratch_ptr(ctx->device));
    });
}

int fmmh_field_fmm(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                   const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, const int32_t* exclusions,
                   int64_t ld_exclusions, int32_t nexcl, double theta, int32_t ncrit, int32_t expansion_order,
                   int32_t field_order, double* field, int64_t ld_field, int32_t damp_type,
                   const double* damping_factors, int64_t n_damping, int32_t is_periodic, const double* box_matrix) {
    (void)comm;
    return field_fmm_impl(ctx, true, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, 1, exclusions,
                          ld_exclusions, nexcl, theta, ncrit, expansion_order, field_order, field, ld_field,
                          damp_type, damping_factors, n_damping, is_periodic, box_matrix);
}

int fmmh_field_fmm_dev(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                       const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                       const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, double theta, int32_t ncrit,
                       int32_t expansion_order, int32_t field_order, double* field, int64_t ld_field,
                       int32_t damp_type, const double* damping_factors, int64_t n_damping, int32_t is_periodic,
                       const double* box_matrix) {
    (void)comm;
    return field_fmm_impl(ctx, false, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, 1, exclusions,
                          ld_exclusions, nexcl, theta, ncrit, expansion_order, field_order, field, ld_field,
                          damp_type, damping_factors, n_damping, is_periodic, box_matrix);
}

int fmmh_field_direct_sets(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                           const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                           const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, int32_t field_order,
                           double* field, int64_t ld_field, int32_t damp_type, const double* damping_factors,
                           int64_t n_damping) {
    (void)comm;
    return field_direct_impl(ctx, true, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, nsets,
                             exclusions, ld_exclusions, nexcl, field_order, field, ld_field, damp_type,
                             damping_factors, n_damping);
}

int fmmh_field_direct_sets_dev(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                               const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                               int32_t nsets, const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                               int32_t field_order, double* field, int64_t ld_field, int32_t damp_type,
                               const double* damping_factors, int64_t n_damping) {
    (void)comm;
    return field_direct_impl(ctx, false, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, nsets,
                             exclusions, ld_exclusions, nexcl, field_order, field, ld_field, damp_type,
                             damping_factors, n_damping);
}

int fmmh_field_fmm_sets(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                        const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                        const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, double theta, int32_t ncrit,
                        int32_t expansion_order, int32_t field_order, double* field, int64_t ld_field,
                        int32_t damp_type, const double* damping_factors, int64_t n_damping, int32_t is_periodic,
                        const double* box_matrix) {
    (void)comm;
    return field_fmm_impl(ctx, true, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, nsets,
                          exclusions, ld_exclusions, nexcl, theta, ncrit, expansion_order, field_order, field,
                          ld_field, damp_type, damping_factors, n_damping, is_periodic, box_matrix);
}

int fmmh_field_fmm_sets_dev(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates,
                            const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, int32_t nsets,
                            const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl, double theta,
                            int32_t ncrit, int32_t expansion_order, int32_t field_order, double* field,
                            int64_t ld_field, int32_t damp_type, const double* damping_factors, int64_t n_damping,
                            int32_t is_periodic, const double* box_matrix) {
    (void)comm;
    return field_fmm_impl(ctx, false, coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, nsets,
                          exclusions, ld_exclusions, nexcl, theta, ncrit, expansion_order, field_order, field,
                          ld_field, damp_type, damping_factors, n_damping, is_periodic, box_matrix);
}

int fmmh_field_direct_targets_dev(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates,
                                  const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim,
                                  const int32_t* exclusions, int64_t ld_exclusions, int32_t nexcl,
                                  const int32_t* targets, int32_t ntargets, int32_t field_order, double* field,
                                  int64_t ld_field, int32_t damp_type, const double* damping_factors,
                                  int64_t n_damping) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (ntargets < 1) throw Error(FMMH_ERR_ARG, "ntargets must be >= 1");
        if (ld_field < ntargets) throw Error(FMMH_ERR_ARG, "ld_field smaller than ntargets");
        Call k = make_call(coordinates, ld_coordinates, multipoles, ld_multipoles, n, dim, exclusions, ld_exclusions,
                           nexcl, field_order, field, n, damp_type, damping_factors);
        check_common(k, n_damping);
        k.ldf = ld_field;
        c->begin_timing();
        run_direct_targets(c, k, targets, ntargets);
        c->end_timing();
    });
}

int fmmh_get_boxes(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates, int32_t n,
                   const double* target_coordinates, int64_t ld_targets, int32_t ntargets, double theta,
                   int32_t ncrit, int32_t expansion_order, int64_t* sizes) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        check_qm(n, ld_coordinates, ntargets, ld_targets, theta, ncrit, expansion_order);
        const double* X = stage_host(c, "h_X", coordinates, ld_coordinates, n, 3);
        const double* T = stage_host(c, "q_T", target_coordinates, ld_targets, ntargets, 3);
        TreeOut tr;
        qm_classify(c, X, n, T, ntargets, theta, ncrit, tr);
        c->qm_multipoles = false;
        sizes[0] = c->qm_num_near;
        sizes[1] = c->qm_num_far;
    });
}

// Shared export of the near-field sites and far boxes of the last classification.
static void qm_export_common(fmmh_context* c, int32_t* idx_near_field, double* box_coordinates, int32_t* far_ids) {
    const int nn = c->qm_num_nodes, nfar = c->qm_num_far, nnear = c->qm_num_near;
    auto get = [&](void* dst, const char* name, size_t bytes) {
        auto it = c->bufs.find(name);
        if (it == c->bufs.end()) throw Error(FMMH_ERR_INTERNAL, std::string("missing buffer ") + name);
        FMMH_HIP(hipMemcpy(dst, it->second.p, bytes, hipMemcpyDeviceToHost));
    };
    if (nnear > 0 && idx_near_field) get(idx_near_field, "q_near", nnear * sizeof(int));
    std::vector<int> ids(std::max(nfar, 1));
    if (nfar > 0) get(ids.data(), "q_far", nfar * sizeof(int));
    if (box_coordinates && nfar > 0) {
        std::vector<double> cx(nn), cy(nn), cz(nn);
        get(cx.data(), "n_cx", nn * sizeof(double));
        get(cy.data(), "n_cy", nn * sizeof(double));
        get(cz.data(), "n_cz", nn * sizeof(double));
        for (int k = 0; k < nfar; ++k) {
            box_coordinates[k] = cx[ids[k]];
            box_coordinates[nfar + k] = cy[ids[k]];
            box_coordinates[2 * (size_t)nfar + k] = cz[ids[k]];
        }
    }
    if (far_ids)
        for (int k = 0; k < nfar; ++k) far_ids[k] = ids[k];
}

int fmmh_get_boxes_export(fmmh_context* ctx, int32_t* idx_near_field, double* box_coordinates, int32_t* box_indices) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (c->qm_num_near < 0 || c->qm_multipoles) throw Error(FMMH_ERR_ARG, "no get_boxes result on this context");
        qm_export_common(c, idx_near_field, box_coordinates, box_indices);
        if (box_indices)
            for (int k = 0; k < c->qm_num_far; ++k) box_indices[k] += 1;   // 1-based node index
    });
}

int fmmh_get_multipoles(fmmh_context* ctx, int32_t comm, const double* coordinates, int64_t ld_coordinates, int32_t n,
                        const double* multipoles, int64_t ld_multipoles, int32_t dim, int32_t ndens,
                        const double* target_coordinates, int64_t ld_targets, int32_t ntargets, double theta,
                        int32_t ncrit, int32_t expansion_order, int64_t* sizes) {
    (void)comm;
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        check_qm(n, ld_coordinates, ntargets, ld_targets, theta, ncrit, expansion_order);
        if (ndens < 1) throw Error(FMMH_ERR_ARG, "ndens must be >= 1");
        if (ld_multipoles < n) throw Error(FMMH_ERR_ARG, "leading dimension of multipoles smaller than n");
        const int m = order_from_dim(dim);
        if (m < 0 || ncomp(m) != dim || m > NEAR_MMAX)
            throw Error(FMMH_ERR_UNSUPPORTED, "multipole dim must be (m+1)(m+2)(m+3)/6 with m <= " +
                                                  std::to_string(NEAR_MMAX));
        hipStream_t s = c->stream;
        const double* X = stage_host(c, "h_X", coordinates, ld_coordinates, n, 3);
        const double* T = stage_host(c, "q_T", target_coordinates, ld_targets, ntargets, 3);
        const double* M = stage_host(c, "h_M", multipoles, ld_multipoles, n, dim * ndens);
        TreeOut tr;
        qm_classify(c, X, n, T, ntargets, theta, ncrit, tr);
        // cell multipoles of every node for every density set (P2M + M2M)
        const int p = expansion_order, nm = ncomp(p), nn = tr.num_nodes;
        Tables& tb = get_tables(c, p, m);
        double* xs = c->buf<double>("s_xs", n);
        double* ys = c->buf<double>("s_ys", n);
        double* zs = c->buf<double>("s_zs", n);
        double* ms_raw = c->buf<double>("s_ms_raw", (size_t)n * dim * ndens);
        double* msc = c->buf<double>("s_ms", (size_t)n * dim * ndens);
        double* dfs = c->buf<double>("s_dfs", n);
        int* excl = c->buf<int>("s_excl", n);
        SiteArgs g{X, n, M, n, dim, nullptr, n, 0, nullptr, tr.t.perm, n, xs, ys, zs, ms_raw, msc, dfs, excl,
                   tb.scale, ndens};
        launch_gather(g, s);
        double* CM = c->buf<double>("e_cm", (size_t)nn * nm * ndens);
        UpArgs up{xs, ys, zs, ms_raw, n, tr.t.leaf_node, tr.t.start, tr.t.count, tr.t.cx, tr.t.cy, tr.t.cz,
                  tr.t.children, CM, nm, p, tb.nsrc, tb.p2m, tb.m2m, tb.mono_parent, tb.mono_axis, nn, ndens, dim, tb.xyz};
        launch_p2m(up, tr.num_leaves, s);
        for (int lev = tr.max_depth - 1; lev >= 1; --lev) {
            int o0 = tr.level_off[lev], o1 = tr.level_off[lev + 1];
            launch_m2m(up, tr.t.level_nodes + o0, o1 - o0, s);
        }
        c->sync();
        c->qm_multipoles = true;
        c->qm_nm = nm;
        c->qm_ndens = ndens;
        sizes[0] = c->qm_num_near;
        sizes[1] = c->qm_num_far;
    });
}

int fmmh_get_multipoles_export(fmmh_context* ctx, int32_t* idx_near_field, double* box_coordinates,
                               double* box_multipoles) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (c->qm_num_near < 0 || !c->qm_multipoles)
            throw Error(FMMH_ERR_ARG, "no get_multipoles result on this context");
        const int nfar = c->qm_num_far, nm = c->qm_nm, nd = c->qm_ndens, nn = c->qm_num_nodes;
        std::vector<int> ids(std::max(nfar, 1));
        qm_export_common(c, idx_near_field, box_coordinates, ids.data());
        if (box_multipoles && nfar > 0) {
            std::vector<double> cm((size_t)nn * nm * nd);
            auto it = c->bufs.find("e_cm");
            FMMH_HIP(hipMemcpy(cm.data(), it->second.p, cm.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (int d = 0; d < nd; ++d)
                for (int t = 0; t < nm; ++t)
                    for (int k = 0; k < nfar; ++k)
                        box_multipoles[((size_t)d * nm + t) * nfar + k] = cm[((size_t)d * nn + ids[k]) * nm + t];
        }
    });
}

int fmmh_interpolate_electron_fields(fmmh_context* ctx, int32_t comm, const double* coordinates,
                                     int64_t ld_coordinates, int32_t n, const double* local_expansion,
                                     int64_t ld_local_expansion, int32_t nbox, int32_t ndens,
                                     const int32_t* box_indices, double theta, int32_t ncrit,
                                     int32_t expansion_order, int32_t field_order, double* field,
                                     int64_t ld_field) {
    (void)comm;
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        check_qm(n, ld_coordinates, 0, 0, theta, ncrit, expansion_order);
        if (ndens < 1 || nbox < 0) throw Error(FMMH_ERR_ARG, "ndens must be >= 1 and nbox >= 0");
        if (nbox > 0 && ld_local_expansion < nbox) throw Error(FMMH_ERR_ARG, "leading dimension of local_expansion");
        if (ld_field < n) throw Error(FMMH_ERR_ARG, "leading dimension of field smaller than n");
        if (field_order < 0 || field_order > NEAR_FMAX)
            throw Error(FMMH_ERR_UNSUPPORTED, "field_order must be 0.." + std::to_string(NEAR_FMAX));
        hipStream_t s = c->stream;
        const double* X = stage_host(c, "h_X", coordinates, ld_coordinates, n, 3);
        c->cache_valid = false;
        TreeOut tr;
        build_tree(c, X, n, n, ncrit, tr);
        const int p = expansion_order, nm = ncomp(p), nn = tr.num_nodes, nf = ncomp(field_order);
        // local expansions: zero except the given boxes (src/fmm.F90:1322-1328), later rows win
        std::vector<double> le((size_t)nn * nm * ndens, 0.0);
        for (int i = 0; i < nbox; ++i) {
            const int b = box_indices[i];
            if (b < 1 || b > nn)
                throw Error(FMMH_ERR_ARG, "box_indices(" + std::to_string(i + 1) + ") = " + std::to_string(b) +
                                              " is outside 1..num_nodes (" + std::to_string(nn) + ")");
            for (int d = 0; d < ndens; ++d)
                for (int t = 0; t < nm; ++t)
                    le[((size_t)d * nn + (b - 1)) * nm + t] =
                        local_expansion[((size_t)d * nm + t) * ld_local_expansion + i];
        }
        double* LE = c->buf<double>("e_le", le.size());
        FMMH_HIP(hipMemcpyAsync(LE, le.data(), le.size() * sizeof(double), hipMemcpyHostToDevice, s));
        Tables& tb = get_tables(c, p, 0);
        double* xs = c->buf<double>("s_xs", n);
        double* ys = c->buf<double>("s_ys", n);
        double* zs = c->buf<double>("s_zs", n);
        double* dfs = c->buf<double>("s_dfs", n);
        int* excl = c->buf<int>("s_excl", n);
        SiteArgs g{X, n, nullptr, n, 0, nullptr, n, 0, nullptr, tr.t.perm, n, xs, ys, zs, nullptr, nullptr, dfs,
                   excl, tb.scale, 1};
        launch_gather(g, s);
        double* near = c->buf<double>("e_near", (size_t)n * nf * ndens);
        FMMH_HIP(hipMemsetAsync(near, 0, (size_t)n * nf * ndens * sizeof(double), s));
        double* F = c->buf<double>("h_F", (size_t)n * nf * ndens);
        FMMH_HIP(hipMemcpy2DAsync(F, n * sizeof(double), field, ld_field * sizeof(double), n * sizeof(double),
                                  (size_t)nf * ndens, hipMemcpyHostToDevice, s));
        DownArgs da{tr.t.cx, tr.t.cy, tr.t.cz, tr.t.parent, LE, nm, p, tb.l2l, tb.l2l, tb.mono_parent, tb.mono_axis,
                    tr.t.leaf_node, tr.t.start, tr.t.count, xs, ys, zs, tr.t.perm, near, n, F, n, nf, nn, ndens, tb.xyz};
        for (int lev = 2; lev <= tr.max_depth; ++lev) {
            int o0 = tr.level_off[lev], o1 = tr.level_off[lev + 1];
            launch_l2l(da, tr.t.level_nodes + o0, o1 - o0, s);
        }
        launch_l2p(da, tr.num_leaves, s);
        FMMH_HIP(hipMemcpy2DAsync(field, ld_field * sizeof(double), F, n * sizeof(double), n * sizeof(double),
                                  (size_t)nf * ndens, hipMemcpyDeviceToHost, s));
        c->sync();
    });
}

int fmmh_tree_build(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates, int32_t n, int32_t ncrit,
                    double theta, int64_t* info) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (n < 1 || ld_coordinates < n) throw Error(FMMH_ERR_ARG, "bad n / leading dimension");
        if (n < ncrit) throw Error(FMMH_ERR_ARG, "the tree needs n >= ncrit");
        check_fmm_args(theta, ncrit, 0, 0, nullptr);
        double* X = c->buf<double>("h_X", (size_t)n * 3);
        FMMH_HIP(hipMemcpy2DAsync(X, n * sizeof(double), coordinates, ld_coordinates * sizeof(double),
                                  n * sizeof(double), 3, hipMemcpyHostToDevice, c->stream));
        c->cache_valid = false;
        TreeOut tr;
        build_tree(c, X, n, n, ncrit, tr);
        ListsOut L;
        build_lists(c, tr, theta, L, 1, upload_disp(c, Periodic{}), [] {});
        c->sync();
        c->num_nodes = tr.num_nodes; c->num_leaves = tr.num_leaves; c->max_depth = tr.max_depth;
        c->n_m2l = L.n_m2l; c->n_p2p = L.n_p2p; c->n_sites = n;
        info[0] = tr.num_nodes; info[1] = tr.num_leaves; info[2] = tr.max_depth; info[3] = L.n_m2l;
        info[4] = L.n_p2p;
    });
}

int fmmh_tree_export(fmmh_context* ctx, double* centers, double* r, double* rmax, int32_t* depth, int32_t* parent,
                     int32_t* is_leaf, int32_t* nleaf, int32_t* mem_off, int32_t* members, int32_t* m2l_off,
                     int32_t* m2l_src, int32_t* p2p_off, int32_t* p2p_src) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        const int nn = c->num_nodes, n = c->n_sites;
        if (nn <= 0) throw Error(FMMH_ERR_ARG, "no tree has been built on this context");
        auto get = [&](void* dst, const char* name, size_t bytes) {
            auto it = c->bufs.find(name);
            if (it == c->bufs.end()) throw Error(FMMH_ERR_INTERNAL, std::string("missing buffer ") + name);
            FMMH_HIP(hipMemcpy(dst, it->second.p, bytes, hipMemcpyDeviceToHost));
        };
        get(centers, "n_cx", nn * sizeof(double));
        get(centers + nn, "n_cy", nn * sizeof(double));
        get(centers + 2 * (size_t)nn, "n_cz", nn * sizeof(double));
        get(r, "n_r", nn * sizeof(double));
        get(rmax, "n_rmax", nn * sizeof(double));
        get(depth, "n_level", nn * sizeof(int));
        get(parent, "n_parent", nn * sizeof(int));
        get(is_leaf, "n_is_leaf", nn * sizeof(int));
        get(nleaf, "n_count", nn * sizeof(int));
        std::vector<int> start(nn), perm(n);
        get(start.data(), "n_start", nn * sizeof(int));
        get(perm.data(), "t_perm", n * sizeof(int));
        int mo = 0;
        for (int id = 0; id < nn; ++id) {
            mem_off[id] = mo;
            if (is_leaf[id])
                for (int q = 0; q < nleaf[id]; ++q) members[mo++] = perm[start[id] + q];
        }
        mem_off[nn] = mo;
        get(m2l_off, "l_m2l_off", (nn + 1) * sizeof(int));
        get(p2p_off, "l_p2p_off", (nn + 1) * sizeof(int));
        if (c->n_m2l) get(m2l_src, "l_m2l_src", c->n_m2l * sizeof(int));
        if (c->n_p2p) get(p2p_src, "l_p2p_src", c->n_p2p * sizeof(int));
    });
}

int fmmh_set_option(fmmh_context* ctx, int32_t key, double value) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        ++c->cache_gen;   // any option change: captured graphs are not replayed again
        if (key == FMMH_OPT_TIMING) c->timing = value != 0.0;
        else if (key == FMMH_OPT_LIST_CAPACITY) c->list_cap_scale = value > 0.0 ? value : 1.0;
        else if (key == FMMH_OPT_REUSE_TREE) {
            c->reuse_tree = value != 0.0;
            if (!c->reuse_tree) c->cache_valid = false;
        } else if (key == FMMH_OPT_PRECISION) {
            if (value != 0.0 && value != 1.0) throw Error(FMMH_ERR_ARG, "precision must be 0 (FP64) or 1 (mixed)");
            c->precision = (int)value;
        } else if (key == FMMH_OPT_ORDER_TOL) {
            if (!(value >= 0.0)) throw Error(FMMH_ERR_ARG, "order tolerance must be >= 0");
            c->order_tol = value;
            c->sel_adapt = 1.0;
        } else if (key == FMMH_OPT_ORDER_MARGIN) {
            if (!(value >= -1.0 && value <= 11.0)) throw Error(FMMH_ERR_ARG, "order margin must be in -1..11");
            c->sel_margin = (float)value;
        } else if (key == FMMH_OPT_SPOT_CHECKS) {
            if (!(value >= 0.0 && value <= 65536.0)) throw Error(FMMH_ERR_ARG, "spot checks must be in 0..65536");
            c->spot_checks = (int)value;
        } else if (key == FMMH_OPT_ORDER_SAFETY) {
            if (!(value > 0.0)) throw Error(FMMH_ERR_ARG, "order safety factor must be > 0");
            c->sel_kappa = value;
            c->sel_adapt = 1.0;
        } else if (key == FMMH_OPT_NEAR_CHUNK) {
            if (!(value >= 0.0 && value < 1e15)) throw Error(FMMH_ERR_ARG, "near chunk cap must be >= 0");
            c->near_chunk_cap = (long long)value;
            c->cache_work_valid = false;   // the cached block order follows the cap
        } else if (key == FMMH_OPT_GRAPH) {
            c->graph_opt = value != 0.0 ? 1 : 0;
        } else if (key == FMMH_OPT_AUTOTUNE) {
            c->autotune = value != 0.0 ? 1 : 0;
            if (value < 0.0) c->tuned.clear();   // forget earlier choices
        } else if (key == FMMH_OPT_ORDER_FORCE_FULL) {
            c->sel_force_full = value != 0.0 ? 1 : 0;
        } else if (key == FMMH_OPT_M2L_MFMA) {
            if (value != 0.0 && value != 1.0 && value != 2.0 && value != 3.0 && value != 4.0)
                throw Error(FMMH_ERR_ARG, "M2L MFMA mode must be 0, 1, 2, 3 or 4");
            c->m2l_mfma = (int)value;
        }
        else throw Error(FMMH_ERR_ARG, "unknown option key " + std::to_string(key));
    });
}

int fmmh_phase_times(fmmh_context* ctx, double* ms, int32_t cap, char* names, int32_t name_cap) {
    int count = 0;
    int rc = guarded([&] {
        fmmh_context* c = resolve(ctx);
        std::string all;
        for (auto& kv : c->phase_ms) {
            if (count < cap) ms[count] = kv.second;
            ++count;
            if (!all.empty()) all += ";";
            all += kv.first;
        }
        if (names && name_cap > 0) {
            std::strncpy(names, all.c_str(), (size_t)name_cap - 1);
            names[name_cap - 1] = '\0';
        }
    });
    return rc == FMMH_OK ? std::min(count, (int)cap) : -rc;
}

int fmmh_debug_buffer(fmmh_context* ctx, const char* name, void* host, int64_t bytes) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        auto it = c->bufs.find(name);
        if (it == c->bufs.end()) throw Error(FMMH_ERR_ARG, std::string("unknown buffer ") + name);
        if ((size_t)bytes > it->second.cap) throw Error(FMMH_ERR_ARG, std::string("buffer too small: ") + name);
        FMMH_HIP(hipStreamSynchronize(c->stream));
        FMMH_HIP(hipMemcpy(host, it->second.p, (size_t)bytes, hipMemcpyDeviceToHost));
    });
}

int fmmh_stats(fmmh_context* ctx, int64_t* out, int32_t cap) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        for (int i = 0; i < std::min(cap, 12); ++i) out[i] = c->stats[i];
    });
}

}  // extern "C"

// ---------------------------------------------------------------- learned order selection
namespace {

// Training examples from the measured truncation errors: per sample, NTAU
// tolerances spread (stratified, jittered) over log10 of its error curve;
// the label is the fractional order where the suffix-maximum error curve
// E_q = max_{q' >= q} err_{q'} crosses tau (log-linear between integer
// orders), cap + 1 where even cap misses tau.
constexpr int SEL_NTAU = 16;

struct TrainSet {
    std::vector<float> X, Y;
    float rel_lo = 1e30f, rel_hi = -1e30f;
};

TrainSet build_examples(const std::vector<float>& feat, const std::vector<double>& err, int ns, int cap,
                        unsigned long long seed) {
    TrainSet t;
    std::mt19937_64 rng(seed ^ 0x5eedull);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> env(cap + 1);
    for (int s = 0; s < ns; ++s) {
        const double* e = &err[(size_t)s * (cap + 1)];
        bool ok = std::isfinite(feat[(size_t)s * SEL_IN + 2]);
        for (int q = 0; q <= cap; ++q) ok = ok && std::isfinite(e[q]) && e[q] > 0.0;
        if (!ok) continue;
        double mx = 0.0;
        for (int q = cap; q >= 0; --q) { mx = std::max(mx, e[q]); env[q] = mx; }
        const double lo = std::log10(env[cap]) - 1.0, hi = std::log10(env[0]) + 0.5;
        for (int k = 0; k < SEL_NTAU; ++k) {
            const double lt = lo + (hi - lo) * (k + U(rng)) / SEL_NTAU;
            double y = cap + 1.0;
            for (int q = 0; q <= cap; ++q) {
                if (std::log10(env[q]) <= lt) {
                    if (q == 0) y = 0.0;
                    else {
                        const double l1 = std::log10(env[q - 1]), l2 = std::log10(env[q]);
                        y = (q - 1) + (l1 > l2 ? (l1 - lt) / (l1 - l2) : 1.0);
                    }
                    break;
                }
            }
            const float* f = &feat[(size_t)s * SEL_IN];
            t.X.insert(t.X.end(), {f[0], f[1], f[2], (float)lt});
            t.Y.push_back((float)y);
            const float rel = f[2] - (float)lt;
            t.rel_lo = std::min(t.rel_lo, rel);
            t.rel_hi = std::max(t.rel_hi, rel);
        }
    }
    return t;
}

}  // namespace

extern "C" {

int fmmh_selector_train(fmmh_context* ctx, const double* coordinates, int64_t ld_coordinates,
                        const double* multipoles, int64_t ld_multipoles, int32_t n, int32_t dim, double theta,
                        int32_t ncrit, int32_t cap, int32_t nsamples, int32_t epochs, uint64_t seed, double* info) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        hipStream_t s = c->stream;
        const int m = order_from_dim(dim);
        if (m < 0 || ncomp(m) != dim || m > SPOT_MMAX) throw Error(FMMH_ERR_ARG, "dim must be 1, 4 or 10");
        check_fmm_args(theta, ncrit, cap, 0, nullptr);
        if (cap < 1 || cap > M2L_GEMM_PMAX) throw Error(FMMH_ERR_ARG, "cap must be in 1..10");
        if (n <= ncrit || ld_coordinates < n || ld_multipoles < n) throw Error(FMMH_ERR_ARG, "need n > ncrit");
        if (nsamples < 1 || nsamples > (1 << 16)) throw Error(FMMH_ERR_ARG, "nsamples must be in 1..65536");
        if (epochs < 1 || epochs > 100000) throw Error(FMMH_ERR_ARG, "epochs must be in 1..100000");
        auto t0 = std::chrono::steady_clock::now();
        double* X = c->buf<double>("h_X", (size_t)n * 3);
        double* M = c->buf<double>("h_M", (size_t)n * dim);
        FMMH_HIP(hipMemcpy2DAsync(X, n * sizeof(double), coordinates, ld_coordinates * sizeof(double),
                                  n * sizeof(double), 3, hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpy2DAsync(M, n * sizeof(double), multipoles, ld_multipoles * sizeof(double),
                                  n * sizeof(double), dim, hipMemcpyHostToDevice, s));
        c->cache_valid = false;
        TreeOut tr;
        build_tree(c, X, n, n, ncrit, tr);
        ListsOut L;
        build_lists(c, tr, theta, L, 1, upload_disp(c, Periodic{}), [] {});
        if (L.n_m2l == 0) throw Error(FMMH_ERR_ARG, "the tree has no M2L pairs to learn from");
        Tables& tb = get_tables(c, cap, m);
        const int nn = tr.num_nodes, nm = ncomp(cap);
        double* xs = c->buf<double>("s_xs", n);
        double* ys = c->buf<double>("s_ys", n);
        double* zs = c->buf<double>("s_zs", n);
        double* ms_raw = c->buf<double>("s_ms_raw", (size_t)n * dim);
        double* ms = c->buf<double>("s_ms", (size_t)n * dim);
        double* dfs = c->buf<double>("s_dfs", n);
        int* excl = c->buf<int>("s_excl", n);
        SiteArgs g{X, n, M, n, dim, nullptr, n, 0, nullptr, tr.t.perm, n, xs, ys, zs, ms_raw, ms, dfs, excl,
                   tb.scale, 1};
        launch_gather(g, s);
        double* CM = c->buf<double>("e_cm", (size_t)nn * nm);
        UpArgs up{xs, ys, zs, ms_raw, n, tr.t.leaf_node, tr.t.start, tr.t.count, tr.t.cx, tr.t.cy, tr.t.cz,
                  tr.t.children, CM, nm, cap, tb.nsrc, tb.p2m, tb.m2m, tb.mono_parent, tb.mono_axis, nn, 1, dim, tb.xyz};
        launch_p2m(up, tr.num_leaves, s);
        for (int lev = tr.max_depth - 1; lev >= 1; --lev)
            launch_m2m(up, tr.t.level_nodes + tr.level_off[lev], tr.level_off[lev + 1] - tr.level_off[lev], s);
        double* node_norm = c->buf<double>("sel_node_norm", nn);
        launch_sel_node_norm(CM, tr.t.r, tb.xyz, nn, nm, 1, node_norm, s);
        // sampled pairs: uniform over the M2L lists (seeded)
        std::mt19937_64 rng(seed);
        std::vector<int> pick(nsamples);
        for (auto& v : pick) v = (int)(rng() % (unsigned long long)L.n_m2l);
        int* dpick = c->buf<int>("sel_pick", nsamples);
        FMMH_HIP(hipMemcpyAsync(dpick, pick.data(), nsamples * sizeof(int), hipMemcpyHostToDevice, s));
        float* feat = c->buf<float>("sel_feat", (size_t)nsamples * SEL_IN);
        double* err = c->buf<double>("sel_err", (size_t)nsamples * (cap + 1));
        double* fref = c->buf<double>("sel_fref", nsamples);
        SelSampleArgs sa{L.m2l_keys, dpick, nsamples, tr.t.cx, tr.t.cy, tr.t.cz, tr.t.r, tr.t.start, tr.t.count,
                         xs, ys, zs, ms, n, CM, nm, cap, tb.cs, tb.xyz, tb.level_base, node_norm, feat, err, fref};
        launch_sel_sample(sa, m, s);
        std::vector<unsigned long long> keys(nsamples);
        c->train_feat.assign((size_t)nsamples * SEL_IN, 0.f);
        c->train_err.assign((size_t)nsamples * (cap + 1), 0.0);
        c->train_fref.assign(nsamples, 0.0);
        std::vector<unsigned long long> all_keys(L.n_m2l);
        FMMH_HIP(hipMemcpyAsync(c->train_feat.data(), feat, c->train_feat.size() * sizeof(float),
                                hipMemcpyDeviceToHost, s));
        FMMH_HIP(hipMemcpyAsync(c->train_err.data(), err, c->train_err.size() * sizeof(double),
                                hipMemcpyDeviceToHost, s));
        FMMH_HIP(hipMemcpyAsync(c->train_fref.data(), fref, nsamples * sizeof(double), hipMemcpyDeviceToHost, s));
        FMMH_HIP(hipMemcpyAsync(all_keys.data(), L.m2l_keys, L.n_m2l * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, s));
        c->sync();
        c->train_pairs.assign((size_t)nsamples * 2, 0);
        for (int k = 0; k < nsamples; ++k) {
            c->train_pairs[2 * k] = (int)(all_keys[pick[k]] >> 32);
            c->train_pairs[2 * k + 1] = src_node((int)(all_keys[pick[k]] & 0xffffffffull));
        }
        c->train_nsamples = nsamples;
        c->train_cap = cap;
        auto t1 = std::chrono::steady_clock::now();
        TrainSet ts = build_examples(c->train_feat, c->train_err, nsamples, cap, seed);
        const int ne = (int)ts.Y.size();
        if (ne < 64) throw Error(FMMH_ERR_ARG, "too few usable training samples");
        // feature standardisation (model constants)
        float norm[2 * SEL_IN];
        for (int f = 0; f < SEL_IN; ++f) {
            double mu = 0.0, m2 = 0.0;
            for (int e = 0; e < ne; ++e) mu += ts.X[(size_t)e * SEL_IN + f];
            mu /= ne;
            for (int e = 0; e < ne; ++e) { const double d = ts.X[(size_t)e * SEL_IN + f] - mu; m2 += d * d; }
            const double sd = std::sqrt(m2 / ne);
            norm[f] = (float)mu;
            norm[SEL_IN + f] = (float)(sd > 1e-12 ? 1.0 / sd : 1.0);
        }
        // He-uniform initial weights; the output bias starts at the mean label
        std::vector<float> W(SEL_NPARAM, 0.f);
        auto fill = [&](int off, int count, int fan_in) {
            const double lim = std::sqrt(6.0 / fan_in);
            std::uniform_real_distribution<double> u(-lim, lim);
            for (int x = 0; x < count; ++x) W[off + x] = (float)u(rng);
        };
        fill(0, SEL_H * SEL_IN, SEL_IN);
        fill(SEL_H * SEL_IN + SEL_H, SEL_H * SEL_H, SEL_H);
        fill(SEL_H * SEL_IN + SEL_H + SEL_H * SEL_H + SEL_H, SEL_H, SEL_H);
        double ym = 0.0;
        for (float y : ts.Y) ym += y;
        W[SEL_NPARAM - 1] = (float)(ym / ne);
        const int batch = std::min(1024, (ne / 64) * 64);
        const int nblk = batch / 64;
        const int steps_per_epoch = std::max(1, ne / batch);
        const int steps = epochs * steps_per_epoch;
        float* dX = c->buf<float>("sel_tX", (size_t)ne * SEL_IN);
        float* dY = c->buf<float>("sel_tY", ne);
        float* dW = c->buf<float>("sel_W", SEL_NPARAM);
        float* dnorm = c->buf<float>("sel_norm", 2 * SEL_IN);
        float* part = c->buf<float>("sel_part", (size_t)nblk * (SEL_NPARAM + 1));
        float* grad = c->buf<float>("sel_grad", SEL_NPARAM + 1);
        float* am = c->buf<float>("sel_adam_m", SEL_NPARAM);
        float* av = c->buf<float>("sel_adam_v", SEL_NPARAM);
        int* step = c->buf<int>("sel_step", 1);
        float* loss_log = c->buf<float>("sel_loss", steps);
        FMMH_HIP(hipMemcpyAsync(dX, ts.X.data(), ts.X.size() * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dY, ts.Y.data(), ts.Y.size() * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dW, W.data(), SEL_NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dnorm, norm, sizeof(norm), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemsetAsync(am, 0, SEL_NPARAM * sizeof(float), s));
        FMMH_HIP(hipMemsetAsync(av, 0, SEL_NPARAM * sizeof(float), s));
        FMMH_HIP(hipMemsetAsync(step, 0, sizeof(int), s));
        c->sync();
        // a multiplier coprime to ne for the per-epoch sample permutation
        unsigned long long pa = 2654435761ull % (unsigned long long)ne;
        auto gcd = [](unsigned long long a, unsigned long long b) { while (b) { unsigned long long t = a % b; a = b; b = t; } return a; };
        while (pa < 2 || gcd(pa, (unsigned long long)ne) != 1) ++pa;
        MlpArgs ma{dX, dY, ne, dnorm, dW, part, grad, am, av, step, loss_log, batch, nblk, 1, pa,
                   (unsigned long long)(seed % (unsigned long long)ne), 3e-3f, 0.9f, 0.999f, 1e-8f, 4.0f,
                   std::max(1, steps / 4)};
        // the training steps run as replays of a captured graph of `chunk` steps
        // (the step counter lives on the device, so every replay advances it)
        const int chunk = std::min(steps, 32);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        FMMH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < chunk; ++k) launch_mlp_train_step(ma, s);
        FMMH_HIP(hipStreamEndCapture(s, &graph));
        FMMH_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        const int replays = steps / chunk;
        for (int r = 0; r < replays; ++r) FMMH_HIP(hipGraphLaunch(exec, s));
        for (int k = replays * chunk; k < steps; ++k) launch_mlp_train_step(ma, s);
        c->sync();
        FMMH_HIP(hipGraphExecDestroy(exec));
        FMMH_HIP(hipGraphDestroy(graph));
        auto t2 = std::chrono::steady_clock::now();
        c->train_loss.assign(steps, 0.f);
        FMMH_HIP(hipMemcpy(c->train_loss.data(), loss_log, steps * sizeof(float), hipMemcpyDeviceToHost));
        c->sel_trained = true;
        c->sel_adapt = 1.0;
        c->sel_cap = cap;
        c->sel_rel_lo = ts.rel_lo;
        c->sel_rel_hi = ts.rel_hi;
        double tail = 0.0;
        const int nt = std::min(steps, 16);
        for (int k = steps - nt; k < steps; ++k) tail += c->train_loss[k];
        if (info) {
            info[0] = ne;
            info[1] = nsamples;
            info[2] = c->train_loss[0];
            info[3] = tail / nt;
            info[4] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            info[5] = std::chrono::duration<double, std::milli>(t2 - t1).count();
            info[6] = steps;
            info[7] = replays;
        }
    });
}

int fmmh_selector_samples(fmmh_context* ctx, int32_t* pairs, float* features, double* errors, double* field_rms) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (c->train_nsamples == 0) throw Error(FMMH_ERR_ARG, "no training samples on this context");
        if (pairs) std::copy(c->train_pairs.begin(), c->train_pairs.end(), pairs);
        if (features) std::copy(c->train_feat.begin(), c->train_feat.end(), features);
        if (errors) std::copy(c->train_err.begin(), c->train_err.end(), errors);
        if (field_rms) std::copy(c->train_fref.begin(), c->train_fref.end(), field_rms);
    });
}

int fmmh_selector_loss(fmmh_context* ctx, float* loss, int32_t cap) {
    int count = 0;
    int rc = guarded([&] {
        fmmh_context* c = resolve(ctx);
        count = (int)c->train_loss.size();
        for (int k = 0; k < std::min(count, (int)cap); ++k) loss[k] = c->train_loss[k];
    });
    return rc == FMMH_OK ? count : -rc;
}

int fmmh_selector_get(fmmh_context* ctx, float* weights, float* norm, double* meta) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (!c->sel_trained) throw Error(FMMH_ERR_ARG, "no trained order selector on this context");
        c->sync();
        FMMH_HIP(hipMemcpy(weights, c->buf<float>("sel_W", SEL_NPARAM), SEL_NPARAM * sizeof(float),
                           hipMemcpyDeviceToHost));
        FMMH_HIP(hipMemcpy(norm, c->buf<float>("sel_norm", 2 * SEL_IN), 2 * SEL_IN * sizeof(float),
                           hipMemcpyDeviceToHost));
        meta[0] = c->sel_cap;
        meta[1] = c->sel_rel_lo;
        meta[2] = c->sel_rel_hi;
        meta[3] = c->sel_margin;
    });
}

int fmmh_selector_set(fmmh_context* ctx, const float* weights, const float* norm, const double* meta) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (!(meta[0] >= 1 && meta[0] <= PMAX) || !(meta[1] <= meta[2]))
            throw Error(FMMH_ERR_ARG, "selector meta must be {cap 1..10, rel_lo <= rel_hi, margin}");
        FMMH_HIP(hipMemcpy(c->buf<float>("sel_W", SEL_NPARAM), weights, SEL_NPARAM * sizeof(float),
                           hipMemcpyHostToDevice));
        FMMH_HIP(hipMemcpy(c->buf<float>("sel_norm", 2 * SEL_IN), norm, 2 * SEL_IN * sizeof(float),
                           hipMemcpyHostToDevice));
        c->sel_cap = (int)meta[0];
        c->sel_rel_lo = (float)meta[1];
        c->sel_rel_hi = (float)meta[2];
        c->sel_margin = (float)meta[3];
        c->sel_trained = true;
        c->sel_adapt = 1.0;
    });
}

int fmmh_selector_grad(fmmh_context* ctx, const float* weights, const float* norm, const float* features,
                       const float* targets, int32_t ns, float asym, float* grad_loss) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (ns < 1 || ns > 65536) throw Error(FMMH_ERR_ARG, "ns must be in 1..65536");
        hipStream_t s = c->stream;
        const int nblk = (ns + 63) / 64;
        float* dX = c->buf<float>("chk_X", (size_t)ns * SEL_IN);
        float* dY = c->buf<float>("chk_Y", ns);
        float* dW = c->buf<float>("chk_W", SEL_NPARAM);
        float* dn = c->buf<float>("chk_norm", 2 * SEL_IN);
        float* part = c->buf<float>("chk_part", (size_t)nblk * (SEL_NPARAM + 1));
        float* grad = c->buf<float>("chk_grad", SEL_NPARAM + 1);
        FMMH_HIP(hipMemcpyAsync(dX, features, (size_t)ns * SEL_IN * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dY, targets, ns * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dW, weights, SEL_NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dn, norm, 2 * SEL_IN * sizeof(float), hipMemcpyHostToDevice, s));
        MlpArgs ma{dX, dY, ns, dn, dW, part, grad, nullptr, nullptr, nullptr, nullptr, ns, nblk, 0, 0, 0,
                   0.f, 0.f, 0.f, 0.f, asym, 1};
        launch_mlp_grad(ma, s);
        FMMH_HIP(hipMemcpyAsync(grad_loss, grad, (SEL_NPARAM + 1) * sizeof(float), hipMemcpyDeviceToHost, s));
        c->sync();
    });
}

int fmmh_selector_predict(fmmh_context* ctx, const float* weights, const float* norm, const float* features,
                          int32_t ns, float* out) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        if (ns < 1) throw Error(FMMH_ERR_ARG, "ns must be >= 1");
        hipStream_t s = c->stream;
        float* dX = c->buf<float>("chk_X", (size_t)ns * SEL_IN);
        float* dY = c->buf<float>("chk_Y", ns);
        float* dW = c->buf<float>("chk_W", SEL_NPARAM);
        float* dn = c->buf<float>("chk_norm", 2 * SEL_IN);
        FMMH_HIP(hipMemcpyAsync(dX, features, (size_t)ns * SEL_IN * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dW, weights, SEL_NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
        FMMH_HIP(hipMemcpyAsync(dn, norm, 2 * SEL_IN * sizeof(float), hipMemcpyHostToDevice, s));
        launch_sel_predict(dW, dn, dX, ns, dY, s);
        FMMH_HIP(hipMemcpyAsync(out, dY, ns * sizeof(float), hipMemcpyDeviceToHost, s));
        c->sync();
    });
}

int fmmh_selection_stats(fmmh_context* ctx, double* out, int32_t cap) {
    return guarded([&] {
        fmmh_context* c = resolve(ctx);
        for (int i = 0; i < std::min(cap, 8); ++i) out[i] = c->sel_stats[i];
        for (int q = 0; q <= PMAX && 8 + q < cap; ++q) out[8 + q] = (double)c->sel_hist[q];
        if (cap > 9 + PMAX) out[9 + PMAX] = c->sel_adapt;
        if (cap > 10 + PMAX) out[10 + PMAX] = c->sel_full_err;
    });
}

int32_t fmmh_selector_nparam(void) { return SEL_NPARAM; }

int fmmh_tuning(fmmh_context* ctx, int64_t* keys, int32_t* variants, double* last_ms, int32_t cap) {
    int count = 0;
    int rc = guarded([&] {
        fmmh_context* c = resolve(ctx);
        for (auto& kv : c->tuned) {
            if (count < cap) {
                if (keys) keys[count] = kv.first;
                if (variants) variants[count] = kv.second;
            }
            ++count;
        }
        if (last_ms)
            for (int v = 0; v < M2L_MFMA_VARIANTS; ++v)
                last_ms[v] = v < (int)c->tune_ms.size() ? c->tune_ms[v] : -1.0;
    });
    return rc == FMMH_OK ? count : -rc;
}

}  // extern "C"
