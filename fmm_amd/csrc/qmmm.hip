// QM/MM embedding helpers of peter-reinholdt/fmm (src/fmm.F90:1134-1336):
// classification of the octree against target points for get_multipoles and
// get_boxes, and the compaction of its near-field sites and far boxes.
#include "common.h"
#include "kernels.h"

namespace fmmh {

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// flang NORM2 of a 3-vector: running maximum with a scaled sum of squares
// (the same emulation as the dual traversal's MAC).
__device__ __forceinline__ double qm_norm2(double v0, double v1, double v2) {
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

constexpr int CLS_TILE = 256;

// close[node] = any target with norm2(target - centre) * theta <= rmax
// (src/fmm.F90:1148-1154).  Targets are staged through LDS in tiles.
__global__ __launch_bounds__(256) void k_classify_close(const double* __restrict__ tg, int64_t ldt, int nt,
                                                        double theta, TreeDev t, int num_nodes, int* __restrict__ close) {
#pragma clang fp contract(off)
    __shared__ double sx[CLS_TILE], sy[CLS_TILE], sz[CLS_TILE];
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = node < num_nodes;
    const double cx = live ? t.cx[node] : 0.0, cy = live ? t.cy[node] : 0.0, cz = live ? t.cz[node] : 0.0;
    const double rmax = live ? t.rmax[node] : 0.0;
    int c = 0;
    for (int b0 = 0; b0 < nt; b0 += CLS_TILE) {
        const int cnt = min(CLS_TILE, nt - b0);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            sx[threadIdx.x] = tg[b0 + threadIdx.x];
            sy[threadIdx.x] = tg[ldt + b0 + threadIdx.x];
            sz[threadIdx.x] = tg[2 * ldt + b0 + threadIdx.x];
        }
        __syncthreads();
        if (live && !c)
            for (int q = 0; q < cnt; ++q) {
                const double r = qm_norm2(sx[q] - cx, sy[q] - cy, sz[q] - cz);
                if (r * theta <= rmax) { c = 1; break; }
            }
    }
    if (live) close[node] = c;
}

// Visited iff every ancestor is close (src/fmm.F90:1156-1168).
__global__ __launch_bounds__(256) void k_classify_type(TreeDev t, int num_nodes, const int* __restrict__ close,
                                                       int* __restrict__ type) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= num_nodes) return;
    bool visited = true;
    for (int a = t.parent[node]; a >= 0; a = t.parent[a])
        if (!close[a]) { visited = false; break; }
    int ty = 0;
    if (visited) ty = close[node] ? (t.is_leaf[node] ? 2 : 0) : 1;
    type[node] = ty;
}

__global__ __launch_bounds__(256) void k_qmmm_counts(const int* __restrict__ type, const int* __restrict__ count,
                                                     int num_nodes, int* __restrict__ near_cnt,
                                                     int* __restrict__ far_cnt) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node < num_nodes) {
        near_cnt[node] = type[node] == 2 ? count[node] : 0;
        far_cnt[node] = type[node] == 1 ? 1 : 0;
    } else if (node == num_nodes) {
        near_cnt[node] = 0;
        far_cnt[node] = 0;
    }
}

__global__ __launch_bounds__(256) void k_qmmm_compact(const int* __restrict__ type, TreeDev t, int num_nodes,
                                                      const int* __restrict__ near_off, const int* __restrict__ far_off,
                                                      int* __restrict__ near_sites, int* __restrict__ far_nodes) {
    const int node = blockIdx.x;
    if (node >= num_nodes) return;
    const int ty = type[node];
    if (ty == 2) {
        const int o = near_off[node], st = t.start[node], c = t.count[node];
        for (int q = threadIdx.x; q < c; q += blockDim.x) near_sites[o + q] = t.perm[st + q] + 1;
    } else if (ty == 1 && threadIdx.x == 0) {
        far_nodes[far_off[node]] = node;
    }
}

void launch_classify(const double* targets, int64_t ldt, int ntargets, double theta, const TreeDev& t, int num_nodes,
                     int* close, int* type, hipStream_t s) {
    hipLaunchKernelGGL(k_classify_close, dim3(nblk(num_nodes, 256)), dim3(256), 0, s, targets, ldt, ntargets, theta, t,
                       num_nodes, close);
    hipLaunchKernelGGL(k_classify_type, dim3(nblk(num_nodes, 256)), dim3(256), 0, s, t, num_nodes, close, type);
    FMMH_HIP(hipGetLastError());
}

void launch_qmmm_counts(const int* type, const int* count, int num_nodes, int* near_cnt, int* far_cnt, hipStream_t s) {
    hipLaunchKernelGGL(k_qmmm_counts, dim3(nblk(num_nodes + 1, 256)), dim3(256), 0, s, type, count, num_nodes, near_cnt,
                       far_cnt);
    FMMH_HIP(hipGetLastError());
}

void launch_qmmm_compact(const int* type, const TreeDev& t, int num_nodes, const int* near_off, const int* far_off,
                         int* near_sites, int* far_nodes, hipStream_t s) {
    hipLaunchKernelGGL(k_qmmm_compact, dim3(num_nodes), dim3(64), 0, s, type, t, num_nodes, near_off, far_off,
                       near_sites, far_nodes);
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
