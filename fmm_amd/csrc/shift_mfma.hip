// M2M and L2L on FP32 MFMA, for the mixed-precision mode (FMMH_OPT_PRECISION = 1).
//
// In the octree every child sits at d = r_c u from its parent, with r_c the child's half-width and
// u = (+-1, +-1, +-1) fixed by the octant o (bit 0: x, bit 1: y, bit 2: z).  In box-scaled form,
// M~(s) = M(s) r^-|s| for multipoles and L~(t) = L(t) r^|t| for local expansions, the two shifts of
// the reference (M2M src/fmm.F90:387-435, L2L src/fmm.F90:981-1044; the binomial / Taylor
// coefficients c(t, s) of the FP64 shift tables) become one matrix per octant, the same on every
// level because r_parent = 2 r_c exactly:
//   M2M   M~_parent(t) = sum_o sum_s c(t, s) u_o^(t-s) 2^-|t| M~_child_o(s)
//   L2L   L~_child_o(t) += sum_s c(t, s) u_o^(s-t) 2^-|s| L~_parent(s)
// The matrices (built once per order on the host, kept in MFMA fragment order) have entries of
// magnitude <= 1, and the scaled expansions stay in FP32 range whatever the length unit.
//
// A workgroup (four wavefronts) takes 32 columns -- (internal node, set) pairs, set fastest -- of one
// level; the columns' source expansions are scaled and rounded to FP32 into LDS, and each wavefront
// runs its 16-row target tiles on v_mfma_f32_16x16x4f32, FP32 accumulation over runs of 32 source
// coefficients, each run's FP32 tile added to FP64 accumulators (compensated summation of the products
// into the expansions).  M2M sums the eight octants that way and writes the parent unscaled in FP64;
// L2L adds each octant's sum, unscaled in FP64, to the child's FP64 local expansion.  The columns come from per-level lists of
// internal nodes (k_internal_nodes), whose counts stay on the device.
//
// MFMA operand layouts (measured on gfx950, f32_16x16x4f32): A lane l -> A[l % 16][l / 16],
// B lane l -> B[l / 16][l % 16], accumulator register q of lane l -> D[4 (l / 16) + q][l % 16].
#include "common.h"
#include "kernels.h"

namespace fmmh {

using v4f = float __attribute__((ext_vector_type(4)));

__global__ void k_fill_cnt(int* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

constexpr int SHIFT_NT = 2;   // 16-column tiles per workgroup (each A fragment feeds SHIFT_NT MFMAs)
constexpr int KC = 8;         // source steps whose A fragments are loaded together
// Row-tile groups: G workgroups split a column block's 16-row target tiles (at most two per wavefront), so
// that the top levels, with a handful of nodes, do not serialise all tiles of all octants on four waves;
// L2L also takes one octant per workgroup.  Every output element is still computed by one wavefront in
// the same order.
constexpr int shift_row_groups(int nrows) { return (((nrows + 15) / 16) + 7) / 8; }

// RED (L2L only): the local expansions are harmonic and the shift runs on their (p + 1)^2 components with
// t_z <= 1 (read and written at their packed positions a.tfull[k]; the other components are neither read
// nor kept up to date) with the matrices R S U -- the full shift S between the unfold U (components with
// t_z >= 2 from the reduced ones, exact for a harmonic expansion) and the restriction R to the reduced
// rows; a harmonic parent gives a harmonic child, so the reduced child is exact.
template <int P, bool L2L, bool RED>
__global__ __launch_bounds__(256) void k_shift_mfma(ShiftMfmaArgs a) {
    static_assert(!RED || L2L, "the reduced form is for local expansions");
    constexpr int NMS = ncomp(P);               // row length of the expansions in memory
    constexpr int NM = RED ? rdim(P) : NMS;     // rows and k of the shift
    constexpr int MT = (NM + 15) / 16;          // 16-row target tiles
    constexpr int KS = (NM + 3) / 4;            // 4-deep source steps
    constexpr int KP = KS * 4;
    constexpr int G = shift_row_groups(NM);     // workgroups sharing the row tiles of one column block
    constexpr int MW = (MT + 4 * G - 1) / (4 * G);   // target tiles per wavefront
    constexpr int NT = SHIFT_NT;
    constexpr int NC = 16 * NT;                 // columns per workgroup
    __shared__ float Bs[NC][KP + 1];
    __shared__ double rpw[NC][P + 1];           // r_parent^k
    __shared__ double rin[NC][P + 1];           // r_child^-k = (2 / r_parent)^k
    __shared__ int cnode[NC], cset[NC];
    __shared__ int deg[KP];
    __shared__ int pos[RED ? KP : 1];           // packed position of reduced component k
    const int count = *a.count;                 // internal nodes of the level (device count)
    const int ncol = count * a.nsets;
    const int col0 = blockIdx.x * NC;
    if (col0 >= ncol) return;                   // the grid is sized for every node of the level
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c16 = lane & 15, g4 = lane >> 4;
    const int mrow = 4 * (int)blockIdx.y + w;   // this wavefront's first row tile; then every 4 G-th
    if (tid < NC) {
        const int col = col0 + tid;
        const int node = col < ncol ? a.nodes[col / a.nsets] : -1;
        cnode[tid] = node;
        cset[tid] = col < ncol ? col % a.nsets : 0;
        const double r = node >= 0 ? a.r[node] : 1.0;
        const double ri = 2.0 / r;
        double u = 1.0, v = 1.0;
        for (int k = 0; k <= P; ++k) {
            rpw[tid][k] = u;
            rin[tid][k] = v;
            u *= r;
            v *= ri;
        }
    }
    for (int k = tid; k < KP; k += 256) {
        if constexpr (RED) {
            const int f = k < NM ? a.tfull[k] : 0;
            const int x = a.xyz[f];
            deg[k] = k < NM ? (x & 255) + ((x >> 8) & 255) + (x >> 16) : 0;
            pos[k] = f;
        } else {
            const int x = k < NM ? a.xyz[k] : 0;
            deg[k] = (x & 255) + ((x >> 8) & 255) + (x >> 16);
        }
    }
    __syncthreads();
    // the columns' source expansions, box-scaled, FP32: M2M M_child(s) r_child^-|s|, L2L L_parent(s) r_parent^|s|
    auto stage = [&](int o) {
        for (int x = tid; x < NC * KP; x += 256) {
            const int c = x / KP, k = x - c * KP;
            const int node = cnode[c];
            float v = 0.f;
            if (node >= 0 && k < NM) {
                const int sn = L2L ? node : a.children[node * 8 + o];
                if (sn >= 0) {
                    const int kk = RED ? pos[k] : k;
                    const double e = a.src[(int64_t)cset[c] * a.set_stride + (int64_t)sn * NMS + kk];
                    v = (float)(e * (L2L ? rpw[c][deg[k]] : rin[c][deg[k]]));
                }
            }
            Bs[c][k] = v;
        }
    };
    double acc64[MW][NT][4];
#pragma unroll
    for (int mi = 0; mi < MW; ++mi)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc64[mi][nt][q] = 0.0;
    if (L2L) {
        stage(0);
        __syncthreads();
    }
    const int o_first = L2L ? (int)blockIdx.z : 0, o_end = L2L ? o_first + 1 : 8;
    for (int o = o_first; o < o_end; ++o) {
        if (!L2L) {
            if (o > o_first) __syncthreads();    // readers of the previous octant's columns are done
            stage(o);
            __syncthreads();
        }
        const float* __restrict__ A = a.A + (size_t)o * MT * KS * 64;
        if (L2L) {
#pragma unroll
            for (int mi = 0; mi < MW; ++mi)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc64[mi][nt][q] = 0.0;
        }
        // A fragments come from global memory (L2): KC source steps' worth are loaded together, then
        // their MFMAs run, so one wait covers KC x MW loads; the FP32 accumulators cover those KC steps
        // only and are then added to the FP64 ones
        for (int k0 = 0; k0 < KS; k0 += KC) {
            v4f acc[MW][NT];
#pragma unroll
            for (int mi = 0; mi < MW; ++mi)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[mi][nt] = v4f{0.f, 0.f, 0.f, 0.f};
            float av[KC][MW];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int mi = 0; mi < MW; ++mi) {
                    const int ks = k0 + kk, mt = mrow + 4 * G * mi;
                    av[kk][mi] = (ks < KS && mt < MT) ? A[((size_t)mt * KS + ks) * 64 + lane] : 0.f;
                }
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int ks = k0 + kk;
                if (ks >= KS) break;
                float b[NT];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) b[nt] = Bs[nt * 16 + c16][ks * 4 + g4];
#pragma unroll
                for (int mi = 0; mi < MW; ++mi) {
                    if (mrow + 4 * G * mi < MT) {
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[mi][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[kk][mi], b[nt], acc[mi][nt], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int mi = 0; mi < MW; ++mi)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc64[mi][nt][q] += (double)acc[mi][nt][q];
        }
        if (L2L) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int c = nt * 16 + c16;
                const int node = cnode[c];
                const int ch = node >= 0 ? a.children[node * 8 + o] : -1;
                if (ch < 0) continue;
                double* __restrict__ dst = a.dst + (int64_t)cset[c] * a.set_stride + (int64_t)ch * NMS;
#pragma unroll
                for (int mi = 0; mi < MW; ++mi) {
                    const int mt = mrow + 4 * G * mi;
                    if (mt >= MT) continue;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int t = mt * 16 + 4 * g4 + q;
                        if (t < NM) dst[RED ? pos[t] : t] += acc64[mi][nt][q] * rin[c][deg[t]];
                    }
                }
            }
        }
    }
    if (!L2L) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int c = nt * 16 + c16;
            const int node = cnode[c];
            if (node < 0) continue;
            double* __restrict__ dst = a.dst + (int64_t)cset[c] * a.set_stride + (int64_t)node * NMS;
#pragma unroll
            for (int mi = 0; mi < MW; ++mi) {
                const int mt = mrow + 4 * G * mi;
                if (mt >= MT) continue;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int t = mt * 16 + 4 * g4 + q;
                    if (t < NM) dst[t] = acc64[mi][nt][q] * rpw[c][deg[t]];
                }
            }
        }
    }
}

// Internal nodes of every level (any order within a level): per-workgroup LDS counters and one global
// atomic per level and workgroup.  out[level_off[L] + i], i < cnt[L].
__global__ __launch_bounds__(256) void k_internal_nodes(int num_nodes, const int* __restrict__ level,
                                                        const int* __restrict__ children, LevelStarts ls,
                                                        int* __restrict__ out, int* __restrict__ cnt) {
    __shared__ int lc[KEY_LEVELS + 2];
    __shared__ int base[KEY_LEVELS + 2];
    if (threadIdx.x < KEY_LEVELS + 2) lc[threadIdx.x] = 0;
    __syncthreads();
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    int L = 0, pos = -1;
    if (id < num_nodes) {
        bool any = false;
#pragma unroll
        for (int o = 0; o < 8; ++o) any |= children[id * 8 + o] >= 0;
        if (any) {
            L = level[id];
            pos = atomicAdd(&lc[L], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < KEY_LEVELS + 2 && lc[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cnt[threadIdx.x], lc[threadIdx.x]);
    __syncthreads();
    if (pos >= 0) out[ls.off[L] + base[L] + pos] = id;
}

void launch_internal_nodes(int num_nodes, const int* level, const int* children, const LevelStarts& ls, int* out,
                           int* cnt, hipStream_t s) {
    if (num_nodes <= 0) return;
    hipLaunchKernelGGL(k_fill_cnt, dim3(1), dim3(64), 0, s, cnt, KEY_LEVELS + 2);
    hipLaunchKernelGGL(k_internal_nodes, dim3((num_nodes + 255) / 256), dim3(256), 0, s, num_nodes, level, children,
                       ls, out, cnt);
    FMMH_HIP(hipGetLastError());
}

int shift_mfma_frag_floats(int p, bool reduced) {
    const int nm = reduced ? rdim(p) : ncomp(p);
    return ((nm + 15) / 16) * ((nm + 3) / 4) * 64;
}

template <bool L2L, bool RED>
static void launch_p(const ShiftMfmaArgs& a, int p, hipStream_t s) {
    const dim3 g((unsigned)(((long long)a.max_count * a.nsets + 16 * SHIFT_NT - 1) / (16 * SHIFT_NT)),
                 shift_row_groups(RED ? rdim(p) : ncomp(p)), L2L ? 8 : 1);
    switch (p) {
        case 1: hipLaunchKernelGGL((k_shift_mfma<1, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_shift_mfma<2, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_shift_mfma<3, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_shift_mfma<4, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 5: hipLaunchKernelGGL((k_shift_mfma<5, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 6: hipLaunchKernelGGL((k_shift_mfma<6, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 7: hipLaunchKernelGGL((k_shift_mfma<7, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_shift_mfma<8, L2L, RED>), g, dim3(256), 0, s, a); break;
        case 9: hipLaunchKernelGGL((k_shift_mfma<9, L2L, RED>), g, dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((k_shift_mfma<10, L2L, RED>), g, dim3(256), 0, s, a); break;
    }
}

void launch_shift_mfma(const ShiftMfmaArgs& a, int p, bool l2l, hipStream_t s) {
    if (a.max_count <= 0 || p < 1 || p > PMAX) return;
    if (l2l && a.tfull) launch_p<true, true>(a, p, s);
    else if (l2l) launch_p<true, false>(a, p, s);
    else launch_p<false, false>(a, p, s);
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
