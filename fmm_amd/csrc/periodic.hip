// Periodic boundaries of field_fmm: the lattice sum beyond the 27 explicit
// cells (periodic_sum, src/fmm.F90:884-944) and the dipole correction
// (dipole_correction, src/fmm.F90:946-979).  The 27 cells themselves go
// through the ordinary interaction lists with image displacements.
#include "common.h"
#include "hermite.h"
#include "kernels.h"

namespace fmmh {

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// 3^l as the reference forms it (3.0_rp ** level): exact while 3^l < 2^53.
__device__ __forceinline__ double pow3(int l) {
    double v = 1.0;
    for (int q = 0; q < l; ++q) v *= 3.0;
    return v;
}

// matmul(box_matrix, a), column by column (src/fmm.F90:914, 923)
__device__ __forceinline__ void box_times(const double* __restrict__ box, double a0, double a1, double a2, double* d) {
#pragma clang fp contract(off)
    for (int r = 0; r < 3; ++r) d[r] = (box[r] * a0 + box[r + 3] * a1) + box[r + 6] * a2;
}

// Neighbour id 0..25 -> (ai, aj, ak) in the reference's loop order without the centre.
__device__ __forceinline__ void neighbour(int id, int& ai, int& aj, int& ak) {
    const int q = id < 13 ? id : id + 1;
    ai = q / 9 - 1;
    aj = (q / 3) % 3 - 1;
    ak = q % 3 - 1;
}

// One workgroup per (level, neighbour supercell): sum of T(d) over its 27
// sub-cells, d = matmul(box, a 3^l) + matmul(box, b 3^{l-1}).
__global__ __launch_bounds__(64) void k_periodic_shells(PeriodicArgs a) {
#pragma clang fp contract(off)
    extern __shared__ double lds[];
    const int K = 2 * a.p;
    const int nT = ncomp(K);
    double* A = lds;
    double* sum = lds + hermite_aux_size(K);
    const int level = blockIdx.x / 26 + 1;
    const int nb = blockIdx.x % 26;
    int ai, aj, ak;
    neighbour(nb, ai, aj, ak);
    const double s3 = pow3(level), s2 = pow3(level - 1);
    double da[3];
    box_times(a.box, ai * s3, aj * s3, ak * s3, da);
    for (int idx = threadIdx.x; idx < nT; idx += 64) sum[idx] = 0.0;
    for (int b = 0; b < 27; ++b) {
        const int bi = b / 9 - 1, bj = (b / 3) % 3 - 1, bk = b % 3 - 1;
        double db[3];
        box_times(a.box, bi * s2, bj * s2, bk * s2, db);
        __syncthreads();
        hermite_aux(da[0] + db[0], da[1] + db[1], da[2] + db[2], K, A, a.xyz, a.level_base);
        for (int idx = threadIdx.x; idx < nT; idx += 64) sum[idx] += hermite_T(A, idx, a.xyz, a.level_base);
    }
    for (int idx = threadIdx.x; idx < nT; idx += 64)
        a.shells[((int64_t)(level - 1) * 26 + nb) * nT + idx] = sum[idx];
}

constexpr int PS_NM = ncomp(PMAX);
constexpr int PS_NT = ncomp(2 * PMAX);

// One workgroup per set: the level loop of periodic_sum on the root expansions.
__global__ __launch_bounds__(256) void k_periodic_sum(PeriodicArgs a) {
    __shared__ double M[PS_NM], Mn[PS_NM], mono[PS_NM], contrib[PS_NM];
    __shared__ double Sl[PS_NT];
    __shared__ double red[256];
    __shared__ int done;
    const int set = blockIdx.x;
    const int nm = ncomp(a.p), nT = ncomp(2 * a.p);
    const int tid = threadIdx.x;
    const double* __restrict__ cm = a.CM + (int64_t)set * a.num_nodes * nm;   // root = node 0
    double* __restrict__ le = a.LE + (int64_t)set * a.num_nodes * nm;
    for (int t = tid; t < nm; t += 256) {
        const double v = t == 0 ? 0.0 : cm[t];     // charge neutrality (src/fmm.F90:897)
        M[t] = v;
        Mn[t] = v;
    }
    if (tid == 0) done = 0;
    __syncthreads();
    int level = 1;
    for (; level <= PSUM_LEVELS; ++level) {
        for (int idx = tid; idx < nT; idx += 256) {
            double v = 0.0;
            for (int nb = 0; nb < 26; ++nb) v += a.shells[((int64_t)(level - 1) * 26 + nb) * nT + idx];
            Sl[idx] = v;
        }
        __syncthreads();
        // contribution(t) = sum_s c_s M(s) S_l[s + t]   (multipole_field at order p)
        double sq = 0.0;
        for (int t = tid; t < nm; t += 256) {
            const int pt = a.xyz[t];
            const int tx = pt & 255, ty = (pt >> 8) & 255, tz = pt >> 16;
            double v = 0.0;
            for (int s = 0; s < nm; ++s) {
                const int ps = a.xyz[s];
                v = __builtin_fma(a.cs[s] * M[s], Sl[xyz2idx(tx + (ps & 255), ty + ((ps >> 8) & 255), tz + (ps >> 16))], v);
            }
            contrib[t] = v;
            le[t] += v;
            if (t >= 1) sq += v * v;
        }
        red[tid] = sq;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        const bool converged = sqrt(red[0]) < 1.0e-8;
        // next supercell multipole: M + sum over the 26 neighbours of M shifted by matmul(box, a 3^l) / 3
        for (int nb = 0; nb < 26; ++nb) {
            int ai, aj, ak;
            neighbour(nb, ai, aj, ak);
            const double s3 = pow3(level);
            if (tid == 0) {
                double d[3];
                box_times(a.box, ai * s3, aj * s3, ak * s3, d);
                const double dx = d[0] / 3.0, dy = d[1] / 3.0, dz = d[2] / 3.0;
                mono[0] = 1.0;
                for (int e = 1; e < nm; ++e) {
                    const int ax = a.mono_axis[e];
                    mono[e] = mono[a.mono_parent[e]] * (ax == 0 ? dx : (ax == 1 ? dy : dz));
                }
            }
            __syncthreads();
            for (int t = tid; t < nm; t += 256) {
                double v = 0.0;
                for (int b = a.m2m.off[t]; b < a.m2m.off[t + 1]; ++b)
                    v = __builtin_fma(a.m2m.coef[b] * mono[a.m2m.e[b]], M[a.m2m.s[b]], v);
                Mn[t] += v;
            }
            __syncthreads();
        }
        for (int t = tid; t < nm; t += 256) M[t] = Mn[t];
        __syncthreads();
        if (converged) break;
    }
    if (tid == 0 && a.levels_used) a.levels_used[set] = min(level, PSUM_LEVELS);
}

// dipole_correction: potential and field of the cell's dipole (and the
// quadrupole trace term of the potential), with the reference's constants
// and operation order.
__global__ __launch_bounds__(256) void k_dipole_correction(PeriodicArgs a) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int set = blockIdx.y;
    if (i >= a.n) return;
    const int nm = ncomp(a.p);
    const double* __restrict__ root = a.CM + (int64_t)set * a.num_nodes * nm;
    const double d1 = root[1], d2 = root[2], d3 = root[3];
    const double qtr = (root[4] + root[7]) + root[9];
    const double x1 = a.X[i], x2 = a.X[a.ldx + i], x3 = a.X[2 * a.ldx + i];
    const double sdx = (d1 * x1 + d2 * x2) + d3 * x3;
    double* f = a.field + (int64_t)set * a.nf * a.ldf;
    f[i] = (f[i] + (a.fourpi * sdx) / a.threeV) + (a.twopi * qtr) / a.threeV;
    if (a.nf >= 4) {
        f[a.ldf + i] = f[a.ldf + i] + (a.fourpi * d1) / a.threeV;
        f[2 * a.ldf + i] = f[2 * a.ldf + i] + (a.fourpi * d2) / a.threeV;
        f[3 * a.ldf + i] = f[3 * a.ldf + i] + (a.fourpi * d3) / a.threeV;
    }
}

void launch_periodic_shells(const PeriodicArgs& a, hipStream_t s) {
    const int K = 2 * a.p;
    const size_t lds = (size_t)(hermite_aux_size(K) + ncomp(K)) * sizeof(double);
    FMMH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_periodic_shells),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_periodic_shells, dim3(PSUM_LEVELS * 26), dim3(64), lds, s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_periodic_sum(const PeriodicArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_periodic_sum, dim3(a.nsets), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_dipole_correction(const PeriodicArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_dipole_correction, dim3(nblk(a.n, 256), a.nsets), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
