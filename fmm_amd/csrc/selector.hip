// Learned expansion order per M2L interaction (no counterpart in the
// reference, whose M2L uses one expansion_order for every pair,
// src/fmm.F90:826-882).
//
// A pair (i, j) of the M2L lists contributes L_i(t) += sum_s c_s M_j(s)
// T_{s+t}(c_i - c_j) with |s|, |t| <= p.  Truncating that one pair at an
// order q <= p (|s|, |t| <= q) changes the field at the sites of box i by
// the pair's truncation error e_q; L2L and L2P are exact for polynomials of
// degree <= p, so e_q is exactly what the final field loses.  A small MLP
// predicts, per pair, the smallest q whose error stays below a per-pair
// tolerance tau:
//   inputs  x0 = r_i / |R|, x1 = r_j / |R|            (box sizes over distance)
//           x2 = log10(A_j / |R|^2)                     (source strength as a field at |R|;
//                A_j = |(M_j(s) r_j^{-|s|})_s|_2 over |s| <= p, max over sets)
//           x3 = log10(tau)                              (absolute field tolerance of the pair)
//   hidden  2 x 64 ReLU, FP32;  output  a real order q^ -> q = ceil(q^ + margin), clamped to [0, p].
// The network is trained on the engine's own measurements (k_sel_sample:
// per sampled pair, M2L at every order q <= cap plus exact evaluation at up
// to SEL_NS sites of box i, against the direct sum over the sites of box j),
// by the forward / backward / Adam kernels below.
#include <climits>

#include "common.h"
#include "kernels.h"
#include "hermite.h"
#include "tensors.h"

namespace fmmh {

constexpr int SEL_CT_AUX = hermite_aux_size(HERMITE_KMAX);
constexpr int SEL_MAX_NM = ncomp(PMAX);

// ------------------------------------------------------------------ training data
// One 64-lane workgroup per sampled pair.  Exact field (potential + field,
// undamped, every site of box j: the far field of the reference carries
// neither damping nor exclusions, src/fmm.F90:872-873) at up to SEL_NS sites
// of box i, then for q = 0..cap the pair's local expansion truncated at q,
// evaluated at the same sites; err[q] = RMS over the sites of |f_q - f|
// (field block).
template <int M>
__global__ __launch_bounds__(64) void k_sel_sample(SelSampleArgs a) {
    constexpr int NM = ncomp(M);
    __shared__ double A[SEL_CT_AUX];
    __shared__ double Mt[SEL_MAX_NM];
    __shared__ double L[SEL_MAX_NM];
    __shared__ double ex[2][SEL_NS][4];
    __shared__ double sq[SEL_NS];
    __shared__ double fact[PMAX + 1];
    const int smp = blockIdx.x;
    const int tid = threadIdx.x;
    const unsigned long long key = a.pair_keys[a.sample_pair[smp]];
    const int i = (int)(key >> 32);
    const int j = src_node((int)(key & 0xffffffffull));
    const double Rx = a.cx[i] - a.cx[j], Ry = a.cy[i] - a.cy[j], Rz = a.cz[i] - a.cz[j];
    const int cap = a.cap, nm = a.nm;
    hermite_aux(Rx, Ry, Rz, 2 * cap, A, a.xyz, a.level_base);
    for (int s = tid; s < nm; s += 64) Mt[s] = a.cs[s] * a.CM[(int64_t)j * nm + s];
    if (tid == 0) {
        double f = 1.0;
        fact[0] = 1.0;
        for (int n = 1; n <= PMAX; ++n) { f *= (double)n; fact[n] = 1.0 / f; }
    }
    const int ci = a.count[i], si = a.start[i];
    const int nt = min(SEL_NS, ci);
    const int k = tid & (SEL_NS - 1), h = tid >> 5;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    if (k < nt) {
        const int site = si + (int)(((long long)k * ci) / nt);
        tx = a.xs[site]; ty = a.ys[site]; tz = a.zs[site];
    }
    {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (k < nt) {
            const int s0 = a.start[j], s1 = s0 + a.count[j];
            for (int jj = s0 + h; jj < s1; jj += 2) {
                const double dx = tx - a.xs[jj], dy = ty - a.ys[jj], dz = tz - a.zs[jj];
                double g[M + 2];
                radial<M + 1, DAMP_NONE>(dx * dx + dy * dy + dz * dz, 0.0, g);
                double T[ncomp(M + 1)];
                md_tensor<M + 1>(dx, dy, dz, g, T);
                double m[NM];
#pragma unroll
                for (int s = 0; s < NM; ++s) m[s] = a.ms[(int64_t)s * a.ldm + jj];
                contract<M, 1>(m, T, acc);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) ex[h][k][c] = acc[c];
    }
    __syncthreads();
    if (tid < SEL_NS) {
#pragma unroll
        for (int c = 0; c < 4; ++c) ex[0][tid][c] += ex[1][tid][c];
    }
    __syncthreads();
    const double dx = tx - a.cx[i], dy = ty - a.cy[i], dz = tz - a.cz[i];
    for (int q = 0; q <= cap; ++q) {
        const int nq = ncomp(q);
        for (int t = tid; t < nq; t += 64) {
            const int pt = a.xyz[t];
            const int ttx = pt & 255, tty = (pt >> 8) & 255, ttz = pt >> 16;
            double sum = 0.0;
            for (int s = 0; s < nq; ++s) {
                const int ps = a.xyz[s];
                const int idx = xyz2idx(ttx + (ps & 255), tty + ((ps >> 8) & 255), ttz + (ps >> 16));
                sum = __builtin_fma(Mt[s], hermite_T(A, idx, a.xyz, a.level_base), sum);
            }
            L[t] = sum;
        }
        __syncthreads();
        if (tid < SEL_NS) {
            double f[4] = {0.0, 0.0, 0.0, 0.0};
            if (k < nt) {
                double px = 1.0;
                for (int e1 = 0; e1 <= q; ++e1) {
                    double pxy = px;
                    for (int e2 = 0; e1 + e2 <= q; ++e2) {
                        double mm = pxy * fact[e1] * fact[e2];
                        for (int e3 = 0; e1 + e2 + e3 <= q; ++e3) {
                            const double mz = mm * fact[e3];
                            const int ne = e1 + e2 + e3;
                            f[0] = __builtin_fma(mz, L[xyz2idx(e1, e2, e3)], f[0]);
                            if (ne < q) {
                                f[1] = __builtin_fma(mz, L[xyz2idx(e1 + 1, e2, e3)], f[1]);
                                f[2] = __builtin_fma(mz, L[xyz2idx(e1, e2 + 1, e3)], f[2]);
                                f[3] = __builtin_fma(mz, L[xyz2idx(e1, e2, e3 + 1)], f[3]);
                            }
                            mm *= dz;
                        }
                        pxy *= dy;
                    }
                    px *= dx;
                }
                double e2s = 0.0;
#pragma unroll
                for (int c = 1; c < 4; ++c) {
                    const double d = f[c] - ex[0][k][c];
                    e2s += d * d;
                }
                sq[k] = e2s;
            }
        }
        __syncthreads();
        if (tid == 0) {
            double s2 = 0.0;
            for (int kk = 0; kk < nt; ++kk) s2 += sq[kk];
            a.err[(int64_t)smp * (cap + 1) + q] = sqrt(s2 / nt);
        }
        __syncthreads();
    }
    if (tid == 0) {
        double s2 = 0.0;
        for (int kk = 0; kk < nt; ++kk)
            for (int c = 1; c < 4; ++c) s2 += ex[0][kk][c] * ex[0][kk][c];
        const double R = sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
        a.fref[smp] = sqrt(s2 / nt);
        float* x = a.feat + (int64_t)smp * SEL_IN;
        x[0] = (float)(a.r[i] / R);
        x[1] = (float)(a.r[j] / R);
        x[2] = (float)log10(a.node_norm[j] / (R * R));
        x[3] = 0.0f;
    }
}

void launch_sel_sample(const SelSampleArgs& a, int m, hipStream_t s) {
    if (a.nsamples <= 0) return;
    switch (m) {
        case 0: hipLaunchKernelGGL(k_sel_sample<0>, dim3(a.nsamples), dim3(64), 0, s, a); break;
        case 1: hipLaunchKernelGGL(k_sel_sample<1>, dim3(a.nsamples), dim3(64), 0, s, a); break;
        default: hipLaunchKernelGGL(k_sel_sample<2>, dim3(a.nsamples), dim3(64), 0, s, a); break;
    }
    FMMH_HIP(hipGetLastError());
}

// A_j = |(M_j(s) r_j^{-|s|})_s|_2 over the nm expansion components, max over sets.
// One wavefront per node: lane l takes the coefficients s = l, l + 64, ... of every set (their box
// scale r^-|s| computed once), the squares of a set are summed across the wavefront in a fixed order,
// and the node keeps the largest set norm.
constexpr int NORM_PER_LANE = (ncomp(PMAX) + 63) / 64;

__global__ __launch_bounds__(64) void k_sel_node_norm(const double* __restrict__ CM, const double* __restrict__ r,
                                                      const int* __restrict__ xyz, int num_nodes, int nm, int nsets,
                                                      double* __restrict__ out) {
    const int j = blockIdx.x;
    if (j >= num_nodes) return;
    const int lane = threadIdx.x;
    const double rinv = 1.0 / r[j];
    double f[NORM_PER_LANE];
#pragma unroll
    for (int q = 0; q < NORM_PER_LANE; ++q) {
        const int s = lane + 64 * q;
        double v = 0.0;
        if (s < nm) {
            const int ps = xyz[s];
            const int d = (ps & 255) + ((ps >> 8) & 255) + (ps >> 16);
            v = 1.0;
            for (int k = 0; k < d; ++k) v *= rinv;
        }
        f[q] = v;
    }
    double best = 0.0;
    for (int set = 0; set < nsets; ++set) {
        const double* c = CM + ((int64_t)set * num_nodes + j) * nm;
        double s2 = 0.0;
#pragma unroll
        for (int q = 0; q < NORM_PER_LANE; ++q) {
            const int s = lane + 64 * q;
            const double v = s < nm ? c[s] * f[q] : 0.0;
            s2 = __builtin_fma(v, v, s2);
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) s2 += __shfl_xor(s2, d);
        best = fmax(best, s2);
    }
    if (lane == 0) out[j] = sqrt(best);
}

void launch_sel_node_norm(const double* CM, const double* r, const int* xyz, int num_nodes, int nm, int nsets,
                          double* out, hipStream_t s) {
    if (num_nodes <= 0) return;
    hipLaunchKernelGGL(k_sel_node_norm, dim3(num_nodes), dim3(64), 0, s, CM, r, xyz, num_nodes, nm, nsets, out);
    FMMH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ MLP
// Parameter layout (floats): W1[64][4], b1[64], W2[64][64], b2[64], W3[64], b3.
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + SEL_H * SEL_IN;
constexpr int OFF_W2 = OFF_B1 + SEL_H;
constexpr int OFF_B2 = OFF_W2 + SEL_H * SEL_H;
constexpr int OFF_W3 = OFF_B2 + SEL_H;
constexpr int OFF_B3 = OFF_W3 + SEL_H;
static_assert(OFF_B3 + 1 == SEL_NPARAM, "parameter layout");

// Forward pass of one sample, weights in LDS; h1 lives in registers and the
// second layer is folded into the output sum, so only one unit of layer 2
// is alive at a time.
__device__ __forceinline__ float mlp_forward(const float* __restrict__ W, const float (&xn)[SEL_IN]) {
    float h1[SEL_H];
#pragma unroll
    for (int u = 0; u < SEL_H; ++u) {
        float z = W[OFF_B1 + u];
#pragma unroll
        for (int f = 0; f < SEL_IN; ++f) z = __builtin_fmaf(W[OFF_W1 + u * SEL_IN + f], xn[f], z);
        h1[u] = fmaxf(z, 0.0f);
    }
    float y = W[OFF_B3];
    for (int u = 0; u < SEL_H; ++u) {
        float z = W[OFF_B2 + u];
#pragma unroll
        for (int k2 = 0; k2 < SEL_H; ++k2) z = __builtin_fmaf(W[OFF_W2 + u * SEL_H + k2], h1[k2], z);
        y = __builtin_fmaf(W[OFF_W3 + u], fmaxf(z, 0.0f), y);
    }
    return y;
}

// Sample index of minibatch slot `l` at step t: a fixed permutation
// (a k + c) mod ns of the running position k = t * batch + l (a coprime to
// ns), so an epoch visits every sample once when batch divides ns.
__device__ __forceinline__ int mlp_sample(const MlpArgs& a, int t, int l) {
    if (!a.use_perm) return l;
    const unsigned long long k = ((unsigned long long)t * a.batch + l) % (unsigned long long)a.ns;
    return (int)((a.perm_a * k + a.perm_c) % (unsigned long long)a.ns);
}

// One workgroup (256 lanes) per 64 samples of the minibatch.  Activations of
// the 64 samples are matrices in LDS; lane (s, g) = (tid / 4, tid % 4)
// computes units 16 g .. 16 g + 15 of sample s.  Loss (per sample)
// w (y^ - y)^2 with w = asym where y^ < y (under-predicted orders cost
// accuracy, over-predicted ones only time), averaged over the minibatch.
// Each workgroup writes its parameter-gradient partial sums (and loss) to
// part[blockIdx.x]; k_mlp_reduce sums them in a fixed order.
constexpr int MLP_LD = SEL_H + 1;
__global__ __launch_bounds__(256) void k_mlp_grad(MlpArgs a) {
    __shared__ float W[SEL_NPARAM];
    __shared__ float Xn[64][SEL_IN];
    __shared__ float H1[64][MLP_LD];
    __shared__ float H2[64][MLP_LD];
    __shared__ float D2[64][MLP_LD];
    __shared__ float D1[64][MLP_LD];
    __shared__ float G[64];
    __shared__ float lsum[64];
    const int tid = threadIdx.x;
    const int t = a.step ? *a.step : 0;
    for (int x = tid; x < SEL_NPARAM; x += 256) W[x] = a.W[x];
    if (tid < 64) {
        const int l = blockIdx.x * 64 + tid;
        const bool valid = l < a.batch && (a.use_perm || l < a.ns);
        const int idx = valid ? mlp_sample(a, t, l) : 0;
#pragma unroll
        for (int f = 0; f < SEL_IN; ++f)
            Xn[tid][f] = valid ? (a.X[(int64_t)idx * SEL_IN + f] - a.norm[f]) * a.norm[SEL_IN + f] : 0.0f;
        G[tid] = valid ? a.Y[idx] : 0.0f;   // target for now
        lsum[tid] = valid ? 1.0f : 0.0f;    // validity for now
    }
    __syncthreads();
    const int s = tid >> 2, g = tid & 3, u0 = 16 * g;
    // layer 1
#pragma unroll
    for (int uu = 0; uu < 16; ++uu) {
        const int u = u0 + uu;
        float z = W[OFF_B1 + u];
#pragma unroll
        for (int f = 0; f < SEL_IN; ++f) z = __builtin_fmaf(W[OFF_W1 + u * SEL_IN + f], Xn[s][f], z);
        H1[s][u] = fmaxf(z, 0.0f);
    }
    __syncthreads();
    // layer 2 and the output
    float part = 0.0f;
#pragma unroll 4
    for (int uu = 0; uu < 16; ++uu) {
        const int u = u0 + uu;
        float z = W[OFF_B2 + u];
        for (int k2 = 0; k2 < SEL_H; ++k2) z = __builtin_fmaf(W[OFF_W2 + u * SEL_H + k2], H1[s][k2], z);
        const float h = fmaxf(z, 0.0f);
        H2[s][u] = h;
        part = __builtin_fmaf(W[OFF_W3 + u], h, part);
    }
    part += __shfl_xor(part, 1);
    part += __shfl_xor(part, 2);
    const float yhat = part + W[OFF_B3];
    const float y = G[s];
    const float valid = lsum[s];
    __syncthreads();
    const float diff = yhat - y;
    const float w = diff < 0.0f ? a.asym : 1.0f;
    const float gs = valid * 2.0f * w * diff / (float)a.batch;
    if (g == 0) {
        G[s] = gs;
        lsum[s] = valid * w * diff * diff;
    }
    // dL/dz2 = g W3 relu'(z2)
#pragma unroll
    for (int uu = 0; uu < 16; ++uu) {
        const int u = u0 + uu;
        D2[s][u] = H2[s][u] > 0.0f ? gs * W[OFF_W3 + u] : 0.0f;
    }
    __syncthreads();
    // dL/dz1 = (W2^T dz2) relu'(z1)
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
        const int k2 = u0 + kk;
        float v = 0.0f;
        for (int u = 0; u < SEL_H; ++u) v = __builtin_fmaf(W[OFF_W2 + u * SEL_H + k2], D2[s][u], v);
        D1[s][k2] = H1[s][k2] > 0.0f ? v : 0.0f;
    }
    __syncthreads();
    // parameter gradients: sums over the 64 samples of this workgroup
    float* out = a.part + (int64_t)blockIdx.x * (SEL_NPARAM + 1);
    for (int x = tid; x < SEL_H * SEL_H; x += 256) {      // W2[u][k2]
        const int u = x / SEL_H, k2 = x - u * SEL_H;
        float v = 0.0f;
        for (int q = 0; q < 64; ++q) v = __builtin_fmaf(D2[q][u], H1[q][k2], v);
        out[OFF_W2 + x] = v;
    }
    if (tid < SEL_H * SEL_IN) {                            // W1[u][f]
        const int u = tid / SEL_IN, f = tid - u * SEL_IN;
        float v = 0.0f;
        for (int q = 0; q < 64; ++q) v = __builtin_fmaf(D1[q][u], Xn[q][f], v);
        out[OFF_W1 + tid] = v;
    }
    if (tid < SEL_H) {
        float b1 = 0.0f, b2 = 0.0f, w3 = 0.0f;
        for (int q = 0; q < 64; ++q) {
            b1 += D1[q][tid];
            b2 += D2[q][tid];
            w3 = __builtin_fmaf(G[q], H2[q][tid], w3);
        }
        out[OFF_B1 + tid] = b1;
        out[OFF_B2 + tid] = b2;
        out[OFF_W3 + tid] = w3;
    }
    if (tid == 0) {
        float b3 = 0.0f, ls = 0.0f;
        for (int q = 0; q < 64; ++q) { b3 += G[q]; ls += lsum[q]; }
        out[OFF_B3] = b3;
        out[SEL_NPARAM] = ls / (float)a.batch;
    }
}

// grad[x] = sum over workgroups of part[b][x], in workgroup order.
__global__ __launch_bounds__(256) void k_mlp_reduce(MlpArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x > SEL_NPARAM) return;
    float v = 0.0f;
    for (int b = 0; b < a.nblk; ++b) v += a.part[(int64_t)b * (SEL_NPARAM + 1) + x];
    a.grad[x] = v;
}

// Adam (Kingma & Ba 2015) with bias correction; step counter t on the device.
__global__ __launch_bounds__(256) void k_mlp_adam(MlpArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= SEL_NPARAM) return;
    const int t = *a.step + 1;
    const float gr = a.grad[x];
    const float m = a.beta1 * a.adam_m[x] + (1.0f - a.beta1) * gr;
    const float v = a.beta2 * a.adam_v[x] + (1.0f - a.beta2) * gr * gr;
    a.adam_m[x] = m;
    a.adam_v[x] = v;
    const float mh = m / (1.0f - powf(a.beta1, (float)t));
    const float vh = v / (1.0f - powf(a.beta2, (float)t));
    // step-size schedule: lr halves every lr_half steps
    const float lr = a.lr * exp2f(-(float)(t - 1) / (float)a.lr_half);
    a.W[x] -= lr * mh / (sqrtf(vh) + a.eps);
}

__global__ void k_mlp_step(MlpArgs a) {
    a.loss_log[*a.step] = a.grad[SEL_NPARAM];
    *a.step += 1;
}

void launch_mlp_grad(const MlpArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(a.nblk), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((SEL_NPARAM + 1 + 255) / 256), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

void launch_mlp_train_step(const MlpArgs& a, hipStream_t s) {
    launch_mlp_grad(a, s);
    hipLaunchKernelGGL(k_mlp_adam, dim3((SEL_NPARAM + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mlp_step, dim3(1), dim3(1), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ inference
__device__ __forceinline__ void load_weights(const float* __restrict__ src, float* W) {
    for (int x = threadIdx.x; x < SEL_NPARAM; x += blockDim.x) W[x] = src[x];
}

// Predictions y^ for given raw features (checks of the inference path).
__global__ __launch_bounds__(256) void k_sel_predict(const float* __restrict__ Wg, const float* __restrict__ norm,
                                                     const float* __restrict__ X, int ns, float* __restrict__ Y) {
    __shared__ float W[SEL_NPARAM];
    load_weights(Wg, W);
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    float xn[SEL_IN];
#pragma unroll
    for (int f = 0; f < SEL_IN; ++f) xn[f] = (X[(int64_t)k * SEL_IN + f] - norm[f]) * norm[SEL_IN + f];
    Y[k] = mlp_forward(W, xn);
}

void launch_sel_predict(const float* W, const float* norm, const float* X, int ns, float* Y, hipStream_t s) {
    if (ns <= 0) return;
    hipLaunchKernelGGL(k_sel_predict, dim3((ns + 255) / 256), dim3(256), 0, s, W, norm, X, ns, Y);
    FMMH_HIP(hipGetLastError());
}

// Per-pair order (CSR position e): features of (i, j), forward pass, then
// q = ceil(y^ + margin) in [0, p]; full order p where x2 - x3 lies above the
// training range or the prediction exceeds the trained cap (fall back to
// expansion_order).  Also the order histogram.
__global__ __launch_bounds__(256) void k_sel_orders(SelOrderArgs a) {
    __shared__ float W[SEL_NPARAM];
    __shared__ int hist[PMAX + 1];
    load_weights(a.W, W);
    if (threadIdx.x <= PMAX) hist[threadIdx.x] = 0;
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < a.n_pairs) {
        const unsigned long long key = a.pair_keys[e];
        const int i = (int)(key >> 32);
        const int j = src_node((int)(key & 0xffffffffull));
        const double Rx = a.cx[i] - a.cx[j], Ry = a.cy[i] - a.cy[j], Rz = a.cz[i] - a.cz[j];
        const double R = sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
        const double tau = *a.tau;
        float x[SEL_IN];
        x[0] = (float)(a.r[i] / R);
        x[1] = (float)(a.r[j] / R);
        x[2] = (float)log10(a.node_norm[j] / (R * R));
        x[3] = (float)log10(tau);
        int q = a.p;
        const float rel = x[2] - x[3];
        // above the training range (a tolerance tighter than any the network saw): expansion_order;
        // below it the labels saturate at order 0, so the network's prediction is kept
        if (isfinite(rel) && rel <= a.rel_hi && tau > 0.0) {
            float xn[SEL_IN];
#pragma unroll
            for (int f = 0; f < SEL_IN; ++f) xn[f] = (x[f] - a.norm[f]) * a.norm[SEL_IN + f];
            const float y = mlp_forward(W, xn);
            if (isfinite(y) && y <= (float)a.cap_train) q = min(a.p, max(0, (int)ceilf(y + a.margin)));
        }
        if (a.force_p) q = a.p;
        a.order[e] = (unsigned char)q;
        atomicAdd(&hist[q], 1);
    }
    __syncthreads();
    if (threadIdx.x <= PMAX && hist[threadIdx.x]) atomicAdd(&a.hist[threadIdx.x], (unsigned long long)hist[threadIdx.x]);
}

void launch_sel_orders(const SelOrderArgs& a, hipStream_t s) {
    if (a.n_pairs <= 0) return;
    hipLaunchKernelGGL(k_sel_orders, dim3((a.n_pairs + 255) / 256), dim3(256), 0, s, a);
    FMMH_HIP(hipGetLastError());
}

// Sort keys of the class-grouped M2L with per-pair orders: class-sorted
// position k -> (class << 4 | order), value = CSR position.
__global__ __launch_bounds__(256) void k_sel_group_keys(const int* __restrict__ pair_csr,
                                                        const int* __restrict__ pair_class,
                                                        const unsigned char* __restrict__ order, int n,
                                                        unsigned long long* __restrict__ keys,
                                                        int* __restrict__ vals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int e = pair_csr[k];
    keys[k] = ((unsigned long long)pair_class[e] << 4) | order[e];
    vals[k] = e;
}

__global__ __launch_bounds__(256) void k_sel_group_flags(const unsigned long long* __restrict__ keys, int n,
                                                         int* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) flag[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

// group id g (exclusive scan of the flags at a start) -> class and order
__global__ __launch_bounds__(256) void k_sel_group_info(const unsigned long long* __restrict__ keys,
                                                        const int* __restrict__ flag, const int* __restrict__ gid,
                                                        int n, int* __restrict__ grp_class, int* __restrict__ grp_order,
                                                        int* __restrict__ ngroups) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (flag[k]) {
        grp_class[gid[k]] = (int)(keys[k] >> 4);
        grp_order[gid[k]] = (int)(keys[k] & 15ull);
    }
    if (k == n - 1) *ngroups = gid[k] + flag[k];
}

void launch_sel_groups(const int* pair_csr, const int* pair_class, const unsigned char* order, int n,
                       unsigned long long* keys, int* vals, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sel_group_keys, dim3((n + 255) / 256), dim3(256), 0, s, pair_csr, pair_class, order, n, keys,
                       vals);
    FMMH_HIP(hipGetLastError());
}

void launch_sel_group_flags(const unsigned long long* keys, int n, int* flag, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sel_group_flags, dim3((n + 255) / 256), dim3(256), 0, s, keys, n, flag);
    FMMH_HIP(hipGetLastError());
}

void launch_sel_group_info(const unsigned long long* keys, const int* flag, const int* gid, int n, int* grp_class,
                           int* grp_order, int* ngroups, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sel_group_info, dim3((n + 255) / 256), dim3(256), 0, s, keys, flag, gid, n, grp_class,
                       grp_order, ngroups);
    FMMH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ per-call tolerance
// Sum over sites of |near field|^2 (field block, set 0) into acc[0], and
// sum over leaves of count * (M2L list sizes along the path to the root)
// into acc[1]: the per-pair tolerance is
//   tau = eps * kappa * sqrt(acc[0] / n) / sqrt(acc[1] / n),
// the relative RMS budget spread over the pairs a site sees.
// grid.y = set: the near-field sum of squares of each set (field block; the potential for field order 0);
// row 0 also sums the M2L list sizes along each site's path.
__global__ __launch_bounds__(256) void k_sel_budget(const double* __restrict__ near, int64_t ldn, int n, int nf,
                                                    const int* __restrict__ leaf_node, const int* __restrict__ count,
                                                    const int* __restrict__ parent, const int* __restrict__ m2l_off,
                                                    int num_leaves, double* __restrict__ acc) {
    __shared__ double red[2][256];
    const int set = blockIdx.y;
    const double* __restrict__ nr = near + (int64_t)set * nf * ldn;
    double f2 = 0.0, paths = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (nf >= 4) {
            for (int c = 1; c < 4; ++c) { const double v = nr[(int64_t)c * ldn + k]; f2 += v * v; }
        } else {
            const double v = nr[k];
            f2 += v * v;
        }
    }
    if (set == 0)
        for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < num_leaves; l += gridDim.x * blockDim.x) {
            int node = leaf_node[l];
            long long s = 0;
            while (node >= 0) { s += m2l_off[node + 1] - m2l_off[node]; node = parent[node]; }
            paths += (double)count[leaf_node[l]] * (double)s;
        }
    red[0][threadIdx.x] = f2;
    red[1][threadIdx.x] = paths;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&acc[4 + set], red[0][0]);
        if (set == 0) atomicAdd(&acc[1], red[1][0]);
    }
}

// tau from the smallest non-zero near-field RMS over the sets (a set without near field has no far
// field to keep either)
__global__ void k_sel_tau(double* __restrict__ acc, int n, int nsets, double eps_kappa) {
    double f2 = 0.0;
    for (int s = 0; s < nsets; ++s) {
        const double v = acc[4 + s];
        if (v > 0.0 && (f2 == 0.0 || v < f2)) f2 = v;
    }
    acc[0] = f2;
    const double frms = sqrt(f2 / n);
    const double neff = fmax(acc[1] / n, 1.0);
    acc[2] = eps_kappa * frms / sqrt(neff);
}

void launch_sel_budget(const double* near, int64_t ldn, int n, int nf, int nsets, const int* leaf_node,
                       const int* count, const int* parent, const int* m2l_off, int num_leaves, double* acc,
                       double eps_kappa, hipStream_t s) {
    hipLaunchKernelGGL(k_sel_budget, dim3(64, nsets), dim3(256), 0, s, near, ldn, n, nf, leaf_node, count, parent,
                       m2l_off, num_leaves, acc);
    hipLaunchKernelGGL(k_sel_tau, dim3(1), dim3(1), 0, s, acc, n, nsets, eps_kappa);
    FMMH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ spot checks
// Far-field part of the direct sum at spot sites, with the reference's far
// semantics: every source whose leaf is not in the P2P list of the site's
// leaf, undamped and without exclusions.  The spot sites are evenly spaced in
// the sorted (octree) site order, spot k at sorted site (2k + 1) n / (2 nspots):
// a systematic sample over all sites, spread over the box, so that the
// relative RMS over the spots estimates the one over all sites.
// The sum over the sources is a GEMM on the FP64 matrix cores: for field
// component c (potential and field) and source component a,
//   D_c[spot][set] += sum_j T_{a+c}(x_spot - x_j) m~_j(a)[set]
// (the contraction of contract<M, 1>).  A workgroup takes SP_W 16-row tiles
// of spot sites (one per wavefront), 16 NTL multipole sets and one chunk of
// the sources, staged in LDS SP_TJ<M> at a time; in each 4-source step lane
// (row, k) computes the tensor of pair (spot row, source k) and feeds it to
// the 4 ncomp(M) NTL MFMAs of the step.  The tensor is zero at R = 0 and for
// sources of the spot leaf's P2P list: the list's leaves are in site order, so
// each lane keeps a cursor on the first list range not yet passed.  Partial
// sums per chunk, out[chunk][set][spot][4], added in chunk order by the
// comparison (deterministic).
using v4d = double __attribute__((ext_vector_type(4)));
constexpr int SP_W = 4;   // wavefronts per workgroup

template <int M>
constexpr int sp_tj() { return M >= 2 ? 16 : 32; }   // sources per staged tile

int spot_tile(int m) { return m >= 2 ? sp_tj<2>() : sp_tj<1>(); }
int spot_row_blocks(int nspots) { return (nspots + 16 * SP_W - 1) / (16 * SP_W); }

template <int M, int NTL>
__global__ __launch_bounds__(64 * SP_W) void k_spot_far(SpotArgs a) {
    constexpr int NM = ncomp(M), NC = ncomp(M + 1);
    constexpr int TJ = sp_tj<M>();
    constexpr int BW = 16 * NTL, BWP = BW + 1;   // odd row stride: the staging stores hit distinct banks
    __shared__ double sx[TJ], sy[TJ], sz[TJ];
    __shared__ double sb[NM * TJ * BWP];        // [component][source][set]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int spot = (blockIdx.x * SP_W + w) * 16 + row;
    const bool valid = spot < a.nspots;
    const int site = (int)(((2ll * (valid ? spot : 0) + 1) * a.n) / (2ll * a.nspots));
    int leaf = 0;   // the leaf holding the site: the last one starting at or before it (leaves are in site order)
    {
        int hi = a.num_leaves;
        while (hi - leaf > 1) {
            const int mid = (leaf + hi) >> 1;
            if (a.start[a.leaf_node[mid]] <= site) leaf = mid;
            else hi = mid;
        }
    }
    const int node = a.leaf_node[leaf];
    const double x = a.xs[site], y = a.ys[site], z = a.zs[site];
    const int set0 = blockIdx.y * BW;
    const int j0 = blockIdx.z * a.chunk, j1 = min(a.n, j0 + a.chunk);
    // cursor: the first P2P-list entry whose site range ends after this lane's first source
    const int eend = a.p2p_off[node + 1];
    int e = a.p2p_off[node];
    {
        int hi = eend;
        const int jf = j0 + kq;
        while (e < hi) {
            const int mid = (e + hi) >> 1;
            const int sn = src_node(a.p2p_src[mid]);
            if (a.start[sn] + a.count[sn] <= jf) e = mid + 1;
            else hi = mid;
        }
    }
    int rlo = INT_MAX, rhi = INT_MAX;   // the entry's site range [rlo, rhi)
    auto load_range = [&]() {
        if (e < eend) {
            const int sn = src_node(a.p2p_src[e]);
            rlo = a.start[sn];
            rhi = rlo + a.count[sn];
        } else {
            rlo = rhi = INT_MAX;
        }
    };
    load_range();
    v4d acc[4][NTL];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) acc[c][nt] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int tb = j0; tb < j1; tb += TJ) {
        const int cn = min(TJ, j1 - tb);
        __syncthreads();   // readers of the previous tile are done
        for (int q = tid; q < TJ; q += 64 * SP_W) {
            const bool ok = q < cn;
            sx[q] = ok ? a.xs[tb + q] : 0.0;
            sy[q] = ok ? a.ys[tb + q] : 0.0;
            sz[q] = ok ? a.zs[tb + q] : 0.0;
        }
        for (int q = tid; q < NM * BW * TJ; q += 64 * SP_W) {
            const int jl = q % TJ, r = q / TJ, comp = r % NM, u = r / NM;   // source fastest: coalesced loads
            const int set = set0 + u;
            sb[(comp * TJ + jl) * BWP + u] =
                (jl < cn && set < a.nsets) ? a.ms[((int64_t)set * NM + comp) * a.ldm + tb + jl] : 0.0;
        }
        __syncthreads();
#pragma unroll 2
        for (int k4 = 0; k4 < TJ; k4 += 4) {
            const int jl = k4 + kq, j = tb + jl;
            while (j >= rhi) {
                ++e;
                load_range();
            }
            const bool use = valid && jl < cn && j < rlo;
            const double dx = x - sx[jl], dy = y - sy[jl], dz = z - sz[jl];
            const double R2 = dx * dx + dy * dy + dz * dz;
            double g[M + 2];
            radial_inv<M + 1, DAMP_NONE>(R2, (use && R2 > 0.0) ? rsqrt_pos(R2) : 0.0, 0.0, g);
            double T[NC];
            md_tensor<M + 1>(dx, dy, dz, g, T);
            double bv[NM][NTL];
#pragma unroll
            for (int comp = 0; comp < NM; ++comp)
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt) bv[comp][nt] = sb[(comp * TJ + jl) * BWP + 16 * nt + row];
#pragma unroll
            for (int comp = 0; comp < NM; ++comp) {
                const Exps es = idx2xyz(comp);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double av = T[xyz2idx(es.x + (c == 1), es.y + (c == 2), es.z + (c == 3))];
#pragma unroll
                    for (int nt = 0; nt < NTL; ++nt)
                        acc[c][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[comp][nt], acc[c][nt], 0, 0, 0);
                }
            }
        }
    }
    // accumulator register q of lane l holds D[spot row 4 q + l / 16][set l % 16]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int sp = (blockIdx.x * SP_W + w) * 16 + 4 * q + kq;
        if (sp >= a.nspots) continue;
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) {
            const int set = set0 + 16 * nt + row;
            if (set >= a.nsets) continue;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                a.out[(((int64_t)blockIdx.z * a.nsets + set) * a.nspots + sp) * 4 + c] = acc[c][nt][q];
        }
    }
    if (valid && blockIdx.y == 0 && blockIdx.z == 0 && kq == 0) a.site[spot] = site;
}

template <int M>
static void spot_far_m(const SpotArgs& a, hipStream_t s) {
    const int ntl = a.nsets > 16 ? 2 : 1;
    const dim3 g(spot_row_blocks(a.nspots), (a.nsets + 16 * ntl - 1) / (16 * ntl), a.nchunks);
    if (ntl == 2) hipLaunchKernelGGL((k_spot_far<M, 2>), g, dim3(64 * SP_W), 0, s, a);
    else hipLaunchKernelGGL((k_spot_far<M, 1>), g, dim3(64 * SP_W), 0, s, a);
}

void launch_spot_far(const SpotArgs& a, int m, hipStream_t s) {
    if (a.nspots <= 0) return;
    switch (m) {
        case 0: spot_far_m<0>(a, s); break;
        case 1: spot_far_m<1>(a, s); break;
        default: spot_far_m<2>(a, s); break;
    }
    FMMH_HIP(hipGetLastError());
}

}  // namespace fmmh
