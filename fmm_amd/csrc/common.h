// Shared definitions of the MI355X (gfx950) field engine.
//
// Packing follows peter-reinholdt/fmm: a Cartesian component (x,y,z) of
// total degree k sits at k(k+1)(k+2)/6 + (y^2 + 2yz + y + z^2 + 3z)/2
// (src/fmm.F90:80-89, 0-based here), i.e. q, x, y, z, xx, xy, xz, yy, yz, zz...
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "fmm_hip.h"

namespace fmmh {

constexpr int KEY_LEVELS = 21;   // 3-bit octant digits in a 63-bit key
constexpr int PMAX = 10;         // largest expansion order accepted
// Interaction-list sources carry a periodic image id in their top bits:
// src = image << IMG_SHIFT | node, image 0 = the central cell (the only one
// without periodic boundaries), 1..26 = the neighbour cells.
constexpr int IMG_SHIFT = 27;
constexpr int NODE_MASK = (1 << IMG_SHIFT) - 1;
constexpr int NUM_IMAGES = 27;
// (the encoded id uses all 32 bits: decode it unsigned)
__host__ __device__ __forceinline__ int src_node(int enc) { return enc & NODE_MASK; }
__host__ __device__ __forceinline__ int src_image(int enc) { return (int)((unsigned)enc >> IMG_SHIFT); }
__host__ __device__ __forceinline__ int src_encode(int img, int node) {
    return (int)(((unsigned)img << IMG_SHIFT) | (unsigned)node);
}
constexpr int NUM_XCD = 8;       // MI355X: workgroup g runs on XCD g mod 8, each XCD with its own L2
constexpr int KMAX_NEAR = 6;     // largest near-field tensor order (damped tensors stop at 6)

enum Damp { DAMP_NONE = 0, DAMP_THOLE = 1, DAMP_AMOEBA = 2, DAMP_ERF = 3 };

__host__ __device__ constexpr int ncomp(int k) { return (k + 1) * (k + 2) * (k + 3) / 6; }
// Harmonic (traceless) components of degree <= q: those with z <= 1, 2n + 1 of degree n.
__host__ __device__ constexpr int rdim(int q) { return (q + 1) * (q + 1); }
__host__ __device__ constexpr int ndeg(int k) { return (k + 1) * (k + 2) / 2; }
__host__ __device__ constexpr int xyz2idx(int x, int y, int z) {
    return (x + y + z) * (x + y + z + 1) * (x + y + z + 2) / 6 + (y * y + 2 * y * z + y + z * z + 3 * z) / 2;
}

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace fmmh

#define FMMH_HIP(x)                                                                                 \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw ::fmmh::Error(FMMH_ERR_HIP, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
    } while (0)
