// Device-wide sort and scan primitives (rocPRIM), isolated in one
// translation unit because the templates are slow to compile.  Used for the
// site-key sort of the tree build, the interaction-list sort and the
// node-numbering scans -- integer work off the arithmetic hot path.
#include <cstring>
#include <map>
#include <mutex>
#include <rocprim/rocprim.hpp>
#include "common.h"
#include "kernels.h"

namespace fmmh {

// Temporary-storage sizes depend on the element count only; the library's size queries cost tens of
// microseconds of host time each, which the GPU would spend idle in a tree-building call, so they are
// asked once per count (per process).
template <class F>
static size_t cached_temp(int which, int n, F query) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, size_t> sizes;
    std::lock_guard<std::mutex> lk(mu);
    auto it = sizes.find({which, n});
    if (it != sizes.end()) return it->second;
    const size_t b = query();
    sizes.emplace(std::make_pair(which, n), b);
    return b;
}

size_t sort_pairs_u64_temp(int n) {
    return cached_temp(0, n, [n] {
        size_t b = 0;
        FMMH_HIP(rocprim::radix_sort_pairs(nullptr, b, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr, (const int*)nullptr, (int*)nullptr,
                                           (unsigned)n, 0, 63));
        return b;
    });
}

void sort_pairs_u64(void* temp, size_t temp_bytes, const unsigned long long* kin, unsigned long long* kout,
                    const int* vin, int* vout, int n, int end_bit, hipStream_t s) {
    FMMH_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, kin, kout, vin, vout, (unsigned)n, 0, end_bit, s));
}

size_t sort_keys_u64_temp(int n) {
    return cached_temp(1, n, [n] {
        size_t b = 0;
        FMMH_HIP(rocprim::radix_sort_keys(nullptr, b, (const unsigned long long*)nullptr,
                                          (unsigned long long*)nullptr, (unsigned)n, 0, 64));
        return b;
    });
}

void sort_keys_u64(void* temp, size_t temp_bytes, const unsigned long long* kin, unsigned long long* kout, int n,
                   int end_bit, hipStream_t s) {
    FMMH_HIP(rocprim::radix_sort_keys(temp, temp_bytes, kin, kout, (unsigned)n, 0, end_bit, s));
}

size_t exscan_i32_temp(int n) {
    return cached_temp(2, n, [n] {
        size_t b = 0;
        FMMH_HIP(rocprim::exclusive_scan(nullptr, b, (const int*)nullptr, (int*)nullptr, 0, (size_t)n,
                                         rocprim::plus<int>()));
        return b;
    });
}

void exscan_i32(void* temp, size_t temp_bytes, const int* in, int* out, int n, hipStream_t s) {
    FMMH_HIP(rocprim::exclusive_scan(temp, temp_bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), s));
}

}  // namespace fmmh
