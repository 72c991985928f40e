// Radix sort of (key, value) u32 pairs by the keys' low `bits` bits (rocPRIM's device radix sort,
// stable), used to build the reverse lists of a neighbours-of-neighbours pass: the pairs
// (target id, point) sorted by target id are the lists themselves, in point order within each.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "knng_kernels.hpp"

namespace knng {

size_t rsort_temp_bytes(size_t count, uint32_t bits) {
    size_t tb = 0;
    rocprim::double_buffer<uint32_t> dk(nullptr, nullptr), dv(nullptr, nullptr);
    if (rocprim::radix_sort_pairs(nullptr, tb, dk, dv, count, 0, bits) != hipSuccess) return 0;
    return tb;
}

hipError_t rsort_pairs(void* temp, size_t temp_bytes, uint32_t* keys_a, uint32_t* keys_b, uint32_t* vals_a,
                       uint32_t* vals_b, size_t count, uint32_t bits, hipStream_t st, bool* in_b) {
    rocprim::double_buffer<uint32_t> dk(keys_a, keys_b), dv(vals_a, vals_b);
    size_t tb = temp_bytes;
    const hipError_t e = rocprim::radix_sort_pairs(temp, tb, dk, dv, count, 0, bits, st);
    *in_b = dv.current() == vals_b;
    return e;
}

}  // namespace knng
