// rocPRIM radix sorts and scans used by the lookup build (fiber order and
// row order) and the sample deduplication.  Kept in their
// own translation unit: the templates are slow to compile.
//
// Each wrapper returns the temporary-storage size it needs; called with
// temp == nullptr it only queries that size.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"

namespace exafac {
namespace k {

size_t sort_pairs_u64_u32(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int bits) {
    size_t need = 0;
    EX_HIP(rocprim::radix_sort_pairs(nullptr, need, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
    if (temp) {
        EX_CHECK(temp_bytes >= need, 3, "radix sort scratch too small");
        EX_HIP(rocprim::radix_sort_pairs(temp, need, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
    }
    return need;
}

size_t sort_pairs_u32_u64(hipStream_t s, void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint64_t* vin, uint64_t* vout, int64_t n, int bits) {
    size_t need = 0;
    EX_HIP(rocprim::radix_sort_pairs(nullptr, need, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
    if (temp) {
        EX_CHECK(temp_bytes >= need, 3, "radix sort scratch too small");
        EX_HIP(rocprim::radix_sort_pairs(temp, need, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
    }
    return need;
}

size_t sort_pairs_u32_u32(hipStream_t s, void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int bits) {
    size_t need = 0;
    EX_HIP(rocprim::radix_sort_pairs(nullptr, need, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
    if (temp) {
        EX_CHECK(temp_bytes >= need, 3, "radix sort scratch too small");
        EX_HIP(rocprim::radix_sort_pairs(temp, need, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
    }
    return need;
}

size_t exclusive_scan_u32(hipStream_t s, void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out, int64_t n) {
    size_t need = 0;
    EX_HIP(rocprim::exclusive_scan(nullptr, need, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
    if (temp) {
        EX_CHECK(temp_bytes >= need, 3, "scan scratch too small");
        EX_HIP(rocprim::exclusive_scan(temp, need, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
    }
    return need;
}

size_t exclusive_scan_u64(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* in, uint64_t* out, int64_t n) {
    size_t need = 0;
    EX_HIP(rocprim::exclusive_scan(nullptr, need, in, out, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), s));
    if (temp) {
        EX_CHECK(temp_bytes >= need, 3, "scan scratch too small");
        EX_HIP(rocprim::exclusive_scan(temp, need, in, out, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), s));
    }
    return need;
}

}  // namespace k
}  // namespace exafac
