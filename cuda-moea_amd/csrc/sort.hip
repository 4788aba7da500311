// Stable radix sort of (u64 key, i32 value) pairs over key bits [0, end_bit)
// (all 64 by default), the
// ordering of CUB DeviceRadixSort::SortPairs that the reference's niching and
// RVEA elitism rely on (core/cuda/cuda_ref_segment_ops.cu:51-83).  Built on
// rocPRIM's device radix sort through hipCUB.
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "moea_internal.hpp"

namespace moea {

size_t sort_pairs_temp_bytes(int n, int end_bit)
{
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<const uint64_t*>(nullptr),
                                       static_cast<uint64_t*>(nullptr),
                                       static_cast<const int32_t*>(nullptr),
                                       static_cast<int32_t*>(nullptr), n, 0, end_bit, hipStream_t{});
    return bytes;
}

int sort_pairs(void* temp, size_t temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
               const int32_t* vals_in, int32_t* vals_out, int n, hipStream_t stream, int end_bit)
{
    if (n <= 0) return kOk;
    size_t bytes = temp_bytes;
    MOEA_HIP(hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys_in, keys_out, vals_in, vals_out, n,
                                                0, end_bit, stream));
    return kOk;
}

}  // namespace moea
