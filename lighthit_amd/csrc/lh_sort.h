// Stable key-value radix sort of the library (rocPRIM onesweep).
//
// Every sort of the library goes through lh::sort_pairs, which fixes one
// onesweep configuration: 9 radix bits per pass in workgroups of 1024 threads
// with 8 keys each and the match-based rank. On gfx950 it sorts 4-byte keys
// with 4-byte values 1.3-1.7x faster than the library default (8 bits per
// pass), e.g. 9.6M keys of 27 bits in 3 passes instead of 4. Arrays of at most
// 200,000 items take rocPRIM's merge sort, which is faster there; above that
// onesweep is (measured: 600k keys 68 against 130 us, 1M 75 against 156 us;
// rocPRIM's default switches at 2^20).
// The size query and the sort must use the same key type, item count and bit
// range, so both go through this one function.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

namespace lh {

using SortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>,
                                        rocprim::kernel_config<1024, 8>, 9,
                                        rocprim::block_radix_rank_algorithm::match>,
    200000>;

// Sorts (keys_in, values_in) by bits [begin_bit, end_bit) of the keys into
// (keys_out, values_out); with temp == nullptr only sets temp_bytes.
template <class Key, class Value>
inline hipError_t sort_pairs(void* temp, size_t& temp_bytes, const Key* keys_in, Key* keys_out,
                             const Value* values_in, Value* values_out, long long items,
                             int begin_bit, int end_bit, hipStream_t stream) {
    return rocprim::radix_sort_pairs<SortConfig>(temp, temp_bytes, keys_in, keys_out, values_in,
                                                 values_out, (size_t)items, (unsigned)begin_bit,
                                                 (unsigned)end_bit, stream);
}

}  // namespace lh
