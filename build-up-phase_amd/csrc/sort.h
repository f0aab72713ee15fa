// sort.h — stable radix sort of (key, uint32 value) pairs on a HIP stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtx {

// Scratch bytes needed to sort n pairs with key_bytes-byte keys (4 or 8).
uint64_t sort_scratch_bytes(uint32_t n, int key_bytes);

// Sort keys[0..n) ascending, carrying vals; input values are ignored and replaced by the original indices
// (the sorted permutation).  Stable.  err_flags receives kErrSortTimeout if a look-back spin ran out.
hipError_t sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t n, void* scratch, uint32_t* err_flags,
                          hipStream_t stream);
hipError_t sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint32_t n, void* scratch, uint32_t* err_flags,
                          hipStream_t stream);

}  // namespace rtx
