// Device exclusive scan and stable radix sort of (key, value) pairs
// (rm_sort.hip), used by the GPU tree build.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace rmh {

// Scratch bytes for exclusive_scan of n elements.
size_t scan_tmp_bytes(int n);
// out[i] = in[0] + ... + in[i - 1]; out may be in.
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int n, void* tmp, hipStream_t st);

// Scratch bytes for sort_pairs_63 of n pairs.
size_t sort_tmp_bytes(int n);
// Stable sort of n (key, value) pairs by key bits [0, 63) (bit 63 must be
// zero): keys_out / vals_out receive the sorted pairs; keys_in / vals_in are
// used as the other buffer of the passes and overwritten.
hipError_t sort_pairs_63(uint64_t* keys_in, uint32_t* vals_in, uint64_t* keys_out, uint32_t* vals_out, int n,
                         void* tmp, hipStream_t st);

}  // namespace rmh
