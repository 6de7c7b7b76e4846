// Stable radix sort of 32-bit (key, value) pairs (sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

// Bytes of work space radix_sort_pairs needs for n pairs.
size_t radix_sort_worksize(int64_t n);

// Sorts the pairs (keys[i], vals[i]), i < n, by the low `endbit` bits of
// the key, stably; keys2 and vals2 are the second buffers of the passes.
// Returns the buffer that holds the sorted keys (keys or keys2) and sets
// *vals_out to the one holding the values.  Launches on stream s.
uint32_t *radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys2, uint32_t *vals2,
                           int64_t n, int endbit, void *work, hipStream_t s, uint32_t **vals_out);
