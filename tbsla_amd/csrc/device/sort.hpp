// Device sort and scan primitives (convert.hip), for HIP sources.
#pragma once

#include <hip/hip_runtime.h>

namespace tbsla {
namespace hip {

// In-place exclusive scan of the ints a[0, m) on stream s.
void exclusive_scan(int *a, long m, hipStream_t s);

// Stable LSD radix sort of the pairs (key, val)[0, n) by key bits [0, bits)
// on stream s (8-bit digits); the result is in (key, val) again, k2 and v2
// are ping-pong buffers of n entries each.
void radix_sort_pairs(unsigned long long *key, int *val, unsigned long long *k2, int *v2, long n, int bits,
                      hipStream_t s);

}  // namespace hip
}  // namespace tbsla
