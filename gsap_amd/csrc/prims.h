// Device-wide scan and radix sort (see prims.hip).
#pragma once
#include "common.h"

namespace gsap {

// words of temporary storage needed by scan_exclusive on n items
size_t scan_temp_words(int64_t n);
// out[i] = sum(in[0..i)), *total = sum(in) (total may be nullptr); in == out allowed
void scan_exclusive(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total, uint32_t *tmp,
                    hipStream_t st);

// words of temporary storage needed by radix_sort_pairs on n items
size_t radix_temp_words(int64_t n);
// Stable LSD sort of (keys, vals) on the low `bits` bits.  Buffers ping-pong: on
// return `keys`/`vals` point at the sorted data.  Returns the number of passes.
template <class K>
int radix_sort_pairs(K *&keys, uint32_t *&vals, K *&keys_alt, uint32_t *&vals_alt, int64_t n,
                     int bits, uint32_t *tmp, hipStream_t st);

}  // namespace gsap
