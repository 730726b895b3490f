// Project-owned device primitives: exclusive scan, stable LSD radix sort,
// wave-level reductions. No CUB/rocPRIM; wave64 ballots and LDS.
#pragma once
#include "ctx.hpp"

namespace pfg {

// exclusive prefix sum; out may alias in only for the u64 variant.
// d_total (nullable, device) receives the sum of all elements.
void scan_u32_to_u64(Ctx &c, const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *d_total);
void scan_u64(Ctx &c, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *d_total);

// stable LSD radix sort on the low nbits of the keys (8-bit digits), with an
// optional uint32 payload. Sorts in place (temporaries from the context).
void radix_sort_u32_pairs(Ctx &c, uint32_t *keys, uint32_t *vals, uint64_t n, int nbits);
// u64 keys sorted on bits [lo, lo + nbits); returns the buffer that holds the
// result: keys or alt (n keys) when alt is given, else keys or the context's radix
// temporary (valid until the next sort)
uint64_t *radix_sort_u64_keys(Ctx &c, uint64_t *keys, uint64_t n, int nbits, int lo = 0, uint64_t *alt = nullptr);

// max of a u32 array (device), returned to host (synchronizes)
uint32_t reduce_max_u32(Ctx &c, const uint32_t *in, uint64_t n);

}  // namespace pfg
