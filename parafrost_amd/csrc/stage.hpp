// LDS staging of an elected variable's occurrence lists for the per-variable
// kernels (BVE detection, ERE, SUB).
//
// A variable's clauses are read from the pool once, into a lane-private region of
// the block's LDS arena, and the per-variable logic then re-reads them from LDS.
// Region layout (words):
//   [table: one entry per list entry, positive list then negative list]
//   [(STAGE_TAB2) a second copy of the table]
//   per staged clause: [ref_lo, ref_hi, word0, sig, size, lits.. (, list size of each lit)]
// A table entry is the arena offset of the clause's word0, so a staged clause
// pointer has the pool layout (word0, sig, size, lits) and the pool reference two
// words before it.
//
// The lanes of a 64-lane block are staged in rounds: each round takes the pending
// lanes, in lane order, whose regions fit in the arena together; a variable larger
// than the whole arena is handed to the caller unstaged.
#pragma once
#include "device.hpp"

namespace pfg {

enum { STAGE_SIZES = 1, STAGE_TAB2 = 2 };   // stage_* flags

struct StageSrc {
    const uint32_t *mem;
    const OL *ot;
    const uint64_t *occ;
};

// words a variable's region needs; STAGE_SIZES adds one word per literal
__device__ __forceinline__ uint32_t stage_need(const StageSrc &S, uint32_t v, int flags)
{
    const bool with_sizes = flags & STAGE_SIZES;
    const OL lp = S.ot[2 * v], ln = S.ot[2 * v + 1];
    uint32_t need = (flags & STAGE_TAB2) ? 2 * (lp.sz + ln.sz) : lp.sz + ln.sz;
    for (int s = 0; s < 2; s++) {
        const OL l = s ? ln : lp;
        for (uint32_t i = 0; i < l.sz; i++) {
            const uint32_t sz = S.mem[S.occ[l.off + i] + 2];
            need += 5 + (with_sizes ? 2 * sz : sz);
        }
    }
    return need;
}

// copy the variable's clauses into [base, base + need); returns the table
__device__ __forceinline__ uint32_t *stage_copy(const StageSrc &S, uint32_t *arena, uint32_t base, uint32_t v,
                                                int flags)
{
    const bool with_sizes = flags & STAGE_SIZES;
    uint32_t *tab = arena + base;
    const OL lp = S.ot[2 * v], ln = S.ot[2 * v + 1];
    const uint32_t ne = lp.sz + ln.sz;
    uint32_t w = base + ((flags & STAGE_TAB2) ? 2 * ne : ne), k = 0;
    for (int s = 0; s < 2; s++) {
        const OL l = s ? ln : lp;
        for (uint32_t i = 0; i < l.sz; i++) {
            const uint64_t ref = S.occ[l.off + i];
            const uint32_t *g = S.mem + ref;
            const uint32_t sz = g[2];
            arena[w] = (uint32_t)ref;
            arena[w + 1] = (uint32_t)(ref >> 32);
            for (uint32_t t = 0; t < 3 + sz; t++) arena[w + 2 + t] = g[t];
            if (with_sizes)
                for (uint32_t t = 0; t < sz; t++) arena[w + 5 + sz + t] = S.ot[g[3 + t]].sz;
            if (flags & STAGE_TAB2) tab[ne + k] = w + 2;
            tab[k++] = w + 2;
            w += 5 + (with_sizes ? 2 * sz : sz);
        }
    }
    return tab;
}

// pool reference of a staged clause; bit 31 of ref_hi marks a modified copy
constexpr uint32_t STAGE_DIRTY = 0x80000000u;
__device__ __forceinline__ uint64_t staged_ref(const uint32_t *c)
{
    return (uint64_t)c[-2] | ((uint64_t)(c[-1] & ~STAGE_DIRTY) << 32);
}
__device__ __forceinline__ void staged_dirty(uint32_t *c) { c[-1] |= STAGE_DIRTY; }
__device__ __forceinline__ bool staged_is_dirty(const uint32_t *c) { return (c[-1] & STAGE_DIRTY) != 0; }

// Run body(tab) for the lane's variable with its region staged (tab != nullptr)
// or, for a variable larger than the arena, unstaged (tab == nullptr). Every lane
// of the 64-lane block calls it (active or not); the arena is free again on return.
template <typename Body>
__device__ __forceinline__ void stage_rounds(const StageSrc &S, uint32_t *arena, uint32_t arena_words, bool active,
                                             uint32_t v, int flags, Body body)
{
    const int lane = __lane_id();
    const uint32_t need = active ? stage_need(S, v, flags) : 0u;
    bool pending = active;
    while (__ballot(pending)) {
        uint32_t incl = pending ? need : 0u;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        const bool first = pending && incl == need;   // first pending lane of the round
        if (pending && incl <= arena_words) {
            body(stage_copy(S, arena, incl - need, v, flags));
            pending = false;
        }
        else if (first) {
            body((uint32_t *)nullptr);
            pending = false;
        }
        __syncthreads();
    }
}

}  // namespace pfg
