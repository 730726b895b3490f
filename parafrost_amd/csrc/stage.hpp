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

// words a variable's region needs; STAGE_SIZES adds one word per literal.
// Entries are taken four at a time so that four independent reads are in flight.
__device__ __forceinline__ uint32_t stage_need(const StageSrc &S, uint32_t v, int flags)
{
    const uint32_t per = (flags & STAGE_SIZES) ? 2u : 1u;
    const OL lp = S.ot[2 * v], ln = S.ot[2 * v + 1];
    uint32_t need = (flags & STAGE_TAB2) ? 2 * (lp.sz + ln.sz) : lp.sz + ln.sz;
    for (int s = 0; s < 2; s++) {
        const OL l = s ? ln : lp;
        const uint64_t *o = S.occ + l.off;
        uint32_t i = 0;
        for (; i + 4 <= l.sz; i += 4) {
            const uint64_t r0 = o[i], r1 = o[i + 1], r2 = o[i + 2], r3 = o[i + 3];
            const uint32_t z0 = S.mem[r0 + 2], z1 = S.mem[r1 + 2], z2 = S.mem[r2 + 2], z3 = S.mem[r3 + 2];
            need += 20 + per * (z0 + z1 + z2 + z3);
        }
        for (; i < l.sz; i++) need += 5 + per * S.mem[o[i] + 2];
    }
    return need;
}

__device__ __forceinline__ void stage_one(const StageSrc &S, uint32_t *arena, uint32_t &w, uint64_t ref, uint32_t sz,
                                          bool with_sizes)
{
    const uint32_t *g = S.mem + ref;
    arena[w] = (uint32_t)ref;
    arena[w + 1] = (uint32_t)(ref >> 32);
    arena[w + 2] = g[0];
    arena[w + 3] = g[1];
    arena[w + 4] = sz;
    uint32_t t = 0;
    for (; t + 4 <= sz; t += 4) {
        const uint32_t a = g[3 + t], b = g[4 + t], c = g[5 + t], d = g[6 + t];
        arena[w + 5 + t] = a;
        arena[w + 6 + t] = b;
        arena[w + 7 + t] = c;
        arena[w + 8 + t] = d;
    }
    for (; t < sz; t++) arena[w + 5 + t] = g[3 + t];
    if (with_sizes)
        for (t = 0; t < sz; t++) arena[w + 5 + sz + t] = S.ot[arena[w + 5 + t]].sz;
    w += 5 + (with_sizes ? 2 * sz : sz);
}

// copy the variable's clauses into [base, base + need); returns the table
__device__ __forceinline__ uint32_t *stage_copy(const StageSrc &S, uint32_t *arena, uint32_t base, uint32_t v,
                                                int flags)
{
    const bool with_sizes = flags & STAGE_SIZES, tab2 = flags & STAGE_TAB2;
    uint32_t *tab = arena + base;
    const OL lp = S.ot[2 * v], ln = S.ot[2 * v + 1];
    const uint32_t ne = lp.sz + ln.sz;
    uint32_t w = base + (tab2 ? 2 * ne : ne), k = 0;
    for (int s = 0; s < 2; s++) {
        const OL l = s ? ln : lp;
        const uint64_t *o = S.occ + l.off;
        uint32_t i = 0;
        for (; i + 4 <= l.sz; i += 4) {
            const uint64_t r0 = o[i], r1 = o[i + 1], r2 = o[i + 2], r3 = o[i + 3];
            const uint32_t z0 = S.mem[r0 + 2], z1 = S.mem[r1 + 2], z2 = S.mem[r2 + 2], z3 = S.mem[r3 + 2];
            const uint64_t rr[4] = {r0, r1, r2, r3};
            const uint32_t zz[4] = {z0, z1, z2, z3};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (tab2) tab[ne + k] = w + 2;
                tab[k++] = w + 2;
                stage_one(S, arena, w, rr[u], zz[u], with_sizes);
            }
        }
        for (; i < l.sz; i++) {
            const uint64_t r = o[i];
            if (tab2) tab[ne + k] = w + 2;
            tab[k++] = w + 2;
            stage_one(S, arena, w, r, S.mem[r + 2], with_sizes);
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
