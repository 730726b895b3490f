// Device-side clause model and small helpers shared by the simplifier kernels.
// Clause image (SCLAUSE, ParaFROST src/gpu/sclause.h:37-186):
//   word0 = status:2 | molten:1 | added:1 | usage:2 | lbd:26, word1 = signature,
//   word2 = size, words 3.. = literals (ascending).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "ctx.hpp"

namespace pfg {

constexpr uint32_t ST_ORIGINAL = 0u, ST_LEARNT = 1u, ST_DELETED = 2u;
constexpr uint8_t MELTED_M = 1u, FROZEN_M = 2u;
constexpr int HSE_MAX_CL_SIZE = 1000;  // subsume.h:25

__device__ __forceinline__ uint32_t absv(uint32_t l) { return l >> 1; }
__device__ __forceinline__ uint32_t flip(uint32_t l) { return l ^ 1u; }
__device__ __forceinline__ uint32_t maphash(uint32_t l) { return 1u << (l & 31u); }
// V2DEC(vorg[ABS(lit)], SIGN(lit)): a literal in the caller's original numbering
// (model.cpp:263-280, proof.cpp:123); vorg null = identity
__device__ __forceinline__ uint32_t org_lit(const uint32_t *vorg, uint32_t l)
{
    return vorg ? ((vorg[l >> 1] << 1) | (l & 1u)) : l;
}

__device__ __forceinline__ uint32_t c_st(const uint32_t *c) { return c[0] & 3u; }
__device__ __forceinline__ bool c_orig(const uint32_t *c) { return (c[0] & 3u) == 0; }
__device__ __forceinline__ bool c_del(const uint32_t *c) { return (c[0] & ST_DELETED) != 0; }
__device__ __forceinline__ bool c_learnt(const uint32_t *c) { return (c[0] & ST_LEARNT) != 0; }
__device__ __forceinline__ bool c_molten(const uint32_t *c) { return (c[0] >> 2) & 1u; }
__device__ __forceinline__ void c_melt(uint32_t *c) { c[0] |= 4u; }
__device__ __forceinline__ void c_freeze(uint32_t *c) { c[0] &= ~4u; }
__device__ __forceinline__ void c_markdel(uint32_t *c) { c[0] = (c[0] & ~3u) | ST_DELETED; }
__device__ __forceinline__ void c_setst(uint32_t *c, uint32_t st) { c[0] = (c[0] & ~3u) | st; }
__device__ __forceinline__ int c_size(const uint32_t *c) { return (int)c[2]; }

// a clause rewritten in place with its size kept: its bit (reference >> 2) in the
// occurrence-table update's rewrite map, which covers the pool of the last build
__device__ __forceinline__ void mark_rewritten(uint32_t *map, uint64_t words, uint64_t ref)
{
    const uint64_t b = ref >> 2;
    if (map && (b >> 5) < words) atomicOr(&map[b >> 5], 1u << (b & 31u));
}

// literal lit removed from the clause at ref in place: its key goes to the
// occurrence-table update's removal log (cnt: its count; past cap the count alone
// grows, and the next table is sorted in full)
__device__ __forceinline__ void log_removed(uint64_t *log, unsigned long long *cnt, uint64_t cap, uint32_t lit, uint64_t ref)
{
    if (!log) return;
    const unsigned long long k = atomicAdd(cnt, 1ull);
    if (k < cap) log[k] = ((uint64_t)lit << 32) | ref;
}

// signature filters (src/cpu/simplify.h:81-87)
__device__ __forceinline__ bool sig_sub(uint32_t A, uint32_t B) { return !(A & ~B); }
__device__ __forceinline__ bool sig_selfsub(uint32_t A, uint32_t B)
{
    const uint32_t t = B | ((B & 0xAAAAAAAAu) >> 1) | ((B & 0x55555555u) << 1);
    return !(A & ~t);
}

// grid-stride helpers
__device__ __forceinline__ uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round robin. For a
// grid-stride loop over items in memory order, xcd_block() renumbers the blocks so
// that every XCD takes a contiguous share of each stride step: XCD x gets logical
// blocks [x, x + 1) * grid / 8. With a grid of resident size the XCDs' windows move
// through the items together, so that data indexed by nearby items (one instance's
// variables in a batch) stays in the XCD's L2. The grid must be a multiple of 8
// (xcd_grid).
constexpr unsigned NUM_XCD = 8;
__device__ __forceinline__ uint64_t xcd_block()
{
    return (uint64_t)(blockIdx.x % NUM_XCD) * (gridDim.x / NUM_XCD) + blockIdx.x / NUM_XCD;
}

inline unsigned xcd_grid(uint64_t n, unsigned block, unsigned maxblocks)
{
    uint64_t g = (n + block - 1) / block;
    if (g > maxblocks) g = maxblocks;
    g = (g + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    return (unsigned)(g ? g : NUM_XCD);
}

inline unsigned grid_for(uint64_t n, unsigned block, unsigned maxblocks = 8192)
{
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > maxblocks) g = maxblocks;
    return (unsigned)g;
}

// resolvent length without writing it; 0 on tautology (src/cpu/simplify.h:168-193)
__device__ __forceinline__ int merge_len(uint32_t x, const uint32_t *c1, const uint32_t *c2)
{
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c1 + 3, *b = c2 + 3;
    int it1 = 0, it2 = 0, len = n1 + n2 - 2;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = l1 >> 1, v2 = l2 >> 1;
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if ((l1 ^ l2) == 1u) return 0;
        else if (v1 < v2) it1++;
        else if (v2 < v1) it2++;
        else { it1++; it2++; len--; }
    }
    return len;
}

// resolvent written to out; false on tautology (src/cpu/simplify.h:131-166)
__device__ __forceinline__ bool merge_out(uint32_t x, const uint32_t *c1, const uint32_t *c2, uint32_t *out, int &n)
{
    n = 0;
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c1 + 3, *b = c2 + 3;
    int it1 = 0, it2 = 0;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = l1 >> 1, v2 = l2 >> 1;
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if ((l1 ^ l2) == 1u) return false;
        else if (v1 < v2) { it1++; out[n++] = l1; }
        else if (v2 < v1) { it2++; out[n++] = l2; }
        else { it1++; it2++; out[n++] = l1; }
    }
    while (it1 < n1) { const uint32_t l1 = a[it1++]; if ((l1 >> 1) != x) out[n++] = l1; }
    while (it2 < n2) { const uint32_t l2 = b[it2++]; if ((l2 >> 1) != x) out[n++] = l2; }
    return true;
}

__device__ __forceinline__ void sort_u32_small(uint32_t *d, int n)
{
    for (int i = 1; i < n; i++) {
        uint32_t t = d[i];
        int j = i;
        while (j > 0 && d[j - 1] > t) { d[j] = d[j - 1]; j--; }
        d[j] = t;
    }
}


__device__ __forceinline__ void report_conflict(unsigned long long *cnt, uint32_t var)
{
    atomicOr(&cnt[CNT_FLAGS], (unsigned long long)FLAG_UNSAT);
    const unsigned long long p = atomicAdd(&cnt[CNT_CONF], 1ull);
    if (p < CONF_CAP) cnt[CNT_CONF + 1 + p] = var;
}

// instance owning global variable v: the last i with ibase[i] < v (ibase has ninst + 1 entries)
__device__ __forceinline__ uint32_t find_inst(const uint32_t *ibase, uint32_t ninst, uint32_t v)
{
    uint32_t lo = 0, hi = ninst;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ibase[mid] < v) lo = mid;
        else hi = mid;
    }
    return lo;
}

// returns 1 if lit became true, 0 if it was true already, -1 if it was false
__device__ __forceinline__ int try_assign(int8_t *value, uint32_t lit)
{
    uint32_t *w = (uint32_t *)(value + (lit & ~3u));
    const int sl = (int)(lit & 3u) * 8, sn = (int)((lit ^ 1u) & 3u) * 8;
    uint32_t old = __atomic_load_n(w, __ATOMIC_RELAXED), assumed;
    do {
        const int8_t cur = (int8_t)((old >> sl) & 0xFFu);
        if (cur >= 0) return cur ? 0 : -1;
        uint32_t nw = old & ~(0xFFu << sl) & ~(0xFFu << sn);
        nw |= (1u << sl);
        assumed = old;
        old = atomicCAS(w, assumed, nw);
    } while (old != assumed);
    return 1;
}

__device__ __forceinline__ void enqueue_dev(int8_t *value, uint8_t *vstate, uint32_t *trail, unsigned long long *cnt, uint32_t lit)
{
    const int r = try_assign(value, lit);
    if (r == 1) {
        vstate[lit >> 1] = FROZEN_M;   // learnUnit -> markFrozen (src/cpu/solve.h:475-479)
        const unsigned long long pos = atomicAdd(&cnt[CNT_NTRAIL], 1ull);
        trail[pos] = lit;
    }
    else if (r < 0) report_conflict(cnt, lit >> 1);
}

// wave-aggregated append of flagged lanes (one atomic per wave and list kind)
__device__ __forceinline__ void wave_append(bool flag, uint32_t value, uint32_t *out, unsigned long long *counter)
{
    const uint64_t m = __ballot(flag);
    if (!m) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (flag) out[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// Block-wide reservation of output slots: every thread asks for `want` slots and
// gets the offset of its own run; one atomic per block and call. Every thread of
// the block must call it (blockDim.x a multiple of 64, at most 1024).
__device__ __forceinline__ uint64_t block_reserve(uint64_t want, unsigned long long *counter)
{
    __shared__ unsigned long long s_wave[16];
    __shared__ unsigned long long s_base;
    const int lane = __lane_id(), w = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    uint64_t incl = want;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int q = 0; q < nwaves; q++) {
            const unsigned long long t = s_wave[q];
            s_wave[q] = run;
            run += t;
        }
        s_base = run ? atomicAdd(counter, run) : 0ull;
    }
    __syncthreads();
    const uint64_t off = s_base + s_wave[w] + incl - want;
    __syncthreads();   // s_wave / s_base may be reused by the next call
    return off;
}

// block-wide sum; the result is valid in thread 0
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v)
{
    __shared__ unsigned long long s_part[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (__lane_id() == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += s_part[q];
    __syncthreads();
    return t;
}

__device__ __forceinline__ void block_append(bool flag, uint32_t value, uint32_t *out, unsigned long long *counter)
{
    const uint64_t o = block_reserve(flag ? 1ull : 0ull, counter);
    if (flag) out[o] = value;
}

}  // namespace pfg
