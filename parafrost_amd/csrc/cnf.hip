// Clause database kernels: upload, histogram, occurrence table (build, reduce,
// per-list ordering), unit propagation, counting, garbage collection, PV
// filtering and the PV-list sort by the reference clause key.
//
// Order contract: clause references are kept in clause ("scnf") order and the
// pool is laid out in that order, so a reference's value orders clauses the way
// ParaFROST's scnf vector does (src/cpu/simplify.cpp:95-112, bve.cpp:214-218).
// Every occurrence list is left sorted by reference, which is the order in
// which createOT pushes (src/cpu/simplify.cpp:34-53).
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"
#include <vector>
#include <algorithm>

namespace pfg {

// ------------------------------------------------------------------ upload
// one thread per clause: header, sorted literals, signature (prep_cnf_k, kernels.cu:61-70)
__global__ void k_upload(const uint32_t *__restrict__ lits, const uint64_t *__restrict__ offs, uint64_t ncls,
                         const uint32_t *__restrict__ hdr0, uint32_t *__restrict__ mem, uint64_t *__restrict__ refs)
{
    for (uint64_t i = gtid(); i < ncls; i += gstride()) {
        const uint64_t b = offs[i], e = offs[i + 1];
        const int n = (int)(e - b);
        const uint64_t r = b + 3 * i;
        uint32_t *c = mem + r;
        uint32_t *out = c + 3;
        uint32_t sig = 0;
        if (n <= 16) {
            uint32_t t[16];
            for (int k = 0; k < n; k++) t[k] = lits[b + k];
            sort_u32_small(t, n);
            for (int k = 0; k < n; k++) { out[k] = t[k]; sig |= maphash(t[k]); }
        }
        else {
            for (int k = 0; k < n; k++) out[k] = lits[b + k];
            sort_u32_small(out, n);
            for (int k = 0; k < n; k++) sig |= maphash(out[k]);
        }
        c[0] = hdr0 ? hdr0[i] : ST_ORIGINAL;
        c[1] = sig;
        c[2] = (uint32_t)n;
        refs[i] = r;
    }
}

void upload(Ctx &c, const uint32_t *d_lits, const uint64_t *d_offs, uint64_t ncls, uint64_t nlits, const uint32_t *d_hdr0)
{
    const uint64_t words = nlits + 3 * ncls;
    ensure(c, c.mem, words + words / 2 + 4096);
    ensure(c, c.refs, ncls + ncls / 2 + 1024);
    {
        KTimer kt(c, K_UPLOAD, 8.0 * nlits + 28.0 * ncls);
        k_upload<<<grid_for(ncls, 256, 16384), 256, 0, c.stream>>>(d_lits, d_offs, ncls, d_hdr0, c.mem.p, c.refs.p);
    }
    c.msz = words;
    c.nrefs = ncls;
    c.nclauses = (int64_t)ncls;
    c.nliterals = (int64_t)nlits;
    c.ot_valid = false;
    c.otk_valid = false;
    c.hist_live = false;
}

// ------------------------------------------------------------------ histogram
// thread per clause over live (non-deleted) clauses (histSimp, histogram.cu:34-51)
// clause-streaming passes with per-literal atomics walk the clauses in XCD windows
// (xcd_block): in a batch, each XCD's window then touches one stretch of instances,
// whose counters and lists stay in its L2
__global__ void k_hist(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                       uint32_t *__restrict__ hist)
{
    for (uint64_t i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += gstride()) {
        const uint32_t *c = mem + refs[i];
        if (c_del(c)) continue;
        const int sz = c_size(c);
        for (int k = 0; k < sz; k++) atomicAdd(&hist[c[3 + k]], 1u);
    }
}

void hist(Ctx &c)
{
    ensure(c, c.hist, c.nlits2);
    PFG_HIP(hipMemsetAsync(c.hist.p, 0, c.nlits2 * 4, c.stream));
    if (c.nrefs) {
        KTimer kt(c, K_HIST, 20.0 * c.nrefs + 4.0 * c.nliterals + 4.0 * c.nlits2);
        k_hist<<<xcd_grid(c.nrefs, 256, c.tune.clause_grid), 256, 0, c.stream>>>(c.mem.p, c.refs.p, c.nrefs, c.hist.p);
    }
}

// ------------------------------------------------------------------ occurrence table
__global__ void k_ot_init(OL *__restrict__ ot, const uint64_t *__restrict__ segs, const uint32_t *__restrict__ hist, uint64_t n)
{
    for (uint64_t l = gtid(); l < n; l += gstride()) {
        OL o;
        o.off = segs[l];
        o.sz = 0;
        o.cap = hist[l];
        ot[l] = o;
    }
}

// createOT: every live clause pushes its reference to the lists of its literals
__global__ void k_ot_fill(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                          OL *__restrict__ ot, uint64_t *__restrict__ occ)
{
    for (uint64_t i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += gstride()) {
        const uint64_t r = refs[i];
        const uint32_t *c = mem + r;
        if (c_del(c)) continue;
        const int sz = c_size(c);
        for (int k = 0; k < sz; k++) {
            OL *o = ot + c[3 + k];
            const uint32_t p = atomicAdd(&o->sz, 1u);
            occ[o->off + p] = r;
        }
    }
}

__device__ __forceinline__ void cswap64(uint64_t &a, uint64_t &b)
{
    const uint64_t x = a < b ? a : b, y = a < b ? b : a;
    a = x;
    b = y;
}

// one thread per list: lists of up to 8 references are sorted in registers by a
// fixed network, up to 32 by insertion in place; longer lists are handed on
__global__ void __launch_bounds__(256) k_sort_small(const OL *__restrict__ ot, uint64_t n, uint64_t *__restrict__ occ,
                                                    uint32_t *__restrict__ mid_ids, uint32_t *__restrict__ big_ids,
                                                    unsigned long long *cnt)
{
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x; b < n; b += gstride()) {
        const uint64_t l = b + threadIdx.x;
        const uint32_t sz = l < n ? ot[l].sz : 0u;
        if (sz >= 2 && sz <= 8) {
            uint64_t *d = occ + ot[l].off;
            uint64_t t[8];
#pragma unroll
            for (int k = 0; k < 8; k++) t[k] = k < (int)sz ? d[k] : ~0ull;
#pragma unroll
            for (int i = 0; i < 8; i++) {
#pragma unroll
                for (int k = (i & 1); k + 1 < 8; k += 2) cswap64(t[k], t[k + 1]);
            }
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k < (int)sz) d[k] = t[k];
        }
        else if (sz > 8 && sz <= 32) {
            uint64_t *d = occ + ot[l].off;
            for (uint32_t i = 1; i < sz; i++) {
                const uint64_t v = d[i];
                uint32_t j = i;
                while (j > 0 && d[j - 1] > v) { d[j] = d[j - 1]; j--; }
                d[j] = v;
            }
        }
        wave_append(sz > 32 && sz <= 2048, (uint32_t)l, mid_ids, &cnt[CNT_B]);
        wave_append(sz > 2048, (uint32_t)l, big_ids, &cnt[CNT_C]);
    }
}

// block per list, bitonic sort of up to 2048 references in LDS
__global__ void __launch_bounds__(256) k_sort_mid(const uint32_t *__restrict__ ids, const OL *__restrict__ ot, uint64_t *__restrict__ occ)
{
    __shared__ uint64_t s[2048];
    const OL o = ot[ids[blockIdx.x]];
    uint64_t *d = occ + o.off;
    const int sz = (int)o.sz;
    int N = 64;
    while (N < sz) N <<= 1;
    for (int i = threadIdx.x; i < N; i += blockDim.x) s[i] = i < sz ? d[i] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= N; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < N; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const uint64_t a = s[i], b = s[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { s[i] = b; s[l] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < sz; i += blockDim.x) d[i] = s[i];
}

__global__ void k_big_gather(const uint32_t *__restrict__ ids, const uint64_t *__restrict__ base, uint32_t nbig,
                             const OL *__restrict__ ot, const uint64_t *__restrict__ occ, uint64_t *__restrict__ keys)
{
    for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const OL o = ot[ids[b]];
        for (uint32_t k = threadIdx.x; k < o.sz; k += blockDim.x)
            keys[base[b] + k] = ((uint64_t)b << 32) | occ[o.off + k];
    }
}

__global__ void k_big_scatter(const uint32_t *__restrict__ ids, const uint64_t *__restrict__ base, uint32_t nbig,
                              const OL *__restrict__ ot, uint64_t *__restrict__ occ, const uint64_t *__restrict__ keys)
{
    for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
        const OL o = ot[ids[b]];
        for (uint32_t k = threadIdx.x; k < o.sz; k += blockDim.x)
            occ[o.off + k] = keys[base[b] + k] & 0xFFFFFFFFull;
    }
}

static int bits_for(uint64_t v)
{
    int b = 0;
    while (b < 64 && (v >> b)) b++;
    return b;
}

// order every list by reference (= clause order), the createOT push order
static void order_lists_by_ref(Ctx &c)
{
    const uint64_t n = c.nlits2;
    ensure(c, c.tmpu32a, n);
    ensure(c, c.tmpu32b, n);
    ensure(c, c.eligible, n);
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_A, 0, 3 * 8, c.stream));
    k_sort_small<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.ot.p, n, c.occ.p, c.tmpu32b.p, c.eligible.p,
                                                                (unsigned long long *)c.cnt.p);
    read_counters(c, 4);
    const uint64_t nm = c.hcnt[CNT_B], nb = c.hcnt[CNT_C];
    if (nm) k_sort_mid<<<(unsigned)nm, 256, 0, c.stream>>>(c.tmpu32b.p, c.ot.p, c.occ.p);
    if (nb) {
        if (c.msz >= (1ull << 32)) throw Error(PFG_EINTERNAL, "clause pool beyond 2^32 words");
        std::vector<uint32_t> ids(nb);
        PFG_HIP(hipMemcpyAsync(ids.data(), c.eligible.p, nb * 4, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
        std::vector<OL> heads(nb);
        for (uint64_t b = 0; b < nb; b++)
            PFG_HIP(hipMemcpyAsync(&heads[b], c.ot.p + ids[b], sizeof(OL), hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
        std::vector<uint64_t> base(nb);
        uint64_t tot = 0;
        for (uint64_t b = 0; b < nb; b++) { base[b] = tot; tot += heads[b].sz; }
        ensure(c, c.tmpu64b, tot);
        ensure(c, c.tmpu64c, nb + 1);
        PFG_HIP(hipMemcpyAsync(c.tmpu64c.p, base.data(), nb * 8, hipMemcpyHostToDevice, c.stream));
        k_big_gather<<<(unsigned)std::min<uint64_t>(nb, 4096), 256, 0, c.stream>>>(c.eligible.p, c.tmpu64c.p, (uint32_t)nb,
                                                                                 c.ot.p, c.occ.p, c.tmpu64b.p);
        const uint64_t *sorted = radix_sort_u64_keys(c, c.tmpu64b.p, tot, 32 + bits_for(nb));
        k_big_scatter<<<(unsigned)std::min<uint64_t>(nb, 4096), 256, 0, c.stream>>>(c.eligible.p, c.tmpu64c.p, (uint32_t)nb,
                                                                                  c.ot.p, c.occ.p, sorted);
        PFG_HIP(hipStreamSynchronize(c.stream));  // host vectors go out of scope
    }
}

// ---- occurrence table by a stable radix sort of (literal, reference) pairs:
// pairs are laid out in clause order, so sorting them by literal leaves every list
// in clause order (the createOT push order) without atomics; list bounds and the
// live-literal histogram fall out of the sorted keys.
__global__ void k_live_sizes(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                             uint32_t *__restrict__ out)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *c = mem + refs[i];
        out[i] = c_del(c) ? 0u : (uint32_t)c_size(c);
    }
}

// One wave per 64 consecutive clauses: their keys form one contiguous run (positions
// are the scan of the live sizes), written in order with one key per lane and step,
// so stores and literal loads are contiguous across the wave; a lane finds its clause
// by a binary search over the wave's prefix of sizes (the last clause whose prefix is
// not past the key, which skips deleted and empty clauses).
// With szs given, clause i contributes its first szs[i] literals (0: none) instead of
// its live size.
__global__ void __launch_bounds__(256) k_ot_expand(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs,
                                                  uint64_t n, const uint32_t *__restrict__ szs,
                                                  const uint64_t *__restrict__ pos, uint64_t *__restrict__ keys)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t base = wave * 64; base < n; base += nwaves * 64) {
        const uint64_t i = base + lane;
        uint64_t r = 0;
        uint32_t sz = 0;
        if (i < n) {
            r = refs[i];
            if (szs) sz = szs[i];
            else {
                const uint32_t *c = mem + r;
                sz = c_del(c) ? 0u : (uint32_t)c_size(c);
            }
        }
        uint32_t incl = sz;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if ((int)lane >= o) incl += t;
        }
        const uint32_t pre = incl - sz, total = __shfl(incl, 63, 64);
        const uint64_t out0 = pos[base];
        const uint32_t rlo = (uint32_t)r, rhi = (uint32_t)(r >> 32);
        for (uint32_t ob = 0; ob < total; ob += 64) {
            const uint32_t o = ob + lane;
            uint32_t m = 0;
#pragma unroll
            for (uint32_t step = 32; step > 0; step >>= 1) {
                const uint32_t cand = m + step;   // at most 63
                if (__shfl(pre, (int)cand, 64) <= o) m = cand;
            }
            const uint32_t pm = __shfl(pre, (int)m, 64);
            const uint64_t rm = ((uint64_t)(uint32_t)__shfl(rhi, (int)m, 64) << 32) | (uint32_t)__shfl(rlo, (int)m, 64);
            if (o < total) keys[out0 + o] = ((uint64_t)mem[rm + 3 + (o - pm)] << 32) | rm;
        }
    }
}

// segment starts go to OL.off, segment ends (exclusive) to OL.cap for now
__global__ void k_ot_bounds(const uint64_t *__restrict__ keys, uint64_t L, OL *__restrict__ ot, uint64_t *__restrict__ occ)
{
    for (uint64_t k = gtid(); k < L; k += gstride()) {
        const uint64_t key = keys[k];
        const uint32_t lit = (uint32_t)(key >> 32);
        occ[k] = key & 0xFFFFFFFFull;
        if (k == 0 || (uint32_t)(keys[k - 1] >> 32) != lit) ot[lit].off = k;
        if (k + 1 == L || (uint32_t)(keys[k + 1] >> 32) != lit) ot[lit].cap = (uint32_t)(k + 1);
    }
}

__global__ void k_ot_finish(OL *__restrict__ ot, uint32_t *__restrict__ hist, uint64_t n)
{
    for (uint64_t l = gtid(); l < n; l += gstride()) {
        OL o = ot[l];
        const uint32_t sz = o.cap ? (uint32_t)(o.cap - o.off) : 0u;
        o.sz = o.cap = sz;
        ot[l] = o;
        hist[l] = sz;
    }
}

// ---- update of the last radix-built table (tuning ot_build 0 and 3)
// A table is the sorted array of (literal << 32 | reference) keys of the live
// clauses. Between two builds clauses are deleted, shrunk (strengthening, false
// literals removed by propagation), rewritten in place (BVE's equivalence
// substitution) and appended (resolvents); references do not move (compaction and
// new uploads end the update chain). A clause of the last build whose live size
// changed or that was rewritten (its bit in the rewrite map) is "changed": its old
// keys are among the keys of its current literals (a deleted clause keeps its
// literals) and of the literals removed from it in place (the removal log). So the
// next array is the old array without those keys, merged with the keys of the
// changed clauses still live and of the appended ones, sorted by literal: the same
// array a full sort gives, so every list is the same. Few clauses change per round,
// so the dropped keys are found by binary search and marked by position, and the
// old array is read once.

// Per clause: its live size now (nsz), the literals to add (asz: every live literal
// of a changed or appended clause) and to drop (dsz: every literal of a changed
// clause that had keys at the last build).
__global__ void k_otu_sizes(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n, uint64_t n0,
                            const uint32_t *__restrict__ osz, const uint32_t *__restrict__ rewritten,
                            uint32_t *__restrict__ nsz, uint32_t *__restrict__ asz, uint32_t *__restrict__ dsz)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint64_t r = refs[i];
        const uint32_t *c = mem + r;
        const uint32_t sz = c_del(c) ? 0u : (uint32_t)c_size(c);
        nsz[i] = sz;
        bool changed = true;
        uint32_t d = 0;
        if (i < n0) {
            const uint64_t b = r >> 2;
            const uint32_t o = osz[i];
            changed = sz != o || ((rewritten[b >> 5] >> (b & 31u)) & 1u);
            if (changed && o) d = (uint32_t)c_size(c);
        }
        asz[i] = changed ? sz : 0u;
        dsz[i] = d;
    }
}

constexpr uint32_t OTU_BLOCK = 256, OTU_TILE = 2048, OTU_ITEMS = OTU_TILE / OTU_BLOCK, OTU_WORDS = OTU_TILE / 32;

// first position in a[lo, hi) whose key is not below key
__device__ __forceinline__ uint64_t lower_bound_u64(const uint64_t *__restrict__ a, uint64_t lo, uint64_t hi, uint64_t key)
{
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// a key to drop: its position in the old array gets its bit (once; keys that are not
// there, like a substituted literal, are passed over) and its tile one dropped key
__global__ void k_otu_locate(const uint64_t *__restrict__ okeys, uint64_t L0, const uint64_t *__restrict__ dkeys, uint64_t nd,
                             uint32_t *__restrict__ pbits, unsigned long long *__restrict__ dcnt)
{
    for (uint64_t j = gtid(); j < nd; j += gstride()) {
        const uint64_t key = dkeys[j];
        const uint64_t p = lower_bound_u64(okeys, 0, L0, key);
        if (p >= L0 || okeys[p] != key) continue;
        const uint32_t bit = 1u << (p & 31u);
        if (!(atomicOr(&pbits[p >> 5], bit) & bit)) atomicAdd(&dcnt[p / OTU_TILE], 1ull);
    }
}

// nb[t] = the first new key not below tile t's first old key (nb[0] = 0): new keys
// [nb[t], nb[t + 1]) are merged into tile t, the ones past the last tile's first key
// into the last tile
__global__ void k_otu_splits(const uint64_t *__restrict__ okeys, uint64_t ntiles, const uint64_t *__restrict__ knew,
                             uint64_t M, uint64_t *__restrict__ nb)
{
    for (uint64_t t = gtid(); t < ntiles; t += gstride()) nb[t] = t ? lower_bound_u64(knew, 0, M, okeys[t * OTU_TILE]) : 0;
}

// keys out of tile t: its old keys less the dropped ones, plus its new keys
__global__ void k_otu_tot(uint64_t L0, const unsigned long long *__restrict__ dcnt, const uint64_t *__restrict__ nb,
                          uint64_t ntiles, uint64_t M, uint64_t *__restrict__ tot)
{
    for (uint64_t t = gtid(); t < ntiles; t += gstride()) {
        const uint64_t tn = L0 - t * OTU_TILE < OTU_TILE ? L0 - t * OTU_TILE : OTU_TILE;
        tot[t] = tn - dcnt[t] + (t + 1 < ntiles ? nb[t + 1] : M) - nb[t];
    }
}

// One block per tile: the tile's old keys and drop bits in LDS, the kept keys'
// exclusive ranks (and the tile's new keys, up to OTU_NEW of them); every kept key
// goes to outbase[t] + its rank + the tile's new keys below it, and every new key of
// the tile to outbase[t] + its index among them + the kept keys below it. Kept and
// new keys never coincide (a changed clause's old keys are all dropped), so the
// positions are distinct. The keys go out (the next update's array) with their
// references (the occurrence lists); a tile whose new keys were staged also merges
// in LDS, so that it writes the list bounds of the literals that start (OL.off) or
// end (OL.cap, one past) inside its output; k_otu_edges does the tile's two ends and
// the whole output of a tile that was not staged.
constexpr uint32_t OTU_NEW = 1024, OTU_NITEMS = OTU_NEW / OTU_BLOCK, OTU_OUT = OTU_TILE + OTU_NEW;
__global__ void __launch_bounds__(OTU_BLOCK) k_otu_merge(const uint64_t *__restrict__ okeys, uint64_t L0,
                                                         const uint32_t *__restrict__ pbits, const uint64_t *__restrict__ knew,
                                                         uint64_t M, const uint64_t *__restrict__ nb, uint64_t ntiles,
                                                         const uint64_t *__restrict__ outbase, uint64_t *__restrict__ out,
                                                         uint64_t *__restrict__ occ, OL *__restrict__ ot)
{
    __shared__ uint64_t s_key[OTU_OUT];   // the tile's old keys, then its merged output
    __shared__ uint64_t s_new[OTU_NEW];
    __shared__ uint16_t s_rank[OTU_TILE + 1];
    __shared__ uint32_t s_bits[OTU_WORDS];
    __shared__ uint32_t s_w[OTU_BLOCK / 64];
    const uint32_t tid = threadIdx.x;
    const uint64_t t = blockIdx.x, base = t * OTU_TILE;
    const uint32_t tn = (uint32_t)(L0 - base < (uint64_t)OTU_TILE ? L0 - base : (uint64_t)OTU_TILE);
    const uint64_t n0 = nb[t], n1 = t + 1 < ntiles ? nb[t + 1] : M, ob = outbase[t];
    const bool staged = n1 - n0 <= OTU_NEW;
    for (uint32_t p = tid; p < OTU_TILE; p += OTU_BLOCK) s_key[p] = p < tn ? okeys[base + p] : ~0ull;
    if (tid < OTU_WORDS) s_bits[tid] = tid * 32 < tn ? pbits[(base >> 5) + tid] : 0u;
    if (staged)
        for (uint32_t j = tid; j < (uint32_t)(n1 - n0); j += OTU_BLOCK) s_new[j] = knew[n0 + j];
    __syncthreads();
    // thread tid ranks positions [tid * OTU_ITEMS, (tid + 1) * OTU_ITEMS): bits of one drop word
    const uint32_t p0 = tid * OTU_ITEMS;
    const uint32_t dropw = (s_bits[p0 >> 5] >> (p0 & 31u)) & ((1u << OTU_ITEMS) - 1u);
    const uint32_t valid = p0 >= tn ? 0u : tn - p0 >= OTU_ITEMS ? (1u << OTU_ITEMS) - 1u : (1u << (tn - p0)) - 1u;
    const uint32_t flags = valid & ~dropw;
    const uint32_t cnt = (uint32_t)__popc(flags);
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = (uint32_t)__shfl_up(incl, o, 64);
        if ((int)__lane_id() >= o) incl += x;
    }
    if (__lane_id() == 63) s_w[tid >> 6] = incl;
    __syncthreads();
    uint32_t r = incl - cnt;
    for (uint32_t w = 0; w < (tid >> 6); w++) r += s_w[w];
#pragma unroll
    for (uint32_t u = 0; u < OTU_ITEMS; u++) {
        s_rank[p0 + u] = (uint16_t)r;
        r += (flags >> u) & 1u;
    }
    if (tid == OTU_BLOCK - 1) s_rank[OTU_TILE] = (uint16_t)r;
    __syncthreads();
    const uint32_t nnew = (uint32_t)(n1 - n0);
    if (!staged) {   // keys and references straight out; k_otu_edges finds the bounds
        for (uint32_t p = tid; p < tn; p += OTU_BLOCK) {
            if (s_rank[p + 1] == s_rank[p]) continue;   // dropped
            const uint64_t key = s_key[p];
            const uint64_t q = ob + s_rank[p] + lower_bound_u64(knew, n0, n1, key) - n0;
            out[q] = key;
            occ[q] = key & 0xFFFFFFFFull;
        }
        for (uint64_t j = n0 + tid; j < n1; j += OTU_BLOCK) {
            const uint64_t key = knew[j];
            uint32_t lo = 0, hi = tn;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_key[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            const uint64_t q = ob + (j - n0) + s_rank[lo];
            out[q] = key;
            occ[q] = key & 0xFFFFFFFFull;
        }
        return;
    }
    // positions in the tile's output, kept in registers until the old keys are read
    uint64_t kk[OTU_ITEMS], nk[OTU_NITEMS];
    uint32_t kq[OTU_ITEMS], nq[OTU_NITEMS];
#pragma unroll
    for (uint32_t u = 0; u < OTU_ITEMS; u++) {
        const uint32_t p = tid + u * OTU_BLOCK;
        kq[u] = 0xFFFFFFFFu;
        kk[u] = 0;
        if (p < tn && s_rank[p + 1] != s_rank[p]) {
            const uint64_t key = s_key[p];
            uint32_t lo = 0, hi = nnew;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_new[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            kk[u] = key;
            kq[u] = s_rank[p] + lo;
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < OTU_NITEMS; u++) {
        const uint32_t j = tid + u * OTU_BLOCK;
        nq[u] = 0xFFFFFFFFu;
        nk[u] = 0;
        if (j < nnew) {
            const uint64_t key = s_new[j];
            uint32_t lo = 0, hi = tn;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_key[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            nk[u] = key;
            nq[u] = j + s_rank[lo];
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < OTU_ITEMS; u++)
        if (kq[u] != 0xFFFFFFFFu) s_key[kq[u]] = kk[u];
#pragma unroll
    for (uint32_t u = 0; u < OTU_NITEMS; u++)
        if (nq[u] != 0xFFFFFFFFu) s_key[nq[u]] = nk[u];
    __syncthreads();
    const uint32_t T = s_rank[OTU_TILE] + nnew;
    for (uint32_t q = tid; q < T; q += OTU_BLOCK) {
        const uint64_t key = s_key[q];
        const uint32_t lit = (uint32_t)(key >> 32);
        out[ob + q] = key;
        occ[ob + q] = key & 0xFFFFFFFFull;
        if (q > 0 && (uint32_t)(s_key[q - 1] >> 32) != lit) ot[lit].off = ob + q;
        if (q + 1 < T && (uint32_t)(s_key[q + 1] >> 32) != lit) ot[lit].cap = (uint32_t)(ob + q + 1);
    }
}

// the bounds k_otu_merge left: a staged tile's first and last output key against the
// neighbouring tiles' keys; every key of a tile that was not staged (as k_ot_bounds)
__global__ void __launch_bounds__(256) k_otu_edges(const uint64_t *__restrict__ out, uint64_t L, const uint64_t *__restrict__ nb,
                                                   uint64_t ntiles, uint64_t M, const uint64_t *__restrict__ outbase,
                                                   const uint64_t *__restrict__ tot, OL *__restrict__ ot)
{
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t T = tot[t], ob = outbase[t];
        if (!T) continue;
        const uint64_t n0 = nb[t], n1 = t + 1 < ntiles ? nb[t + 1] : M;
        const bool staged = n1 - n0 <= OTU_NEW;
        // staged: thread 0 takes the first key, thread 1 the last (thread 0 both if one)
        const uint64_t qend = staged ? (threadIdx.x < 2 ? ob + threadIdx.x + 1 : ob) : ob + T;
        for (uint64_t q0 = ob + threadIdx.x; q0 < qend; q0 += blockDim.x) {
            const uint64_t q = staged && threadIdx.x == 1 ? ob + T - 1 : q0;
            if (staged && threadIdx.x == 1 && T == 1) break;
            const uint32_t lit = (uint32_t)(out[q] >> 32);
            const bool first = !staged || q == ob, last = !staged || q == ob + T - 1;
            if (first && (q == 0 || (uint32_t)(out[q - 1] >> 32) != lit)) ot[lit].off = q;
            if (last && (q + 1 == L || (uint32_t)(out[q + 1] >> 32) != lit)) ot[lit].cap = (uint32_t)(q + 1);
        }
    }
}

__global__ void k_otu_diff(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b, uint64_t n, unsigned long long *cnt)
{
    uint32_t d = 0;
    for (uint64_t i = gtid(); i < n; i += gstride()) d += a[i] != b[i];
    if (d) atomicAdd(cnt, (unsigned long long)d);
}

static int bits_of(uint64_t v)
{
    int b = 0;
    while (b < 64 && (v >> b)) b++;
    return b;
}

// live size of every clause into sizes, their exclusive scan into pos; returns the total
static uint64_t live_sizes_scan(Ctx &c, uint32_t *sizes, uint64_t *pos)
{
    const uint64_t n = c.nrefs;
    k_live_sizes<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, sizes);
    scan_u32_to_u64(c, sizes, pos, n, c.cnt.p + CNT_TOTAL);
    read_counters(c, 1);
    return c.hcnt[CNT_TOTAL];
}

// the keys of every clause (szs: of its first szs[i] literals) at pos, in clause order
static void expand_keys(Ctx &c, const uint32_t *szs, const uint64_t *pos, uint64_t L, uint64_t *kbuf)
{
    const uint64_t n = c.nrefs;
    KTimer kt(c, K_OT_FILL, 20.0 * n + 12.0 * L);
    k_ot_expand<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, szs, pos, kbuf);
}

// ... sorted by literal: returns kbuf or alt (the radix temporary when alt is null),
// whichever holds the L sorted keys
static uint64_t *expand_sort(Ctx &c, const uint32_t *szs, const uint64_t *pos, uint64_t L, uint64_t *kbuf, uint64_t *alt)
{
    expand_keys(c, szs, pos, L, kbuf);
    KTimer kt(c, K_OT_SORT, 32.0 * L * ((bits_of(c.nlits2 - 1) + 7) / 8));
    return radix_sort_u64_keys(c, kbuf, L, bits_of(c.nlits2 - 1), 32, alt);
}

// The update: the new sorted keys in c.otk (their count in L); false when the removal
// log overflowed or so much changed that sorting every key is the cheaper way (a
// quarter of the old keys; nothing is changed then).
static bool update_keys(Ctx &c, uint64_t &L)
{
    const uint64_t n = c.nrefs, n0 = c.otk_nrefs, L0 = c.otk_n;
    ensure(c, c.tmpu32a, n + 1);
    ensure(c, c.tmpu32b, n + 1);
    ensure(c, c.tmpu32c, n + 1);
    ensure(c, c.tmpu64a, n + 1);
    ensure(c, c.tmpu64c, n + 1);
    k_otu_sizes<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, n0, c.otsz.p, c.otdirty.p,
                                                                c.tmpu32a.p, c.tmpu32b.p, c.tmpu32c.p);
    scan_u32_to_u64(c, c.tmpu32b.p, c.tmpu64a.p, n, c.cnt.p + CNT_A);   // keys to add
    scan_u32_to_u64(c, c.tmpu32c.p, c.tmpu64c.p, n, c.cnt.p + CNT_B);   // keys of changed clauses to drop
    PFG_HIP(hipMemcpyAsync(c.hcnt, c.cnt.p, 8 * (CNT_OTRM + 1), hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    const uint64_t M = c.hcnt[CNT_A], MD = c.hcnt[CNT_B], NR = c.hcnt[CNT_OTRM];
    // the checked setting updates whatever changed, so that the update itself is tested
    if (NR > c.otrm_cap || (c.tune.ot_build != 3 && 4 * (M + MD + NR) > L0)) return false;
    ensure(c, c.tmpu64b, std::max(M, MD) + 1024);
    const uint64_t ntiles = (L0 + OTU_TILE - 1) / OTU_TILE;
    ensure(c, c.otpos, L0 / 32 + 1);
    ensure(c, c.otcnt, ntiles + 1);   // per tile: dropped keys
    unsigned long long *dcnt = (unsigned long long *)c.otcnt.p;
    PFG_HIP(hipMemsetAsync(c.otpos.p, 0, (L0 / 32 + 1) * 4, c.stream));
    PFG_HIP(hipMemsetAsync(dcnt, 0, 8 * (ntiles + 1), c.stream));
    if (MD) {   // the changed clauses' current literals
        expand_keys(c, c.tmpu32c.p, c.tmpu64c.p, MD, c.tmpu64b.p);
        k_otu_locate<<<grid_for(MD, 256, 8192), 256, 0, c.stream>>>(c.otk.p, L0, c.tmpu64b.p, MD, c.otpos.p, dcnt);
    }
    if (NR)     // the literals removed from them in place
        k_otu_locate<<<grid_for(NR, 256, 8192), 256, 0, c.stream>>>(c.otk.p, L0, c.otrm.p, NR, c.otpos.p, dcnt);
    const uint64_t *knew = M ? expand_sort(c, c.tmpu32b.p, c.tmpu64a.p, M, c.tmpu64b.p, nullptr) : c.tmpu64b.p;
    ensure(c, c.tmpu64c, 3 * (ntiles + 1));
    uint64_t *nb = c.tmpu64c.p, *tot = nb + ntiles + 1, *ob = tot + ntiles + 1;
    k_otu_splits<<<grid_for(ntiles, 256, 4096), 256, 0, c.stream>>>(c.otk.p, ntiles, knew, M, nb);
    k_otu_tot<<<grid_for(ntiles, 256, 4096), 256, 0, c.stream>>>(L0, dcnt, nb, ntiles, M, tot);
    scan_u64(c, tot, ob, ntiles, c.cnt.p + CNT_TOTAL);
    read_counters(c, 1);
    L = c.hcnt[CNT_TOTAL];
    ensure(c, c.otk2, L + 1024);
    ensure(c, c.occ, L + 1024);
    PFG_HIP(hipMemsetAsync(c.ot.p, 0, c.nlits2 * sizeof(OL), c.stream));
    {
        KTimer kt(c, K_OT_UPDATE, 8.0 * ((double)L0 + 2.0 * (double)L + (double)M));
        k_otu_merge<<<(unsigned)ntiles, OTU_BLOCK, 0, c.stream>>>(c.otk.p, L0, c.otpos.p, knew, M, nb, ntiles, ob,
                                                                 c.otk2.p, c.occ.p, c.ot.p);
    }
    k_otu_edges<<<(unsigned)std::min<uint64_t>(ntiles, 65536), 256, 0, c.stream>>>(c.otk2.p, L, nb, ntiles, M, ob, tot,
                                                                                   c.ot.p);
    std::swap(c.otk, c.otk2);
    std::swap(c.otsz, c.tmpu32a);   // live sizes at this build
    c.path[PATH_OT_UPDATE]++;
    return true;
}

static void create_ot_radix(Ctx &c)
{
    const uint64_t n = c.nrefs;
    const bool keep_keys = c.tune.ot_build == 0 || c.tune.ot_build == 3;
    ensure(c, c.ot, c.nlits2);
    ensure(c, c.hist, c.nlits2);
    uint64_t L = 0;
    bool updated = false;
    if (keep_keys && c.otk_valid && c.otk_n && n >= c.otk_nrefs) updated = update_keys(c, L);
    if (updated && c.tune.ot_build == 3) {   // checked against a full sort
        ensure(c, c.tmpu32c, n + 1);
        ensure(c, c.tmpu64c, n + 1);
        const uint64_t Lf = live_sizes_scan(c, c.tmpu32c.p, c.tmpu64c.p);
        if (Lf != L) throw Error(PFG_EINTERNAL, "occurrence-table update: key count differs from a full build");
        ensure(c, c.tmpu64b, L + 1024);
        ensure(c, c.otk2, L + 1024);
        const uint64_t *full = L ? expand_sort(c, nullptr, c.tmpu64c.p, L, c.tmpu64b.p, c.otk2.p) : c.tmpu64b.p;
        PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_A, 0, 8, c.stream));
        k_otu_diff<<<grid_for(L, 256, 4096), 256, 0, c.stream>>>(full, c.otk.p, L, (unsigned long long *)c.cnt.p + CNT_A);
        read_counters(c, 2);
        if (c.hcnt[CNT_A]) throw Error(PFG_EINTERNAL, "occurrence-table update differs from a full build");
        // the lists and list bounds the merge wrote against those the full build derives
        const uint64_t lo = (L + 2) & ~1ull;
        ensure(c, c.otchk, lo + 2 * c.nlits2);
        OL *ot2 = (OL *)(c.otchk.p + lo);
        PFG_HIP(hipMemsetAsync(ot2, 0, c.nlits2 * sizeof(OL), c.stream));
        if (L) k_ot_bounds<<<grid_for(L, 256, 16384), 256, 0, c.stream>>>(c.otk.p, L, ot2, c.otchk.p);
        k_otu_diff<<<grid_for(L, 256, 4096), 256, 0, c.stream>>>(c.otchk.p, c.occ.p, L, (unsigned long long *)c.cnt.p + CNT_A);
        k_otu_diff<<<grid_for(2 * c.nlits2, 256, 4096), 256, 0, c.stream>>>((const uint64_t *)ot2, (const uint64_t *)c.ot.p,
                                                                           2 * c.nlits2, (unsigned long long *)c.cnt.p + CNT_A);
        read_counters(c, 2);
        if (c.hcnt[CNT_A]) throw Error(PFG_EINTERNAL, "occurrence-table update: lists differ from a full build");
    }
    if (!updated) {
        ensure(c, c.otsz, n + 1);
        ensure(c, c.tmpu64a, n + 1);
        L = live_sizes_scan(c, c.otsz.p, c.tmpu64a.p);
        ensure(c, c.otk, L + 1024);
        ensure(c, c.otk2, L + 1024);
        if (L) {
            const uint64_t *sorted = expand_sort(c, nullptr, c.tmpu64a.p, L, c.otk2.p, c.otk.p);
            if (sorted == c.otk2.p) std::swap(c.otk, c.otk2);
        }
    }
    if (!updated) {   // the update's merge wrote the lists and their bounds itself
        ensure(c, c.occ, L + 1024);
        PFG_HIP(hipMemsetAsync(c.ot.p, 0, c.nlits2 * sizeof(OL), c.stream));
        if (L) k_ot_bounds<<<grid_for(L, 256, 16384), 256, 0, c.stream>>>(c.otk.p, L, c.ot.p, c.occ.p);
    }
    k_ot_finish<<<grid_for(c.nlits2, 256, 8192), 256, 0, c.stream>>>(c.ot.p, c.hist.p, c.nlits2);
    // what the next build updates from; the rewrite map and the removal log start empty
    c.otk_n = L;
    c.otk_nrefs = n;
    c.otk_valid = keep_keys;
    c.otdirty_words = (c.msz >> 7) + 1;
    ensure(c, c.otdirty, c.otdirty_words);
    PFG_HIP(hipMemsetAsync(c.otdirty.p, 0, c.otdirty_words * 4, c.stream));
    c.otrm_cap = std::max<uint64_t>(L / 4, 1u << 16);
    ensure(c, c.otrm, c.otrm_cap);
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_OTRM, 0, 8, c.stream));
    c.ot_valid = true;
    c.hist_live = true;
}

// Two ways to the same table (every list in clause order, histogram of live literals):
// an LSD radix sort of (literal, reference) keys, or counting -- the histogram, list
// starts by a scan, an atomic fill and a sort of every list by reference (the fill
// order is not the clause order). The tuning picks one; references beyond 32 bits
// always take the counting way.
void create_ot(Ctx &c)
{
    if (c.tune.ot_build != 1 && c.msz < (1ull << 32) && c.msz + 3 * c.nrefs < (1ull << 32)) {
        create_ot_radix(c);
        return;
    }
    c.otk_valid = false;
    if (!c.hist_live) hist(c);
    ensure(c, c.ot, c.nlits2);
    ensure(c, c.tmpu64a, c.nlits2);
    scan_u32_to_u64(c, c.hist.p, c.tmpu64a.p, c.nlits2, c.cnt.p + CNT_TOTAL);
    read_counters(c, 1);
    const uint64_t total = c.hcnt[CNT_TOTAL];
    ensure(c, c.occ, total + 1024);
    k_ot_init<<<grid_for(c.nlits2, 256, 8192), 256, 0, c.stream>>>(c.ot.p, c.tmpu64a.p, c.hist.p, c.nlits2);
    if (c.nrefs) {
        KTimer kt(c, K_OT_FILL, 20.0 * c.nrefs + 12.0 * total + 16.0 * c.nlits2);
        k_ot_fill<<<xcd_grid(c.nrefs, 256, c.tune.clause_grid), 256, 0, c.stream>>>(c.mem.p, c.refs.p, c.nrefs, c.ot.p,
                                                                                    c.occ.p);
    }
    {
        KTimer kt(c, K_OT_SORT, 16.0 * total);
        order_lists_by_ref(c);
    }
    c.ot_valid = true;
    c.hist_live = true;
}

// reduceOL (src/cpu/simplify.cpp:55-65): in-order filter of deleted references
__global__ void k_reduce_ot(const uint32_t *__restrict__ mem, OL *__restrict__ ot, uint64_t *__restrict__ occ, uint64_t n)
{
    for (uint64_t l = gtid(); l < n; l += gstride()) {
        OL o = ot[l];
        uint64_t *d = occ + o.off;
        uint32_t j = 0;
        for (uint32_t i = 0; i < o.sz; i++) {
            const uint64_t r = d[i];
            if (!c_del(mem + r)) d[j++] = r;
        }
        ot[l].sz = j;
    }
}

void reduce_ot(Ctx &c)
{
    k_reduce_ot<<<grid_for(c.nlits2, 256, 8192), 256, 0, c.stream>>>(c.mem.p, c.ot.p, c.occ.p, c.nlits2);
}

// ------------------------------------------------------------------ unit propagation
// prop (src/cpu/elimination.cpp:64-97) in batches: all clauses on the lists of the
// batch's units (and of their negations) are claimed once and brought to the state
// the sequential loop leaves them in: satisfied -> deleted, false literals removed
// (order kept, signature recomputed), new units assigned and appended to the trail.
// claim the clause at r; fresh claims go out with one atomic per wave
__device__ __forceinline__ void prop_claim(bool valid, uint64_t r, uint32_t *__restrict__ claim, uint64_t *__restrict__ cand,
                                           unsigned long long *cnt)
{
    bool fresh = false;
    if (valid) {
        const uint64_t key = r >> 1;
        const uint32_t bit = 1u << (key & 31u);
        fresh = !(atomicOr(&claim[key >> 5], bit) & bit);
    }
    const uint64_t fm = __ballot(fresh);
    unsigned long long at = 0;
    if (__lane_id() == 0 && fm) at = atomicAdd(&cnt[CNT_A], (unsigned long long)__popcll(fm));
    at = __shfl(at, 0, 64);
    if (fresh) cand[at + __popcll(fm & ((1ull << __lane_id()) - 1ull))] = r;
}

// A wave takes 64 lists (a unit's literal and its negation, for 32 units): their
// entries are walked as one run, a lane per entry and step (a lane finds its entry's
// list by a binary search over the wave's prefix of list sizes), so short lists do
// not leave lanes idle. Lists longer than PROP_LONG go to a worklist (count in
// CNT_B) that k_prop_collect_long walks a block per list.
constexpr uint32_t PROP_LONG = 1024;
__global__ void __launch_bounds__(256) k_prop_collect(const uint32_t *__restrict__ trail, uint32_t b0, uint32_t b1,
                                                     const OL *__restrict__ ot, const uint64_t *__restrict__ occ,
                                                     uint32_t *__restrict__ claim, uint64_t *__restrict__ cand,
                                                     unsigned long long *cnt, uint32_t *__restrict__ longq)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    const uint64_t nq = 2ull * (b1 - b0);
    for (uint64_t base = wave * 64; base < nq; base += nwaves * 64) {
        const uint64_t q = base + lane;
        uint64_t off = 0;
        uint32_t sz = 0;
        if (q < nq) {
            const uint32_t u = trail[b0 + (uint32_t)(q >> 1)];
            const OL o = ot[(q & 1) ? u : (u ^ 1u)];
            off = o.off;
            sz = o.sz;
        }
        wave_append(sz > PROP_LONG, (uint32_t)q, longq, &cnt[CNT_B]);
        if (sz > PROP_LONG) sz = 0;
        uint32_t incl = sz;
#pragma unroll
        for (int k = 1; k < 64; k <<= 1) {
            const uint32_t t = __shfl_up(incl, k, 64);
            if ((int)lane >= k) incl += t;
        }
        const uint32_t pre = incl - sz, total = __shfl(incl, 63, 64);
        const uint32_t olo = (uint32_t)off, ohi = (uint32_t)(off >> 32);
        for (uint32_t t0 = 0; t0 < total; t0 += 64) {
            const uint32_t e = t0 + lane;
            uint32_t m = 0;   // the last list whose prefix is at most e
#pragma unroll
            for (uint32_t step = 32; step > 0; step >>= 1)
                if ((uint32_t)__shfl((int)pre, (int)(m + step), 64) <= e) m += step;
            const uint32_t pm = (uint32_t)__shfl((int)pre, (int)m, 64);
            const uint64_t om = ((uint64_t)(uint32_t)__shfl((int)ohi, (int)m, 64) << 32) |
                                (uint32_t)__shfl((int)olo, (int)m, 64);
            const uint64_t r = e < total ? occ[om + (e - pm)] : 0ull;
            prop_claim(e < total, r, claim, cand, cnt);
        }
    }
}

// the long lists, a block each (the worklist's count is read on the device)
__global__ void __launch_bounds__(256) k_prop_collect_long(const uint32_t *__restrict__ trail, uint32_t b0,
                                                          const OL *__restrict__ ot, const uint64_t *__restrict__ occ,
                                                          uint32_t *__restrict__ claim, uint64_t *__restrict__ cand,
                                                          unsigned long long *cnt, const uint32_t *__restrict__ longq)
{
    const uint64_t nl = cnt[CNT_B];
    for (uint64_t k = blockIdx.x; k < nl; k += gridDim.x) {
        const uint32_t q = longq[k];
        const uint32_t u = trail[b0 + (q >> 1)];
        const OL o = ot[(q & 1) ? u : (u ^ 1u)];
        for (uint32_t e0 = 0; e0 < o.sz; e0 += blockDim.x) {
            const uint32_t e = e0 + threadIdx.x;
            prop_claim(e < o.sz, e < o.sz ? occ[o.off + e] : 0ull, claim, cand, cnt);
        }
    }
}

// hist (when live) stays the live-occurrence histogram for every variable LCVE reads:
// a deleted clause's unassigned literals are taken off it. Assigned variables (the
// false literals removed, the true ones) are passed over by LCVE, so theirs are left.
__global__ void k_prop_apply(uint32_t *__restrict__ mem, const uint64_t *__restrict__ cand, uint64_t n, int8_t *value,
                             uint8_t *vstate, uint32_t *trail, unsigned long long *cnt, uint32_t *__restrict__ claim,
                             uint32_t *__restrict__ hist, uint64_t *__restrict__ rmlog, uint64_t rmcap)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint64_t r = cand[i];
        const uint64_t key = r >> 1;
        atomicAnd(&claim[key >> 5], ~(1u << (key & 31u)));
        uint32_t *c = mem + r;
        if (c_del(c)) continue;
        const int sz = c_size(c);
        uint32_t *lits = c + 3;
        bool sat = false;
        for (int k = 0; k < sz; k++)
            if (__atomic_load_n(&value[lits[k]], __ATOMIC_RELAXED) > 0) { sat = true; break; }
        if (sat) {
            c_markdel(c);
            if (hist)
                for (int k = 0; k < sz; k++)
                    if (__atomic_load_n(&value[lits[k]], __ATOMIC_RELAXED) < 0) atomicSub(&hist[lits[k]], 1u);
            continue;
        }
        int j = 0;
        uint32_t sig = 0;
        for (int k = 0; k < sz; k++) {
            const uint32_t l = lits[k];
            if (__atomic_load_n(&value[l], __ATOMIC_RELAXED) == 0) {
                log_removed(rmlog, &cnt[CNT_OTRM], rmcap, l, r);
                continue;
            }
            lits[j++] = l;
            sig |= maphash(l);
        }
        c[1] = sig;
        c[2] = (uint32_t)j;
        if (j == 0) report_conflict(cnt, lits[0] >> 1);   // every literal false: lits[] untouched
        else if (j == 1) enqueue_dev(value, vstate, trail, cnt, lits[0]);
    }
}

// empties the units' lists; resets the candidate and long-list counts for the next batch
__global__ void k_prop_clear(const uint32_t *__restrict__ trail, uint32_t b0, uint32_t b1, OL *__restrict__ ot,
                             unsigned long long *cnt)
{
    if (gtid() == 0) {
        cnt[CNT_A] = 0;
        cnt[CNT_B] = 0;
    }
    for (uint64_t q = gtid() + b0; q < b1; q += gstride()) {
        const uint32_t u = trail[q];
        ot[u].sz = 0;
        ot[u ^ 1u].sz = 0;
    }
}

bool prop(Ctx &c)
{
    if (c.propagated >= c.ntrail) return !c.unsat;
    const uint32_t before = c.propagated;
    ensure(c, c.claim, (c.mem.cap >> 6) + 2);
    PFG_HIP(hipMemsetAsync(c.claim.p, 0, c.claim.cap * 4, c.stream));
    ensure(c, c.tmpu64b, c.occ.cap);
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_A, 0, 16, c.stream));   // CNT_A candidates, CNT_B long lists (then k_prop_clear)
    while (c.propagated < c.ntrail) {
        const uint32_t b0 = c.propagated, b1 = c.ntrail;
        ensure(c, c.plong, 2ull * (b1 - b0));
        k_prop_collect<<<grid_for(2ull * (b1 - b0), 256, 8192), 256, 0, c.stream>>>(
            c.trail.p, b0, b1, c.ot.p, c.occ.p, c.claim.p, c.tmpu64b.p, (unsigned long long *)c.cnt.p, c.plong.p);
        k_prop_collect_long<<<(unsigned)std::min<uint64_t>(2ull * (b1 - b0), 1024), 256, 0, c.stream>>>(
            c.trail.p, b0, c.ot.p, c.occ.p, c.claim.p, c.tmpu64b.p, (unsigned long long *)c.cnt.p, c.plong.p);
        read_counters(c, 2);
        const uint64_t nc = c.hcnt[CNT_A];
        if (nc) k_prop_apply<<<grid_for(nc, 256, 8192), 256, 0, c.stream>>>(c.mem.p, c.tmpu64b.p, nc, c.value.p, c.vstate.p,
                                                                             c.trail.p, (unsigned long long *)c.cnt.p, c.claim.p,
                                                                             c.hist_live ? c.hist.p : nullptr, c.otrm.p,
                                                                             c.otrm_cap);
        k_prop_clear<<<grid_for(b1 - b0, 256, 1024), 256, 0, c.stream>>>(c.trail.p, b0, b1, c.ot.p,
                                                                          (unsigned long long *)c.cnt.p);
        read_counters(c, 10);
        c.propagated = b1;
        c.ntrail = (uint32_t)c.hcnt[CNT_NTRAIL];
        if (c.hcnt[CNT_FLAGS] & FLAG_UNSAT) {
            c.unsat = true;
            if (!c.batch) return false;   // a batch propagates the other instances to the end
        }
    }
    if (c.propagated != before && !c.opts.sub_en) reduce_ot(c);
    return !(!c.batch && c.unsat);
}

// ------------------------------------------------------------------ batches
// live clauses and literals per instance (instance of a clause = instance of its
// first literal's variable; empty clauses, left only in refuted instances, are not
// counted). A lane keeps the variable range of the instance it counts for: clauses of
// one instance are mostly contiguous in the pool, so the binary search over the
// instance bases runs once per run of clauses. A lane's sums go out when its instance
// changes inside a tile, and after every tile as one atomic pair per wave when all the
// wave's counting lanes share the instance (the usual case).
constexpr int CI_BLOCK = 256, CI_ITEMS = 16;

__device__ __forceinline__ void ci_flush_wave(unsigned long long *icnt, uint32_t cur, unsigned long long &nc,
                                              unsigned long long &nl)
{
    const uint64_t act = __ballot(nc != 0);
    if (!act) return;
    const uint32_t i0 = (uint32_t)__shfl((int)cur, __ffsll((long long)act) - 1, 64);
    if (__ballot(nc != 0 && cur != i0) == 0) {
        unsigned long long a = nc, b = nl;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, 64);
            b += __shfl_down(b, o, 64);
        }
        if (__lane_id() == 0) {
            atomicAdd(&icnt[2 * i0], a);
            atomicAdd(&icnt[2 * i0 + 1], b);
        }
    }
    else if (nc) {
        atomicAdd(&icnt[2 * cur], nc);
        atomicAdd(&icnt[2 * cur + 1], nl);
    }
    nc = nl = 0;
}

__global__ void __launch_bounds__(CI_BLOCK) k_count_inst(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs,
                                                        uint64_t n, const uint32_t *__restrict__ ibase, uint32_t ninst,
                                                        unsigned long long *__restrict__ icnt, uint64_t msz)
{
    uint32_t cur = 0, lo = 1, hi = 0;   // instance counted for; its variables are (lo, hi]
    unsigned long long nc = 0, nl = 0;
    for (uint64_t t0 = (uint64_t)blockIdx.x * CI_BLOCK * CI_ITEMS; t0 < n; t0 += (uint64_t)gridDim.x * CI_BLOCK * CI_ITEMS) {
        // every item's reference, then every item's header words, in flight together
        uint64_t r[CI_ITEMS];
#pragma unroll
        for (int k = 0; k < CI_ITEMS; k++) {
            const uint64_t i = t0 + (uint64_t)k * CI_BLOCK + threadIdx.x;
            r[k] = i < n ? refs[i] : ~0ull;
        }
        uint32_t h0[CI_ITEMS], h2[CI_ITEMS], h3[CI_ITEMS];
#pragma unroll
        for (int k = 0; k < CI_ITEMS; k++) {
            h0[k] = ST_DELETED;
            h2[k] = h3[k] = 0;
            if (r[k] != ~0ull) {
                const uint32_t *cl = mem + r[k];
                h0[k] = cl[0];
                h2[k] = cl[2];
                // the first literal; only an empty image at the end of the pool has no
                // word after its header
                if (r[k] + 3 < msz) h3[k] = cl[3];
            }
        }
#pragma unroll
        for (int k = 0; k < CI_ITEMS; k++) {
            // deleted, past the end, or empty (a refuted instance's empty clause names no instance)
            if ((h0[k] & ST_DELETED) || !h2[k]) continue;
            const uint32_t v = h3[k] >> 1;
            if (v <= lo || v > hi) {
                if (nc) {
                    atomicAdd(&icnt[2 * cur], nc);
                    atomicAdd(&icnt[2 * cur + 1], nl);
                    nc = nl = 0;
                }
                cur = find_inst(ibase, ninst, v);
                lo = ibase[cur];
                hi = cur + 1 < ninst ? ibase[cur + 1] : 0xFFFFFFFFu;
            }
            nc++;
            nl += h2[k];
        }
        ci_flush_wave(icnt, cur, nc, nl);
    }
}

void count_inst(Ctx &c, std::vector<int64_t> &cls, std::vector<int64_t> &lits)
{
    PFG_HIP(hipMemsetAsync(c.icnt.p, 0, 16ull * c.ninst, c.stream));
    if (c.nrefs) {
        KTimer kt(c, K_COUNT_INST, 20.0 * c.nrefs + 4.0 * c.nliterals);
        k_count_inst<<<grid_for((c.nrefs + CI_ITEMS - 1) / CI_ITEMS, CI_BLOCK, 8192), CI_BLOCK, 0, c.stream>>>(
            c.mem.p, c.refs.p, c.nrefs, c.ibase.p, c.ninst, (unsigned long long *)c.icnt.p, c.msz);
    }
    std::vector<uint64_t> h(2ull * c.ninst);
    PFG_HIP(hipMemcpyAsync(h.data(), c.icnt.p, 16ull * c.ninst, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    cls.resize(c.ninst);
    lits.resize(c.ninst);
    for (uint32_t i = 0; i < c.ninst; i++) { cls[i] = (int64_t)h[2 * i]; lits[i] = (int64_t)h[2 * i + 1]; }
}

// Clause images grouped by instance (clause order kept within an instance), literals
// in each instance's own numbering: instance keys per clause, a stable sort of the
// clause indices by key, the image sizes scanned in that order, then the copy.
// An empty clause (left only in a refuted instance) has no literal to name its
// instance: it takes key ninst and sorts after every instance's range.
__global__ void k_clause_inst(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                              const uint32_t *__restrict__ ibase, uint32_t ninst, uint32_t *__restrict__ keys,
                              uint32_t *__restrict__ idx)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *cl = mem + refs[i];
        keys[i] = cl[2] ? find_inst(ibase, ninst, cl[3] >> 1) : ninst;
        idx[i] = (uint32_t)i;
    }
}

__global__ void k_image_sizes(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs,
                              const uint32_t *__restrict__ idx, uint64_t n, uint32_t *__restrict__ sizes)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) sizes[i] = 3u + mem[refs[idx[i]] + 2];
}

__global__ void k_image_write(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs,
                              const uint32_t *__restrict__ idx, const uint32_t *__restrict__ keys,
                              const uint64_t *__restrict__ woff, uint64_t n, const uint32_t *__restrict__ ibase,
                              uint32_t *__restrict__ out)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *cl = mem + refs[idx[i]];
        uint32_t *o = out + woff[i];
        const uint32_t sz = cl[2], shift = sz ? 2u * ibase[keys[i]] : 0u;
        o[0] = cl[0];
        o[1] = cl[1];
        o[2] = sz;
        for (uint32_t k = 0; k < sz; k++) o[3 + k] = cl[3 + k] - shift;
    }
}

// first sorted position with key >= k, for k = 0..ninst: start[k] = woff at that
// position (total past the end); cstart[k] = the position itself
__global__ void k_image_starts(const uint32_t *__restrict__ keys, const uint64_t *__restrict__ woff, uint64_t n,
                               uint32_t ninst, uint64_t total, uint64_t *__restrict__ start, uint64_t *__restrict__ cstart)
{
    for (uint64_t k = gtid(); k <= ninst; k += gstride()) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (keys[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        start[k] = lo < n ? woff[lo] : total;
        if (cstart) cstart[k] = lo;
    }
}

// out: device words (at least the live words); start: per-instance word starts (ninst + 1);
// cstart: per-instance first image index (ninst + 1). c.img_off keeps each image's word
// offset in out (grouped order) for exports.
void batch_images(Ctx &c, uint32_t *out, std::vector<uint64_t> &start, std::vector<uint64_t> &cstart)
{
    const uint64_t n = c.nrefs;
    start.assign(c.ninst + 1, 0);
    cstart.assign(c.ninst + 1, 0);
    if (!n) return;
    ensure(c, c.tmpu32a, n + 1);
    ensure(c, c.tmpu32b, n + 1);
    ensure(c, c.tmpu32c, n + 1);
    ensure(c, c.img_off, n + 1);
    ensure(c, c.tmpu64c, 2ull * c.ninst + 4);
    uint32_t *keys = c.tmpu32a.p, *idx = c.tmpu32b.p, *sizes = c.tmpu32c.p;
    k_clause_inst<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, c.ibase.p, c.ninst, keys, idx);
    int bits = 0;
    while (bits < 32 && (c.ninst >> bits)) bits++;   // keys 0..ninst
    if (bits) radix_sort_u32_pairs(c, keys, idx, n, ((bits + 7) / 8) * 8);
    k_image_sizes<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(c.mem.p, c.refs.p, idx, n, sizes);
    scan_u32_to_u64(c, sizes, c.img_off.p, n, c.cnt.p + CNT_TOTAL);
    k_image_write<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(c.mem.p, c.refs.p, idx, keys, c.img_off.p, n, c.ibase.p, out);
    read_counters(c, 1);
    const uint64_t total = c.hcnt[CNT_TOTAL];
    uint64_t *d_start = c.tmpu64c.p, *d_cstart = d_start + c.ninst + 1;
    k_image_starts<<<grid_for((uint64_t)c.ninst + 1, 256, 1024), 256, 0, c.stream>>>(keys, c.img_off.p, n, c.ninst, total,
                                                                                      d_start, d_cstart);
    PFG_HIP(hipMemcpyAsync(start.data(), d_start, 8ull * (c.ninst + 1), hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipMemcpyAsync(cstart.data(), d_cstart, 8ull * (c.ninst + 1), hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
}

// Instance clauses [c0, c0 + ncls) of the grouped images as CSR in the instance's numbering:
// clause q's literals at offs[q] = (its image offset - the first image's) - 3 q.
__global__ void k_export_csr(const uint32_t *__restrict__ img, const uint64_t *__restrict__ img_off, uint64_t c0,
                             uint64_t ncls, uint64_t nlits, uint32_t *__restrict__ lits, uint64_t *__restrict__ offs)
{
    const uint64_t w0 = img_off[c0];
    for (uint64_t q = gtid(); q < ncls; q += gstride()) {
        const uint64_t w = img_off[c0 + q];
        const uint32_t *cl = img + w;
        const uint64_t o = (w - w0) - 3 * q;
        const uint32_t sz = cl[2];
        offs[q] = o;
        for (uint32_t k = 0; k < sz; k++) lits[o + k] = cl[3 + k];
        if (q + 1 == ncls) offs[ncls] = nlits;
    }
}

void batch_export(Ctx &c, const uint32_t *img, uint64_t c0, uint64_t ncls, uint64_t nlits, uint32_t *d_lits,
                  uint64_t *d_offs)
{
    if (!ncls) {
        const uint64_t zero = 0;
        PFG_HIP(hipMemcpyAsync(d_offs, &zero, 8, hipMemcpyHostToDevice, c.stream));
    }
    else
        k_export_csr<<<grid_for(ncls, 256, 8192), 256, 0, c.stream>>>(img, c.img_off.p, c0, ncls, nlits, d_lits, d_offs);
    PFG_HIP(hipStreamSynchronize(c.stream));
}

// Model records grouped by instance (stack order kept within an instance), literals in
// each instance's own numbering. The unit is an eliminated variable's chunk of records
// (contiguous, one instance, starting with a literal); the chunk list holds their starts
// in stack order, so a stable sort of chunk indices by instance gives every instance's
// records oldest first.
__global__ void k_chunk_inst(const uint32_t *__restrict__ res, const uint64_t *__restrict__ chunk, uint64_t n,
                             const uint32_t *__restrict__ ibase, uint32_t ninst, uint32_t *__restrict__ keys,
                             uint32_t *__restrict__ idx)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        keys[i] = find_inst(ibase, ninst, res[chunk[i]] >> 1);
        idx[i] = (uint32_t)i;
    }
}

__global__ void k_chunk_sizes(const uint64_t *__restrict__ chunk, const uint32_t *__restrict__ idx, uint64_t n,
                              uint64_t nres, uint32_t *__restrict__ sizes)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t j = idx[i];
        sizes[i] = (uint32_t)((j + 1 < n ? chunk[j + 1] : nres) - chunk[j]);
    }
}

// a chunk is copied from its end: the last word of a record is its literal count
__global__ void k_chunk_write(const uint32_t *__restrict__ res, const uint64_t *__restrict__ chunk,
                              const uint32_t *__restrict__ idx, const uint32_t *__restrict__ keys,
                              const uint64_t *__restrict__ woff, uint64_t n, uint64_t nres,
                              const uint32_t *__restrict__ ibase, uint32_t *__restrict__ out)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t j = idx[i];
        const uint64_t s0 = chunk[j], s1 = j + 1 < n ? chunk[j + 1] : nres;
        const uint32_t *src = res + s0;
        uint32_t *o = out + woff[i];
        const uint32_t shift = 2u * ibase[keys[i]];
        int64_t e = (int64_t)(s1 - s0) - 1;
        while (e >= 0) {
            const uint32_t k = src[e];
            o[e] = k;
            const int64_t b = e - (int64_t)k;
            for (int64_t t = b < 0 ? 0 : b; t < e; t++) o[t] = src[t] - shift;
            e = b - 1;
        }
    }
}

// out: device words (at least nres); start: per-instance word starts (ninst + 1)
void batch_records(Ctx &c, uint32_t *out, std::vector<uint64_t> &start)
{
    const uint64_t n = c.nchunk;
    start.assign(c.ninst + 1, 0);
    if (!n) return;
    ensure(c, c.tmpu32a, n + 1);
    ensure(c, c.tmpu32b, n + 1);
    ensure(c, c.tmpu32c, n + 1);
    ensure(c, c.tmpu64a, n + 1);
    ensure(c, c.tmpu64c, (uint64_t)c.ninst + 2);
    uint32_t *keys = c.tmpu32a.p, *idx = c.tmpu32b.p, *sizes = c.tmpu32c.p;
    k_chunk_inst<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(c.resolved.p, c.mchunk.p, n, c.ibase.p, c.ninst, keys, idx);
    int bits = 0;
    while (bits < 32 && ((c.ninst - 1) >> bits)) bits++;
    if (bits) radix_sort_u32_pairs(c, keys, idx, n, ((bits + 7) / 8) * 8);
    k_chunk_sizes<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(c.mchunk.p, idx, n, c.nres, sizes);
    scan_u32_to_u64(c, sizes, c.tmpu64a.p, n, c.cnt.p + CNT_TOTAL);
    k_chunk_write<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(c.resolved.p, c.mchunk.p, idx, keys, c.tmpu64a.p, n,
                                                                c.nres, c.ibase.p, out);
    k_image_starts<<<grid_for((uint64_t)c.ninst + 1, 256, 1024), 256, 0, c.stream>>>(keys, c.tmpu64a.p, n, c.ninst,
                                                                                      c.nres, c.tmpu64c.p, nullptr);
    PFG_HIP(hipMemcpyAsync(start.data(), c.tmpu64c.p, 8ull * (c.ninst + 1), hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
}

// ------------------------------------------------------------------ variable mapping
// ParaFROST::map(true) after the simplifier (src/cpu/vmap.cpp:64-131, vmap.h:44-171):
// variables without a state keep their order; the first variable that has a state
// and a value (a level-0 unit) is kept too and stands for every assigned literal.
__global__ void k_vmap_flags(const uint8_t *__restrict__ vstate, const int8_t *__restrict__ value, uint32_t nvars,
                             uint32_t *__restrict__ flag, unsigned long long *first)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) {
        const uint32_t v = (uint32_t)i + 1;
        const bool keep = !vstate[v];
        flag[i] = keep;
        if (!keep && value[2 * v] >= 0) atomicMin(first, (unsigned long long)v);
    }
}

__global__ void k_vmap_first(uint32_t *__restrict__ flag, const unsigned long long *first)
{
    if (gtid() == 0 && *first != ~0ull) flag[*first - 1] = 1;
}

// mapped[v] = new variable (0: dropped); vorg[new] = v; vstate of the kept variables
__global__ void k_vmap_ids(const uint32_t *__restrict__ flag, const uint64_t *__restrict__ pos, uint32_t nvars,
                           const uint8_t *__restrict__ vstate, uint32_t *__restrict__ mapped, uint32_t *__restrict__ vorg,
                           uint8_t *__restrict__ vs_out)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) {
        const uint32_t v = (uint32_t)i + 1;
        if (flag[i]) {
            const uint32_t m = (uint32_t)pos[i] + 1;
            mapped[v] = m;
            vorg[m] = v;
            vs_out[m] = vstate[v];
        }
        else mapped[v] = 0;
    }
}

// mapLit (vmap.h:44-60)
__device__ __forceinline__ uint32_t vmap_lit(uint32_t lit, const uint32_t *mapped, const int8_t *value, uint32_t first,
                                             uint32_t mfirst, int8_t vfirst)
{
    const uint32_t v = lit >> 1, m = mapped[v];
    if (m && v != first) return 2 * m + (lit & 1u);
    const int8_t val = value[lit];
    if (val < 0) return 0;
    return (2 * mfirst) ^ (vfirst != val ? 1u : 0u);
}

__global__ void k_vmap_sizes(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                             uint32_t *__restrict__ sizes)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *c = mem + refs[i];
        sizes[i] = c_del(c) ? 0u : 3u + (uint32_t)c_size(c);
    }
}

// mapped clause images in clause order (mapClause; signature of the new literals)
__global__ void k_vmap_write(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                             const uint64_t *__restrict__ woff, const uint32_t *__restrict__ mapped,
                             const int8_t *__restrict__ value, const unsigned long long *first_p,
                             uint32_t *__restrict__ out)
{
    const uint32_t first = *first_p == ~0ull ? 0u : (uint32_t)*first_p;
    const uint32_t mfirst = first ? mapped[first] : 0u;
    const int8_t vfirst = first ? value[2 * first] : (int8_t)-1;
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *c = mem + refs[i];
        if (c_del(c)) continue;
        uint32_t *o = out + woff[i];
        const int sz = c_size(c);
        uint32_t sig = 0;
        for (int k = 0; k < sz; k++) {
            const uint32_t l = vmap_lit(c[3 + k], mapped, value, first, mfirst, vfirst);
            o[3 + k] = l;
            sig |= maphash(l);
        }
        o[0] = c[0];
        o[1] = sig;
        o[2] = (uint32_t)sz;
    }
}

// mapShrinkLits (vmap.h:162-171) of the trail: the trail holds assigned variables, of
// which only the first one with a state is kept, so at most one literal survives and
// its position needs no order
__global__ void k_vmap_trail(const uint32_t *__restrict__ trail, uint32_t n, const uint32_t *__restrict__ mapped,
                             uint32_t *__restrict__ out, unsigned long long *cnt)
{
    for (uint64_t t = gtid(); t < n; t += gstride()) {
        const uint32_t lit = trail[t], m = mapped[lit >> 1];
        if (m) out[atomicAdd(cnt, 1ull)] = 2 * m + (lit & 1u);
    }
}

void vmap_download(Ctx &c, std::vector<uint32_t> &words, uint64_t &ncls, std::vector<uint32_t> &vorg, uint32_t &nvars_new,
                   std::vector<uint32_t> &trail, std::vector<uint8_t> &vstate)
{
    const uint32_t nv = c.nvars;
    const uint64_t n = c.nrefs;
    c.cache_gen = ~0ull;   // mem2 is reused: per-instance downloads must regroup
    ensure(c, c.tmpu32a, (uint64_t)nv + 2);
    ensure(c, c.tmpu32b, (uint64_t)nv + 2);
    ensure(c, c.tmpu32c, (uint64_t)nv + 2 + c.ntrail);
    ensure(c, c.tmpu8, (uint64_t)nv + 2);
    ensure(c, c.tmpu64a, (uint64_t)std::max<uint64_t>(nv, n) + 2);
    ensure(c, c.tmpu64c, 4);
    unsigned long long *first = (unsigned long long *)c.tmpu64c.p;
    PFG_HIP(hipMemsetAsync(first, 0xFF, 8, c.stream));
    uint32_t *flag = c.tmpu32a.p, *mapped = c.tmpu32b.p, *d_vorg = c.tmpu32c.p, *d_trail = c.tmpu32c.p + nv + 2;
    uint8_t *d_vs = c.tmpu8.p;
    PFG_HIP(hipMemsetAsync(d_vorg, 0, 4, c.stream));
    PFG_HIP(hipMemsetAsync(d_vs, 0, 1, c.stream));
    PFG_HIP(hipMemsetAsync(mapped, 0, 4, c.stream));
    if (nv) {
        k_vmap_flags<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.vstate.p, c.value.p, nv, flag, first);
        k_vmap_first<<<1, 64, 0, c.stream>>>(flag, first);
        scan_u32_to_u64(c, flag, c.tmpu64a.p, nv, c.cnt.p + CNT_TOTAL);
        k_vmap_ids<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(flag, c.tmpu64a.p, nv, c.vstate.p, mapped, d_vorg, d_vs);
    }
    read_counters(c, 1);
    nvars_new = nv ? (uint32_t)c.hcnt[CNT_TOTAL] : 0u;
    // clause images
    ensure(c, c.tmpu32d, n + 1);
    ensure(c, c.mem2, c.msz + 1);
    uint64_t nw = 0;
    ncls = 0;
    if (n) {
        k_vmap_sizes<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, c.tmpu32d.p);
        scan_u32_to_u64(c, c.tmpu32d.p, c.tmpu64a.p, n, c.cnt.p + CNT_TOTAL);
        k_vmap_write<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, c.tmpu64a.p, mapped, c.value.p,
                                                                    first, c.mem2.p);
        read_counters(c, 1);
        nw = c.hcnt[CNT_TOTAL];
    }
    unsigned long long *tcnt = (unsigned long long *)(c.tmpu64c.p + 1);
    PFG_HIP(hipMemsetAsync(tcnt, 0, 8, c.stream));
    if (c.ntrail) k_vmap_trail<<<grid_for(c.ntrail, 256, 1024), 256, 0, c.stream>>>(c.trail.p, c.ntrail, mapped, d_trail, tcnt);
    uint64_t nt = 0;
    PFG_HIP(hipMemcpyAsync(&nt, tcnt, 8, hipMemcpyDeviceToHost, c.stream));
    words.resize(nw);
    vorg.resize((uint64_t)nvars_new + 1);
    vstate.resize((uint64_t)nvars_new + 1);
    PFG_HIP(hipStreamSynchronize(c.stream));
    trail.resize(nt);
    if (nw) PFG_HIP(hipMemcpyAsync(words.data(), c.mem2.p, nw * 4, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipMemcpyAsync(vorg.data(), d_vorg, 4ull * (nvars_new + 1), hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipMemcpyAsync(vstate.data(), d_vs, (uint64_t)nvars_new + 1, hipMemcpyDeviceToHost, c.stream));
    if (nt) PFG_HIP(hipMemcpyAsync(trail.data(), d_trail, 4 * nt, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    for (uint64_t w = 0; w < nw; w += 3ull + words[w + 2]) ncls++;
}

// keep the elected variables of instances with cls 1, then those with cls 2, each
// group in elected order; returns the two group sizes
__global__ void k_pv_class(const uint32_t *__restrict__ pvs, uint32_t n, const uint32_t *__restrict__ ibase, uint32_t ninst,
                           const uint8_t *__restrict__ icls, uint32_t *__restrict__ fa, uint32_t *__restrict__ fb)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint8_t k = icls[find_inst(ibase, ninst, pvs[i])];
        fa[i] = k == 1;
        fb[i] = k == 2;
    }
}

void partition_pvs(Ctx &c, const std::vector<uint8_t> &cls, uint32_t &na, uint32_t &nb)
{
    na = nb = 0;
    if (!c.npvs) return;
    const uint32_t n = c.npvs;
    ensure(c, c.tmpu8, c.ninst + 8);
    ensure(c, c.tmpu32a, 2ull * n + 2);
    ensure(c, c.tmpu32b, 2ull * n + 2);
    PFG_HIP(hipMemcpyAsync(c.tmpu8.p, cls.data(), c.ninst, hipMemcpyHostToDevice, c.stream));
    uint32_t *fa = c.tmpu32a.p, *fb = c.tmpu32a.p + n + 1;
    k_pv_class<<<grid_for(n, 256, 4096), 256, 0, c.stream>>>(c.pvs.p, n, c.ibase.p, c.ninst, c.tmpu8.p, fa, fb);
    na = (uint32_t)compact_u32(c, c.pvs.p, fa, n, c.tmpu32b.p);
    nb = (uint32_t)compact_u32(c, c.pvs.p, fb, n, c.tmpu32b.p + na);
    if (na + nb) PFG_HIP(hipMemcpyAsync(c.pvs.p, c.tmpu32b.p, 4ull * (na + nb), hipMemcpyDeviceToDevice, c.stream));
    c.npvs = na + nb;
}

// ------------------------------------------------------------------ counting
// countAll; optionally also the literal histogram of the live clauses, which the
// next round's occurrence table then reuses
__global__ void k_count(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n, unsigned long long *cnt,
                        uint32_t *__restrict__ hist)
{
    unsigned long long nc = 0, nl = 0;
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *c = mem + refs[i];
        if (!c_del(c)) {
            const int sz = c_size(c);
            nc++;
            nl += (unsigned)sz;
            if (hist)
                for (int k = 0; k < sz; k++) atomicAdd(&hist[c[3 + k]], 1u);
        }
    }
    nc = block_sum(nc);
    nl = block_sum(nl);
    if (threadIdx.x == 0) { atomicAdd(&cnt[CNT_A], nc); atomicAdd(&cnt[CNT_B], nl); }
}

void count_all(Ctx &c, int64_t &cls, int64_t &lits, bool with_hist)
{
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_A, 0, 16, c.stream));
    if (with_hist) {
        ensure(c, c.hist, c.nlits2);
        PFG_HIP(hipMemsetAsync(c.hist.p, 0, c.nlits2 * 4, c.stream));
    }
    if (c.nrefs) {
        KTimer kt(c, K_COUNT, 20.0 * c.nrefs + 4.0 * c.nliterals);
        k_count<<<grid_for(c.nrefs, 256, 1024), 256, 0, c.stream>>>(c.mem.p, c.refs.p, c.nrefs, (unsigned long long *)c.cnt.p,
                                                                      with_hist ? c.hist.p : nullptr);
    }
    read_counters(c, 3);
    cls = (int64_t)c.hcnt[CNT_A];
    lits = (int64_t)c.hcnt[CNT_B];
    if (with_hist) c.hist_live = true;
}

// ------------------------------------------------------------------ garbage collection
// stable compaction of live clauses into the second pool (compactCNF, recycle.cu:71-97)
__global__ void k_gc_sizes(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                           uint64_t *__restrict__ words, uint32_t *__restrict__ live)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *c = mem + refs[i];
        const bool l = !c_del(c);
        words[i] = l ? 3ull + (uint64_t)c_size(c) : 0ull;
        live[i] = l ? 1u : 0u;
    }
}

// One wave per 64 consecutive clauses: their live images form one contiguous run of
// the new pool (woff is the scan of the live image sizes), copied one word per lane and
// step, so loads and stores are contiguous across the wave; a lane finds the clause of
// its word by a binary search over the wave's prefix of image sizes.
__global__ void __launch_bounds__(256) k_gc_move(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs,
                                                uint64_t n, const uint64_t *__restrict__ woff,
                                                const uint64_t *__restrict__ idx, const uint32_t *__restrict__ live,
                                                uint32_t *__restrict__ mem2, uint64_t *__restrict__ refs2)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t base = wave * 64; base < n; base += nwaves * 64) {
        const uint64_t i = base + lane;
        uint64_t r = 0;
        uint32_t words = 0;
        if (i < n && live[i]) {
            r = refs[i];
            words = 3u + (uint32_t)c_size(mem + r);
            refs2[idx[i]] = woff[i];
        }
        uint32_t incl = words;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if ((int)lane >= o) incl += t;
        }
        const uint32_t pre = incl - words, total = __shfl(incl, 63, 64);
        const uint64_t out0 = woff[base];
        const uint32_t rlo = (uint32_t)r, rhi = (uint32_t)(r >> 32);
        for (uint32_t ob = 0; ob < total; ob += 64) {
            const uint32_t o = ob + lane;
            uint32_t m = 0;
#pragma unroll
            for (uint32_t step = 32; step > 0; step >>= 1) {
                const uint32_t cand = m + step;   // at most 63
                if (__shfl(pre, (int)cand, 64) <= o) m = cand;
            }
            const uint32_t pm = __shfl(pre, (int)m, 64);
            const uint64_t rm = ((uint64_t)(uint32_t)__shfl(rhi, (int)m, 64) << 32) | (uint32_t)__shfl(rlo, (int)m, 64);
            if (o < total) mem2[out0 + o] = mem[rm + (o - pm)];
        }
    }
}

void shrink(Ctx &c)
{
    const uint64_t n = c.nrefs;
    if (!n) return;
    ensure(c, c.tmpu64a, n);
    ensure(c, c.tmpu64b, n);
    ensure(c, c.tmpu64c, n);
    ensure(c, c.tmpu32a, n);
    k_gc_sizes<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, c.tmpu64a.p, c.tmpu32a.p);
    scan_u64(c, c.tmpu64a.p, c.tmpu64b.p, n, c.cnt.p + CNT_A);
    scan_u32_to_u64(c, c.tmpu32a.p, c.tmpu64c.p, n, c.cnt.p + CNT_B);
    read_counters(c, 3);
    const uint64_t words = c.hcnt[CNT_A], live = c.hcnt[CNT_B];
    ensure(c, c.mem2, std::max<uint64_t>(c.mem.cap, words + 1));
    ensure(c, c.refs2, std::max<uint64_t>(c.refs.cap, live + 1));
    {
        KTimer kt(c, K_GC, 8.0 * words + 36.0 * n);
        k_gc_move<<<grid_for(n, 256, 16384), 256, 0, c.stream>>>(c.mem.p, c.refs.p, n, c.tmpu64b.p, c.tmpu64c.p,
                                                                       c.tmpu32a.p, c.mem2.p, c.refs2.p);
    }
    std::swap(c.mem, c.mem2);
    std::swap(c.refs, c.refs2);
    c.msz = words;
    c.nrefs = live;
    c.ot_valid = false;
    c.otk_valid = false;   // references moved: the next table is sorted in full
}

// ------------------------------------------------------------------ ordered compaction
__global__ void k_compact_scatter(const uint32_t *__restrict__ in, const uint32_t *__restrict__ flags,
                                  const uint64_t *__restrict__ pos, uint64_t n, uint32_t *__restrict__ out)
{
    for (uint64_t i = gtid(); i < n; i += gstride())
        if (flags[i]) out[pos[i]] = in[i];
}

uint64_t compact_u32(Ctx &c, const uint32_t *in, const uint32_t *flags, uint64_t n, uint32_t *out)
{
    if (!n) return 0;
    ensure(c, c.tmpu64c, n);
    scan_u32_to_u64(c, flags, c.tmpu64c.p, n, c.cnt.p + CNT_A);
    k_compact_scatter<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(in, flags, c.tmpu64c.p, n, out);
    read_counters(c, 2);
    return c.hcnt[CNT_A];
}

// filterPVs (src/cpu/solve.h:697-707)
__global__ void k_pv_flags(const uint32_t *__restrict__ pvs, uint64_t n, const uint8_t *__restrict__ vstate, uint32_t *__restrict__ f)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t x = pvs[i];
        f[i] = (x && !vstate[x]) ? 1u : 0u;
    }
}

void filter_pvs(Ctx &c)
{
    if (!c.npvs) return;
    ensure(c, c.tmpu32a, c.npvs);
    ensure(c, c.tmpu32b, c.npvs);
    k_pv_flags<<<grid_for(c.npvs, 256, 4096), 256, 0, c.stream>>>(c.pvs.p, c.npvs, c.vstate.p, c.tmpu32a.p);
    const uint64_t n = compact_u32(c, c.pvs.p, c.tmpu32a.p, c.npvs, c.tmpu32b.p);
    if (n) PFG_HIP(hipMemcpyAsync(c.pvs.p, c.tmpu32b.p, n * 4, hipMemcpyDeviceToDevice, c.stream));
    c.npvs = (uint32_t)n;
}

// ------------------------------------------------------------------ units from passes
// root units given by the caller: each is assigned by an atomic exchange on its
// variable's value, so a repeated unit is kept once and a unit whose negation is also
// given reports a contradiction, as in the sequential enqueue; only the order in
// which they enter the trail differs (no pass reads it)
__global__ void k_assign_host_units(const uint32_t *__restrict__ lits, uint32_t n, int8_t *value, uint8_t *vstate,
                                    uint32_t *trail, unsigned long long *cnt)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) enqueue_dev(value, vstate, trail, cnt, lits[i]);
}

__global__ void k_set_ifrozen(const uint32_t *__restrict__ vars, uint32_t n, uint8_t on, uint8_t *__restrict__ ifrozen)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) ifrozen[vars[i]] = on;
}

// ifreeze / iunfreeze (src/cpu/assume.cpp:45-61) of host-given variables
void set_ifrozen(Ctx &c, const uint32_t *vars, uint32_t n, bool on)
{
    if (!n) return;
    ensure(c, c.tmpu32a, n);
    PFG_HIP(hipMemcpyAsync(c.tmpu32a.p, vars, n * 4ull, hipMemcpyHostToDevice, c.stream));
    k_set_ifrozen<<<grid_for(n, 256, 1024), 256, 0, c.stream>>>(c.tmpu32a.p, n, on ? 1 : 0, c.ifrozen.p);
    PFG_HIP(hipStreamSynchronize(c.stream));
}

void enqueue_units_host(Ctx &c, const uint32_t *lits, uint32_t n)
{
    if (!n) return;
    ensure(c, c.tmpu32a, n);
    PFG_HIP(hipMemcpyAsync(c.tmpu32a.p, lits, n * 4, hipMemcpyHostToDevice, c.stream));
    c.hist_live = false;
    k_assign_host_units<<<grid_for(n, 256, 4096), 256, 0, c.stream>>>(c.tmpu32a.p, n, c.value.p, c.vstate.p, c.trail.p,
                                                                      (unsigned long long *)c.cnt.p);
    read_counters(c, 10);
    c.ntrail = (uint32_t)c.hcnt[CNT_NTRAIL];
    if (c.hcnt[CNT_FLAGS] & FLAG_UNSAT) c.unsat = true;
}

// Units found by BVE (newResolvent / substitute_single, bve.cpp:199-211, equivalence.h:55-84)
// in PV order: the first occurrence of a variable is enqueued if unassigned; a later
// occurrence of its negation, or a unit that is already false, proves a contradiction.
__global__ void k_units_first(const uint32_t *__restrict__ u, uint64_t n, uint32_t *__restrict__ first)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) atomicMin(&first[u[i] >> 1], (uint32_t)i);
}

__global__ void k_units_keep(const uint32_t *__restrict__ u, uint64_t n, const uint32_t *__restrict__ first,
                             const int8_t *__restrict__ value, uint32_t *__restrict__ keep, unsigned long long *cnt)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t l = u[i];
        const int8_t v = value[l];
        uint32_t k = 0;
        if (v >= 0) { if (v == 0) report_conflict(cnt, l >> 1); }
        else {
            const uint32_t f = first[l >> 1];
            if (f == (uint32_t)i) k = 1;
            else if (u[f] != l) report_conflict(cnt, l >> 1);
        }
        keep[i] = k;
    }
}

__global__ void k_units_append(const uint32_t *__restrict__ kept, uint64_t n, uint32_t base, int8_t *value, uint8_t *vstate,
                               uint32_t *trail, uint32_t *__restrict__ first)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t l = kept[i];
        value[l] = 1;
        value[l ^ 1u] = 0;
        vstate[l >> 1] = FROZEN_M;
        trail[base + i] = l;
    }
}

__global__ void k_units_reset(const uint32_t *__restrict__ u, uint64_t n, uint32_t *__restrict__ first)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) first[u[i] >> 1] = 0xFFFFFFFFu;
}

void merge_units(Ctx &c, const uint32_t *d_units, uint64_t n)
{
    if (!n) return;
    // ufirst: per-variable first occurrence, all-ones outside this function
    ensure(c, c.tmpu32a, n);
    ensure(c, c.tmpu32b, n);
    k_units_first<<<grid_for(n, 256, 4096), 256, 0, c.stream>>>(d_units, n, c.ufirst.p);
    k_units_keep<<<grid_for(n, 256, 4096), 256, 0, c.stream>>>(d_units, n, c.ufirst.p, c.value.p, c.tmpu32a.p,
                                                                (unsigned long long *)c.cnt.p);
    k_units_reset<<<grid_for(n, 256, 4096), 256, 0, c.stream>>>(d_units, n, c.ufirst.p);
    const uint64_t nk = compact_u32(c, d_units, c.tmpu32a.p, n, c.tmpu32b.p);
    if (nk) k_units_append<<<grid_for(nk, 256, 4096), 256, 0, c.stream>>>(c.tmpu32b.p, nk, c.ntrail, c.value.p, c.vstate.p,
                                                                          c.trail.p, c.ufirst.p);
    c.ntrail += (uint32_t)nk;
    {
        const unsigned long long nt = c.ntrail;
        // keep the device trail counter in step with the host
        PFG_HIP(hipMemcpyAsync(c.cnt.p + CNT_NTRAIL, &nt, 8, hipMemcpyHostToDevice, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    read_counters(c, 10);
    if (c.hcnt[CNT_FLAGS] & FLAG_UNSAT) c.unsat = true;
}

// ------------------------------------------------------------------ PV list sort (sortOT)
// CNF_CMP_KEY (src/cpu/simplify.h:29-43): size, first literal, last literal, signature.
struct KeyLess {
    const uint32_t *mem;
    __device__ bool operator()(uint64_t x, uint64_t y) const
    {
        const uint32_t *a = mem + x, *b = mem + y;
        const int xs = (int)a[2], ys = (int)b[2];
        if (xs != ys) return xs < ys;
        if (a[3] != b[3]) return a[3] < b[3];
        const uint32_t xb = a[3 + xs - 1], yb = b[3 + ys - 1];
        if (xb != yb) return xb < yb;
        return a[1] < b[1];
    }
};

// libstdc++ std::sort (introsort) order, restated for the device
template <typename T, typename LT>
__device__ void is_adjust_heap(T *f, int64_t hole, int64_t len, T value, LT lt)
{
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (lt(f[child], f[child - 1])) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    int64_t parent = (hole - 1) / 2;
    while (hole > top && lt(f[parent], value)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = value;
}

template <typename T, typename LT>
__device__ void is_heapsort(T *f, int64_t len, LT lt)
{
    if (len >= 2) {
        for (int64_t parent = (len - 2) / 2;; parent--) {
            is_adjust_heap(f, parent, len, f[parent], lt);
            if (parent == 0) break;
        }
    }
    while (len > 1) {
        len--;
        const T v = f[len];
        f[len] = f[0];
        is_adjust_heap(f, 0, len, v, lt);
    }
}

template <typename T, typename LT>
__device__ void is_linear_insert(T *last, LT lt)
{
    const T v = *last;
    T *next = last - 1;
    while (lt(v, *next)) { *last = *next; last = next; --next; }
    *last = v;
}

template <typename T, typename LT>
__device__ void is_insertion(T *first, T *last, LT lt)
{
    if (first == last) return;
    for (T *i = first + 1; i != last; ++i) {
        if (lt(*i, *first)) {
            const T v = *i;
            for (T *p = i; p != first; --p) *p = *(p - 1);
            *first = v;
        }
        else is_linear_insert(i, lt);
    }
}

template <typename T, typename LT>
__device__ void introsort_dev(T *d, int64_t n, LT lt)
{
    if (n < 2) return;
    struct Frame { uint32_t f, l, depth; };   // lists hold fewer than 2^32 entries
    Frame st[72];
    int sp = 0;
    int64_t lg = 63 - __clzll((long long)n);
    st[sp++] = {0u, (uint32_t)n, (uint32_t)(2 * lg)};
    while (sp) {
        Frame fr = st[--sp];
        int64_t f = fr.f, l = fr.l, depth = fr.depth;
        while (l - f > 16) {
            if (depth == 0) { is_heapsort(d + f, l - f, lt); break; }
            --depth;
            T *first = d + f, *last = d + l;
            T *a = first + 1, *b = first + (l - f) / 2, *cc = last - 1, *pick;
            if (lt(*a, *b)) {
                if (lt(*b, *cc)) pick = b;
                else if (lt(*a, *cc)) pick = cc;
                else pick = a;
            }
            else if (lt(*a, *cc)) pick = a;
            else if (lt(*b, *cc)) pick = cc;
            else pick = b;
            { const T t = *first; *first = *pick; *pick = t; }
            T *lo = first + 1, *hi = last;
            for (;;) {
                while (lt(*lo, *first)) ++lo;
                --hi;
                while (lt(*first, *hi)) --hi;
                if (!(lo < hi)) break;
                const T t = *lo; *lo = *hi; *hi = t;
                ++lo;
            }
            const int64_t cut = lo - d;
            st[sp++] = {(uint32_t)cut, (uint32_t)l, (uint32_t)depth};   // the right part is sorted independently
            l = cut;
        }
    }
    if (n > 16) {
        is_insertion(d, d + 16, lt);
        for (T *i = d + 16; i != d + n; ++i) is_linear_insert(i, lt);
    }
    else is_insertion(d, d + n, lt);
}

template <typename LT>
__device__ void stable_sort_dev(uint64_t *d, uint64_t *tmp, int64_t n, LT lt)
{
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            const int64_t mid = min(lo + w, n), hi = min(lo + 2 * w, n);
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = lt(d[j], d[i]) ? d[j++] : d[i++];
            while (i < mid) tmp[k++] = d[i++];
            while (j < hi) tmp[k++] = d[j++];
        }
        for (int64_t k = 0; k < n; k++) d[k] = tmp[k];
    }
}

// A list of up to SORT_SHORT entries is sorted in LDS: each entry's CNF_CMP_KEY is
// read from its clause once, as two words compared in order (size, first literal |
// last literal, signature), so the insertion sort compares registers instead of
// gathering clause headers per comparison; the comparisons and their outcomes are
// those of KeyLess, so the order is the same.
constexpr int SORT_SHORT = 20;

__device__ __forceinline__ void short_key(const uint32_t *mem, uint64_t r, uint64_t &hi, uint64_t &lo)
{
    const uint32_t *c = mem + r;
    const uint32_t sz = c[2];
    const uint32_t first = sz ? c[3] : 0u, last = sz ? c[3 + sz - 1] : 0u;
    hi = ((uint64_t)sz << 32) | first;
    lo = ((uint64_t)last << 32) | c[1];
}

// A longer list is sorted by the introsort over entries that carry their key (the two
// words of short_key) next to the reference, in a region of a scratch buffer taken by
// an atomic: the introsort makes the comparisons KeyLess would make, with the same
// outcomes, so it moves the entries exactly as it would move the references, while a
// comparison reads the entries instead of four scattered clause words each.
struct SortEnt {
    uint64_t hi, lo, ref;
};
struct EntLess {
    __device__ bool operator()(const SortEnt &a, const SortEnt &b) const
    {
        return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
    }
};

// work (optional): the list indices to sort, *nwork of them (read on the device);
// without it every elected variable's two lists
__global__ void __launch_bounds__(64) k_sort_pv_lists(const uint32_t *__restrict__ pvs, uint32_t npvs,
                                                     const OL *__restrict__ ot, uint64_t *occ,
                                                     const uint32_t *__restrict__ mem, int stable, uint64_t *tmp,
                                                     SortEnt *ents, unsigned long long *nents,
                                                     const uint32_t *__restrict__ work,
                                                     const unsigned long long *__restrict__ nwork,
                                                     unsigned long long *nexact)
{
    __shared__ uint64_t s_hi[SORT_SHORT][64], s_lo[SORT_SHORT][64], s_ref[SORT_SHORT][64];
    const int lane = threadIdx.x;
    const KeyLess lt{mem};
    const uint64_t nq = work ? *nwork : 2ull * npvs;
    if (work && gtid() == 0 && nq) atomicAdd(nexact, nq);   // lists sorted here, summed over calls
    for (uint64_t k = gtid(); k < nq; k += gstride()) {
        const uint64_t q = work ? work[k] : k;
        const uint32_t lit = 2 * pvs[q >> 1] + (uint32_t)(q & 1);
        const OL o = ot[lit];
        uint64_t *d = occ + o.off;
        const int64_t n = o.sz;
        if (n <= 1) continue;
        if (n <= SORT_SHORT) {   // insertionSort (src/cpu/sort.h:41-54) on (key, reference)
            int i = 0;
            for (; i + 4 <= (int)n; i += 4) {   // four entries' loads in flight together
                const uint64_t r0 = d[i], r1 = d[i + 1], r2 = d[i + 2], r3 = d[i + 3];
                const uint32_t *c0 = mem + r0, *c1 = mem + r1, *c2 = mem + r2, *c3 = mem + r3;
                const uint32_t z0 = c0[2], z1 = c1[2], z2 = c2[2], z3 = c3[2];
                const uint32_t f0 = z0 ? c0[3] : 0u, f1 = z1 ? c1[3] : 0u, f2 = z2 ? c2[3] : 0u, f3 = z3 ? c3[3] : 0u;
                const uint32_t g0 = c0[1], g1 = c1[1], g2 = c2[1], g3 = c3[1];
                const uint32_t l0 = z0 ? c0[2 + z0] : 0u, l1 = z1 ? c1[2 + z1] : 0u, l2 = z2 ? c2[2 + z2] : 0u,
                               l3 = z3 ? c3[2 + z3] : 0u;
                s_hi[i][lane] = ((uint64_t)z0 << 32) | f0;
                s_hi[i + 1][lane] = ((uint64_t)z1 << 32) | f1;
                s_hi[i + 2][lane] = ((uint64_t)z2 << 32) | f2;
                s_hi[i + 3][lane] = ((uint64_t)z3 << 32) | f3;
                s_lo[i][lane] = ((uint64_t)l0 << 32) | g0;
                s_lo[i + 1][lane] = ((uint64_t)l1 << 32) | g1;
                s_lo[i + 2][lane] = ((uint64_t)l2 << 32) | g2;
                s_lo[i + 3][lane] = ((uint64_t)l3 << 32) | g3;
                s_ref[i][lane] = r0;
                s_ref[i + 1][lane] = r1;
                s_ref[i + 2][lane] = r2;
                s_ref[i + 3][lane] = r3;
            }
            for (; i < (int)n; i++) {
                const uint64_t r = d[i];
                short_key(mem, r, s_hi[i][lane], s_lo[i][lane]);
                s_ref[i][lane] = r;
            }
            auto less = [&](uint64_t ah, uint64_t al, int j) {
                const uint64_t bh = s_hi[j][lane];
                return ah < bh || (ah == bh && al < s_lo[j][lane]);
            };
            if (n == 2) {
                if (less(s_hi[1][lane], s_lo[1][lane], 0)) { d[0] = s_ref[1][lane]; d[1] = s_ref[0][lane]; }
                continue;
            }
            for (int i = 1; i < (int)n; i++) {
                const uint64_t th = s_hi[i][lane], tl = s_lo[i][lane], tr = s_ref[i][lane];
                int j;
                for (j = i; j > 0 && less(th, tl, j - 1); j--) {
                    s_hi[j][lane] = s_hi[j - 1][lane];
                    s_lo[j][lane] = s_lo[j - 1][lane];
                    s_ref[j][lane] = s_ref[j - 1][lane];
                }
                s_hi[j][lane] = th;
                s_lo[j][lane] = tl;
                s_ref[j][lane] = tr;
            }
            for (int i = 0; i < (int)n; i++) d[i] = s_ref[i][lane];
        }
        else if (stable) stable_sort_dev(d, tmp + o.off, n, lt);
        else {
            SortEnt *e = ents + atomicAdd(nents, (unsigned long long)n);
            for (int64_t i = 0; i < n; i++) {
                const uint64_t r = d[i];
                uint64_t hi, lo;
                short_key(mem, r, hi, lo);
                e[i] = SortEnt{hi, lo, r};
            }
            introsort_dev(e, n, EntLess{});
            for (int64_t i = 0; i < n; i++) d[i] = e[i].ref;
        }
    }
}

// Wave-cooperative sortOT. A wave takes 64 elected lists (one per lane), stages the
// keys (the two words of short_key) and references of as many of them as fit in
// SPW_CAP entries into its LDS (cooperative loads: a lane per entry, the entry's list
// found by a binary search over the wave's prefix of list sizes), and places every
// entry at its stable rank in its list: the entries with a smaller key plus the equal
// ones before it. That is the reference's insertion sort for lists of up to
// SORT_SHORT entries (stable) and, when no two keys of a list are equal, the one
// order any sort gives; a longer list with equal keys (where std::sort's own tie
// order counts), or one longer than SPW_MAX, goes to the work list for
// k_sort_pv_lists. With stable_sort on, stable ranks are the order for every list.
constexpr uint32_t SPW_CAP = 512, SPW_MAX = 256;

__global__ void __launch_bounds__(64) k_sort_pv_wave(const uint32_t *__restrict__ pvs, uint32_t npvs,
                                                    const OL *__restrict__ ot, uint64_t *occ,
                                                    const uint32_t *__restrict__ mem, int stable,
                                                    uint32_t *__restrict__ work, unsigned long long *nwork)
{
    __shared__ uint64_t s_hi[SPW_CAP], s_lo[SPW_CAP], s_ref[SPW_CAP];
    __shared__ uint32_t s_tie[64];
    const uint32_t lane = threadIdx.x;
    const uint64_t nq = 2ull * npvs;
    for (uint64_t base = (uint64_t)blockIdx.x * 64; base < nq; base += (uint64_t)gridDim.x * 64) {
        const uint64_t q = base + lane;
        OL o{0, 0, 0};
        if (q < nq) o = ot[2 * pvs[q >> 1] + (uint32_t)(q & 1)];
        const bool toolong = o.sz > SPW_MAX;
        wave_append(toolong, (uint32_t)q, work, nwork);
        bool pending = o.sz > 1 && !toolong;
        while (__ballot(pending)) {
            // this round: the pending lists, in lane order, whose entries fit together
            const uint32_t want = pending ? o.sz : 0u;
            uint32_t incl = want;
#pragma unroll
            for (int k = 1; k < 64; k <<= 1) {
                const uint32_t t = __shfl_up(incl, k, 64);
                if ((int)lane >= k) incl += t;
            }
            const bool take = pending && incl <= SPW_CAP;
            const uint32_t n = take ? o.sz : 0u;
            uint32_t pre = n;
#pragma unroll
            for (int k = 1; k < 64; k <<= 1) {
                const uint32_t t = __shfl_up(pre, k, 64);
                if ((int)lane >= k) pre += t;
            }
            const uint32_t total = __shfl(pre, 63, 64);
            pre -= n;
            const uint32_t olo = (uint32_t)o.off, ohi = (uint32_t)(o.off >> 32);
            s_tie[lane] = 0;
            // stage keys and references
            for (uint32_t e0 = 0; e0 < total; e0 += 64) {
                const uint32_t e = e0 + lane;
                uint32_t m = 0;
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1)
                    if ((uint32_t)__shfl((int)pre, (int)(m + step), 64) <= e) m += step;
                const uint32_t pm = (uint32_t)__shfl((int)pre, (int)m, 64);
                const uint64_t om = ((uint64_t)(uint32_t)__shfl((int)ohi, (int)m, 64) << 32) |
                                    (uint32_t)__shfl((int)olo, (int)m, 64);
                if (e < total) {
                    const uint64_t r = occ[om + (e - pm)];
                    short_key(mem, r, s_hi[e], s_lo[e]);
                    s_ref[e] = r;
                }
            }
            __syncthreads();
            // stable ranks (two passes over the staged entries: ties first, then the writes)
            uint32_t rk[SPW_CAP / 64];
#pragma unroll
            for (uint32_t k = 0; k < SPW_CAP / 64; k++) {
                const uint32_t e = k * 64 + lane;
                uint32_t m = 0;
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1)
                    if ((uint32_t)__shfl((int)pre, (int)(m + step), 64) <= e) m += step;
                const uint32_t pm = (uint32_t)__shfl((int)pre, (int)m, 64);
                const uint32_t nm = (uint32_t)__shfl((int)n, (int)m, 64);
                rk[k] = 0;
                if (e < total) {
                    const uint64_t h = s_hi[e], l = s_lo[e];
                    uint32_t r = 0;
                    bool tie = false;
                    for (uint32_t j = pm; j < pm + nm; j++) {
                        const uint64_t hj = s_hi[j], lj = s_lo[j];
                        if (hj < h || (hj == h && lj < l)) r++;
                        else if (hj == h && lj == l && j != e) {
                            tie = true;
                            if (j < e) r++;
                        }
                    }
                    rk[k] = r;
                    if (tie && !stable && nm > (uint32_t)SORT_SHORT) s_tie[m] = 1;
                }
            }
            __syncthreads();
            // lists with equal keys past SORT_SHORT entries go to the exact introsort
            wave_append(take && s_tie[lane] != 0, (uint32_t)q, work, nwork);
#pragma unroll
            for (uint32_t k = 0; k < SPW_CAP / 64; k++) {
                const uint32_t e = k * 64 + lane;
                uint32_t m = 0;
#pragma unroll
                for (uint32_t step = 32; step > 0; step >>= 1)
                    if ((uint32_t)__shfl((int)pre, (int)(m + step), 64) <= e) m += step;
                const uint64_t om = ((uint64_t)(uint32_t)__shfl((int)ohi, (int)m, 64) << 32) |
                                    (uint32_t)__shfl((int)olo, (int)m, 64);
                if (e < total && s_tie[m] == 0) occ[om + rk[k]] = s_ref[e];
            }
            __syncthreads();
            if (take) pending = false;
        }
    }
}

// entries of the lists longer than SORT_SHORT (the introsort's scratch entries)
__global__ void k_long_entries(const uint32_t *__restrict__ pvs, uint32_t npvs, const OL *__restrict__ ot,
                               unsigned long long *out)
{
    unsigned long long acc = 0;
    for (uint64_t q = gtid(); q < 2ull * npvs; q += gstride()) {
        const uint32_t sz = ot[2 * pvs[q >> 1] + (uint32_t)(q & 1)].sz;
        if (sz > (uint32_t)SORT_SHORT) acc += sz;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0 && acc) atomicAdd(out, acc);
}

__global__ void k_pv_bytes(const uint32_t *__restrict__ pvs, uint32_t b, uint32_t e, const OL *__restrict__ ot,
                           const uint64_t *__restrict__ occ, const uint32_t *__restrict__ mem, int with_clauses,
                           unsigned long long *__restrict__ out)
{
    unsigned long long acc = 0;
    for (uint64_t i = 2ull * b + gtid(); i < 2ull * e; i += gstride()) {
        const uint32_t lit = 2u * pvs[i >> 1] + (uint32_t)(i & 1u);
        const OL l = ot[lit];
        acc += 16ull + 8ull * l.sz;
        if (with_clauses)
            for (uint32_t j = 0; j < l.sz; j++) acc += 12ull + 4ull * mem[occ[l.off + j] + 2];
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0 && acc) atomicAdd(out, acc);
}

double pv_bytes(Ctx &c, uint32_t b, uint32_t e, bool with_clauses)
{
    if (!c.account || e <= b) return 0.0;
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_ACCT, 0, 8, c.stream));
    k_pv_bytes<<<grid_for(2ull * (e - b), 256, 4096), 256, 0, c.stream>>>(c.pvs.p, b, e, c.ot.p, c.occ.p, c.mem.p,
                                                                         with_clauses ? 1 : 0,
                                                                         (unsigned long long *)c.cnt.p + CNT_ACCT);
    uint64_t v = 0;
    PFG_HIP(hipMemcpyAsync(&v, c.cnt.p + CNT_ACCT, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    return (double)v;
}

void sort_ot(Ctx &c)
{
    if (!c.npvs) return;
    uint64_t *tmp = nullptr;
    SortEnt *ents = nullptr;
    if (c.opts.stable_sort) { ensure(c, c.tmpu64b, c.occ.cap); tmp = c.tmpu64b.p; }
    else {
        PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_A, 0, 16, c.stream));
        k_long_entries<<<grid_for(2ull * c.npvs, 256, 2048), 256, 0, c.stream>>>(c.pvs.p, c.npvs, c.ot.p,
                                                                                (unsigned long long *)c.cnt.p + CNT_A);
        read_counters(c, CNT_A + 1);
        const uint64_t ne = c.hcnt[CNT_A];
        if (ne) {
            ensure(c, c.tmpu64b, 3 * ne);   // SortEnt = 3 words
            ents = (SortEnt *)c.tmpu64b.p;
        }
    }
    ensure(c, c.tmpu32a, 2ull * c.npvs + 1);   // lists left to k_sort_pv_lists, count in CNT_C
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_C, 0, 8, c.stream));
    {
        KTimer kt(c, K_SORT_PV, 2.0 * pv_bytes(c, 0, c.npvs, false));   // lists read and written once
        k_sort_pv_wave<<<grid_for(2ull * c.npvs, 64, 16384), 64, 0, c.stream>>>(c.pvs.p, c.npvs, c.ot.p, c.occ.p, c.mem.p,
                                                                                c.opts.stable_sort, c.tmpu32a.p,
                                                                                (unsigned long long *)c.cnt.p + CNT_C);
    }
    k_sort_pv_lists<<<grid_for(2ull * c.npvs, 64, 2048), 64, 0, c.stream>>>(c.pvs.p, c.npvs, c.ot.p, c.occ.p, c.mem.p,
                                                                             c.opts.stable_sort, tmp, ents,
                                                                             (unsigned long long *)c.cnt.p + CNT_B,
                                                                             c.tmpu32a.p, (unsigned long long *)c.cnt.p + CNT_C,
                                                                             (unsigned long long *)c.cnt.p + CNT_SORTX);
}

}  // namespace pfg
