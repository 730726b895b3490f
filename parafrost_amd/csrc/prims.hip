// Exclusive scan and stable LSD radix sort for wave64 (CDNA4).
//  scan : reduce-then-scan over 2048-element tiles (256 threads x 8 items),
//         coalesced tile loads through LDS, wave scans by __shfl_up.
//  radix: digits of at most 8 bits, balanced over the passes, over 8192-element tiles
//         (512 threads x 16 items); per-tile LDS histograms (wave-aggregated),
//         stored tile-major, column scan to per-tile digit offsets, stable scatter ranked
//         per wave with match-any ballots, tile held in registers.
#include "prims.hpp"
#include "device.hpp"

namespace pfg {

constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int RS_BLOCK = 512, RS_ITEMS = 16, RS_TILE = RS_BLOCK * RS_ITEMS;

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <typename Tin>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_sums(const Tin *__restrict__ in, uint64_t n, uint64_t *__restrict__ sums)
{
    __shared__ uint64_t ws[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const uint64_t i = base + (uint64_t)k * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += (uint64_t)in[i];
    }
    s = wave_sum(s);
    if (__lane_id() == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; w++) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

template <typename Tin>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(const Tin *in, uint64_t *out, uint64_t n,
                                                           const uint64_t *__restrict__ offs, uint64_t *d_total)
{
    __shared__ uint64_t tile[SCAN_TILE];
    __shared__ uint64_t ws[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const uint64_t i = base + (uint64_t)k * SCAN_BLOCK + threadIdx.x;
        tile[k * SCAN_BLOCK + threadIdx.x] = i < n ? (uint64_t)in[i] : 0;
    }
    __syncthreads();
    uint64_t v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) { v[j] = tile[threadIdx.x * SCAN_ITEMS + j]; s += v[j]; }
    const uint64_t inc = wave_incl_scan(s);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 63) ws[w] = inc;
    __syncthreads();
    uint64_t pre = 0, tot = 0;
    for (int q = 0; q < SCAN_BLOCK / 64; q++) { if (q < w) pre += ws[q]; tot += ws[q]; }
    const uint64_t off = offs ? offs[blockIdx.x] : 0;
    uint64_t run = off + pre + inc - s;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) { tile[threadIdx.x * SCAN_ITEMS + j] = run; run += v[j]; }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const uint64_t i = base + (uint64_t)k * SCAN_BLOCK + threadIdx.x;
        if (i < n) out[i] = tile[k * SCAN_BLOCK + threadIdx.x];
    }
    if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *d_total = off + tot;
}

static uint64_t scan_scratch_words(uint64_t n)
{
    uint64_t w = 0;
    uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    while (nb > 1) { w += nb; nb = (nb + SCAN_TILE - 1) / SCAN_TILE; }
    return w + 8;
}

template <typename Tin>
static void scan_rec(Ctx &c, const Tin *in, uint64_t *out, uint64_t n, uint64_t *d_total, uint64_t *scratch)
{
    if (n == 0) {
        if (d_total) PFG_HIP(hipMemsetAsync(d_total, 0, sizeof(uint64_t), c.stream));
        return;
    }
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb == 1) {
        k_scan_apply<Tin><<<1, SCAN_BLOCK, 0, c.stream>>>(in, out, n, nullptr, d_total);
        return;
    }
    uint64_t *sums = scratch;
    k_scan_sums<Tin><<<(unsigned)nb, SCAN_BLOCK, 0, c.stream>>>(in, n, sums);
    scan_rec<uint64_t>(c, sums, sums, nb, nullptr, scratch + nb);
    k_scan_apply<Tin><<<(unsigned)nb, SCAN_BLOCK, 0, c.stream>>>(in, out, n, sums, d_total);
}

void scan_u32_to_u64(Ctx &c, const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *d_total)
{
    ensure(c, c.scan_tmp, scan_scratch_words(n) * 8);
    scan_rec<uint32_t>(c, in, out, n, d_total, (uint64_t *)c.scan_tmp.p);
}

void scan_u64(Ctx &c, const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *d_total)
{
    ensure(c, c.scan_tmp, scan_scratch_words(n) * 8);
    scan_rec<uint64_t>(c, in, out, n, d_total, (uint64_t *)c.scan_tmp.p);
}

// match-any on a DB-bit digit: the valid lanes of the wave holding the same digit
template <int DB>
__device__ __forceinline__ uint64_t match_digit(bool valid, uint32_t d)
{
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < DB; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(valid && bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

// Per-tile digit counts (NB = 2^DB digits). All loads of the tile are issued first;
// lanes of a wave that share a digit are counted by their lowest lane with one LDS
// atomic (skewed digits would otherwise serialize on one counter).
template <typename K, int DB>
__global__ void __launch_bounds__(RS_BLOCK) k_radix_hist(const K *__restrict__ keys, uint64_t n, int shift,
                                                         uint32_t *__restrict__ counts, uint32_t ch,
                                                         unsigned long long *__restrict__ part)
{
    constexpr int NB = 1 << DB;
    __shared__ uint32_t h[NB];
    for (int d = threadIdx.x; d < NB; d += RS_BLOCK) h[d] = 0;
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    K kk[RS_ITEMS];
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint64_t i = base + (uint64_t)k * RS_BLOCK + threadIdx.x;
        kk[k] = i < n ? keys[i] : (K)0;
    }
    __syncthreads();
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const bool valid = base + (uint64_t)k * RS_BLOCK + threadIdx.x < n;
        const uint32_t d = (uint32_t)(kk[k] >> shift) & (NB - 1u);
        const uint64_t m = match_digit<DB>(valid, d);
        if (valid && (m & lt) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += RS_BLOCK) {
        const uint32_t c = h[d];
        counts[(uint64_t)blockIdx.x * NB + d] = c;
        if (c) atomicAdd(&part[(uint64_t)(blockIdx.x / ch) * NB + d], (unsigned long long)c);
    }
}

// Digit offsets of every tile from the tile-major counts (counts[t * NB + d]):
// offs[t * NB + d] = (keys of digits < d) + (keys of digit d in tiles < t).
// Tiles form at most 32 chunks of ch tiles; the histogram kernel adds each tile's
// counts into its chunk's row (part), one block scans the chunk rows, and one block
// per chunk walks its tiles (thread d down column d).
// one block of max(NB, 64) threads: chunk sums -> chunk starts (per digit, then the digit base added)
template <int DB>
__global__ void __launch_bounds__(256) k_radix_coltop(uint64_t *__restrict__ part, uint32_t nchunks)
{
    constexpr int NB = 1 << DB;
    __shared__ uint64_t wsum[NB / 64 > 0 ? NB / 64 : 1];
    const uint32_t d = threadIdx.x, w = d >> 6;
    uint64_t s = 0;
    for (uint32_t b = 0; b < nchunks && d < (uint32_t)NB; b++) {
        const uint64_t v = part[(uint64_t)b * NB + d];
        part[(uint64_t)b * NB + d] = s;
        s += v;
    }
    const uint64_t incl = wave_incl_scan(s);
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    uint64_t base = incl - s;
    for (uint32_t q = 0; q < w; q++) base += wsum[q];
    for (uint32_t b = 0; b < nchunks && d < (uint32_t)NB; b++) part[(uint64_t)b * NB + d] += base;
}

// one block of NB threads per chunk
template <int DB>
__global__ void __launch_bounds__(256) k_radix_colapply(const uint32_t *__restrict__ counts, uint32_t ntiles, uint32_t ch,
                                                        const uint64_t *__restrict__ part, uint64_t *__restrict__ offs)
{
    constexpr int NB = 1 << DB;
    const uint32_t d = threadIdx.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * ch, t1 = t0 + ch < ntiles ? t0 + ch : ntiles;
    uint64_t run = part[(uint64_t)blockIdx.x * NB + d];
    uint64_t t = t0;
    for (; t + 4 <= t1; t += 4) {   // four independent loads in flight
        const uint32_t c0 = counts[t * NB + d], c1 = counts[(t + 1) * NB + d];
        const uint32_t c2 = counts[(t + 2) * NB + d], c3 = counts[(t + 3) * NB + d];
        offs[t * NB + d] = run;
        offs[(t + 1) * NB + d] = run + c0;
        offs[(t + 2) * NB + d] = run + c0 + c1;
        offs[(t + 3) * NB + d] = run + c0 + c1 + c2;
        run += (uint64_t)c0 + c1 + c2 + c3;
    }
    for (; t < t1; t++) {
        offs[t * NB + d] = run;
        run += counts[t * NB + d];
    }
}

// orders a wave's LDS accesses across its lanes (the wave runs in lockstep; this keeps
// the compiler from moving LDS accesses across the point)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stable scatter of one tile. Wave w owns the contiguous slice [w WT, (w + 1) WT) of
// the tile and holds it in registers. (1) each wave counts its digits (match-any
// ballots, one LDS add per distinct digit, wave-private counters); (2) per digit, the
// tile start (exclusive scan of the tile histogram; thread t holds digits t R ..
// t R + R - 1) plus the counts of the earlier waves give every wave its first slot;
// (3) each wave places its elements in LDS in order, advancing its own counters; (4)
// the tile goes out as contiguous runs per digit, so global writes are coalesced.
// Three block barriers in all; elements of wave w precede those of wave w + 1 within a
// digit, so the pass is stable. The per-lane word of step (1) holds the digit (NB = no
// element, bits 0 .. DB), its rank among equal digits (7 bits) and, for the lowest
// such lane, the number of lanes with that digit (7 bits).
template <typename K, bool V, int DB>
__global__ void __launch_bounds__(RS_BLOCK) k_radix_scatter(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                            K *__restrict__ kout, uint32_t *__restrict__ vout, uint64_t n,
                                                            int shift, const uint64_t *__restrict__ offs,
                                                            const uint32_t *__restrict__ counts)
{
    constexpr int NB = 1 << DB, NW = RS_BLOCK / 64, WT = RS_TILE / NW, WR = WT / 64;
    constexpr int R = NB >= RS_BLOCK ? NB / RS_BLOCK : 1, DSH = DB + 1, LSH = DSH + 7;
    __shared__ K sk[RS_TILE];
    __shared__ uint32_t sv[V ? RS_TILE : 1];
    __shared__ uint32_t wcnt[NW][NB];   // per wave: digit counts, then next slot per digit
    __shared__ uint32_t lbase[NB], wsum[NW];
    __shared__ uint64_t gbase[NB];
    const int tid = threadIdx.x, w = tid >> 6, lane = __lane_id();
    const bool hasd = tid * R < NB;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE, wbase = base + (uint64_t)w * WT;
    K kk[WR];
    uint32_t vv[V ? WR : 1];
#pragma unroll
    for (int r = 0; r < WR; r++) {
        const uint64_t i = wbase + (uint64_t)r * 64 + lane;
        kk[r] = i < n ? kin[i] : (K)0;
        if (V) vv[r] = i < n ? vin[i] : 0u;
    }
    uint32_t cnt[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        cnt[r] = hasd ? counts[(uint64_t)blockIdx.x * NB + tid * R + r] : 0u;
        if (hasd) gbase[tid * R + r] = offs[(uint64_t)blockIdx.x * NB + tid * R + r];
    }
    for (int q = lane; q < NB; q += 64) wcnt[w][q] = 0;
    wave_sync();
    // (1) wave histogram
    uint32_t pk[WR];
#pragma unroll
    for (int r = 0; r < WR; r++) {
        const bool valid = wbase + (uint64_t)r * 64 + lane < n;
        const uint32_t d = (uint32_t)(kk[r] >> shift) & (NB - 1u);
        const uint64_t m = match_digit<DB>(valid, d);
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        const uint32_t lead = (valid && rank == 0) ? (uint32_t)__popcll(m) : 0u;
        pk[r] = (valid ? d : (uint32_t)NB) | (rank << DSH) | (lead << LSH);
        if (lead) wcnt[w][d] += lead;
        wave_sync();
    }
    // (2) the thread's digits: tile start, then each wave's first slot
    uint32_t tsum = 0;
#pragma unroll
    for (int r = 0; r < R; r++) tsum += cnt[r];
    const uint32_t incl = wave_incl_scan(tsum);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t slot = incl - tsum;
    for (int q = 0; q < w; q++) slot += wsum[q];
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (!hasd) break;
        const int d = tid * R + r;
        lbase[d] = slot;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const uint32_t c = wcnt[q][d];
            wcnt[q][d] = slot;
            slot += c;
        }
    }
    __syncthreads();
    // (3) place the wave's elements in order
#pragma unroll
    for (int r = 0; r < WR; r++) {
        const uint32_t d = pk[r] & ((1u << DSH) - 1u), lead = pk[r] >> LSH;
        if (d < (uint32_t)NB) {
            const uint32_t pos = wcnt[w][d] + ((pk[r] >> DSH) & 127u);
            sk[pos] = kk[r];
            if (V) sv[pos] = vv[r];
        }
        wave_sync();
        if (lead) wcnt[w][d] += lead;
        wave_sync();
    }
    __syncthreads();
    // (4) contiguous runs per digit
    const uint32_t tn = (uint32_t)((n - base) < (uint64_t)RS_TILE ? (n - base) : (uint64_t)RS_TILE);
    for (uint32_t p = tid; p < tn; p += RS_BLOCK) {
        const K key = sk[p];
        const uint32_t d = (uint32_t)(key >> shift) & (NB - 1u);
        const uint64_t g = gbase[d] + (p - lbase[d]);
        kout[g] = key;
        if (V) vout[g] = sv[p];
    }
}

template <typename K, bool V, int DB>
static void radix_pass(Ctx &c, const K *a, const uint32_t *va, K *b, uint32_t *vb, uint64_t n, int shift, uint64_t ntiles,
                       uint32_t ch, uint32_t nchunks, uint32_t *counts, uint64_t *offs, uint64_t *part)
{
    constexpr int NB = 1 << DB;
    PFG_HIP(hipMemsetAsync(part, 0, 8ull * NB * nchunks, c.stream));
    k_radix_hist<K, DB><<<(unsigned)ntiles, RS_BLOCK, 0, c.stream>>>(a, n, shift, counts, ch, (unsigned long long *)part);
    k_radix_coltop<DB><<<1, NB > 64 ? NB : 64, 0, c.stream>>>(part, nchunks);
    k_radix_colapply<DB><<<nchunks, NB, 0, c.stream>>>(counts, (uint32_t)ntiles, ch, part, offs);
    k_radix_scatter<K, V, DB><<<(unsigned)ntiles, RS_BLOCK, 0, c.stream>>>(a, va, b, vb, n, shift, offs, counts);
}

// LSD passes over bits [lo, lo + nbits): ceil(nbits / 8) passes with digit widths as
// even as possible (26 bits: 7, 7, 6, 6): fewer digits per pass make longer runs of
// one digit per tile, so the scatter's writes are longer; a width under 5 takes a
// 5-bit digit (the extra bits belong to later passes, which re-sort them stably).
// Returns the buffer holding the sorted keys (keys or the context's temporary); with
// copy_back the result is always moved into keys/vals.
template <typename K, bool V>
static K *radix_impl(Ctx &c, K *keys, uint32_t *vals, uint64_t n, int nbits, int lo, bool copy_back, K *alt = nullptr)
{
    if (n <= 1 || nbits <= 0) return keys;
    const uint64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    const uint64_t cnt_n = 256 * ntiles;   // digits of at most 8 bits
    auto al = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
    const uint32_t ch = (uint32_t)((ntiles + 31) / 32), nchunks = (uint32_t)((ntiles + ch - 1) / ch);
    const uint64_t bk = alt ? 0 : al(n * sizeof(K)), bv = V ? al(n * 4) : 0, bc = al(cnt_n * 4), bo = al(cnt_n * 8),
                   bp = al(256ull * nchunks * 8);
    ensure(c, c.radix_tmp, bk + bv + bc + bo + bp);
    uint8_t *t = c.radix_tmp.p;
    K *kt = alt ? alt : (K *)t;
    uint32_t *vt = V ? (uint32_t *)(t + bk) : nullptr;
    uint32_t *counts = (uint32_t *)(t + bk + bv);
    uint64_t *offs = (uint64_t *)(t + bk + bv + bc);
    uint64_t *part = (uint64_t *)(t + bk + bv + bc + bo);
    K *a = keys, *b = kt;
    uint32_t *va = vals, *vb = vt;
    const int passes = (nbits + 7) / 8;
    int shift = lo;
    for (int p = 0; p < passes; p++) {
        const int left = lo + nbits - shift, width = (left + (passes - p) - 1) / (passes - p);
        switch (width < 5 ? 5 : width) {
            case 5: radix_pass<K, V, 5>(c, a, va, b, vb, n, shift, ntiles, ch, nchunks, counts, offs, part); break;
            case 6: radix_pass<K, V, 6>(c, a, va, b, vb, n, shift, ntiles, ch, nchunks, counts, offs, part); break;
            case 7: radix_pass<K, V, 7>(c, a, va, b, vb, n, shift, ntiles, ch, nchunks, counts, offs, part); break;
            default: radix_pass<K, V, 8>(c, a, va, b, vb, n, shift, ntiles, ch, nchunks, counts, offs, part); break;
        }
        shift += width;
        K *tk = a; a = b; b = tk;
        uint32_t *tv = va; va = vb; vb = tv;
    }
    if (copy_back && a != keys) {
        PFG_HIP(hipMemcpyAsync(keys, a, n * sizeof(K), hipMemcpyDeviceToDevice, c.stream));
        if (V) PFG_HIP(hipMemcpyAsync(vals, va, n * 4, hipMemcpyDeviceToDevice, c.stream));
        return keys;
    }
    return a;
}

void radix_sort_u32_pairs(Ctx &c, uint32_t *keys, uint32_t *vals, uint64_t n, int nbits)
{
    radix_impl<uint32_t, true>(c, keys, vals, n, nbits, 0, true);
}

uint64_t *radix_sort_u64_keys(Ctx &c, uint64_t *keys, uint64_t n, int nbits, int lo, uint64_t *alt)
{
    return radix_impl<uint64_t, false>(c, keys, nullptr, n, nbits, lo, false, alt);
}

__global__ void k_max_u32(const uint32_t *__restrict__ in, uint64_t n, unsigned long long *out)
{
    __shared__ uint32_t s_part[16];
    uint32_t m = 0;
    for (uint64_t i = gtid(); i < n; i += gstride()) m = max(m, in[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_down(m, o, 64));
    if (__lane_id() == 0) s_part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = max(m, s_part[q]);
        atomicMax(out, (unsigned long long)m);
    }
}

uint32_t reduce_max_u32(Ctx &c, const uint32_t *in, uint64_t n)
{
    PFG_HIP(hipMemsetAsync(c.cnt.p, 0, 8, c.stream));
    k_max_u32<<<grid_for(n, 256, 512), 256, 0, c.stream>>>(in, n, (unsigned long long *)c.cnt.p);
    read_counters(c, 1);
    return (uint32_t)c.hcnt[0];
}

}  // namespace pfg
