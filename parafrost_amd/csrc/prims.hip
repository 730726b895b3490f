// Exclusive scan and stable LSD radix sort for wave64 (CDNA4).
//  scan : reduce-then-scan over 2048-element tiles (256 threads x 8 items),
//         coalesced tile loads through LDS, wave scans by __shfl_up.
//  radix: 8-bit digits; per-tile LDS histograms, digit-major global scan,
//         stable scatter ranked with 8 wave ballots per round (match-any on the digit).
#include "prims.hpp"
#include "device.hpp"

namespace pfg {

constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int RS_BLOCK = 256, RS_ITEMS = 8, RS_TILE = RS_BLOCK * RS_ITEMS;

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

template <typename K>
__global__ void __launch_bounds__(RS_BLOCK) k_radix_hist(const K *__restrict__ keys, uint64_t n, int shift,
                                                         uint32_t *__restrict__ counts, uint32_t ntiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const uint64_t i = base + (uint64_t)k * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile: the tile's keys are ranked by digit (8 wave ballots
// per round), reordered in LDS, and written out as contiguous runs per digit, so
// global writes are coalesced. The tile's digit starts come from its histogram.
template <typename K, bool V>
__global__ void __launch_bounds__(RS_BLOCK) k_radix_scatter(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                            K *__restrict__ kout, uint32_t *__restrict__ vout, uint64_t n,
                                                            int shift, const uint64_t *__restrict__ offs,
                                                            const uint32_t *__restrict__ counts, uint32_t ntiles)
{
    __shared__ K sk[RS_TILE];
    __shared__ uint32_t sv[V ? RS_TILE : 1];
    __shared__ uint32_t lbase[256], lrun[256], wsum[RS_BLOCK / 64];
    __shared__ uint64_t gbase[256];
    __shared__ uint32_t wc[RS_BLOCK / 64][256];
    const int tid = threadIdx.x, w = tid >> 6, lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    // tile-local start of every digit: exclusive scan of the tile's digit counts
    const uint32_t cnt = counts[(uint64_t)tid * ntiles + blockIdx.x];
    const uint32_t incl = wave_incl_scan(cnt);
    if (lane == 63) wsum[w] = incl;
    gbase[tid] = offs[(uint64_t)tid * ntiles + blockIdx.x];
    __syncthreads();
    uint32_t excl = incl - cnt;
    for (int q = 0; q < w; q++) excl += wsum[q];
    lbase[tid] = excl;
    lrun[tid] = excl;
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
#pragma unroll
        for (int q = 0; q < RS_BLOCK / 64; q++) wc[q][tid] = 0;
        __syncthreads();
        const uint64_t i = base + (uint64_t)r * RS_BLOCK + tid;
        const bool valid = i < n;
        const K key = valid ? kin[i] : (K)0;
        const uint32_t d = valid ? ((uint32_t)(key >> shift) & 255u) : 0u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(valid && bit);
            m &= bit ? bal : ~bal;
        }
        uint32_t rank = 0;
        if (valid) {
            rank = (uint32_t)__popcll(m & lt);
            if (rank == 0) wc[w][d] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = lrun[d] + rank;
            for (int q = 0; q < w; q++) pos += wc[q][d];
            sk[pos] = key;
            if (V) sv[pos] = vin[i];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int q = 0; q < RS_BLOCK / 64; q++) add += wc[q][tid];
        lrun[tid] += add;
    }
    __syncthreads();
    const uint32_t tn = (uint32_t)((n - base) < (uint64_t)RS_TILE ? (n - base) : (uint64_t)RS_TILE);
    for (uint32_t p = tid; p < tn; p += RS_BLOCK) {
        const K key = sk[p];
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        const uint64_t g = gbase[d] + (p - lbase[d]);
        kout[g] = key;
        if (V) vout[g] = sv[p];
    }
}

// returns the buffer holding the sorted keys (keys or the context's temporary);
// with copy_back the result is always moved into keys/vals
template <typename K, bool V>
static K *radix_impl(Ctx &c, K *keys, uint32_t *vals, uint64_t n, int nbits, int lo, bool copy_back)
{
    if (n <= 1 || nbits <= 0) return keys;
    const uint64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    const uint64_t cnt_n = 256 * ntiles;
    auto al = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
    const uint64_t bk = al(n * sizeof(K)), bv = V ? al(n * 4) : 0, bc = al(cnt_n * 4), bo = al(cnt_n * 8);
    ensure(c, c.radix_tmp, bk + bv + bc + bo);
    uint8_t *t = c.radix_tmp.p;
    K *kt = (K *)t;
    uint32_t *vt = V ? (uint32_t *)(t + bk) : nullptr;
    uint32_t *counts = (uint32_t *)(t + bk + bv);
    uint64_t *offs = (uint64_t *)(t + bk + bv + bc);
    K *a = keys, *b = kt;
    uint32_t *va = vals, *vb = vt;
    for (int shift = lo; shift < lo + nbits; shift += 8) {
        k_radix_hist<K><<<(unsigned)ntiles, RS_BLOCK, 0, c.stream>>>(a, n, shift, counts, (uint32_t)ntiles);
        scan_u32_to_u64(c, counts, offs, cnt_n, nullptr);
        k_radix_scatter<K, V><<<(unsigned)ntiles, RS_BLOCK, 0, c.stream>>>(a, va, b, vb, n, shift, offs, counts,
                                                                            (uint32_t)ntiles);
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

uint64_t *radix_sort_u64_keys(Ctx &c, uint64_t *keys, uint64_t n, int nbits, int lo)
{
    return radix_impl<uint64_t, false>(c, keys, nullptr, n, nbits, lo, false);
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
