// The batch export (one instance's clauses as CSR in its own numbering, into caller
// device memory) without any allocation, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).
//
// pfg_batch_export groups every instance's clause images first, in buffers the context
// grows on demand. This variant walks the clause pool directly: the pool is cut into
// at most XNA_BLOCKS chunks; one pass counts the instance's clauses and literals per
// chunk, one block turns the counts into chunk offsets, and one pass writes each
// chunk's clauses in pool order (block scans inside a chunk). The per-chunk counts,
// offsets and the totals live in `scratch`, a file-scope device array of fixed size:
// every call on a device shares that device's copy, so calls that overlap on one
// device race. Clause order and literals equal pfg_batch_export's (pool order within
// the instance, an empty clause belongs to no instance).
#include "ctx.hpp"
#include "device.hpp"
#include <mutex>

namespace pfg {

constexpr uint32_t XNA_BLOCKS = 8192;               // chunks (blocks) of one call
constexpr uint64_t XNA_CHUNK_MIN = 1024;            // clauses per chunk, at least
constexpr uint64_t XNA_CHUNK_MAX = 1ull << 20;      // ... and at most
constexpr uint64_t XNA_MAX_CLAUSES = (uint64_t)XNA_BLOCKS * XNA_CHUNK_MAX;   // 2^33 clauses in the pool
constexpr uint32_t XNA_TPB = 256;

// [0, XNA_BLOCKS): clause count then offset per chunk; [XNA_BLOCKS, 2 XNA_BLOCKS): literal
// count then offset per chunk; [2 XNA_BLOCKS], [2 XNA_BLOCKS + 1]: the totals (128 KB + 16 B)
static __device__ __align__(256) uint64_t scratch[2 * XNA_BLOCKS + 2];

static std::mutex xna_mutex;
static uint64_t *xna_ptr[256];   // per device: the address of its copy of scratch

// the device copy of scratch for the current device `dev`, looked up once
static uint64_t *scratch_of(int dev)
{
    if (dev < 0 || dev >= 256) throw Error(PFG_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(xna_mutex);
    if (!xna_ptr[dev]) {
        void *p = nullptr;
        PFG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(PFG_EDEVICE, "no device address for the export scratch");
        xna_ptr[dev] = (uint64_t *)p;
    }
    return xna_ptr[dev];
}

__device__ __forceinline__ bool xna_member(const uint32_t *cl, const uint32_t *ibase, uint32_t ninst, uint32_t inst)
{
    return cl[2] && find_inst(ibase, ninst, cl[3] >> 1) == inst;
}

// exclusive scan over the 256 threads of a block; total in *tot
__device__ __forceinline__ uint64_t xna_block_scan(uint64_t x, uint64_t *wsum, uint64_t *tot)
{
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint64_t incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t t = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint64_t base = 0, all = 0;
    for (uint32_t q = 0; q < XNA_TPB / 64; q++) {
        if (q < w) base += wsum[q];
        all += wsum[q];
    }
    __syncthreads();
    *tot = all;
    return base + incl - x;
}

__global__ void k_xna_clear(uint64_t *S)
{
    for (uint64_t i = gtid(); i < 2ull * XNA_BLOCKS + 2; i += gstride()) S[i] = 0;
}

__global__ void __launch_bounds__(XNA_TPB) k_xna_count(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs,
                                                       uint64_t n, uint64_t chunk, const uint32_t *__restrict__ ibase,
                                                       uint32_t ninst, uint32_t inst, uint64_t *S)
{
    const uint64_t b0 = (uint64_t)blockIdx.x * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
    unsigned long long nc = 0, nl = 0;
    for (uint64_t i = b0 + threadIdx.x; i < b1; i += XNA_TPB) {
        const uint32_t *cl = mem + refs[i];
        if (xna_member(cl, ibase, ninst, inst)) {
            nc++;
            nl += cl[2];
        }
    }
    nc = block_sum(nc);
    nl = block_sum(nl);
    if (threadIdx.x == 0) {
        S[blockIdx.x] = nc;
        S[XNA_BLOCKS + blockIdx.x] = nl;
    }
}

// one block: per-chunk counts -> exclusive offsets, totals after them
__global__ void __launch_bounds__(XNA_TPB) k_xna_scan(uint64_t *S, uint32_t nb)
{
    __shared__ uint64_t wsum[XNA_TPB / 64];
    for (int arr = 0; arr < 2; arr++) {
        uint64_t *a = S + (uint64_t)arr * XNA_BLOCKS, carry = 0;
        for (uint32_t b = 0; b < nb; b += XNA_TPB) {
            const uint32_t i = b + threadIdx.x;
            const uint64_t x = i < nb ? a[i] : 0;
            uint64_t tot;
            const uint64_t ex = xna_block_scan(x, wsum, &tot);
            if (i < nb) a[i] = carry + ex;
            carry += tot;
        }
        if (threadIdx.x == 0) S[2ull * XNA_BLOCKS + arr] = carry;
        __syncthreads();
    }
}

// each chunk's clauses of the instance in pool order: offsets (from 0) and literals in
// the instance's own numbering; the end offset after the last clause
__global__ void __launch_bounds__(XNA_TPB) k_xna_write(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs,
                                                       uint64_t n, uint64_t chunk, const uint32_t *__restrict__ ibase,
                                                       uint32_t ninst, uint32_t inst, const uint64_t *S,
                                                       uint32_t *__restrict__ lits, uint64_t *__restrict__ offs)
{
    __shared__ uint64_t wsum[XNA_TPB / 64];
    const uint64_t b0 = (uint64_t)blockIdx.x * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
    uint64_t cbase = S[blockIdx.x], lbase = S[XNA_BLOCKS + blockIdx.x];
    const uint32_t shift = 2u * ibase[inst];
    for (uint64_t t = b0; t < b1; t += XNA_TPB) {
        const uint64_t i = t + threadIdx.x;
        const uint32_t *cl = nullptr;
        uint32_t in = 0, sz = 0;
        if (i < b1) {
            cl = mem + refs[i];
            if (xna_member(cl, ibase, ninst, inst)) {
                in = 1;
                sz = cl[2];
            }
        }
        uint64_t tc, tl;
        const uint64_t ci = xna_block_scan(in, wsum, &tc);
        const uint64_t li = xna_block_scan(sz, wsum, &tl);
        if (in) {
            const uint64_t o = lbase + li;
            offs[cbase + ci] = o;
            for (uint32_t k = 0; k < sz; k++) lits[o + k] = cl[3 + k] - shift;
        }
        cbase += tc;
        lbase += tl;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) offs[S[2ull * XNA_BLOCKS]] = S[2ull * XNA_BLOCKS + 1];
}

void batch_export_noalloc(Ctx &c, uint32_t inst, uint32_t *d_lits, uint64_t *d_offs, uint64_t *nlits, uint64_t *ncls)
{
    if (inst >= c.ninst) throw Error(PFG_EINVAL, "instance out of range");
    const uint64_t n = c.nrefs;
    if (n > XNA_MAX_CLAUSES)
        throw Error(PFG_EINVAL, "the pool holds more clauses than the no-allocation export serves (2^33)");
    uint64_t *S = scratch_of(c.device);
    uint64_t chunk = (n + XNA_BLOCKS - 1) / XNA_BLOCKS;
    if (chunk < XNA_CHUNK_MIN) chunk = XNA_CHUNK_MIN;
    const uint32_t nb = (uint32_t)((n + chunk - 1) / chunk > 0 ? (n + chunk - 1) / chunk : 1);
    k_xna_clear<<<64, 256, 0, c.stream>>>(S);
    k_xna_count<<<nb, XNA_TPB, 0, c.stream>>>(c.mem.p, c.refs.p, n, chunk, c.ibase.p, c.ninst, inst, S);
    k_xna_scan<<<1, XNA_TPB, 0, c.stream>>>(S, nb);
    if (d_offs) k_xna_write<<<nb, XNA_TPB, 0, c.stream>>>(c.mem.p, c.refs.p, n, chunk, c.ibase.p, c.ninst, inst, S, d_lits, d_offs);
    uint64_t tot[2] = {0, 0};
    PFG_HIP(hipMemcpyAsync(tot, S + 2ull * XNA_BLOCKS, sizeof(tot), hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    if (ncls) *ncls = tot[0];
    if (nlits) *nlits = tot[1];
}

}  // namespace pfg
