// Binary DRAT output: ordering, decisions and encoding of the events a pass logged
// (proof.hpp), appended to the context's proof bytes.
//
// Order: the events of elected variable q take positions evoff[q] + seq, so the
// array is in the sequential loop order. Decisions (passes whose units are
// assigned after the pass, VE): a candidate unit at position i is enqueued if
// its literal was unassigned before the pass and no earlier event assigned
// either polarity; it is a contradiction if its literal was false before the
// pass or its complement was assigned earlier in the pass. A candidate line is
// dropped when one of its literals was true before the pass or was assigned by
// an earlier unit. Everything after the first contradiction is dropped, and the
// empty clause ends the proof (the reference stops there: killSolver).
//
// Encoding (proof.cpp:118-135): 'a' or 'd', each literal V2DEC(vorg[v], sign) in the
// caller's original numbering (identity when no mapping was given) as a
// little-endian base-128 varint (bit 7 = more), 0.
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"
#include "proof.hpp"

namespace pfg {

__device__ __forceinline__ uint32_t varint_len(uint32_t r)
{
    uint32_t n = 1;
    while (r > 127u) { r >>= 7; n++; }
    return n;
}

__device__ __forceinline__ uint8_t *varint_put(uint8_t *o, uint32_t r)
{
    while (r > 127u) { *o++ = (uint8_t)((r & 127u) | 128u); r >>= 7; }
    *o++ = (uint8_t)r;
    return o;
}

__global__ void k_pf_order(const PEv *__restrict__ ev, uint64_t n, const uint64_t *__restrict__ evoff,
                           PEv *__restrict__ ord)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const PEv e = ev[i];
        ord[evoff[e.q] + e.seq] = e;
    }
}

__global__ void k_pf_from_lits(const uint32_t *__restrict__ lits, uint64_t n, PEv *__restrict__ ord)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        PEv e;
        e.q = 0;
        e.seq = (uint32_t)i;
        e.tag = PE_UNIT;
        e.pad = 0;
        e.payload = lits[i];
        ord[i] = e;
    }
}

__global__ void k_pf_first(const PEv *__restrict__ ord, uint64_t n, unsigned long long *first)
{
    for (uint64_t i = gtid(); i < n; i += gstride())
        if (ord[i].tag == PE_CAND_UNIT) atomicMin(first + ord[i].payload, (unsigned long long)i);
}

__global__ void k_pf_reset(const PEv *__restrict__ ord, uint64_t n, uint64_t *first)
{
    for (uint64_t i = gtid(); i < n; i += gstride())
        if (ord[i].tag == PE_CAND_UNIT) first[ord[i].payload] = ~0ull;
}

__global__ void k_pf_conflict(const PEv *__restrict__ ord, uint64_t n, const int8_t *__restrict__ value,
                              const uint64_t *__restrict__ first, unsigned long long *cpos)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const PEv e = ord[i];
        bool conflict = e.tag == PE_CONFLICT;
        if (e.tag == PE_CAND_UNIT) {
            const uint32_t u = (uint32_t)e.payload;
            const int8_t v = value[u];
            conflict = v == 0 || (v < 0 && first[u ^ 1u] < i);
        }
        if (conflict) atomicMin(cpos, (unsigned long long)i);
    }
}

// does event i make a line, and of which literals
__device__ __forceinline__ bool pf_line(const PEv &e, uint64_t i, const uint32_t *mem, const int8_t *value,
                                        const uint64_t *first, uint64_t cpos, const uint32_t *&lits, uint32_t &n,
                                        uint32_t &unit, uint8_t &tag)
{
    if (i >= cpos) return false;
    switch (e.tag) {
    case PE_ADD:
    case PE_DEL:
    case PE_CAND_ADD: {
        const uint32_t *c = mem + e.payload;
        lits = c + 3;
        n = c[2];
        tag = e.tag == PE_DEL ? 'd' : 'a';
        if (e.tag == PE_CAND_ADD)
            for (uint32_t k = 0; k < n; k++)
                if (value[lits[k]] == 1 || first[lits[k]] < i) return false;
        return true;
    }
    case PE_UNIT:
        unit = (uint32_t)e.payload;
        lits = &unit;
        n = 1;
        tag = 'a';
        return true;
    case PE_CAND_UNIT:
        unit = (uint32_t)e.payload;
        if (!(value[unit] < 0 && first[unit] == i)) return false;
        lits = &unit;
        n = 1;
        tag = 'a';
        return true;
    default:
        return false;
    }
}

__global__ void k_pf_bytes(const PEv *__restrict__ ord, uint64_t n, const uint32_t *__restrict__ mem,
                           const int8_t *__restrict__ value, const uint64_t *__restrict__ first,
                           const unsigned long long *cpos, const uint32_t *__restrict__ vorg,
                           uint32_t *__restrict__ bytes)
{
    const uint64_t cp = *cpos;
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *lits;
        uint32_t m = 0, unit = 0, b = 0;
        uint8_t tag;
        if (pf_line(ord[i], i, mem, value, first, cp, lits, m, unit, tag)) {
            b = 2;
            for (uint32_t k = 0; k < m; k++) b += varint_len(org_lit(vorg, lits[k]));
        }
        bytes[i] = b;
    }
}

__global__ void k_pf_write(const PEv *__restrict__ ord, uint64_t n, const uint32_t *__restrict__ mem,
                           const int8_t *__restrict__ value, const uint64_t *__restrict__ first,
                           const unsigned long long *cpos, const uint32_t *__restrict__ vorg,
                           const uint64_t *__restrict__ boff, uint8_t *__restrict__ out)
{
    const uint64_t cp = *cpos;
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *lits;
        uint32_t m = 0, unit = 0;
        uint8_t tag;
        if (!pf_line(ord[i], i, mem, value, first, cp, lits, m, unit, tag)) continue;
        uint8_t *o = out + boff[i];
        *o++ = tag;
        for (uint32_t k = 0; k < m; k++) o = varint_put(o, org_lit(vorg, lits[k]));
        *o = 0;
    }
}

static bool proof_on(const Ctx &c) { return c.opts.proof_en && !c.batch && !c.proof_closed; }

ProofLog proof_begin(Ctx &c, uint32_t np, uint64_t cap)
{
    ProofLog L{nullptr, nullptr, 0, nullptr};
    if (!proof_on(c)) return L;
    cap += 1024;
    ensure(c, c.pev, cap * sizeof(PEv));
    ensure(c, c.pnev, (uint64_t)np + 1);
    ensure(c, c.pcnt, 4);
    PFG_HIP(hipMemsetAsync(c.pnev.p, 0, 4ull * ((uint64_t)np + 1), c.stream));
    PFG_HIP(hipMemsetAsync(c.pcnt.p, 0, 8, c.stream));
    L.ev = (PEv *)c.pev.p;
    L.n = (unsigned long long *)c.pcnt.p;
    L.cap = cap;
    L.nev = c.pnev.p;
    c.pcap = cap;
    return L;
}

void proof_empty(Ctx &c)
{
    if (!proof_on(c)) return;
    ensure(c, c.proof, c.nproof + 2, true);
    static const uint8_t line[2] = {'a', 0};
    PFG_HIP(hipMemcpyAsync(c.proof.p + c.nproof, line, 2, hipMemcpyHostToDevice, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    c.nproof += 2;
    c.proof_closed = true;
}

// decide, size and write the n ordered events in c.pord
static void encode(Ctx &c, uint64_t n, bool decide)
{
    ensure(c, c.pcnt, 4);
    const uint64_t none = ~0ull;
    PFG_HIP(hipMemcpyAsync(c.pcnt.p + 1, &none, 8, hipMemcpyHostToDevice, c.stream));
    const PEv *ord = (const PEv *)c.pord.p;
    if (c.pfirst.cap < c.nlits2) {
        release(c, c.pfirst);
        ensure(c, c.pfirst, c.nlits2);
        PFG_HIP(hipMemsetAsync(c.pfirst.p, 0xFF, c.pfirst.cap * 8, c.stream));
    }
    if (decide) k_pf_first<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(ord, n, (unsigned long long *)c.pfirst.p);
    // contradictions: logged ones (SUB) and decided ones (VE)
    k_pf_conflict<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(ord, n, c.value.p, c.pfirst.p,
                                                                (unsigned long long *)c.pcnt.p + 1);
    ensure(c, c.pbytes, n + 1);
    ensure(c, c.pboff, n + 2);
    const uint32_t *vorg = c.has_vorg ? c.vorg.p : nullptr;
    k_pf_bytes<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(ord, n, c.mem.p, c.value.p, c.pfirst.p,
                                                             (const unsigned long long *)c.pcnt.p + 1, vorg,
                                                             c.pbytes.p);
    scan_u32_to_u64(c, c.pbytes.p, c.pboff.p, n, c.pcnt.p + 2);
    uint64_t h[2];
    PFG_HIP(hipMemcpyAsync(h, c.pcnt.p + 1, 16, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    const uint64_t cpos = h[0], total = h[1];
    if (total) {
        ensure(c, c.proof, c.nproof + total + 2, true);
        k_pf_write<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(ord, n, c.mem.p, c.value.p, c.pfirst.p,
                                                                 (const unsigned long long *)c.pcnt.p + 1, vorg,
                                                                 c.pboff.p, c.proof.p + c.nproof);
        c.nproof += total;
    }
    if (decide) k_pf_reset<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(ord, n, c.pfirst.p);
    if (cpos != ~0ull) proof_empty(c);
}

void proof_end(Ctx &c, uint32_t np, bool decide)
{
    if (!proof_on(c)) return;
    uint64_t n = 0;
    PFG_HIP(hipMemcpyAsync(&n, c.pcnt.p, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    if (n > c.pcap) throw Error(PFG_EINTERNAL, "proof event log overflow");
    if (!n) return;
    ensure(c, c.poff, (uint64_t)np + 2);
    scan_u32_to_u64(c, c.pnev.p, c.poff.p, np, c.poff.p + np);
    ensure(c, c.pord, n * sizeof(PEv));
    k_pf_order<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>((const PEv *)c.pev.p, n, c.poff.p, (PEv *)c.pord.p);
    encode(c, n, decide);
}

void proof_units(Ctx &c, uint32_t from, uint32_t to)
{
    if (!proof_on(c) || to <= from) return;
    const uint64_t n = to - from;
    ensure(c, c.pord, n * sizeof(PEv));
    k_pf_from_lits<<<grid_for(n, 256, 8192), 256, 0, c.stream>>>(c.trail.p + from, n, (PEv *)c.pord.p);
    encode(c, n, false);
}

}  // namespace pfg
