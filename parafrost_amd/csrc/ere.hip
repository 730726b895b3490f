// Eager redundancy elimination on the elected variables of the final round
// (ParaFROST src/cpu/redundancy.h:26-143, elimination.cpp:169-197).
//
// For every non-tautological resolvent of a variable's clauses the reference
// deletes the first clause on the shortest occurrence list (among the literals
// met before one antecedent runs out) that equals it. Targets never contain an
// elected variable, so two variables interact only when they pick the same
// clause. The parallel pass claims each target for its variable; if no clause
// is claimed twice, every variable saw exactly what the sequential loop shows
// it, and the claimed clauses are deleted. Otherwise the pass is redone in
// elected order by a single thread.
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"

namespace pfg {

constexpr uint32_t NO_OWNER = 0xFFFFFFFFu;

struct EreArgs {
    uint32_t *mem;
    const OL *ot;
    const uint64_t *occ;
    int ere_limit;
    int maxsize;
    uint32_t *owner;
    uint64_t *targets;
    unsigned long long *cnt;
};

__device__ bool merge_ere(uint32_t x, const uint32_t *c1, const uint32_t *c2, const uint32_t *target)
{   // merge_ere (redundancy.h:37-85)
    const int n1 = c_size(c1), n2 = c_size(c2), n = c_size(target);
    const uint32_t *a = c1 + 3, *b = c2 + 3, *t = target + 3;
    int it1 = 0, it2 = 0, it = 0;
    while (it < n && it1 < n1 && it2 < n2) {
        const uint32_t lit = t[it];
        if ((lit >> 1) == x) return false;
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = l1 >> 1, v2 = l2 >> 1;
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if (v1 < v2) { if (l1 != lit) return false; it1++; it++; }
        else if (v2 < v1) { if (l2 != lit) return false; it2++; it++; }
        else { if (l1 != lit) return false; it1++; it2++; it++; }
    }
    while (it < n && it1 < n1) { const uint32_t l1 = a[it1]; if ((l1 >> 1) != x && l1 != t[it++]) return false; it1++; }
    while (it < n && it2 < n2) { const uint32_t l2 = b[it2]; if ((l2 >> 1) != x && l2 != t[it++]) return false; it2++; }
    return true;
}

// forward_equ (redundancy.h:87-143). seq: delete in place (sequential replay);
// otherwise claim the target for variable q. Returns false on a claim collision.
__device__ bool forward_equ(const EreArgs &A, uint32_t x, const uint32_t *c1, const uint32_t *c2, bool seq, uint32_t q)
{
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c1 + 3, *b = c2 + 3;
    int len = n1 + n2 - 2, it1 = 0, it2 = 0;
    uint32_t minsize = 0xFFFFFFFFu, msig = 0, best = 0;
    bool any = false;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = l1 >> 1, v2 = l2 >> 1;
        uint32_t lit;
        if (v1 == x) { it1++; continue; }
        else if (v2 == x) { it2++; continue; }
        else if ((l1 ^ l2) == 1u) return true;
        else if (v1 < v2) { it1++; lit = l1; }
        else if (v2 < v1) { it2++; lit = l2; }
        else { it1++; it2++; len--; lit = l1; }
        const uint32_t ls = A.ot[lit].sz;
        if (!any || ls < minsize) { minsize = ls; best = lit; any = true; }
        msig |= maphash(lit);
    }
    if (!(len > 1 && (!A.maxsize || len <= A.maxsize))) return true;
    if (!any) return true;   // unreachable for clauses of size >= 2
    const uint32_t type = (c_learnt(c1) || c_learnt(c2)) ? ST_LEARNT : ST_ORIGINAL;
    const OL o = A.ot[best];
    for (uint32_t i = 0; i < minsize; i++) {
        const uint64_t r = A.occ[o.off + i];
        uint32_t *c = A.mem + r;
        if (!(len == c_size(c) && (c_learnt(c) || c_st(c) == type) && sig_sub(msig, c[1]) && merge_ere(x, c1, c2, c)))
            continue;
        if (seq) { c_markdel(c); return true; }
        uint32_t *slot = A.owner + (r >> 1);
        const uint32_t cur = __atomic_load_n(slot, __ATOMIC_RELAXED);
        if (cur == q) continue;   // deleted by this variable already
        if (cur != NO_OWNER) return false;
        const uint32_t old = atomicCAS(slot, NO_OWNER, q);
        if (old != NO_OWNER) return false;
        A.targets[atomicAdd(&A.cnt[CNT_A], 1ull)] = r;
        return true;
    }
    return true;
}

__device__ bool ere_var(const EreArgs &A, uint32_t x, bool seq, uint32_t q)
{
    uint32_t dx = 2 * x, fx = dx + 1;
    if (A.ot[dx].sz > A.ot[fx].sz) { dx = fx; fx = 2 * x; }
    const OL me = A.ot[dx], other = A.ot[fx];
    if (!(me.sz <= (uint32_t)A.ere_limit && other.sz <= (uint32_t)A.ere_limit)) return true;
    for (uint32_t i = 0; i < me.sz; i++) {
        const uint32_t *ci = A.mem + A.occ[me.off + i];
        if (c_del(ci)) continue;
        for (uint32_t j = 0; j < other.sz; j++) {
            const uint32_t *cj = A.mem + A.occ[other.off + j];
            if (c_del(cj)) continue;
            if (!forward_equ(A, x, ci, cj, seq, q)) return false;
        }
    }
    return true;
}

__global__ void k_ere_claim(EreArgs A, const uint32_t *__restrict__ pvs, uint32_t npvs)
{
    for (uint64_t q = gtid(); q < npvs; q += gstride())
        if (!ere_var(A, pvs[q], false, (uint32_t)q)) atomicOr(&A.cnt[CNT_B], 1ull);
}

__global__ void k_ere_commit(EreArgs A, uint64_t n, int del)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint64_t r = A.targets[i];
        if (del) c_markdel(A.mem + r);
        A.owner[r >> 1] = NO_OWNER;
    }
}

__global__ void k_ere_seq(EreArgs A, const uint32_t *__restrict__ pvs, uint32_t npvs)
{
    if (gtid() != 0) return;
    for (uint32_t q = 0; q < npvs; q++) ere_var(A, pvs[q], true, q);
}

void ere(Ctx &c)
{
    if (!c.opts.ere_en || !c.npvs) return;
    const uint64_t need = (c.mem.cap >> 1) + 2;
    if (c.owner.cap < need) {
        release(c, c.owner);
        ensure(c, c.owner, need);
        PFG_HIP(hipMemsetAsync(c.owner.p, 0xFF, c.owner.cap * 4, c.stream));
    }
    ensure(c, c.tmpu64b, c.occ.cap + 1);
    EreArgs A{c.mem.p, c.ot.p, c.occ.p, c.opts.ere_limit, c.opts.ere_max_resolvent, c.owner.p, c.tmpu64b.p,
              (unsigned long long *)c.cnt.p};
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_A, 0, 16, c.stream));
    {
        KTimer kt(c, K_ERE, pv_bytes(c, 0, c.npvs, true));
        k_ere_claim<<<grid_for(c.npvs, 64, 16384), 64, 0, c.stream>>>(A, c.pvs.p, c.npvs);
    }
    read_counters(c, 3);
    const uint64_t nt = c.hcnt[CNT_A];
    const bool collision = c.hcnt[CNT_B] != 0;
    if (nt) k_ere_commit<<<grid_for(nt, 256, 4096), 256, 0, c.stream>>>(A, nt, collision ? 0 : 1);
    if (collision) k_ere_seq<<<1, 64, 0, c.stream>>>(A, c.pvs.p, c.npvs);
}

}  // namespace pfg
