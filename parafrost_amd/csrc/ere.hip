// Eager redundancy elimination on the elected variables of the final round
// (ParaFROST src/cpu/redundancy.h:26-143, elimination.cpp:169-197).
//
// For every non-tautological resolvent of a variable's clauses the reference
// deletes the first live clause on the shortest occurrence list (among the
// literals met before one antecedent runs out) that equals it; a deleted clause
// never matches (its status is neither learnt nor the resolvent's type). Targets
// never contain an elected variable, so two variables interact only through the
// clauses they delete: a variable that meets a clause an earlier variable deleted
// goes on to the next equal clause.
//
// Rounds of optimistic claims, committed as a prefix in elected order: every
// variable still to run walks its loop against the deletions committed so far plus
// its own, and claims each target with atomicMin on its elected position. The
// first variable of the round cannot be affected by the others, and by induction
// neither can any variable before q_bad, the first one whose target went to an
// earlier variable. Claims of the variables before q_bad are committed (their
// clauses deleted, their proof lines written); the variables from q_bad on run
// again in the next round. Without a shared target the pass is one round.
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"
#include "proof.hpp"
#include "stage.hpp"

namespace pfg {

constexpr uint32_t NO_OWNER = 0xFFFFFFFFu;

struct EreArgs {
    uint32_t *mem;
    const OL *ot;
    const uint64_t *occ;
    int ere_limit;
    int maxsize;
    uint32_t *owner;      // per (ref >> 1): the smallest elected position claiming it this round
    uint64_t *targets;    // claims of the round: target reference,
    uint32_t *tq;         // ... its claimant's elected position,
    uint32_t *tseq;       // ... and the claim's place in that variable's loop
    uint64_t tcap;
    unsigned long long *cnt;
    ProofLog P;   // proof events: the deleted targets (removeClause, redundancy.h:137)
};

__device__ __forceinline__ bool merge_ere(uint32_t x, const uint32_t *c1, const uint32_t *c2, const uint32_t *target)
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

// The resolvent of c1 and c2 on x as forward_equ (redundancy.h:87-143) sees it:
// its length, signature, and the literal with the shortest occurrence list among
// those met before one antecedent runs out. False when the resolvent is a
// tautology or outside the size bounds. ST: c1 and c2 are staged clauses carrying
// the list size of each literal after their literals (sizes do not change during
// the pass).
struct EreRes {
    int len;
    uint32_t msig, best, minsize, type;
};

template <bool ST>
__device__ __forceinline__ bool ere_resolvent(const EreArgs &A, uint32_t x, const uint32_t *c1, const uint32_t *c2, EreRes &R)
{
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c1 + 3, *b = c2 + 3;
    int len = n1 + n2 - 2, it1 = 0, it2 = 0;
    uint32_t minsize = 0xFFFFFFFFu, msig = 0, best = 0;
    bool any = false;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = l1 >> 1, v2 = l2 >> 1;
        uint32_t lit, ls;
        if (v1 == x) { it1++; continue; }
        else if (v2 == x) { it2++; continue; }
        else if ((l1 ^ l2) == 1u) return false;
        else if (v1 < v2) { lit = l1; ls = ST ? a[n1 + it1] : A.ot[lit].sz; it1++; }
        else if (v2 < v1) { lit = l2; ls = ST ? b[n2 + it2] : A.ot[lit].sz; it2++; }
        else { lit = l1; ls = ST ? a[n1 + it1] : A.ot[lit].sz; it1++; it2++; len--; }
        if (!any || ls < minsize) { minsize = ls; best = lit; any = true; }
        msig |= maphash(lit);
    }
    if (!(len > 1 && (!A.maxsize || len <= A.maxsize))) return false;
    if (!any) return false;   // unreachable for clauses of size >= 2
    R.len = len;
    R.msig = msig;
    R.best = best;
    R.minsize = minsize;
    R.type = (c_learnt(c1) || c_learnt(c2)) ? ST_LEARNT : ST_ORIGINAL;
    return true;
}

__device__ __forceinline__ bool ere_match(const EreArgs &A, uint32_t x, const uint32_t *c1, const uint32_t *c2,
                                          const EreRes &R, const uint32_t *c)
{
    return R.len == c_size(c) && (c_learnt(c) || c_st(c) == R.type) && sig_sub(R.msig, c[1]) && merge_ere(x, c1, c2, c);
}

// forward_equ (redundancy.h:87-143): claim the first matching live target for the
// variable at elected position q. A clause this variable claimed before is deleted by
// then in the sequential loop, so the search goes past it.
template <bool ST>
__device__ __forceinline__ void forward_equ(const EreArgs &A, uint32_t x, const uint32_t *c1, const uint32_t *c2,
                                            uint32_t q, uint32_t &pseq)
{
    EreRes R;
    if (!ere_resolvent<ST>(A, x, c1, c2, R)) return;
    const uint64_t off = A.ot[R.best].off;
    for (uint32_t i = 0; i < R.minsize; i++) {
        const uint64_t r = A.occ[off + i];
        const uint32_t *c = A.mem + r;
        if (!ere_match(A, x, c1, c2, R, c)) continue;
        const uint32_t old = atomicMin(A.owner + (r >> 1), q);
        if (old == q) continue;
        const unsigned long long k = atomicAdd(&A.cnt[CNT_A], 1ull);
        if (k < A.tcap) {
            A.targets[k] = r;
            A.tq[k] = q;
            A.tseq[k] = pseq;
        }
        pseq++;
        return;
    }
}

// tab: the variable's staged region (stage.hpp), or null to read the pool
template <bool ST>
__device__ __forceinline__ void ere_var(const EreArgs &A, uint32_t x, uint32_t q, const uint32_t *tab,
                                        const uint32_t *arena, uint32_t &pseq)
{
    uint32_t dx = 2 * x, fx = dx + 1;
    if (A.ot[dx].sz > A.ot[fx].sz) { dx = fx; fx = 2 * x; }
    const OL me = A.ot[dx], other = A.ot[fx];
    if (!(me.sz <= (uint32_t)A.ere_limit && other.sz <= (uint32_t)A.ere_limit)) return;
    const uint32_t ps = A.ot[2 * x].sz;
    const uint32_t mb = (dx & 1u) ? ps : 0u, ob = (fx & 1u) ? ps : 0u;
    for (uint32_t i = 0; i < me.sz; i++) {
        const uint32_t *ci = ST ? arena + tab[mb + i] : A.mem + A.occ[me.off + i];
        if (c_del(ci)) continue;
        for (uint32_t j = 0; j < other.sz; j++) {
            const uint32_t *cj = ST ? arena + tab[ob + j] : A.mem + A.occ[other.off + j];
            if (c_del(cj)) continue;
            forward_equ<ST>(A, x, ci, cj, q, pseq);
        }
    }
}

// Before the rounds, a probe looks at every (variable, clause of the shorter list,
// clause of the other list) pair independently, one thread each, and flags the
// variable when the pair's shortest list holds a live clause equal to the resolvent
// at all. A variable without such a pair deletes nothing, so only flagged
// variables run the claiming rounds; the number of matching pairs bounds the claims
// of a round.

// pairs of variable q (0 when a list exceeds ere_limit)
__global__ void k_ere_pairs(EreArgs A, const uint32_t *__restrict__ pvs, uint32_t npvs, uint64_t *__restrict__ npairs)
{
    for (uint64_t q = gtid(); q < npvs; q += gstride()) {
        const uint32_t x = pvs[q];
        const uint32_t ps = A.ot[2 * x].sz, ns = A.ot[2 * x + 1].sz;
        npairs[q] = (ps <= (uint32_t)A.ere_limit && ns <= (uint32_t)A.ere_limit) ? (uint64_t)ps * ns : 0ull;
    }
}

// last q in [lo, hi) with poff[q] <= g
__device__ __forceinline__ uint32_t pair_owner(const uint64_t *poff, uint32_t lo, uint32_t hi, uint64_t g)
{
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (poff[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// pair g = poff[q] + i * |other| + j of elected variable q
__device__ __forceinline__ void ere_probe_pair(const EreArgs &A, const uint32_t *pvs, const uint64_t *poff, uint32_t q,
                                               uint64_t g, uint32_t *flag, bool account, unsigned long long &bytes)
{
    const uint32_t x = pvs[q];
    uint32_t dx = 2 * x, fx = dx + 1;
    if (A.ot[dx].sz > A.ot[fx].sz) { dx = fx; fx = 2 * x; }
    const OL me = A.ot[dx], other = A.ot[fx];
    const uint64_t r = g - poff[q];
    const uint32_t i = (uint32_t)(r / other.sz), j = (uint32_t)(r % other.sz);
    const uint32_t *ci = A.mem + A.occ[me.off + i];
    const uint32_t *cj = A.mem + A.occ[other.off + j];
    if (c_del(ci) || c_del(cj)) return;
    EreRes R;
    if (!ere_resolvent<false>(A, x, ci, cj, R)) return;
    const uint64_t off = A.ot[R.best].off;
    // the probe asks whether any clause matches, so four candidates' headers are
    // loaded together and checked before the merge
    bool hit = false;
    uint32_t k = 0;
    for (; k + 4 <= R.minsize && !hit; k += 4) {
        const uint32_t *c0 = A.mem + A.occ[off + k], *c1 = A.mem + A.occ[off + k + 1],
                       *c2 = A.mem + A.occ[off + k + 2], *c3 = A.mem + A.occ[off + k + 3];
        const uint32_t *cc[4] = {c0, c1, c2, c3};
        const uint32_t h[4] = {c0[0], c1[0], c2[0], c3[0]}, z[4] = {c0[2], c1[2], c2[2], c3[2]},
                       gs[4] = {c0[1], c1[1], c2[1], c3[1]};
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool head = (int)z[u] == R.len && ((h[u] & ST_LEARNT) || (h[u] & 3u) == R.type) &&
                              sig_sub(R.msig, gs[u]);
            if (account) bytes += 20ull + (!hit && head ? 4ull * (uint64_t)R.len : 0ull);
            if (!hit && head && merge_ere(x, ci, cj, cc[u])) hit = true;
        }
    }
    for (; k < R.minsize && !hit; k++) {
        const uint32_t *t = A.mem + A.occ[off + k];
        hit = ere_match(A, x, ci, cj, R, t);
        if (account) bytes += 20ull + (hit ? 4ull * (uint64_t)R.len : 0ull);
    }
    if (hit) {
        flag[q] = 1u;
        atomicAdd(&A.cnt[CNT_C], 1ull);
    }
}

// A block takes 256 consecutive pairs: two threads find the variables of its first
// and last pair, and every thread then searches only between them (a few variables,
// their offsets cached), not over all of poff.
// acct (accounting only): adds the bytes of the resolvent checks -- per candidate
// examined its list entry (8) and header (12), per candidate whose size, status and
// signature pass the literals compared (4 per literal of the resolvent)
__global__ void __launch_bounds__(256) k_ere_probe(EreArgs A, const uint32_t *__restrict__ pvs, uint32_t npvs,
                                                  const uint64_t *__restrict__ poff, uint64_t total,
                                                  uint32_t *__restrict__ flag, unsigned long long *acct)
{
    __shared__ uint32_t s_q[2];
    unsigned long long bytes = 0;
    for (uint64_t b = (uint64_t)blockIdx.x * 256; b < total; b += (uint64_t)gridDim.x * 256) {
        const uint64_t last = b + 255 < total ? b + 255 : total - 1;
        if (threadIdx.x < 2) s_q[threadIdx.x] = pair_owner(poff, 0, npvs, threadIdx.x ? last : b);
        __syncthreads();
        const uint64_t g = b + threadIdx.x;
        if (g < total) ere_probe_pair(A, pvs, poff, pair_owner(poff, s_q[0], s_q[1] + 1, g), g, flag, acct != nullptr, bytes);
        __syncthreads();
    }
    if (acct) {
        bytes = block_sum(bytes);
        if (threadIdx.x == 0 && bytes) atomicAdd(acct, bytes);
    }
}

// one claiming round: thread per flagged variable from elected position qfrom on
__global__ void __launch_bounds__(64) k_ere_claim(EreArgs A, const uint32_t *__restrict__ pvs, uint32_t qfrom,
                                                 uint32_t npvs, const uint32_t *__restrict__ flag)
{
    for (uint64_t q = qfrom + gtid(); q < npvs; q += gstride()) {
        if (!flag[q]) continue;
        uint32_t pseq = 0;
        ere_var<false>(A, pvs[q], (uint32_t)q, nullptr, nullptr, pseq);
        plog_end(A.P, (uint32_t)q, pseq);
    }
}

// q_bad (CNT_D): the first elected position with a target that went to an earlier one
__global__ void k_ere_lost(EreArgs A, uint64_t n)
{
    for (uint64_t k = gtid(); k < n; k += gstride())
        if (A.owner[A.targets[k] >> 1] != A.tq[k]) atomicMin(&A.cnt[CNT_D], (unsigned long long)A.tq[k]);
}

// claims of positions before q_bad are final: their targets are deleted and their
// proof lines written; every claim is released
__global__ void k_ere_commit(EreArgs A, uint64_t n)
{
    const uint64_t qbad = A.cnt[CNT_D];
    for (uint64_t k = gtid(); k < n; k += gstride()) {
        const uint64_t r = A.targets[k];
        const uint32_t q = A.tq[k];
        if (q < qbad) {
            c_markdel(A.mem + r);
            plog_at(A.P, q, A.tseq[k], PE_DEL, r);
        }
        A.owner[r >> 1] = NO_OWNER;
    }
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
    const uint32_t np = c.npvs;
    EreArgs A{c.mem.p, c.ot.p, c.occ.p, c.opts.ere_limit, c.opts.ere_max_resolvent, c.owner.p, nullptr, nullptr,
              nullptr, 0, (unsigned long long *)c.cnt.p, proof_begin(c, np, c.nrefs)};
    ensure(c, c.tmpu64a, 2ull * np + 2);
    ensure(c, c.tmpu32c, np + 1);
    uint64_t *npairs = c.tmpu64a.p, *poff = c.tmpu64a.p + np + 1;
    uint32_t *flag = c.tmpu32c.p;
    PFG_HIP(hipMemsetAsync(flag, 0, 4ull * np, c.stream));
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_C, 0, 8, c.stream));
    k_ere_pairs<<<grid_for(np, 256, 4096), 256, 0, c.stream>>>(A, c.pvs.p, np, npairs);
    scan_u64(c, npairs, poff, np, poff + np);
    uint64_t total = 0;
    PFG_HIP(hipMemcpyAsync(&total, poff + np, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    if (total) {
        double check_bytes = 0;
        if (c.account) {   // an untimed run of the probe adds up the bytes of the resolvent checks
            PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_ACCT, 0, 8, c.stream));
            k_ere_probe<<<grid_for(total, 256, 65535), 256, 0, c.stream>>>(A, c.pvs.p, np, poff, total, flag,
                                                                        (unsigned long long *)c.cnt.p + CNT_ACCT);
            read_counters(c, CNT_ACCT + 1);
            check_bytes = (double)c.hcnt[CNT_ACCT];
            PFG_HIP(hipMemsetAsync(flag, 0, 4ull * np, c.stream));
            PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_C, 0, 8, c.stream));
        }
        KTimer kt(c, K_ERE, pv_bytes(c, 0, np, true) + check_bytes);
        k_ere_probe<<<grid_for(total, 256, 65535), 256, 0, c.stream>>>(A, c.pvs.p, np, poff, total, flag, nullptr);
    }
    read_counters(c, CNT_C + 1);
    const uint64_t hits = c.hcnt[CNT_C];   // matching pairs: at most one claim each per round
    if (hits) {
        ensure(c, c.tmpu64b, hits + 1);
        ensure(c, c.tmpu32d, 2 * hits + 2);
        A.targets = c.tmpu64b.p;
        A.tq = c.tmpu32d.p;
        A.tseq = c.tmpu32d.p + hits + 1;
        A.tcap = hits;
        uint32_t qfrom = 0;
        for (;;) {
            const uint64_t init[4] = {0, 0, 0, ~0ull};   // CNT_A claims, CNT_B, CNT_C, CNT_D q_bad
            PFG_HIP(hipMemcpyAsync(c.cnt.p + CNT_A, init, sizeof(init), hipMemcpyHostToDevice, c.stream));
            k_ere_claim<<<grid_for(np - qfrom, 64, 16384), 64, 0, c.stream>>>(A, c.pvs.p, qfrom, np, flag);
            read_counters(c, CNT_A + 1);
            const uint64_t nt = c.hcnt[CNT_A];
            if (nt > hits) throw Error(PFG_EINTERNAL, "ERE claimed more targets than matching pairs");
            if (!nt) break;
            k_ere_lost<<<grid_for(nt, 256, 4096), 256, 0, c.stream>>>(A, nt);
            k_ere_commit<<<grid_for(nt, 256, 4096), 256, 0, c.stream>>>(A, nt);
            read_counters(c, CNT_D + 1);
            const uint64_t qbad = c.hcnt[CNT_D];
            if (qbad == ~0ull) break;
            if (qbad <= qfrom) throw Error(PFG_EINTERNAL, "ERE claim round made no progress");
            c.path[PATH_ERE_ROUNDS]++;
            qfrom = (uint32_t)qbad;
        }
    }
    proof_end(c, np, false);
}

}  // namespace pfg
