// LCVE election on the device.
//
// The reference elects variables greedily in (score, var) order, score = ps*ns
// (uint32), skipping frozen / assigned / eliminated / absent variables, stopping at
// the first candidate whose both occurrence lists reach the temperatures, and
// freezing every cold neighbour of an elected variable (src/cpu/lcve.cpp:46-99).
// That sequential greedy is reproduced exactly by a deterministic parallel
// greedy-MIS over candidate ranks:
//   * candidates before the static break position B0 (first non-skipped candidate
//     that is hot by list size and cannot be frozen) are all freezable, and the
//     "shares a live clause" relation between them is symmetric;
//   * a candidate is elected iff no earlier-ranked elected candidate shares a live
//     clause with it; rounds decide a candidate once every earlier neighbour is
//     decided (or an earlier elected neighbour is found);
//   * a candidate that is hot by list size but freezable ends the election where
//     the sequential loop would reach it unfrozen (dynamic break); it never freezes.
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"

namespace pfg {

enum { K_SKIP = 0, K_CAND = 1, K_DYNBREAK = 2, K_BREAK = 3 };
enum { S_UNDEC = 0, S_IN = 1, S_OUT = 2 };

__global__ void k_scores(const uint32_t *__restrict__ hist, uint32_t nvars, uint32_t *__restrict__ keys, uint32_t *__restrict__ vars)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) {
        const uint32_t v = (uint32_t)i + 1;
        keys[i] = hist[2 * v] * hist[2 * v + 1];   // uint32 product as in hist_score (histogram.h:26)
        vars[i] = v;
    }
}

__global__ void k_rank(const uint32_t *__restrict__ elig, uint32_t nvars, uint32_t *__restrict__ rank)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) rank[elig[i]] = (uint32_t)i;
}

__global__ void k_classify(const uint32_t *__restrict__ elig, uint32_t nvars, const uint32_t *__restrict__ hist,
                           const uint8_t *__restrict__ vstate, const OL *__restrict__ ot, uint32_t pt, uint32_t nt,
                           uint8_t *__restrict__ kind, uint8_t *__restrict__ status, unsigned long long *cnt)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) {
        const uint32_t v = elig[i];
        const uint32_t ps = hist[2 * v], ns = hist[2 * v + 1];
        uint8_t k = K_SKIP;
        if (!vstate[v] && (ps || ns)) {
            const bool hot = ot[2 * v].sz >= pt && ot[2 * v + 1].sz >= nt;
            const bool freezable = ps < pt || ns < nt;
            if (hot && !freezable) { k = K_BREAK; atomicMin(&cnt[CNT_B], (unsigned long long)i); }
            else k = hot ? K_DYNBREAK : K_CAND;
        }
        kind[i] = k;
        status[i] = S_UNDEC;
    }
}

__global__ void k_work_init(const uint8_t *__restrict__ kind, uint64_t b0, uint32_t *__restrict__ work, unsigned long long *cnt)
{
    for (uint64_t i = gtid(); i < b0; i += gstride()) {
        const uint8_t k = kind[i];
        if (k == K_CAND || k == K_DYNBREAK) work[atomicAdd(&cnt[CNT_A], 1ull)] = (uint32_t)i;
    }
}

// one wave per undecided candidate; lanes stride over the candidate's live clauses
__global__ void __launch_bounds__(256) k_mis_round(const uint32_t *__restrict__ work, uint64_t nwork,
                                                   const uint32_t *__restrict__ elig, const uint32_t *__restrict__ rank,
                                                   const uint8_t *__restrict__ kind, uint8_t *status, const OL *__restrict__ ot,
                                                   const uint64_t *__restrict__ occ, const uint32_t *__restrict__ mem,
                                                   uint32_t *__restrict__ next, unsigned long long *cnt)
{
    const int lane = __lane_id();
    const uint64_t wid = gtid() >> 6, nw = gstride() >> 6;
    for (uint64_t q = wid; q < nwork; q += nw) {
        const uint32_t i = work[q];
        const uint32_t v = elig[i];
        bool in_found = false, pending = false;
        for (int sg = 0; sg < 2 && !in_found; sg++) {
            const OL o = ot[2 * v + sg];
            for (uint32_t base = 0; base < o.sz; base += 64) {
                const uint32_t k = base + lane;
                if (k < o.sz) {
                    const uint32_t *c = mem + occ[o.off + k];
                    if (!c_del(c)) {
                        const int sz = c_size(c);
                        for (int t = 0; t < sz; t++) {
                            const uint32_t u = c[3 + t] >> 1;
                            if (u == v) continue;
                            const uint32_t j = rank[u];
                            if (j >= i || kind[j] != K_CAND) continue;
                            const uint8_t st = __atomic_load_n(&status[j], __ATOMIC_RELAXED);
                            if (st == S_IN) { in_found = true; break; }
                            if (st == S_UNDEC) pending = true;
                        }
                    }
                }
                if (__ballot(in_found)) { in_found = true; break; }
            }
        }
        pending = __ballot(pending) != 0;
        if (lane == 0) {
            if (in_found) __atomic_store_n(&status[i], (uint8_t)S_OUT, __ATOMIC_RELAXED);
            else if (!pending) __atomic_store_n(&status[i], (uint8_t)S_IN, __ATOMIC_RELAXED);
            else next[atomicAdd(&cnt[CNT_A], 1ull)] = i;
        }
    }
}

__global__ void k_dynbreak(const uint8_t *__restrict__ kind, const uint8_t *__restrict__ status, uint64_t b0, unsigned long long *cnt)
{
    for (uint64_t i = gtid(); i < b0; i += gstride())
        if (kind[i] == K_DYNBREAK && status[i] == S_IN) atomicMin(&cnt[CNT_B], (unsigned long long)i);
}

__global__ void k_elect_flags(const uint8_t *__restrict__ kind, const uint8_t *__restrict__ status, uint64_t brk,
                              uint32_t nvars, uint32_t *__restrict__ flags)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride())
        flags[i] = (i < brk && kind[i] == K_CAND && status[i] == S_IN) ? 1u : 0u;
}

uint32_t lcve(Ctx &c)
{
    const uint32_t nv = c.nvars;
    hist(c);   // histSimp after prop (src/cpu/lcve.cpp:30)
    ensure(c, c.scores, nv + 1);
    ensure(c, c.eligible, std::max<uint64_t>(c.nlits2, nv + 1));
    ensure(c, c.rank, nv + 1);
    ensure(c, c.tmpu8, 2ull * nv + 2);
    ensure(c, c.tmpu32a, nv + 1);
    ensure(c, c.tmpu32b, nv + 1);
    ensure(c, c.pvs, nv + 1);
    k_scores<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.hist.p, nv, c.scores.p, c.eligible.p);
    const uint32_t maxs = reduce_max_u32(c, c.scores.p, nv);
    int bits = 0;
    while (bits < 32 && (maxs >> bits)) bits++;
    {
        KTimer kt(c, K_LCVE_RADIX, 16.0 * nv * ((bits + 7) / 8));
        radix_sort_u32_pairs(c, c.scores.p, c.eligible.p, nv, ((bits + 7) / 8) * 8);
    }
    k_rank<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.eligible.p, nv, c.rank.p);
    uint8_t *kind = c.tmpu8.p, *status = c.tmpu8.p + nv + 1;
    const uint32_t pt = c.opts.mu_pos << c.mu_inc, nt = c.opts.mu_neg << c.mu_inc;
    const unsigned long long init[2] = {0ull, (unsigned long long)nv};
    PFG_HIP(hipMemcpyAsync(c.cnt.p + CNT_A, init, 16, hipMemcpyHostToDevice, c.stream));
    k_classify<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.eligible.p, nv, c.hist.p, c.vstate.p, c.ot.p, pt, nt,
                                                              kind, status, (unsigned long long *)c.cnt.p);
    read_counters(c, 3);
    const uint64_t b0 = c.hcnt[CNT_B];
    k_work_init<<<grid_for(b0, 256, 8192), 256, 0, c.stream>>>(kind, b0, c.tmpu32a.p, (unsigned long long *)c.cnt.p);
    read_counters(c, 2);
    uint64_t nwork = c.hcnt[CNT_A];
    uint32_t *work = c.tmpu32a.p, *next = c.tmpu32b.p;
    int rounds = 0;
    while (nwork) {
        PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_A, 0, 8, c.stream));
        KTimer kt(c, K_LCVE_MIS, 0.0);
        k_mis_round<<<grid_for(nwork * 64, 256, 8192), 256, 0, c.stream>>>(work, nwork, c.eligible.p, c.rank.p, kind, status,
                                                                            c.ot.p, c.occ.p, c.mem.p, next,
                                                                            (unsigned long long *)c.cnt.p);
        read_counters(c, 2);
        nwork = c.hcnt[CNT_A];
        std::swap(work, next);
        if (++rounds > (1 << 22)) throw Error(PFG_EINTERNAL, "LCVE election did not converge");
    }
    const unsigned long long bb = b0;
    PFG_HIP(hipMemcpyAsync(c.cnt.p + CNT_B, &bb, 8, hipMemcpyHostToDevice, c.stream));
    k_dynbreak<<<grid_for(b0, 256, 8192), 256, 0, c.stream>>>(kind, status, b0, (unsigned long long *)c.cnt.p);
    read_counters(c, 3);
    const uint64_t brk = c.hcnt[CNT_B];
    k_elect_flags<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(kind, status, brk, nv, c.tmpu32a.p);
    c.npvs = (uint32_t)compact_u32(c, c.eligible.p, c.tmpu32a.p, nv, c.pvs.p);
    return c.npvs;
}

}  // namespace pfg
