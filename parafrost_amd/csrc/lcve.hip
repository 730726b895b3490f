// LCVE election on the device.
//
// The reference elects variables greedily in (score, var) order, score = ps*ns
// (uint32), skipping caller-frozen (incremental ifreeze), round-frozen, assigned,
// eliminated and absent variables, stopping at
// the first candidate whose both occurrence lists reach the temperatures, and
// freezing every cold neighbour of an elected variable (src/cpu/lcve.cpp:46-99).
// That sequential greedy is reproduced exactly by a deterministic parallel
// greedy maximal independent set over candidate ranks:
//   * candidates before the static break position B0 (first non-skipped candidate
//     that is hot by list size and cannot be frozen) are all freezable, and the
//     "shares a live clause" relation between them is symmetric;
//   * a candidate is elected iff no earlier-ranked elected candidate shares a live
//     clause with it;
//   * a candidate that is hot by list size but freezable ends the election where
//     the sequential loop would reach it unfrozen (dynamic break); it never freezes.
// Rounds are clause-centric streaming passes over the live clauses that hold two
// or more candidates:
//   A. a clause with an elected (freezing) candidate marks its undecided candidates
//      out; otherwise every undecided candidate ranked after the clause's earliest
//      undecided freezing candidate is blocked for this round;
//   B. every undecided candidate that is not blocked is elected.
// The earliest undecided candidate is never blocked, so every round makes progress;
// settled clauses and decided candidates are dropped between rounds.
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"
#include <cstdio>

namespace pfg {

// K_HALF (option lcve_clause_max only): a candidate whose negative list holds a clause
// over the limit; reached unfrozen it freezes the cold neighbours of its positive list
// and is not elected (src/gpu/lcve.cu:105-127)
enum { K_SKIP = 0, K_CAND = 1, K_DYNBREAK = 2, K_BREAK = 3, K_HALF = 4 };
enum { S_UNDEC = 0, S_IN = 1, S_OUT = 2 };
constexpr uint32_t NO_CAND = 0xFFFFFFFFu;

__global__ void k_scores(const uint32_t *__restrict__ hist, uint32_t nvars, uint32_t *__restrict__ keys, uint32_t *__restrict__ vars)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) {
        const uint32_t v = (uint32_t)i + 1;
        keys[i] = hist[2 * v] * hist[2 * v + 1];   // uint32 product as in hist_score (histogram.h:26)
        vars[i] = v;
    }
}


// Per-instance parameters of a batch (a single instance is the one-entry case):
// instance i owns variables (ibase[i], ibase[i + 1]], elects with temperatures
// mu << imu[i] (IMU_OFF: not electing this round), and gets its own static break
// ib0, dynamic break ibrk and elected count inel.
constexpr uint8_t IMU_OFF = 0xFFu;
struct InstArgs {
    const uint32_t *ibase;
    uint32_t ninst;
    const uint8_t *imu;
    uint32_t mu_pos, mu_neg;
    unsigned long long *ib0, *ibrk, *inel;   // inel is counted for batches only
    int batch;
    // the GPU reference's optional limits (src/gpu/lcve.cu:68, 105-127; off when 0 / null):
    // a candidate with more than lmax live occurrences of either literal is a break;
    // lng[lit] marks literals with a live clause over lcve_clause_max, half[v] the
    // K_HALF candidates
    uint32_t lmax;
    const uint8_t *lng;
    uint8_t *half;
};

__global__ void k_classify(const uint32_t *__restrict__ elig, uint32_t nvars, const uint32_t *__restrict__ hist,
                           const uint8_t *__restrict__ vstate, const uint8_t *__restrict__ ifrozen,
                           const OL *__restrict__ ot, InstArgs I,
                           uint8_t *__restrict__ kind, uint8_t *__restrict__ status, uint8_t *__restrict__ blocked)
{
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x; b < nvars; b += gstride()) {
        const uint64_t i = b + threadIdx.x;
        uint8_t k = K_SKIP;
        uint32_t inst = 0;
        if (i < nvars) {
            const uint32_t v = elig[i];
            inst = find_inst(I.ibase, I.ninst, v);
            const uint8_t mi = I.imu[inst];
            const uint32_t ps = hist[2 * v], ns = hist[2 * v + 1];
            // iassumed (lcve.cpp:57) and eliminated / assigned (lcve.cpp:58) variables are passed over
            if (mi != IMU_OFF && !ifrozen[v] && !vstate[v] && (ps || ns)) {
                const uint32_t pt = I.mu_pos << mi, nt = I.mu_neg << mi;
                const bool hot = (ot[2 * v].sz >= pt && ot[2 * v + 1].sz >= nt) ||
                                 (I.lmax && (ps > I.lmax || ns > I.lmax));
                const bool freezable = ps < pt || ns < nt;
                k = (hot && !freezable) ? K_BREAK : hot ? K_DYNBREAK : K_CAND;
                // depFreeze on the positive list, then the negative one, each failing on a
                // long clause: a positive failure leaves nothing, a negative one the
                // positive list's freezes
                if (k == K_CAND && I.lng) k = I.lng[2 * v] ? K_SKIP : I.lng[2 * v + 1] ? K_HALF : K_CAND;
            }
            if (I.half) I.half[v] = k == K_HALF;
            kind[i] = k;
            status[i] = S_UNDEC;
            blocked[i] = 0;
        }
        // static break = first K_BREAK position per instance: lanes hold consecutive
        // positions, so the lowest such lane of each instance in the wave takes the
        // atomic for its group
        uint64_t m = __ballot(k == K_BREAK);
        while (m) {
            const int leader = __ffsll((long long)m) - 1;
            const uint32_t li = __shfl(inst, leader, 64);
            const uint64_t grp = __ballot(k == K_BREAK && inst == li);
            if (__lane_id() == leader) atomicMin(&I.ib0[li], (unsigned long long)i);
            m &= ~grp;
        }
    }
}

// option lcve_clause_max: the literals of every live clause longer than the limit
__global__ void k_long_lits(const uint32_t *__restrict__ mem, const uint64_t *__restrict__ refs, uint64_t n,
                            uint32_t limit, uint8_t *__restrict__ lng)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t *c = mem + refs[i];
        if (c_del(c) || (uint32_t)c_size(c) <= limit) continue;
        for (int t = 0; t < c_size(c); t++) lng[c[3 + t]] = 1;
    }
}

struct ElectArgs {
    const uint32_t *mem;
    const uint32_t *crank;    // per variable: rank << 1 | freezer, or NO_CAND
    const uint8_t *kind;
    uint8_t *status, *blocked;
};

// per variable (indexed by variable, not rank): its rank among the candidates before
// its instance's static break and whether it freezes neighbours, so the clause pass
// needs one gather per literal; the candidates' ranks also form the first work list
__global__ void __launch_bounds__(256) k_crank(const uint32_t *__restrict__ elig, uint32_t nvars,
                                              const uint8_t *__restrict__ kind, InstArgs I, uint32_t *__restrict__ crank,
                                              uint32_t *__restrict__ work, unsigned long long *nwork)
{
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x; b < nvars; b += gstride()) {
        const uint64_t i = b + threadIdx.x;
        bool cand = false;
        if (i < nvars) {
            const uint32_t v = elig[i];
            const uint8_t k = kind[i];
            if (k == K_CAND || k == K_DYNBREAK || k == K_HALF) cand = i < I.ib0[find_inst(I.ibase, I.ninst, v)];
            crank[v] = cand ? (((uint32_t)i << 1) | (k != K_DYNBREAK ? 1u : 0u)) : NO_CAND;
        }
        block_append(cand, (uint32_t)i, work, nwork);
    }
}

// Live clauses with at least two candidates take part in the election. Each is
// written once as a compact record [k, code_1 .. code_k] of its candidates (code =
// rank << 1 | freezes neighbours), so the
// rounds stream records instead of clauses and rank lookups. Round 0's pass A is
// fused in: nothing is elected yet, so every candidate ranked after the clause's
// earliest freezing candidate is blocked. The first ELECT_CACHE candidate codes of
// a clause stay in LDS between the counting and the writing loop. With HALF (option
// lcve_clause_max) the code of a K_HALF candidate's negative literal has no freezer bit:
// its freezing reaches the clauses of its positive list only.
constexpr int ELECT_CACHE = 16;
template <bool HALF>
__global__ void __launch_bounds__(256) k_elect_init(ElectArgs A, const uint64_t *__restrict__ refs, uint64_t n,
                                                   uint32_t *__restrict__ rec, uint64_t *__restrict__ active,
                                                   unsigned long long *nactive, unsigned long long *cnt,
                                                   const uint8_t *__restrict__ half)
{
    auto code = [&](uint32_t a, uint32_t x) -> uint32_t {
        if (HALF && (a & 1u) && x != NO_CAND && half[a >> 1]) return x & ~1u;
        return x;
    };
    __shared__ uint32_t s_cr[ELECT_CACHE][256];
    const unsigned tid = threadIdx.x;
    for (uint64_t b = xcd_block() * blockDim.x; b < n; b += gstride()) {
        const uint64_t i = b + tid;
        const uint32_t *c = nullptr;
        int sz = 0;
        uint32_t k = 0, m = 0xFFFFFFFFu;
        if (i < n) {
            c = A.mem + refs[i];
            if (!c_del(c)) {
                sz = c_size(c);
                auto take = [&](uint32_t x) {
                    if (x == NO_CAND) return;
                    if (k < (uint32_t)ELECT_CACHE) s_cr[k][tid] = x;
                    k++;
                    if ((x & 1u) && (x >> 1) < m) m = x >> 1;
                };
                int t = 0;
                for (; t + 4 <= sz; t += 4) {   // four rank gathers in flight, taken in order
                    const uint32_t a0 = c[3 + t], a1 = c[4 + t], a2 = c[5 + t], a3 = c[6 + t];
                    const uint32_t x0 = code(a0, A.crank[a0 >> 1]), x1 = code(a1, A.crank[a1 >> 1]),
                                   x2 = code(a2, A.crank[a2 >> 1]), x3 = code(a3, A.crank[a3 >> 1]);
                    take(x0);
                    take(x1);
                    take(x2);
                    take(x3);
                }
                for (; t < sz; t++) take(code(c[3 + t], A.crank[c[3 + t] >> 1]));
            }
        }
        const bool keep = k >= 2u;
        const uint64_t off = block_reserve(keep ? 1ull + k : 0ull, &cnt[CNT_D]);
        const uint64_t aoff = block_reserve(keep ? 1ull : 0ull, nactive);
        if (keep) {
            uint32_t *r = rec + off;
            r[0] = k;
            const uint32_t kc = k < (uint32_t)ELECT_CACHE ? k : (uint32_t)ELECT_CACHE;
            for (uint32_t w = 0; w < kc; w++) {
                const uint32_t x = s_cr[w][tid], j = x >> 1;
                r[1 + w] = x;
                if (j > m) A.blocked[j] = 1;
            }
            if (k > (uint32_t)ELECT_CACHE) {
                uint32_t w = 0;
                for (int t = 0; t < sz; t++) {
                    const uint32_t x = code(c[3 + t], A.crank[c[3 + t] >> 1]);
                    if (x == NO_CAND) continue;
                    if (++w <= (uint32_t)ELECT_CACHE) continue;
                    const uint32_t j = x >> 1;
                    r[w] = x;
                    if (j > m) A.blocked[j] = 1;
                }
            }
            active[aoff] = off;
        }
    }
}

// pass A over the active clause records (count read on the device: rounds are
// launched back to back and the host looks at the counts every few rounds)
__device__ __forceinline__ void elect_pass_a(const ElectArgs &A, const uint32_t *__restrict__ rec,
                                             const uint64_t *__restrict__ active, const unsigned long long *nactive,
                                             uint64_t *__restrict__ next, unsigned long long *nnext)
{
    const uint64_t n = *nactive;
    for (uint64_t b = xcd_block() * blockDim.x; b < n; b += gstride()) {
        const uint64_t i = b + threadIdx.x;
        bool keep = false;
        uint64_t r = 0;
        if (i < n) {
            // the records stream through once (non-temporal loads), so the status bytes
            // they gather stay cached
            r = __builtin_nontemporal_load(active + i);
            const uint32_t *rr = rec + r;
            const uint32_t k = __builtin_nontemporal_load(rr);
            // fz: the earliest elected candidate whose code freezes (none: ~0). Only the
            // undecided candidates ranked after it are frozen by it: one ranked before
            // it can only be a candidate that does not freeze through this clause (a
            // dynamic break, the negative literal of a K_HALF candidate), which an
            // earlier freezing candidate of the clause would have blocked otherwise,
            // and which the sequential walk reaches before the freezer.
            uint32_t fz = 0xFFFFFFFFu;
            uint32_t m = 0xFFFFFFFFu;
            uint32_t undec = 0, alive = 0;
            constexpr uint32_t KR = 8;   // undecided candidates of the first KR entries are kept in registers
            uint32_t und[KR];
            uint32_t nund = 0;
            auto take = [&](uint32_t t, uint32_t x, uint8_t st) {
                const uint32_t j = x >> 1;
                const bool freezer = x & 1u;
                if (st == S_IN && freezer && j < fz) fz = j;
                if (st == S_UNDEC) {
                    undec++;
                    if (freezer && j < m) m = j;
                    if (t <= KR) und[nund++] = j;
                }
                if (st == S_UNDEC || (st == S_IN && freezer)) alive++;
            };
            uint32_t t = 1;
            for (; t + 3 <= k; t += 4) {   // four codes and their status gathers in flight, taken in order
                const uint32_t x0 = __builtin_nontemporal_load(rr + t), x1 = __builtin_nontemporal_load(rr + t + 1),
                               x2 = __builtin_nontemporal_load(rr + t + 2), x3 = __builtin_nontemporal_load(rr + t + 3);
                const uint8_t s0 = A.status[x0 >> 1], s1 = A.status[x1 >> 1], s2 = A.status[x2 >> 1],
                              s3 = A.status[x3 >> 1];
                take(t, x0, s0);
                take(t + 1, x1, s1);
                take(t + 2, x2, s2);
                take(t + 3, x3, s3);
            }
            for (; t <= k; t++) {
                const uint32_t x = __builtin_nontemporal_load(rr + t);
                take(t, x, A.status[x >> 1]);
            }
            if (undec) {
                // a candidate seen undecided above that another record has marked out since
                // gets the same mark again, or a blocked mark that pass B ignores
                for (uint32_t q = 0; q < nund; q++) {
                    const uint32_t j = und[q];
                    if (j > fz) A.status[j] = S_OUT;
                    else if (j > m) A.blocked[j] = 1;
                }
                for (uint32_t t = KR + 1; t <= k; t++) {
                    const uint32_t j = rr[t] >> 1;
                    if (A.status[j] != S_UNDEC) continue;
                    if (j > fz) A.status[j] = S_OUT;
                    else if (j > m) A.blocked[j] = 1;
                }
            }
            keep = fz == 0xFFFFFFFFu && undec >= 1u && alive >= 2u;
        }
        const uint64_t o = block_reserve(keep ? 1ull : 0ull, nnext);
        if (keep) next[o] = r;
    }
}

__global__ void __launch_bounds__(1024) k_elect_pass_a(ElectArgs A, const uint32_t *__restrict__ rec,
                                                      const uint64_t *__restrict__ active,
                                                      const unsigned long long *__restrict__ nactive,
                                                      uint64_t *__restrict__ next, unsigned long long *nnext)
{
    elect_pass_a(A, rec, active, nactive, next, nnext);
}

// pass B over the undecided candidates
__device__ __forceinline__ void elect_pass_b(const ElectArgs &A, const uint32_t *__restrict__ work,
                                             const unsigned long long *nwork, uint32_t *__restrict__ next,
                                             unsigned long long *nnext)
{
    const uint64_t n = *nwork;
    for (uint64_t b = xcd_block() * blockDim.x; b < n; b += gstride()) {
        const uint64_t q = b + threadIdx.x;
        bool keep = false;
        uint32_t j = 0;
        if (q < n) {
            j = work[q];
            if (A.status[j] == S_UNDEC) {
                if (A.blocked[j]) { A.blocked[j] = 0; keep = true; }
                else A.status[j] = S_IN;
            }
        }
        block_append(keep, j, next, nnext);
    }
}

__global__ void __launch_bounds__(1024) k_elect_pass_b(ElectArgs A, const uint32_t *__restrict__ work,
                                                      const unsigned long long *__restrict__ nwork,
                                                      uint32_t *__restrict__ next, unsigned long long *nnext)
{
    elect_pass_b(A, work, nwork, next, nnext);
}

// The election rounds as a captured graph (tuning lcve_graph): the same passes, with
// the round number r read from the device (*round), so one captured sequence of rounds
// replays until no candidate is undecided. Round r's pass A reads the active records
// of buffer (r odd ? act0 : act1) and writes the other; pass B reads the work list of
// (r even ? work0 : work1). Rounds past MAXR - 1 do nothing (the host reports it).
struct ElectRounds {
    const uint32_t *rec;
    uint64_t *act0, *act1;
    uint32_t *work0, *work1;
    unsigned long long *rw, *ra, *round;
    uint32_t maxr;
};

__global__ void __launch_bounds__(1024) k_elect_pass_a_g(ElectArgs A, ElectRounds R)
{
    const uint32_t r = (uint32_t)*R.round;
    if (r == 0 || r + 1 >= R.maxr) return;
    elect_pass_a(A, R.rec, (r & 1u) ? R.act0 : R.act1, R.ra + r, (r & 1u) ? R.act1 : R.act0, R.ra + r + 1);
}

__global__ void __launch_bounds__(1024) k_elect_pass_b_g(ElectArgs A, ElectRounds R)
{
    const uint32_t r = (uint32_t)*R.round;
    if (r + 1 >= R.maxr) return;
    elect_pass_b(A, (r & 1u) ? R.work1 : R.work0, R.rw + r, (r & 1u) ? R.work0 : R.work1, R.rw + r + 1);
}

__global__ void k_elect_round_next(ElectRounds R)
{
    if (threadIdx.x == 0 && *R.round + 1 < R.maxr) *R.round += 1;
}

__global__ void k_dynbreak(const uint32_t *__restrict__ elig, uint32_t nvars, const uint8_t *__restrict__ kind,
                           const uint8_t *__restrict__ status, InstArgs I)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride())
        if (kind[i] == K_DYNBREAK && status[i] == S_IN)
            atomicMin(&I.ibrk[find_inst(I.ibase, I.ninst, elig[i])], (unsigned long long)i);
}

__global__ void k_elect_flags(const uint32_t *__restrict__ elig, const uint8_t *__restrict__ kind,
                              const uint8_t *__restrict__ status, uint32_t nvars, InstArgs I, uint32_t *__restrict__ flags)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) {
        uint32_t f = 0;
        if (kind[i] == K_CAND && status[i] == S_IN) {
            const uint32_t inst = find_inst(I.ibase, I.ninst, elig[i]);
            if (i < I.ibrk[inst]) f = 1;
        }
        flags[i] = f;
    }
}

// batch: elected variables per instance. The elected list is in (instance, score,
// variable) order, so an instance's variables form one run: found by two binary searches.
__global__ void k_inst_elected(const uint32_t *__restrict__ pvs, uint32_t npvs, InstArgs I)
{
    for (uint64_t inst = gtid(); inst < I.ninst; inst += gstride()) {
        uint32_t lo = 0, hi = npvs;   // first position of an instance >= inst
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (find_inst(I.ibase, I.ninst, pvs[mid]) < inst) lo = mid + 1;
            else hi = mid;
        }
        uint32_t lo2 = lo, hi2 = npvs;   // first position of an instance > inst
        while (lo2 < hi2) {
            const uint32_t mid = (lo2 + hi2) >> 1;
            if (find_inst(I.ibase, I.ninst, pvs[mid]) <= inst) lo2 = mid + 1;
            else hi2 = mid;
        }
        I.inel[inst] = lo2 - lo;
    }
}

// batch: order candidates by (instance, score, variable): a stable sort of the
// (score, variable) order by instance
__global__ void k_inst_keys(const uint32_t *__restrict__ elig, uint32_t nvars, InstArgs I, uint32_t *__restrict__ keys)
{
    for (uint64_t i = gtid(); i < nvars; i += gstride()) keys[i] = find_inst(I.ibase, I.ninst, elig[i]);
}

// algorithmic bytes of one pass A (accounting only): per active record its offset
// (8), the record (4 (k + 1)) and one status byte per candidate
__global__ void k_elect_bytes(const uint32_t *__restrict__ rec, const uint64_t *__restrict__ active,
                              const unsigned long long *__restrict__ nactive, unsigned long long *__restrict__ out)
{
    const uint64_t n = *nactive;
    unsigned long long acc = 0;
    for (uint64_t i = gtid(); i < n; i += gstride()) acc += 12ull + 5ull * rec[active[i]];
    acc = block_sum(acc);
    if (threadIdx.x == 0 && acc) atomicAdd(out, acc);
}

static double elect_bytes(Ctx &c, const uint64_t *active, const unsigned long long *nactive, uint64_t bound)
{
    if (!c.account || !bound) return 0.0;
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_ACCT, 0, 8, c.stream));
    k_elect_bytes<<<grid_for(bound, 256, 4096), 256, 0, c.stream>>>(c.erec.p, active, nactive,
                                                                     (unsigned long long *)c.cnt.p + CNT_ACCT);
    uint64_t v = 0;
    PFG_HIP(hipMemcpyAsync(&v, c.cnt.p + CNT_ACCT, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    return (double)v;
}

static int bits_of32(uint32_t x)
{
    int b = 0;
    while (b < 32 && (x >> b)) b++;
    return b;
}

static InstArgs inst_args(Ctx &c)
{
    return InstArgs{c.ibase.p, c.ninst, c.imu.p, c.opts.mu_pos, c.opts.mu_neg, (unsigned long long *)c.ib0.p,
                    (unsigned long long *)c.ib0.p + c.ninst, (unsigned long long *)c.inel.p, c.batch ? 1 : 0,
                    c.opts.lcve_max, nullptr, nullptr};
}

uint32_t lcve(Ctx &c)
{
    const uint32_t nv = c.nvars;
    if (!c.hist_live) hist(c);   // histSimp after prop (src/cpu/lcve.cpp:30)
    ensure(c, c.scores, nv + 1);
    ensure(c, c.eligible, std::max<uint64_t>(c.nlits2, nv + 1));
    ensure(c, c.rank, nv + 1);
    const bool lim = c.opts.lcve_clause_max != 0;
    ensure(c, c.tmpu8, (lim ? 6ull : 3ull) * (nv + 1));
    ensure(c, c.tmpu32a, nv + 1);
    ensure(c, c.tmpu32b, nv + 1);
    ensure(c, c.pvs, nv + 1);
    ensure(c, c.tmpu64a, c.nrefs + 1);
    ensure(c, c.tmpu64b, c.nrefs + 1);
    ensure(c, c.erec, c.msz + 1);
    k_scores<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.hist.p, nv, c.scores.p, c.eligible.p);
    const uint32_t maxs = reduce_max_u32(c, c.scores.p, nv);
    int bits = 0;
    while (bits < 32 && (maxs >> bits)) bits++;
    {
        KTimer kt(c, K_LCVE_RADIX, 16.0 * nv * ((bits + 7) / 8));
        radix_sort_u32_pairs(c, c.scores.p, c.eligible.p, nv, ((bits + 7) / 8) * 8);
    }
    InstArgs I = inst_args(c);
    uint8_t *half = nullptr;
    if (lim) {   // lng (2 nv + 2 bytes), then half (nv + 1), after kind / status / blocked
        uint8_t *lng = c.tmpu8.p + 3ull * (nv + 1);
        half = lng + 2ull * (nv + 1);
        PFG_HIP(hipMemsetAsync(lng, 0, 3ull * (nv + 1), c.stream));
        if (c.nrefs)
            k_long_lits<<<grid_for(c.nrefs, 256, 8192), 256, 0, c.stream>>>(c.mem.p, c.refs.p, c.nrefs,
                                                                             c.opts.lcve_clause_max, lng);
        I.lng = lng;
        I.half = half;
    }
    if (c.ninst > 1) {   // (instance, score, variable) order
        k_inst_keys<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.eligible.p, nv, I, c.scores.p);
        radix_sort_u32_pairs(c, c.scores.p, c.eligible.p, nv, bits_of32(c.ninst - 1));
    }
    if (!c.batch) c.imu_host[0] = (uint8_t)c.mu_inc;
    PFG_HIP(hipMemcpyAsync(c.imu.p, c.imu_host.data(), c.ninst, hipMemcpyHostToDevice, c.stream));
    PFG_HIP(hipMemsetAsync(c.ib0.p, 0xFF, 16ull * c.ninst, c.stream));   // ib0 and ibrk: no break
    PFG_HIP(hipMemsetAsync(c.inel.p, 0, 8ull * c.ninst, c.stream));
    uint8_t *kind = c.tmpu8.p, *status = kind + nv + 1, *blocked = status + nv + 1;
    k_classify<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.eligible.p, nv, c.hist.p, c.vstate.p, c.ifrozen.p, c.ot.p, I,
                                                              kind, status, blocked);
    // per-round counts on the device: rw[r] = undecided candidates entering round r,
    // ra[r] = active clause records entering round r's pass A (rounds 1..)
    constexpr int MAXR = 256;
    ensure(c, c.tmpu64c, 2ull * MAXR + 2);
    unsigned long long *rw = (unsigned long long *)c.tmpu64c.p, *ra = rw + MAXR + 1;
    PFG_HIP(hipMemsetAsync(rw, 0, (2ull * MAXR + 2) * 8, c.stream));
    uint32_t *work = c.tmpu32a.p, *work2 = c.tmpu32b.p;
    uint64_t *act = c.tmpu64a.p, *act2 = c.tmpu64b.p;
    k_crank<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.eligible.p, nv, kind, I, c.rank.p, work, rw);
    ElectArgs A{c.mem.p, c.rank.p, kind, status, blocked};
    if (c.nrefs) {
        KTimer kt(c, K_LCVE_INIT, 20.0 * c.nrefs + 8.0 * c.nliterals);
        PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_D, 0, 8, c.stream));
        const unsigned g = xcd_grid(c.nrefs, 256, c.tune.init_grid ? c.tune.init_grid : c.tune.clause_grid);
        if (half)
            k_elect_init<true><<<g, 256, 0, c.stream>>>(A, c.refs.p, c.nrefs, c.erec.p, act, ra + 1,
                                                        (unsigned long long *)c.cnt.p, half);
        else
            k_elect_init<false><<<g, 256, 0, c.stream>>>(A, c.refs.p, c.nrefs, c.erec.p, act, ra + 1,
                                                         (unsigned long long *)c.cnt.p, nullptr);
    }
    uint64_t hw[2] = {0, 0};
    PFG_HIP(hipMemcpyAsync(hw, rw, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipMemcpyAsync(hw + 1, ra + 1, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    uint64_t nwork = hw[0], nact = hw[1];   // upper bounds of later rounds' counts (sizes the grids)
    if (c.trace)
        fprintf(stderr, "[pfg] LCVE candidates %llu clauses %llu\n", (unsigned long long)nwork, (unsigned long long)nact);
    // round 0's pass A ran inside k_elect_init; rounds are launched in batches and
    // the count of undecided candidates is checked after each batch
    int r = 0;
    if (c.tune.lcve_graph && nwork && !c.account) {
        // round 0 (pass B only) on the stream, then captured batches of GRAPH_ROUNDS rounds
        {
            KTimer kt(c, K_LCVE_PASS_B, 0.0);
            k_elect_pass_b<<<xcd_grid(nwork, 1024, std::max(8u, c.tune.clause_grid / 4)), 1024, 0, c.stream>>>(
                A, work, rw, work2, rw + 1);
        }
        r = 1;
        constexpr int GRAPH_ROUNDS = 4;
        const unsigned long long one = 1;
        unsigned long long *round = (unsigned long long *)c.cnt.p + CNT_ROUND;
        PFG_HIP(hipMemcpyAsync(round, &one, 8, hipMemcpyHostToDevice, c.stream));
        const ElectRounds R{c.erec.p, act, act2, work, work2, rw, ra, round, (uint32_t)MAXR};
        const unsigned ga = xcd_grid(nact, 1024, std::max(8u, c.tune.clause_grid / 4));
        const unsigned gb = xcd_grid(nwork, 1024, std::max(8u, c.tune.clause_grid / 4));
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        PFG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < GRAPH_ROUNDS; k++) {
            k_elect_pass_a_g<<<ga, 1024, 0, c.stream>>>(A, R);
            k_elect_pass_b_g<<<gb, 1024, 0, c.stream>>>(A, R);
            k_elect_round_next<<<1, 64, 0, c.stream>>>(R);
        }
        PFG_HIP(hipStreamEndCapture(c.stream, &graph));
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(graph);
            throw Error(PFG_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
        try {
            while (nwork) {
                {
                    KTimer kt(c, K_LCVE_GRAPH, 0.0);
                    PFG_HIP(hipGraphLaunch(exec, c.stream));
                }
                r += GRAPH_ROUNDS;
                if (r >= MAXR - 1) throw Error(PFG_EINTERNAL, "LCVE election did not converge");
                PFG_HIP(hipMemcpyAsync(hw, rw + r, 8, hipMemcpyDeviceToHost, c.stream));
                PFG_HIP(hipMemcpyAsync(hw + 1, ra + r, 8, hipMemcpyDeviceToHost, c.stream));
                PFG_HIP(hipStreamSynchronize(c.stream));
                nwork = hw[0];
                nact = hw[1];
            }
        }
        catch (...) {
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            throw;
        }
        PFG_HIP(hipGraphExecDestroy(exec));
        PFG_HIP(hipGraphDestroy(graph));
        c.path[PATH_LCVE_GRAPH]++;
    }
    while (nwork) {
        const int batch = r == 0 ? 2 : 4;
        for (int k = 0; k < batch && r < MAXR - 1; k++, r++) {
            if (r > 0) {
                KTimer kt(c, K_LCVE_PASS_A, elect_bytes(c, act, ra + r, nact));
                k_elect_pass_a<<<xcd_grid(nact, 1024, std::max(8u, c.tune.clause_grid / 4)), 1024, 0, c.stream>>>(A, c.erec.p, act, ra + r, act2,
                                                                                           ra + r + 1);
                std::swap(act, act2);
            }
            {
                KTimer kt(c, K_LCVE_PASS_B, 0.0);
                k_elect_pass_b<<<xcd_grid(nwork, 1024, std::max(8u, c.tune.clause_grid / 4)), 1024, 0, c.stream>>>(A, work, rw + r, work2,
                                                                                            rw + r + 1);
            }
            std::swap(work, work2);
        }
        if (r >= MAXR - 1) throw Error(PFG_EINTERNAL, "LCVE election did not converge");
        PFG_HIP(hipMemcpyAsync(hw, rw + r, 8, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipMemcpyAsync(hw + 1, ra + r, 8, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
        if (c.trace) fprintf(stderr, "[pfg] LCVE after round %d: undecided %llu, clauses %llu\n", r - 1,
                             (unsigned long long)hw[0], (unsigned long long)hw[1]);
        nwork = hw[0];
        nact = hw[1];
    }
    k_dynbreak<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.eligible.p, nv, kind, status, I);
    k_elect_flags<<<grid_for(nv, 256, 8192), 256, 0, c.stream>>>(c.eligible.p, kind, status, nv, I, c.tmpu32a.p);
    c.npvs = (uint32_t)compact_u32(c, c.eligible.p, c.tmpu32a.p, nv, c.pvs.p);
    if (c.batch) k_inst_elected<<<grid_for(c.ninst, 256, 1024), 256, 0, c.stream>>>(c.pvs.p, c.npvs, I);
    c.hist_live = false;
    return c.npvs;
}

}  // namespace pfg
