// Subsumption with self-subsuming strengthening on the elected variables
// (ParaFROST src/cpu/subsume.h:27-176, elimination.cpp:108-130, 199-231).
//
// Elected variables share no live clause, so each variable's pass touches only
// its own clauses -- except for one effect: strengthening a binary clause into a
// unit enqueues the unit and deletes every clause on its list at once, before the
// later variables run. A unit can only come from a complementary binary pair
// (x|a, ~x|a) present when the variable starts, so variables without such a pair
// ("unit-capable" = no) run fully in parallel, and each unit-capable variable runs
// alone, in elected order, between the parallel ranges around it.
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"
#include "proof.hpp"
#include "stage.hpp"
#include <cstdio>
#include <vector>

namespace pfg {

struct SubArgs {
    uint32_t *mem;
    OL *ot;
    uint64_t *occ;
    int8_t *value;
    uint8_t *vstate;
    uint32_t *trail;
    unsigned long long *cnt;
    int sub_limit;
    int lbd_tier1;
    int allow_units;
    int defer_oblivion;   // a unit's clauses are deleted by the caller (k_sub_oblivion)
    ProofLog P;           // proof events: strengthened clauses, units, subsumed clauses
    uint64_t *rmlog;      // the occurrence-table update's removal log (strengthened literals)
    uint64_t rmcap;
};

__device__ __forceinline__ bool sub_cl(const uint32_t *sm, const uint32_t *lr)
{   // sub(SCLAUSE&, SCLAUSE&) (subsume.h:39-55)
    int it1 = 0, it2 = 0, sub = 0;
    const int n1 = c_size(sm), n2 = c_size(lr);
    const uint32_t *a = sm + 3, *b = lr + 3;
    while (it1 < n1 && it2 < n2) {
        const uint32_t x = a[it1], y = b[it2];
        if (x < y) it1++;
        else if (y < x) it2++;
        else { sub++; it1++; it2++; }
    }
    return sub == n1;
}

__device__ __forceinline__ bool selfsub_cl(uint32_t x, uint32_t fx, const uint32_t *sm, const uint32_t *lr)
{   // selfsub (subsume.h:72-99)
    int it1 = 0, it2 = 0, sub = 0;
    bool self = false;
    const int n1 = c_size(sm), n2 = c_size(lr);
    const uint32_t *a = sm + 3, *b = lr + 3;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2];
        if (l1 == fx) it1++;
        else if (l2 == x) { self = true; it2++; }
        else if (l1 < l2) it1++;
        else if (l2 < l1) it2++;
        else { sub++; it1++; it2++; }
    }
    if (sub + 1 == n1) {
        if (self) return true;
        while (it2 < n2) { if (b[it2] == x) return true; it2++; }
    }
    return false;
}

__device__ __forceinline__ void toblivion_list(const SubArgs &A, uint32_t lit)
{   // toblivion(OL&) (src/cpu/simplify.h:291-297)
    OL *o = A.ot + lit;
    const uint64_t *d = A.occ + o->off;
    for (uint32_t i = 0; i < o->sz; i++) c_markdel(A.mem + d[i]);
    o->sz = 0;
}

// returns the unit it assigned (0 if none). ref: the clause's pool reference (c may
// be a staged copy); q, seq: the proof event position
__device__ __forceinline__ uint32_t strengthen(const SubArgs &A, uint32_t *c, uint32_t me, uint64_t ref, uint32_t q,
                                               uint32_t &seq)
{   // strengthen (src/cpu/elimination.cpp:199-231)
    uint32_t sig = 0;
    int n = 0;
    uint32_t *lits = c + 3;
    const int sz = c_size(c);
    for (int k = 0; k < sz; k++) {
        const uint32_t lit = lits[k];
        if (lit != me) { lits[n++] = lit; sig |= maphash(lit); }
    }
    c[1] = sig;
    c[2] = (uint32_t)(sz - 1);
    log_removed(A.rmlog, &A.cnt[CNT_OTRM], A.rmcap, me, ref);
    if (n == 1) {
        if (!A.allow_units) { atomicOr(&A.cnt[CNT_FLAGS], (unsigned long long)FLAG_INTERNAL); return 0; }
        const uint32_t unit = lits[0];
        const int r = try_assign(A.value, unit);
        if (r == 1) {
            A.vstate[unit >> 1] = FROZEN_M;
            const unsigned long long pos = atomicAdd(&A.cnt[CNT_NTRAIL], 1ull);
            A.trail[pos] = unit;
            plog(A.P, q, seq, PE_UNIT, unit);   // learnUnit
            if (!A.defer_oblivion) toblivion_list(A, unit);
            return unit;
        }
        else if (r < 0) {
            report_conflict(A.cnt, unit >> 1);
            plog(A.P, q, seq, PE_CONFLICT, 0);
        }
    }
    else {
        if (A.P.ev) {   // addResolvent: not for a clause with a true literal
            bool sat = false;
            for (int k = 0; k < n; k++) sat |= A.value[lits[k]] == 1;
            if (!sat) plog(A.P, q, seq, PE_ADD, ref);
        }
        if (c_learnt(c)) {   // bumpShrunken (elimination.cpp:27-39)
            const uint32_t old = c[0] >> 6;
            if ((int)old > A.lbd_tier1) {
                const uint32_t nl = min((uint32_t)(c_size(c) - 1), old);
                if (nl < old) c[0] = (c[0] & 0xFu) | (1u << 4) | (nl << 6);
            }
        }
    }
    return 0;
}

__device__ __forceinline__ void update_ol(const SubArgs &A, uint32_t lit)
{   // updateOL (subsume.h:27-37)
    OL *o = A.ot + lit;
    uint64_t *d = A.occ + o->off;
    uint32_t j = 0;
    for (uint32_t i = 0; i < o->sz; i++) {
        uint32_t *c = A.mem + d[i];
        if (c_molten(c)) c_freeze(c);
        else if (!c_del(c)) d[j++] = d[i];
    }
    o->sz = j;
}

__device__ __forceinline__ void subsume_l(const SubArgs &A, const uint64_t *list, uint32_t end, uint32_t *cand,
                                          uint64_t ref, uint32_t q, uint32_t &seq)
{   // subsume (subsume.h:101-117)
    const int candsz = c_size(cand);
    for (uint32_t j = 0; j < end; j++) {
        uint32_t *sg = A.mem + list[j];
        if (c_del(sg)) continue;
        if (c_molten(cand) && c_size(sg) > candsz) continue;
        if (c_size(sg) > 1 && sig_sub(sg[1], cand[1]) && sub_cl(sg, cand)) {
            if (c_learnt(sg) && c_orig(cand)) c_setst(sg, ST_ORIGINAL);
            if (!c_del(cand)) {   // removeClause (sclause.cpp:77-84)
                c_markdel(cand);
                plog(A.P, q, seq, PE_DEL, ref);
            }
            return;
        }
    }
}

__device__ __forceinline__ void selfsubsume_l(const SubArgs &A, uint32_t x, uint32_t fx, uint32_t listlit, uint32_t *cand,
                                              uint64_t ref, uint32_t q, uint32_t &seq)
{   // selfsubsume (subsume.h:119-138)
    const OL *o = A.ot + listlit;
    const uint64_t *d = A.occ + o->off;
    const int candsz = c_size(cand);
    const uint32_t candsig = cand[1];
    for (uint32_t j = 0; j < o->sz; j++) {
        const uint32_t *sg = A.mem + d[j];
        const int subsize = c_size(sg);
        if (subsize > candsz) break;
        if (c_del(sg)) continue;
        if (subsize > 1 && sig_selfsub(sg[1], candsig) && selfsub_cl(x, fx, sg, cand)) {
            strengthen(A, cand, x, ref, q, seq);
            c_melt(cand);
            return;
        }
    }
}

__device__ __forceinline__ void self_sub_x(const SubArgs &A, uint32_t p, uint32_t q, uint32_t &seq)
{   // self_sub_x (subsume.h:140-176)
    const uint32_t n = p ^ 1u;
    for (int side = 0; side < 2; side++) {
        const uint32_t me = side ? n : p, other = side ? p : n;
        const OL *o = A.ot + me;
        const uint64_t *d = A.occ + o->off;
        const uint32_t sz = o->sz;
        for (uint32_t i = 0; i < sz; i++) {
            uint32_t *cand = A.mem + d[i];
            if (c_size(cand) > HSE_MAX_CL_SIZE) break;
            if (c_del(cand)) continue;
            selfsubsume_l(A, me, other, other, cand, d[i], q, seq);
            subsume_l(A, d, i, cand, d[i], q, seq);
        }
        update_ol(A, me);
    }
}

// The same pass over a staged variable (stage.hpp, with the original-order table):
// lists are table ranges, modified clause copies are marked dirty, and at the end
// the dirty clause images and the compacted lists are written back to the pool.
__device__ __forceinline__ void subsume_staged(const SubArgs &A, uint32_t *arena, const uint32_t *list, uint32_t end,
                                               uint32_t *cand, uint32_t q, uint32_t &seq)
{   // subsume (subsume.h:101-117)
    const int candsz = c_size(cand);
    for (uint32_t j = 0; j < end; j++) {
        uint32_t *sg = arena + list[j];
        if (c_del(sg)) continue;
        if (c_molten(cand) && c_size(sg) > candsz) continue;
        if (c_size(sg) > 1 && sig_sub(sg[1], cand[1]) && sub_cl(sg, cand)) {
            if (c_learnt(sg) && c_orig(cand)) { c_setst(sg, ST_ORIGINAL); staged_dirty(sg); }
            if (!c_del(cand)) {   // removeClause (sclause.cpp:77-84)
                c_markdel(cand);
                staged_dirty(cand);
                plog(A.P, q, seq, PE_DEL, staged_ref(cand));
            }
            return;
        }
    }
}

__device__ __forceinline__ void selfsubsume_staged(const SubArgs &A, uint32_t x, uint32_t fx, uint32_t *arena,
                                                   const uint32_t *list, uint32_t n, uint32_t *cand, uint32_t q,
                                                   uint32_t &seq)
{   // selfsubsume (subsume.h:119-138)
    const int candsz = c_size(cand);
    const uint32_t candsig = cand[1];
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t *sg = arena + list[j];
        const int subsize = c_size(sg);
        if (subsize > candsz) break;
        if (c_del(sg)) continue;
        if (subsize > 1 && sig_selfsub(sg[1], candsig) && selfsub_cl(x, fx, sg, cand)) {
            strengthen(A, cand, x, staged_ref(cand), q, seq);
            c_melt(cand);
            staged_dirty(cand);
            return;
        }
    }
}

__device__ __forceinline__ void self_sub_x_staged(const SubArgs &A, uint32_t p, uint32_t *tab, uint32_t *arena, uint32_t q,
                                                  uint32_t &seq)
{   // self_sub_x (subsume.h:140-176)
    uint32_t len[2] = {A.ot[p].sz, A.ot[p ^ 1u].sz};
    const uint32_t base[2] = {0u, len[0]}, ne = len[0] + len[1];
    for (int side = 0; side < 2; side++) {
        const uint32_t me = p ^ (uint32_t)side, other = me ^ 1u;
        uint32_t *mt = tab + base[side];
        const uint32_t ms = len[side];
        for (uint32_t i = 0; i < ms; i++) {
            uint32_t *cand = arena + mt[i];
            if (c_size(cand) > HSE_MAX_CL_SIZE) break;
            if (c_del(cand)) continue;
            selfsubsume_staged(A, me, other, arena, tab + base[1 - side], len[1 - side], cand, q, seq);
            subsume_staged(A, arena, mt, i, cand, q, seq);
        }
        uint32_t j = 0;   // updateOL (subsume.h:27-37)
        for (uint32_t i = 0; i < ms; i++) {
            uint32_t *c = arena + mt[i];
            if (c_molten(c)) c_freeze(c);
            else if (!c_del(c)) mt[j++] = mt[i];
        }
        len[side] = j;
    }
    for (uint32_t k = 0; k < ne; k++) {
        const uint32_t *c = arena + tab[ne + k];
        if (!staged_is_dirty(c)) continue;
        uint32_t *g = A.mem + staged_ref(c);
        const int n = c_size(c);
        for (int t = 0; t < 3 + n; t++) g[t] = c[t];
    }
    for (int side = 0; side < 2; side++) {
        OL *o = A.ot + (p ^ (uint32_t)side);
        if (len[side] == o->sz) continue;   // nothing left the list, so its order is unchanged
        uint64_t *d = A.occ + o->off;
        for (uint32_t k = 0; k < len[side]; k++) d[k] = staged_ref(arena + tab[base[side] + k]);
        o->sz = len[side];
    }
}

// ---- a lane group per elected variable (the parallel ranges: no unit can arise)
// The 64 lanes of a wave form 64 / G groups; a group takes one elected variable at a
// time. The variable's two lists and the images of their clauses are copied into the
// group's LDS region cooperatively (each lane takes list entries: reference, size,
// image), then self_sub_x (subsume.h:140-176) runs with its candidate scans split
// across the group, G candidates per step: a clause is strengthened if some
// strengthening clause lies before the first longer one, and deleted by the first
// earlier clause that subsumes it (ballots pick what the sequential loop picks).
// Changes are made in LDS and written through to the pool by the lane that makes
// them; the compacted lists go back at the end of each side. Groups of one wave
// exchange nothing, and a group's lanes communicate through LDS in program order
// (one wave), so no workgroup barrier is used. Variables whose lists or images do
// not fit keep rest[q] = 1 for the next, larger configuration.
__device__ __forceinline__ void wave_order() { __builtin_amdgcn_wave_barrier(); }

template <int G>
__device__ __forceinline__ uint32_t group_incl_scan(uint32_t x)
{
    const uint32_t gl = __lane_id() % G;
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        const uint32_t t = __shfl_up(x, o, G);
        if (gl >= (uint32_t)o) x += t;
    }
    return x;
}

// the staged image at c (LDS) written back to its pool copy
__device__ __forceinline__ void sv_write_image(const SubArgs &A, const uint32_t *c)
{
    uint32_t *g = A.mem + staged_ref(c);
    const int n = c_size(c);
    for (int t = 0; t < 3 + n; t++) g[t] = c[t];
}

// first: the first configuration (every variable of the range); otherwise only the
// variables with rest[q] set
template <int G, uint32_t ENT, uint32_t WORDS>
__global__ void __launch_bounds__(64) k_sub_grp(SubArgs A, const uint32_t *__restrict__ pvs, uint32_t b, uint32_t e,
                                                uint32_t words_cap, uint32_t *__restrict__ rest, int first)
{
    constexpr uint32_t NG = 64 / G;
    __shared__ uint32_t img_all[NG][WORDS];
    __shared__ uint16_t pos_all[NG][2][ENT];
    __shared__ uint32_t seq_all[NG];
    const uint32_t lane = __lane_id(), g = lane / G, gl = lane % G;
    uint32_t *img = img_all[g];
    uint16_t(*pos)[ENT] = pos_all[g];
    uint32_t &s_seq = seq_all[g];
    const uint32_t cap = min(words_cap, WORDS);
    const uint32_t leader = g * G;
    auto gballot = [&](bool x) -> uint64_t {
        const uint64_t m = __ballot(x);
        return G == 64 ? m : (m >> leader) & ((1ull << G) - 1ull);
    };
    const uint64_t n = e - b;
    for (uint64_t it = (uint64_t)blockIdx.x * NG + g; it < n; it += (uint64_t)gridDim.x * NG) {
        const uint32_t q = (uint32_t)(e - 1 - it);   // elected variables come in increasing score order: heaviest first
        if (!first && !rest[q]) continue;
        const uint32_t v = pvs[q], p = 2 * v;
        const OL lp = A.ot[p], ln = A.ot[p + 1];
        if (lp.sz > (uint32_t)A.sub_limit || ln.sz > (uint32_t)A.sub_limit) {   // not processed (subsume.h:145)
            if (gl == 0) rest[q] = 0;
            continue;
        }
        bool fits = lp.sz <= ENT && ln.sz <= ENT;
        const uint32_t ne = lp.sz + ln.sz;
        uint32_t total = 0;
        for (uint32_t k0 = 0; fits && k0 < ne; k0 += G) {
            const uint32_t k = k0 + gl;
            uint64_t r = 0;
            uint32_t w = 0;
            if (k < ne) {
                r = k < lp.sz ? A.occ[lp.off + k] : A.occ[ln.off + (k - lp.sz)];
                w = 5u + min(A.mem[r + 2], cap);   // a clause larger than the cap does not fit anyway
            }
            const uint32_t incl = group_incl_scan<G>(w);
            const uint32_t at = total + incl - w;
            if (k < ne && at + w <= cap) {
                const uint32_t *gsrc = A.mem + r;
                img[at] = (uint32_t)r;
                img[at + 1] = (uint32_t)(r >> 32);
                for (uint32_t t = 0; t + 2 < w; t++) img[at + 2 + t] = gsrc[t];
                if (k < lp.sz) pos[0][k] = (uint16_t)(at + 2);
                else pos[1][k - lp.sz] = (uint16_t)(at + 2);
            }
            total += __shfl(incl, G - 1, G);
            fits = total <= cap;
        }
        if (gl == 0) rest[q] = fits ? 0u : 1u;
        if (!fits) continue;
        if (gl == 0) s_seq = 0;
        wave_order();
        uint32_t len[2] = {lp.sz, ln.sz};
        for (int side = 0; side < 2; side++) {
            const uint32_t me = p ^ (uint32_t)side, other = me ^ 1u;
            uint16_t *mp = pos[side];
            const uint16_t *op = pos[1 - side];
            const uint32_t ms = len[side], os = len[1 - side];
            for (uint32_t i = 0; i < ms; i++) {
                uint32_t *cand = img + mp[i];
                const uint32_t csz = cand[2];
                if (csz > (uint32_t)HSE_MAX_CL_SIZE) break;
                if (c_del(cand)) continue;
                {   // selfsubsume (subsume.h:119-138): any strengthening clause before the first longer one
                    const uint32_t candsig = cand[1];
                    bool found = false;
                    for (uint32_t jb = 0; jb < os; jb += G) {
                        const uint32_t j = jb + gl;
                        bool brk = false, hit = false;
                        if (j < os) {
                            const uint32_t *sg = img + op[j];
                            const uint32_t subsize = sg[2];
                            if (subsize > csz) brk = true;
                            else if (!c_del(sg) && subsize > 1 && sig_selfsub(sg[1], candsig))
                                hit = selfsub_cl(me, other, sg, cand);
                        }
                        const uint64_t mb = gballot(brk), mh = gballot(hit);
                        const uint64_t before = mb ? ((1ull << (__ffsll((long long)mb) - 1)) - 1ull) : ~0ull;
                        if (mh & before) { found = true; break; }
                        if (mb) break;
                    }
                    if (found) {
                        if (gl == 0) {
                            uint32_t seq = s_seq;
                            strengthen(A, cand, me, staged_ref(cand), q, seq);
                            s_seq = seq;
                            c_melt(cand);
                            sv_write_image(A, cand);
                        }
                        wave_order();
                    }
                }
                {   // subsume (subsume.h:101-117): the first earlier clause that subsumes it
                    const uint32_t csz2 = cand[2], candsig = cand[1];
                    const bool molten = c_molten(cand);
                    for (uint32_t jb = 0; jb < i; jb += G) {
                        const uint32_t j = jb + gl;
                        bool hit = false;
                        if (j < i) {
                            const uint32_t *sg = img + mp[j];
                            const uint32_t sz = sg[2];
                            if (!c_del(sg) && !(molten && sz > csz2) && sz > 1 && sig_sub(sg[1], candsig))
                                hit = sub_cl(sg, cand);
                        }
                        const uint64_t mh = gballot(hit);
                        if (mh) {
                            if (gl == (uint32_t)(__ffsll((long long)mh) - 1)) {
                                uint32_t *sg = img + mp[j];
                                if (c_learnt(sg) && c_orig(cand)) {
                                    c_setst(sg, ST_ORIGINAL);
                                    A.mem[staged_ref(sg)] = sg[0];
                                }
                                if (!c_del(cand)) {   // removeClause (sclause.cpp:77-84)
                                    c_markdel(cand);
                                    A.mem[staged_ref(cand)] = cand[0];
                                    uint32_t seq = s_seq;
                                    plog(A.P, q, seq, PE_DEL, staged_ref(cand));
                                    s_seq = seq;
                                }
                            }
                            break;
                        }
                    }
                    wave_order();
                }
            }
            // updateOL (subsume.h:27-37): molten clauses are frozen and leave the list,
            // deleted ones leave it, the rest keep their order. Each step reads its G
            // entries before it writes positions below them.
            uint32_t kept = 0;
            for (uint32_t jb = 0; jb < ms; jb += G) {
                const uint32_t j = jb + gl;
                uint16_t at = 0;
                bool keep = false;
                if (j < ms) {
                    at = mp[j];
                    uint32_t *c = img + at;
                    if (c_molten(c)) {
                        c_freeze(c);
                        A.mem[staged_ref(c)] = c[0];
                    }
                    else keep = !c_del(c);
                }
                const uint64_t mk = gballot(keep);
                const uint32_t rank = (uint32_t)__popcll(mk & ((1ull << gl) - 1ull));
                wave_order();
                if (keep) mp[kept + rank] = at;
                kept += (uint32_t)__popcll(mk);
                wave_order();
            }
            if (kept != ms) {   // nothing left the list: its order is unchanged
                OL *o = A.ot + me;
                uint64_t *d = A.occ + o->off;
                for (uint32_t j = gl; j < kept; j += G) d[j] = staged_ref(img + mp[j]);
                if (gl == 0) o->sz = kept;
            }
            len[side] = kept;
            wave_order();
        }
        if (gl == 0) plog_end(A.P, q, s_seq);
        wave_order();
    }
}

constexpr uint32_t SUB_ARENA_WORDS = 8192;   // LDS words per 64-lane block (32 KB)

// parallel range (no unit can arise): one 64-lane block per 64 elected variables, a lane
// per variable; only the variables k_sub_grp left (rest[q] != 0)
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_sub_staged(SubArgs A, const uint32_t *__restrict__ pvs, uint32_t b, uint32_t e, uint32_t arena_words,
             const uint32_t *__restrict__ rest)
{
    __shared__ uint32_t arena[SUB_ARENA_WORDS];
    const StageSrc S{A.mem, A.ot, A.occ};
    // elected variables come in increasing score order: the heaviest blocks go first
    const uint64_t nblk = ((uint64_t)(e - b) + 63) / 64;
    for (uint64_t it = blockIdx.x; it < nblk; it += gridDim.x) {
        const uint64_t q = b + (nblk - 1 - it) * 64 + threadIdx.x;
        const bool active = q < e;
        const uint32_t v = active ? pvs[q] : 0u;
        const bool take = active && rest[q] && A.ot[2 * v].sz <= (uint32_t)A.sub_limit &&
                          A.ot[2 * v + 1].sz <= (uint32_t)A.sub_limit;
        stage_rounds(S, arena, arena_words, take, v, STAGE_TAB2, [&](uint32_t *tab) {
            uint32_t seq = 0;
            if (tab) self_sub_x_staged(A, 2 * v, tab, arena, (uint32_t)q, seq);
            else self_sub_x(A, 2 * v, (uint32_t)q, seq);
            plog_end(A.P, (uint32_t)q, seq);
        });
    }
}

// One unit-capable variable by one wave: self_sub_x (subsume.h:140-176) with its list
// scans split across the 64 lanes. Per list entry of the side's two lists, LDS holds
// the size with a deleted flag, the signature and, for a binary, its other literal;
// the current clause's literals are in LDS too. A binary (x|b) subsumes the clause iff
// b is in it, and a binary (~x|b) strengthens it iff b is in it, so binaries are
// checked in LDS; longer clauses are checked from the pool. Decisions are those of the
// sequential loop: a clause is strengthened if some strengthening clause lies before
// the size break, it is deleted by the first subsuming earlier clause (whose learnt
// status it promotes), and lane 0 applies every change before the next clause. A unit
// deletes the clauses of its literal: the wave deletes those on its own lists at once
// (the only ones its later steps can see), k_sub_oblivion the rest after the variable.
constexpr uint32_t SW_CAP = 3000;   // lists up to this length (the default sub_limit)
constexpr uint32_t SW_CL = 128;     // clause literals held in LDS
constexpr uint16_t SW_DEL = 0x8000u, SW_BIG = 0x7FFFu;
constexpr uint32_t SW_NONE = 0xFFFFFFFFu;

__device__ __forceinline__ bool sw_has(const uint32_t *lits, uint32_t n, uint32_t x)
{   // binary search of a sorted literal array
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (lits[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && lits[lo] == x;
}

// per-entry arrays of the two lists of one side: in LDS for lists of up to SW_CAP
// entries, in a device buffer of 2 * 10 bytes per entry for longer ones
struct SwLists {
    uint16_t *msz, *osz;
    uint32_t *msig, *osig, *mbin, *obin;
};

__device__ __forceinline__ void sub_wave_body(const SubArgs &A, const uint32_t *__restrict__ pvs, uint32_t q,
                                              const SwLists &L, uint32_t *clit, uint32_t &s_unit, uint32_t &s_seq);

__global__ void __launch_bounds__(64) k_sub_wave(SubArgs A, const uint32_t *__restrict__ pvs, uint32_t q)
{
    __shared__ uint16_t msz[SW_CAP], osz[SW_CAP];
    __shared__ uint32_t msig[SW_CAP], osig[SW_CAP], mbin[SW_CAP], obin[SW_CAP];
    __shared__ uint32_t clit[SW_CL];
    __shared__ uint32_t s_unit, s_seq;
    sub_wave_body(A, pvs, q, SwLists{msz, osz, msig, osig, mbin, obin}, clit, s_unit, s_seq);
}

// the same for lists longer than SW_CAP (sub_limit above its default): the per-entry
// arrays live in `big` (cap entries per array)
__global__ void __launch_bounds__(64) k_sub_wave_big(SubArgs A, const uint32_t *__restrict__ pvs, uint32_t q,
                                                     uint32_t *big, uint32_t cap)
{
    __shared__ uint32_t clit[SW_CL];
    __shared__ uint32_t s_unit, s_seq;
    uint16_t *sz16 = (uint16_t *)big;
    uint32_t *w = big + cap;   // cap u16 pairs = cap words
    sub_wave_body(A, pvs, q, SwLists{sz16, sz16 + cap, w, w + cap, w + 2ull * cap, w + 3ull * cap}, clit, s_unit, s_seq);
}

__device__ __forceinline__ void sub_wave_body(const SubArgs &A, const uint32_t *__restrict__ pvs, uint32_t q,
                                              const SwLists &L, uint32_t *clit, uint32_t &s_unit, uint32_t &s_seq)
{
    uint16_t *msz = L.msz, *osz = L.osz;
    uint32_t *msig = L.msig, *osig = L.osig, *mbin = L.mbin, *obin = L.obin;
    const uint32_t lane = threadIdx.x;
    // units this variable assigns are those past this point of the trail (k_sub_oblivion)
    if (lane == 0 && A.defer_oblivion) A.cnt[CNT_SUBFROM] = A.cnt[CNT_NTRAIL];
    const uint32_t v = pvs[q], p = 2 * v;
    if (A.ot[p].sz > (uint32_t)A.sub_limit || A.ot[p + 1].sz > (uint32_t)A.sub_limit) return;
    if (lane == 0) s_seq = 0;   // proof events: lane 0 strengthens, one lane deletes, barriers between
    for (int side = 0; side < 2; side++) {
        const uint32_t me = side ? p + 1 : p, other = me ^ 1u;
        const OL om = A.ot[me], oo = A.ot[other];
        const uint64_t *d = A.occ + om.off, *od = A.occ + oo.off;
        auto entry = [](const uint32_t *c, uint32_t lit, uint16_t &sz, uint32_t &sig, uint32_t &bin) {
            const uint32_t n = c[2];
            sz = (uint16_t)((n < SW_BIG ? n : SW_BIG) | (c_del(c) ? SW_DEL : 0u));
            sig = c[1];
            bin = (n == 2 && !c_molten(c)) ? (c[3] ^ c[4] ^ lit) : SW_NONE;   // a strengthened clause lost lit
        };
        auto stage = [&]() {
            for (uint32_t j = lane; j < om.sz; j += 64) entry(A.mem + d[j], me, msz[j], msig[j], mbin[j]);
            for (uint32_t j = lane; j < oo.sz; j += 64) entry(A.mem + od[j], other, osz[j], osig[j], obin[j]);
        };
        stage();
        if (lane == 0) s_unit = 0;
        __syncthreads();
        for (uint32_t i = 0; i < om.sz; i++) {
            const uint32_t mi = msz[i];
            if ((mi & SW_BIG) > (uint32_t)HSE_MAX_CL_SIZE) break;
            if (mi & SW_DEL) continue;
            uint32_t *cand = A.mem + d[i];
            uint32_t csz = mi & SW_BIG;
            for (uint32_t k = lane; k < csz && k < SW_CL; k += 64) clit[k] = cand[3 + k];
            __syncthreads();
            // membership of a literal in the current clause
            auto in_cand = [&](uint32_t b, uint32_t n) {
                return n <= SW_CL ? sw_has(clit, n, b) : sw_has(cand + 3, n, b);
            };
            {   // selfsubsume (subsume.h:119-138): any hit before the first longer clause
                const uint32_t candsig = msig[i];
                const bool me_in = in_cand(me, csz);
                bool found = false;
                // four chunks of 64 entries per step of the loop, judged in order
                for (uint32_t b = 0; b < oo.sz && !found; b += 256) {
                    bool brk[4], hit[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const uint32_t j = b + 64u * u + lane;
                        brk[u] = hit[u] = false;
                        if (j < oo.sz) {
                            const uint32_t sj = osz[j], subsize = sj & SW_BIG;
                            if (subsize > csz) brk[u] = true;
                            else if (!(sj & SW_DEL) && subsize > 1 && sig_selfsub(osig[j], candsig))
                                hit[u] = obin[j] != SW_NONE ? (me_in && in_cand(obin[j], csz))
                                                            : selfsub_cl(me, other, A.mem + od[j], cand);
                        }
                    }
                    bool stop = false;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        if (stop) continue;
                        const uint64_t mb = __ballot(brk[u]), mh = __ballot(hit[u]);
                        const uint64_t before = mb ? ((1ull << (__ffsll((long long)mb) - 1)) - 1ull) : ~0ull;
                        if (mh & before) { found = true; stop = true; }
                        else if (mb) stop = true;
                    }
                    if (stop) break;
                }
                if (found) {
                    if (lane == 0) {
                        uint32_t seq = s_seq;
                        s_unit = strengthen(A, cand, me, d[i], q, seq);
                        s_seq = seq;
                        c_melt(cand);
                        msz[i] = (uint16_t)c_size(cand);
                        msig[i] = cand[1];
                        mbin[i] = SW_NONE;   // it no longer holds me
                    }
                    __syncthreads();
                    csz = msz[i] & SW_BIG;
                    for (uint32_t k = lane; k < csz && k < SW_CL; k += 64) clit[k] = cand[3 + k];
                    const uint32_t ul = s_unit;
                    if (ul) {
                        // toblivion(ot[unit]) as far as this variable sees it: its clauses that
                        // hold the unit (the unit's variable is never elected, so these are
                        // exactly its clauses on that list); k_sub_oblivion deletes the rest
                        for (uint32_t j = lane; j < om.sz; j += 64) {
                            uint32_t *c = A.mem + d[j];
                            if (!c_del(c) && sw_has(c + 3, c_size(c), ul)) { c_markdel(c); msz[j] |= SW_DEL; }
                        }
                        for (uint32_t j = lane; j < oo.sz; j += 64) {
                            uint32_t *c = A.mem + od[j];
                            if (!c_del(c) && sw_has(c + 3, c_size(c), ul)) { c_markdel(c); osz[j] |= SW_DEL; }
                        }
                    }
                    __syncthreads();
                    if (lane == 0) s_unit = 0;
                }
            }
            {   // subsume (subsume.h:101-117): the first earlier clause that subsumes it
                const uint32_t ci = msz[i];
                const uint32_t candsig = msig[i];
                const bool molten = c_molten(cand);
                const bool me_in = in_cand(me, csz);
                for (uint32_t b = 0; b < i; b += 256) {
                    bool hit[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const uint32_t j = b + 64u * u + lane;
                        hit[u] = false;
                        if (j < i) {
                            const uint32_t sj = msz[j], sz = sj & SW_BIG;
                            if (!(sj & SW_DEL) && !(molten && sz > csz) && sz > 1 && sig_sub(msig[j], candsig))
                                hit[u] = mbin[j] != SW_NONE ? (me_in && in_cand(mbin[j], csz))
                                                            : sub_cl(A.mem + d[j], cand);
                        }
                    }
                    int first = -1;
                    uint64_t fm = 0;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const uint64_t mh = __ballot(hit[u]);
                        if (first < 0 && mh) { first = u; fm = mh; }
                    }
                    if (first >= 0) {
                        if (lane == (uint32_t)(__ffsll((long long)fm) - 1)) {
                            uint32_t *sg = A.mem + d[b + 64u * (uint32_t)first + lane];
                            if (c_learnt(sg) && c_orig(cand)) c_setst(sg, ST_ORIGINAL);
                            if (!c_del(cand)) {   // removeClause (sclause.cpp:77-84)
                                c_markdel(cand);
                                uint32_t seq = s_seq;
                                plog(A.P, q, seq, PE_DEL, d[i]);
                                s_seq = seq;
                            }
                            msz[i] = (uint16_t)(ci | SW_DEL);
                        }
                        break;
                    }
                }
                __syncthreads();
            }
        }
        if (lane == 0) update_ol(A, me);
        __syncthreads();
    }
    if (lane == 0) plog_end(A.P, q, s_seq);
}

// the rest of toblivion(ot[unit]) for the units assigned since CNT_SUBFROM: every
// clause on the unit's list is deleted, then the lists are emptied
__global__ void k_sub_oblivion(SubArgs A)
{
    const uint64_t from = A.cnt[CNT_SUBFROM], to = A.cnt[CNT_NTRAIL];
    for (uint64_t t = from; t < to; t++) {
        const OL o = A.ot[A.trail[t]];
        for (uint64_t e = gtid(); e < o.sz; e += gstride()) c_markdel(A.mem + A.occ[o.off + e]);
    }
}

__global__ void k_sub_oblivion_sz(SubArgs A)
{
    const uint64_t from = A.cnt[CNT_SUBFROM], to = A.cnt[CNT_NTRAIL];
    for (uint64_t t = from + threadIdx.x; t < to; t += blockDim.x) A.ot[A.trail[t]].sz = 0;
}

// next live binary at or after position i of a list (its other literal in *a); the
// binaries of an elected variable's sorted list form its prefix
__device__ __forceinline__ uint32_t next_binary(const uint64_t *list, uint32_t sz, uint32_t i, const uint32_t *mem,
                                                uint32_t lit, uint32_t *a)
{
    for (; i < sz; i++) {
        const uint32_t *c = mem + list[i];
        if (c_size(c) > 2) return sz;
        if (!c_del(c) && c_size(c) == 2) { *a = c[3] ^ c[4] ^ lit; return i; }
    }
    return sz;
}

// unit-capable: a live binary (x|a) on the positive list and a live binary (~x|a) on the negative list
__global__ void k_sub_unitcap(const uint32_t *__restrict__ pvs, uint32_t npvs, const OL *__restrict__ ot,
                              const uint64_t *__restrict__ occ, const uint32_t *__restrict__ mem, int sub_limit,
                              uint32_t *__restrict__ flag)
{
    for (uint64_t q = gtid(); q < npvs; q += gstride()) {
        const uint32_t v = pvs[q], p = 2 * v, n = p + 1;
        const OL op = ot[p], on = ot[n];
        uint32_t f = 0;
        if (op.sz <= (uint32_t)sub_limit && on.sz <= (uint32_t)sub_limit) {
            // Lists sorted by (size, first, last literal) order the binaries (x|a) by a,
            // so a merge of the two binary prefixes finds a common a. If a prefix is out
            // of that order, every pair is compared instead.
            const uint64_t *lp = occ + op.off, *ln = occ + on.off;
            uint32_t a = 0, b = 0, pa = 0, pb = 0;
            bool ordered = true;
            uint32_t i = next_binary(lp, op.sz, 0, mem, p, &a), j = next_binary(ln, on.sz, 0, mem, n, &b);
            bool firsta = true, firstb = true;
            while (i < op.sz && j < on.sz) {
                if ((!firsta && a < pa) || (!firstb && b < pb)) { ordered = false; break; }
                pa = a; pb = b; firsta = firstb = false;
                if (a == b) { f = 1; break; }
                if (a < b) i = next_binary(lp, op.sz, i + 1, mem, p, &a);
                else j = next_binary(ln, on.sz, j + 1, mem, n, &b);
            }
            if (!ordered) f = 0;
            for (uint32_t i = 0; !ordered && i < op.sz && !f; i++) {
                const uint32_t *c = mem + occ[op.off + i];
                if (c_del(c) || c_size(c) != 2) continue;
                const uint32_t a = c[3] ^ c[4] ^ p;
                for (uint32_t j = 0; j < on.sz; j++) {
                    const uint32_t *d = mem + occ[on.off + j];
                    if (c_del(d) || c_size(d) != 2) continue;
                    if ((d[3] ^ d[4] ^ n) == a) { f = 1; break; }
                }
            }
        }
        flag[q] = f;
    }
}

__global__ void k_iota(uint32_t *out, uint32_t n)
{
    for (uint64_t i = gtid(); i < n; i += gstride()) out[i] = (uint32_t)i;
}

void sub(Ctx &c)
{
    if (!c.npvs || !(c.opts.sub_en || c.opts.ve_en)) return;
    const uint32_t np = c.npvs;
    SubArgs A{c.mem.p, c.ot.p, c.occ.p, c.value.p, c.vstate.p, c.trail.p, (unsigned long long *)c.cnt.p,
              c.opts.sub_limit, c.opts.lbd_tier1, 0, 0, proof_begin(c, np, 2 * c.nrefs + 2ull * np), c.otrm.p, c.otrm_cap};
    ensure(c, c.tmpu32a, np);
    ensure(c, c.tmpu32b, np);
    ensure(c, c.scores, np);
    k_sub_unitcap<<<grid_for(np, 64, 16384), 64, 0, c.stream>>>(c.pvs.p, np, c.ot.p, c.occ.p, c.mem.p, c.opts.sub_limit,
                                                                 c.tmpu32a.p);
    k_iota<<<grid_for(np, 256, 4096), 256, 0, c.stream>>>(c.scores.p, np);
    const uint64_t nu = compact_u32(c, c.scores.p, c.tmpu32a.p, np, c.tmpu32b.p);
    std::vector<uint32_t> ucap(nu);
    if (c.trace) fprintf(stderr, "[pfg] SUB: %u elected, %llu unit-capable\n", np, (unsigned long long)nu);
    if (nu) {
        PFG_HIP(hipMemcpyAsync(ucap.data(), c.tmpu32b.p, nu * 4, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    uint32_t b = 0;
    ensure(c, c.tmpu32c, np);
    uint32_t *rest = c.tmpu32c.p;
    auto run_range = [&](uint32_t e) {
        if (e > b) {
            const uint32_t cap = c.stage_words;
            {   // G lanes per variable (tuning) with 4 G list entries and 32 G staged words per group
                KTimer kt(c, K_SUB, pv_bytes(c, b, e, true));
                const uint32_t G = c.tune.sub_group;
                const uint32_t grid = (uint32_t)std::min<uint64_t>((e - b + 64 / G - 1) / (64 / G), 65536);
                if (G == 8) k_sub_grp<8, 32, 256><<<grid, 64, 0, c.stream>>>(A, c.pvs.p, b, e, cap, rest, 1);
                else if (G == 32) k_sub_grp<32, 128, 1024><<<grid, 64, 0, c.stream>>>(A, c.pvs.p, b, e, cap, rest, 1);
                else if (G == 64) k_sub_grp<64, 256, 2048><<<grid, 64, 0, c.stream>>>(A, c.pvs.p, b, e, cap, rest, 1);
                else k_sub_grp<16, 64, 512><<<grid, 64, 0, c.stream>>>(A, c.pvs.p, b, e, cap, rest, 1);
            }
            KTimer kt(c, K_SUB_REST, 0.0);
            // a wave per variable for lists of up to 256 entries and 2048 words, then a lane
            // per variable (larger regions, or from the pool) for what is left
            k_sub_grp<64, 256, 2048><<<(uint32_t)std::min<uint64_t>(e - b, 16384), 64, 0, c.stream>>>(A, c.pvs.p, b, e,
                                                                                                       cap, rest, 0);
            k_sub_staged<<<grid_for(e - b, 64, 16384), 64, 0, c.stream>>>(A, c.pvs.p, b, e,
                                                                           std::min(SUB_ARENA_WORDS, cap), rest);
        }
    };
    for (uint64_t k = 0; k < nu; k++) {
        const uint32_t u = ucap[k];
        A.allow_units = 0;
        run_range(u);
        A.allow_units = 1;
        if (c.opts.sub_limit <= (int)SW_CAP) {
            A.defer_oblivion = 1;
            k_sub_wave<<<1, 64, 0, c.stream>>>(A, c.pvs.p, u);
            k_sub_oblivion<<<256, 256, 0, c.stream>>>(A);
            k_sub_oblivion_sz<<<1, 64, 0, c.stream>>>(A);
            A.defer_oblivion = 0;
        }
        else {   // lists longer than the LDS arrays: the same wave with its arrays in HBM
            const uint32_t cap = (uint32_t)c.opts.sub_limit + 1;
            ensure(c, c.tmpu32d, 5ull * cap);
            A.defer_oblivion = 1;
            k_sub_wave_big<<<1, 64, 0, c.stream>>>(A, c.pvs.p, u, c.tmpu32d.p, cap);
            k_sub_oblivion<<<256, 256, 0, c.stream>>>(A);
            k_sub_oblivion_sz<<<1, 64, 0, c.stream>>>(A);
            A.defer_oblivion = 0;
            c.path[PATH_SUB_WAVE_BIG]++;
        }
        c.path[PATH_SUB_WAVE]++;
        b = u + 1;
    }
    A.allow_units = 0;
    run_range(np);
    read_counters(c, 10);
    c.ntrail = (uint32_t)c.hcnt[CNT_NTRAIL];
    if (c.hcnt[CNT_FLAGS] & FLAG_INTERNAL) throw Error(PFG_EINTERNAL, "SUB produced a unit outside a unit-capable variable");
    if (c.hcnt[CNT_FLAGS] & FLAG_UNSAT) c.unsat = true;
    proof_end(c, np, false);
}

}  // namespace pfg
