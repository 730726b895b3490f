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
#include "stage.hpp"
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

__device__ __forceinline__ void strengthen(const SubArgs &A, uint32_t *c, uint32_t me)
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
    if (n == 1) {
        if (!A.allow_units) { atomicOr(&A.cnt[CNT_FLAGS], (unsigned long long)FLAG_INTERNAL); return; }
        const uint32_t unit = lits[0];
        const int r = try_assign(A.value, unit);
        if (r == 1) {
            A.vstate[unit >> 1] = FROZEN_M;
            const unsigned long long pos = atomicAdd(&A.cnt[CNT_NTRAIL], 1ull);
            A.trail[pos] = unit;
            toblivion_list(A, unit);
        }
        else if (r < 0) report_conflict(A.cnt, unit >> 1);
    }
    else if (c_learnt(c)) {   // bumpShrunken (elimination.cpp:27-39)
        const uint32_t old = c[0] >> 6;
        if ((int)old > A.lbd_tier1) {
            const uint32_t nl = min((uint32_t)(c_size(c) - 1), old);
            if (nl < old) c[0] = (c[0] & 0xFu) | (1u << 4) | (nl << 6);
        }
    }
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

__device__ __forceinline__ void subsume_l(const SubArgs &A, const uint64_t *list, uint32_t end, uint32_t *cand)
{   // subsume (subsume.h:101-117)
    const int candsz = c_size(cand);
    for (uint32_t j = 0; j < end; j++) {
        uint32_t *sg = A.mem + list[j];
        if (c_del(sg)) continue;
        if (c_molten(cand) && c_size(sg) > candsz) continue;
        if (c_size(sg) > 1 && sig_sub(sg[1], cand[1]) && sub_cl(sg, cand)) {
            if (c_learnt(sg) && c_orig(cand)) c_setst(sg, ST_ORIGINAL);
            if (!c_del(cand)) c_markdel(cand);
            return;
        }
    }
}

__device__ __forceinline__ void selfsubsume_l(const SubArgs &A, uint32_t x, uint32_t fx, uint32_t listlit, uint32_t *cand)
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
            strengthen(A, cand, x);
            c_melt(cand);
            return;
        }
    }
}

__device__ __forceinline__ void self_sub_x(const SubArgs &A, uint32_t p)
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
            selfsubsume_l(A, me, other, other, cand);
            subsume_l(A, d, i, cand);
        }
        update_ol(A, me);
    }
}

// The same pass over a staged variable (stage.hpp, with the original-order table):
// lists are table ranges, modified clause copies are marked dirty, and at the end
// the dirty clause images and the compacted lists are written back to the pool.
__device__ __forceinline__ void subsume_staged(uint32_t *arena, const uint32_t *list, uint32_t end, uint32_t *cand)
{   // subsume (subsume.h:101-117)
    const int candsz = c_size(cand);
    for (uint32_t j = 0; j < end; j++) {
        uint32_t *sg = arena + list[j];
        if (c_del(sg)) continue;
        if (c_molten(cand) && c_size(sg) > candsz) continue;
        if (c_size(sg) > 1 && sig_sub(sg[1], cand[1]) && sub_cl(sg, cand)) {
            if (c_learnt(sg) && c_orig(cand)) { c_setst(sg, ST_ORIGINAL); staged_dirty(sg); }
            if (!c_del(cand)) { c_markdel(cand); staged_dirty(cand); }
            return;
        }
    }
}

__device__ __forceinline__ void selfsubsume_staged(const SubArgs &A, uint32_t x, uint32_t fx, uint32_t *arena,
                                                   const uint32_t *list, uint32_t n, uint32_t *cand)
{   // selfsubsume (subsume.h:119-138)
    const int candsz = c_size(cand);
    const uint32_t candsig = cand[1];
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t *sg = arena + list[j];
        const int subsize = c_size(sg);
        if (subsize > candsz) break;
        if (c_del(sg)) continue;
        if (subsize > 1 && sig_selfsub(sg[1], candsig) && selfsub_cl(x, fx, sg, cand)) {
            strengthen(A, cand, x);
            c_melt(cand);
            staged_dirty(cand);
            return;
        }
    }
}

__device__ __forceinline__ void self_sub_x_staged(const SubArgs &A, uint32_t p, uint32_t *tab, uint32_t *arena)
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
            selfsubsume_staged(A, me, other, arena, tab + base[1 - side], len[1 - side], cand);
            subsume_staged(arena, mt, i, cand);
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

constexpr uint32_t SUB_ARENA_WORDS = 8192;   // LDS words per 64-lane block (32 KB)

// parallel range (no unit can arise): one 64-lane block per 64 elected variables
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_sub_staged(SubArgs A, const uint32_t *__restrict__ pvs, uint32_t b, uint32_t e, uint32_t arena_words)
{
    __shared__ uint32_t arena[SUB_ARENA_WORDS];
    const StageSrc S{A.mem, A.ot, A.occ};
    // elected variables come in increasing score order: the heaviest blocks go first
    const uint64_t nblk = ((uint64_t)(e - b) + 63) / 64;
    for (uint64_t it = blockIdx.x; it < nblk; it += gridDim.x) {
        const uint64_t q = b + (nblk - 1 - it) * 64 + threadIdx.x;
        const bool active = q < e;
        const uint32_t v = active ? pvs[q] : 0u;
        const bool take = active && A.ot[2 * v].sz <= (uint32_t)A.sub_limit && A.ot[2 * v + 1].sz <= (uint32_t)A.sub_limit;
        stage_rounds(S, arena, arena_words, take, v, STAGE_TAB2, [&](uint32_t *tab) {
            if (tab) self_sub_x_staged(A, 2 * v, tab, arena);
            else self_sub_x(A, 2 * v);
        });
    }
}

__global__ void k_sub_range(SubArgs A, const uint32_t *__restrict__ pvs, uint32_t b, uint32_t e)
{
    for (uint64_t q = gtid() + b; q < e; q += gstride()) {
        const uint32_t v = pvs[q];
        const uint32_t p = 2 * v;
        if (A.ot[p].sz <= (uint32_t)A.sub_limit && A.ot[p + 1].sz <= (uint32_t)A.sub_limit) self_sub_x(A, p);
    }
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
            for (uint32_t i = 0; i < op.sz && !f; i++) {
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
    SubArgs A{c.mem.p, c.ot.p, c.occ.p, c.value.p, c.vstate.p, c.trail.p, (unsigned long long *)c.cnt.p,
              c.opts.sub_limit, c.opts.lbd_tier1, 0};
    const uint32_t np = c.npvs;
    ensure(c, c.tmpu32a, np);
    ensure(c, c.tmpu32b, np);
    ensure(c, c.scores, np);
    k_sub_unitcap<<<grid_for(np, 64, 16384), 64, 0, c.stream>>>(c.pvs.p, np, c.ot.p, c.occ.p, c.mem.p, c.opts.sub_limit,
                                                                 c.tmpu32a.p);
    k_iota<<<grid_for(np, 256, 4096), 256, 0, c.stream>>>(c.scores.p, np);
    const uint64_t nu = compact_u32(c, c.scores.p, c.tmpu32a.p, np, c.tmpu32b.p);
    std::vector<uint32_t> ucap(nu);
    if (nu) {
        PFG_HIP(hipMemcpyAsync(ucap.data(), c.tmpu32b.p, nu * 4, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    uint32_t b = 0;
    auto run_range = [&](uint32_t e) {
        if (e > b) {
            KTimer kt(c, K_SUB, pv_bytes(c, b, e, true));
            k_sub_staged<<<grid_for(e - b, 64, 16384), 64, 0, c.stream>>>(A, c.pvs.p, b, e,
                                                                           std::min(SUB_ARENA_WORDS, c.stage_words));
        }
    };
    for (uint64_t k = 0; k < nu; k++) {
        const uint32_t u = ucap[k];
        A.allow_units = 0;
        run_range(u);
        A.allow_units = 1;
        k_sub_range<<<1, 1, 0, c.stream>>>(A, c.pvs.p, u, u + 1);
        b = u + 1;
    }
    A.allow_units = 0;
    run_range(np);
    read_counters(c, 10);
    c.ntrail = (uint32_t)c.hcnt[CNT_NTRAIL];
    if (c.hcnt[CNT_FLAGS] & FLAG_INTERNAL) throw Error(PFG_EINTERNAL, "SUB produced a unit outside a unit-capable variable");
    if (c.hcnt[CNT_FLAGS] & FLAG_UNSAT) c.unsat = true;
}

}  // namespace pfg
