// Bounded variable elimination with gate substitution on the elected variables
// (ParaFROST src/cpu/bve.cpp:28-220, simplify.h, and.h, equivalence.h,
// ifthenelse.h, xor.h).
//
// Three phases, as the reference GPU engine does it (kernels.cu:1085-1120), with
// the reference CPU decisions:
//   1. detect (thread per variable): pure literal, equivalence, AND/OR, ITE, XOR
//      gates, else plain resolution under the clause bound; record the kind and
//      the exact number of resolvents, resolvent words, unit resolvents and
//      model-stack words. Gate search marks clauses molten as in the reference.
//   2. exclusive scans over the elected order place every variable's resolvents,
//      clause references, units and model records (deterministic, equal to the
//      sequential append order of the reference).
//   3. apply (thread per variable): write model records and resolvents, rewrite
//      substituted clauses, delete antecedents, mark the variable eliminated.
#include "device.hpp"
#include "passes.hpp"
#include "prims.hpp"
#include "proof.hpp"
#include "stage.hpp"

namespace pfg {

// T_FUNPEND: no gate of the other kinds; the function-table gate (k_ve_fun) decides.
// T_RESPEND: no gate; plain resolution is counted by a lane group (k_ve_res).
enum { T_NONE = 0, T_PURE = 1, T_BN = 2, T_SUBST = 3, T_RES = 4, T_CORE = 5, T_FUNPEND = 6, T_RESPEND = 7 };

// resolvent pairs from which plain resolution is counted by a lane group instead of the detecting lane
constexpr uint64_t RES_WAVE_MIN = 128;

struct VeArgs {
    uint32_t *mem;
    OL *ot;
    uint64_t *occ;
    int rlimit;
    int lbound;
    int xor_max_arity;
    // per-lane LDS stage of the variable's own clauses (detect only; null = global)
    uint32_t v, ps;
    const uint32_t *tab;   // LDS word offsets of the staged clauses, positive list then negative
    uint32_t *arena;       // LDS arena; a staged clause is [ref_lo, ref_hi, word0, sig, size, lits..]
};

__device__ __forceinline__ uint32_t *cl_at(const VeArgs &A, uint32_t lit, uint32_t i)
{
    if (A.tab && (lit >> 1) == A.v) return A.arena + A.tab[(lit & 1u) ? A.ps + i : i];
    return A.mem + A.occ[A.ot[lit].off + i];
}

constexpr uint32_t VE_ARENA_WORDS = 8192;   // LDS words per 64-lane block for staged clauses (32 KB) at most

// Header word writes keep the staged copy and the pool copy equal: a staged clause
// writes through to the pool; a pool clause that is also staged (reached through
// another literal's list) updates its staged copy.
__device__ __forceinline__ void set_w0(const VeArgs &A, uint32_t *c, uint32_t w)
{
    if (c[0] == w) return;
    c[0] = w;
    if (!A.tab) return;
    if (c >= A.arena && c < A.arena + VE_ARENA_WORDS) {
        A.mem[staged_ref(c)] = w;
        return;
    }
    const uint64_t ref = (uint64_t)(c - A.mem);
    const uint32_t n = A.ps + A.ot[2 * A.v + 1].sz;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t *l = A.arena + A.tab[k];
        if (l[-2] == (uint32_t)ref && l[-1] == (uint32_t)(ref >> 32)) { l[0] = w; return; }
    }
}
__device__ __forceinline__ void melt(const VeArgs &A, uint32_t *c) { set_w0(A, c, c[0] | 4u); }
__device__ __forceinline__ void freeze(const VeArgs &A, uint32_t *c) { set_w0(A, c, c[0] & ~4u); }

__device__ __forceinline__ int count_orgs(const VeArgs &A, uint32_t lit)
{
    const uint32_t sz = A.ot[lit].sz;
    int n = 0;
    for (uint32_t i = 0; i < sz; i++) n += c_orig(cl_at(A, lit, i));
    return n;
}

// which resolvent pairs are made: every pair (plain resolution), pairs of one gate and
// one non-gate clause (substitution), pairs not both outside a function table's core
enum { R_ALL = 0, R_SUBST = 1, R_CORE = 2 };

__device__ __forceinline__ bool skip_pair(int rule, bool ci_molten, bool cj_molten)
{
    return rule == R_SUBST ? ci_molten == cj_molten : rule == R_CORE ? ci_molten && cj_molten : false;
}

// countSubstituted / countResolvents (src/cpu/simplify.h:227-287), and for R_CORE the
// GPU reference's countCoreSubstituted (src/gpu/function.cuh:191-259); on success the
// totals needed for placement are returned through words/units.
__device__ __forceinline__ bool count_subst(const VeArgs &A, uint32_t x, int clsbefore, uint32_t me, uint32_t other, int rule,
                            int &nAdded, uint64_t &words, uint32_t &units)
{
    int nAddedLits = 0;
    uint32_t nunit = 0;
    const uint32_t ms = A.ot[me].sz, os = A.ot[other].sz;
    for (uint32_t i = 0; i < ms; i++) {
        const uint32_t *ci = cl_at(A, me, i);
        if (!c_orig(ci)) continue;
        const bool a = c_molten(ci);
        for (uint32_t j = 0; j < os; j++) {
            const uint32_t *cj = cl_at(A, other, j);
            if (!c_orig(cj)) continue;
            if (skip_pair(rule, a, c_molten(cj))) continue;
            const int rsize = merge_len(x, ci, cj);
            if (rsize > 1) {
                if (++nAdded > clsbefore || (A.rlimit && rsize > A.rlimit)) return true;
                nAddedLits += rsize;
            }
            else if (rsize == 1) nunit++;
        }
    }
    if (A.lbound) {
        int before = 0;
        for (uint32_t i = 0; i < ms; i++) { const uint32_t *c = cl_at(A, me, i); if (c_orig(c)) before += c_size(c); }
        for (uint32_t j = 0; j < os; j++) { const uint32_t *c = cl_at(A, other, j); if (c_orig(c)) before += c_size(c); }
        if (nAddedLits > before) return true;
    }
    words = 3ull * (uint64_t)nAdded + (uint64_t)nAddedLits;
    units = nAdded ? nunit : 0;
    return false;
}

__device__ __forceinline__ void freeze_binaries(const VeArgs &A, uint32_t lit)
{
    const uint32_t sz = A.ot[lit].sz;
    for (uint32_t i = 0; i < sz; i++) {
        uint32_t *c = cl_at(A, lit, i);
        if (c_orig(c) && c_size(c) == 2) freeze(A, c);
    }
}

// find_sfanin + find_BN_gate (equivalence.h:88-130)
__device__ __forceinline__ uint32_t find_bn_gate(const VeArgs &A, uint32_t p)
{
    const uint32_t n = p ^ 1u;
    const uint32_t ps = A.ot[p].sz;
    uint32_t imp = 0;
    int nImps = 0;
    uint32_t first = 0;
    bool many = false;
    for (uint32_t i = 0; i < ps; i++) {
        uint32_t *c = cl_at(A, p, i);
        if (c_orig(c) && c_size(c) == 2) {
            imp = (c[3] ^ c[4] ^ p) ^ 1u;
            melt(A, c);
            nImps++;
        }
        if (nImps > 1) { many = true; break; }
    }
    first = many ? 0 : imp;
    if (first) {
        uint32_t second = n, def = first;
        if (second < first) { first = second; second = def; }
        const uint32_t ns = A.ot[n].sz;
        for (uint32_t i = 0; i < ns; i++) {
            uint32_t *c = cl_at(A, n, i);
            if (c_orig(c) && c_size(c) == 2 && c[3] == first && c[4] == second) {
                melt(A, c);
                return def;
            }
        }
    }
    freeze_binaries(A, p);
    return 0;
}

__device__ __forceinline__ bool has_lit(const uint32_t *c, uint32_t lit)
{
    const int sz = c_size(c);
    for (int i = 0; i < sz; i++) if (c[3 + i] == lit) return true;
    return false;
}

// find_AO_gate (and.h:25-85) without materialising the fan-in vector: the fan-out
// clause must have |fanins|+1 distinct literals, all of them fan-ins or ~dx
__device__ __forceinline__ bool find_ao_gate(const VeArgs &A, uint32_t dx, int nOrgCls, int &nAdded, uint64_t &words, uint32_t &units)
{
    const uint32_t x = dx >> 1, f_dx = dx ^ 1u;
    const uint32_t is = A.ot[dx].sz;
    uint32_t sig = 0;
    int nbin = 0;
    for (uint32_t i = 0; i < is; i++) {
        uint32_t *c = cl_at(A, dx, i);
        if (c_orig(c) && c_size(c) == 2) {
            const uint32_t imp = (c[3] ^ c[4] ^ dx) ^ 1u;
            sig |= maphash(imp);
            nbin++;
            melt(A, c);
        }
    }
    if (nbin > 1) {
        sig |= maphash(f_dx);
        const int want = nbin + 1;
        const uint32_t os = A.ot[f_dx].sz;
        for (uint32_t i = 0; i < os; i++) {
            uint32_t *c = cl_at(A, f_dx, i);
            if (!(c_orig(c) && c_size(c) == want && sig_sub(c[1], sig))) continue;
            bool eq = true;
            for (int k = 0; k < want && eq; k++) {
                const uint32_t l = c[3 + k];
                if (l == f_dx) continue;
                bool found = false;
                for (uint32_t t = 0; t < is && !found; t++) {
                    const uint32_t *b = cl_at(A, dx, t);
                    if (c_orig(b) && c_size(b) == 2 && ((b[3] ^ b[4] ^ dx) ^ 1u) == l) found = true;
                }
                eq = found;
            }
            if (!eq) continue;
            melt(A, c);
            nAdded = 0;
            if (count_subst(A, x, nOrgCls, dx, f_dx, R_SUBST, nAdded, words, units)) { freeze(A, c); break; }
            return true;
        }
    }
    freeze_binaries(A, dx);
    return false;
}

// Both searches below look for a clause holding a literal of the variable, so when
// the variable is staged its own list holds every candidate. With exactly one match
// (or none) the result equals the reference's walk of the shortest list; with
// duplicates the reference's order decides, so that walk is taken.

// fast_equality_check (ifthenelse.h:24-40)
__device__ __forceinline__ uint32_t *fast_equality_check(const VeArgs &A, uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t t;
    if (A.tab && (x >> 1) == A.v) {
        uint32_t a = x, b = y, d = z;
        if (b > d) { t = b; b = d; d = t; }
        if (a > d) { t = a; a = d; d = t; }
        if (a > b) { t = a; a = b; b = t; }
        const uint32_t sz = A.ot[x].sz;
        uint32_t *hit = nullptr;
        int nhit = 0;
        for (uint32_t i = 0; i < sz && nhit < 2; i++) {
            uint32_t *c = cl_at(A, x, i);
            if (c_molten(c)) continue;
            if (c_orig(c) && c_size(c) == 3 && c[3] == a && c[4] == b && c[5] == d) { if (!hit) hit = c; nhit++; }
        }
        if (nhit < 2) return hit;
    }
    if (A.ot[y].sz > A.ot[z].sz) { t = y; y = z; z = t; }
    if (A.ot[x].sz > A.ot[y].sz) { t = x; x = y; y = t; }
    const uint32_t lx = x;
    if (y > z) { t = y; y = z; z = t; }
    if (x > z) { t = x; x = z; z = t; }
    if (x > y) { t = x; x = y; y = t; }
    const uint32_t sz = A.ot[lx].sz;
    for (uint32_t i = 0; i < sz; i++) {
        uint32_t *c = cl_at(A, lx, i);
        if (c_molten(c)) continue;
        if (c_orig(c) && c_size(c) == 3 && c[3] == x && c[4] == y && c[5] == z) return c;
    }
    return nullptr;
}

// find_ITE_gate (ifthenelse.h:42-95)
// An ITE gate needs two original ternaries on dx's list and two original,
// non-molten ternaries on ~dx's list; without them the reference's search finds
// nothing and changes nothing.
__device__ __forceinline__ uint32_t count_ternaries(const VeArgs &A, uint32_t lit, bool skip_molten)
{
    const uint32_t sz = A.ot[lit].sz;
    uint32_t n = 0;
    for (uint32_t i = 0; i < sz && n < 2; i++) {
        const uint32_t *c = cl_at(A, lit, i);
        if (c_orig(c) && c_size(c) == 3 && !(skip_molten && c_molten(c))) n++;
    }
    return n;
}

__device__ __forceinline__ bool find_ite_gate(const VeArgs &A, uint32_t dx, int nOrgCls, int &nAdded, uint64_t &words, uint32_t &units)
{
    const uint32_t fx = dx ^ 1u;
    const uint32_t is = A.ot[dx].sz;
    if (count_ternaries(A, dx, false) < 2 || count_ternaries(A, fx, true) < 2) return false;
    for (uint32_t i = 0; i < is; i++) {
        uint32_t *ci = cl_at(A, dx, i);
        if (!(c_orig(ci) && c_size(ci) == 3)) continue;
        uint32_t xi = ci[3], yi = ci[4], zi = ci[5], t;
        if (yi == dx) { t = xi; xi = yi; yi = t; }
        if (zi == dx) { t = xi; xi = zi; zi = t; }
        for (uint32_t j = i + 1; j < is; j++) {
            uint32_t *cj = cl_at(A, dx, j);
            if (!(c_orig(cj) && c_size(cj) == 3)) continue;
            uint32_t xj = cj[3], yj = cj[4], zj = cj[5];
            if (yj == dx) { t = xj; xj = yj; yj = t; }
            if (zj == dx) { t = xj; xj = zj; zj = t; }
            if ((yi >> 1) == (zj >> 1)) { t = yj; yj = zj; zj = t; }
            if ((zi >> 1) == (zj >> 1)) continue;
            if (yi != (yj ^ 1u)) continue;
            uint32_t *d1 = fast_equality_check(A, fx, yi, zi ^ 1u);
            if (!d1) continue;
            uint32_t *d2 = fast_equality_check(A, fx, yj, zj ^ 1u);
            if (!d2) continue;
            melt(A, ci); melt(A, cj); melt(A, d1); melt(A, d2);
            nAdded = 0;
            if (count_subst(A, dx >> 1, nOrgCls, dx, fx, R_SUBST, nAdded, words, units)) {
                freeze(A, ci); freeze(A, cj); freeze(A, d1); freeze(A, d2);
                return false;
            }
            return true;
        }
    }
    return false;
}

// XOR gates (xor.h:30-147)
__device__ __forceinline__ void freeze_arities(const VeArgs &A, uint32_t me, uint32_t other)
{
    for (int s = 0; s < 2; s++) {
        const uint32_t l = s ? other : me, sz = A.ot[l].sz;
        for (uint32_t i = 0; i < sz; i++) {
            uint32_t *c = cl_at(A, l, i);
            if (c_size(c) > 2 && c_molten(c)) freeze(A, c);
        }
    }
}

__device__ __forceinline__ bool make_arity(const VeArgs &A, uint32_t &parity, uint32_t *lits, int size)
{
    const uint32_t old = parity;
    while (__builtin_parity(++parity));
    for (int k = 0; k < size; k++) {
        const uint32_t bit = 1u << k;
        if ((parity & bit) != (old & bit)) lits[k] ^= 1u;
    }
    if (A.tab) {
        uint32_t own = lits[0];
        for (int k = 1; k < size; k++) if ((lits[k] >> 1) == A.v) own = lits[k];
        const uint32_t sz = A.ot[own].sz;
        uint32_t *hit = nullptr;
        int nhit = 0;
        for (uint32_t i = 0; i < sz && nhit < 2; i++) {
            uint32_t *c = cl_at(A, own, i);
            if (!(c_orig(c) && c_size(c) == size)) continue;
            bool ok = true;
            for (int k = 0; k < size && ok; k++) {
                const uint32_t l = c[3 + k];
                int j;
                for (j = 0; j < size; j++) if (l == lits[j]) break;
                ok = j < size;
            }
            if (ok) { if (!hit) hit = c; nhit++; }
        }
        if (nhit == 1) { melt(A, hit); return true; }
        if (nhit == 0) return false;
    }
    uint32_t best = lits[0];
    uint32_t minsize = A.ot[best].sz;
    for (int k = 1; k < size; k++) {
        const uint32_t ls = A.ot[lits[k]].sz;
        if (ls < minsize) { minsize = ls; best = lits[k]; }
    }
    const uint32_t sz = A.ot[best].sz;
    for (uint32_t i = 0; i < sz; i++) {
        uint32_t *c = cl_at(A, best, i);
        if (!(c_orig(c) && c_size(c) == size)) continue;
        bool ok = true;
        for (int k = 0; k < size && ok; k++) {
            const uint32_t l = c[3 + k];
            int j;
            for (j = 0; j < size; j++) if (l == lits[j]) break;
            ok = j < size;
        }
        if (ok) { melt(A, c); return true; }
    }
    return false;
}

// An XOR over a clause of `size` literals needs all 2^(size-1) parity variants,
// each an original clause of that size holding a literal of the variable. With
// fewer such clauses on the two lists the reference's attempt fails, and its only
// lasting effect is freeze_arities, which the caller applies when anything is molten.
__device__ __forceinline__ bool xor_possible(const VeArgs &A, uint32_t dx, uint32_t fx, int size)
{
    const uint32_t need = 1u << (size - 1);
    uint32_t cnt = 0;
    for (int s = 0; s < 2; s++) {
        const uint32_t l = s ? fx : dx, sz = A.ot[l].sz;
        for (uint32_t i = 0; i < sz; i++) {
            const uint32_t *c = cl_at(A, l, i);
            const int cs = c_size(c);
            if (c_orig(c) && cs == size && ++cnt >= need) return true;
        }
    }
    return false;
}

// Every parity variant of ci has ci's variables in ci's order (literals are
// sorted and a literal orders like its variable), so fewer than 2^(size-1) such
// original clauses on the two lists means the attempt fails.
__device__ __forceinline__ bool xor_variants_possible(const VeArgs &A, uint32_t dx, const uint32_t *ci, int size)
{
    const uint32_t need = 1u << (size - 1);
    uint32_t cnt = 0;
    for (int s = 0; s < 2; s++) {
        const uint32_t l = dx ^ (uint32_t)s, sz = A.ot[l].sz;
        for (uint32_t i = 0; i < sz; i++) {
            const uint32_t *c = cl_at(A, l, i);
            if (!c_orig(c) || c_size(c) != size) continue;
            bool same = true;
            for (int k = 0; k < size && same; k++) same = (c[3 + k] >> 1) == (ci[3 + k] >> 1);
            if (same && ++cnt >= need) return true;
        }
    }
    return false;
}

// Original clauses of size >= 3 on both lists counted per hash bucket of (size,
// variable sequence): 64 buckets of 3-bit saturating counters, bit-sliced. A parity
// variant of a clause has its size and variable sequence, so a bucket count is at
// least the number of variants of any clause hashed to it.
struct SeqCounts {
    uint64_t b0 = 0, b1 = 0, b2 = 0;
    __device__ __forceinline__ void add(uint32_t b)
    {
        const uint64_t m = 1ull << b;
        if (b0 & b1 & b2 & m) return;   // saturated at 7
        const uint64_t c0 = b0 & m;
        b0 ^= m;
        const uint64_t c1 = b1 & c0;
        b1 ^= c0;
        b2 ^= c1;
    }
    __device__ __forceinline__ uint32_t get(uint32_t b) const
    {
        return (uint32_t)(((b0 >> b) & 1u) | (((b1 >> b) & 1u) << 1) | (((b2 >> b) & 1u) << 2));
    }
};

__device__ __forceinline__ uint32_t seq_bucket(const uint32_t *c, int size)
{
    uint32_t h = (uint32_t)size * 0x9E3779B1u;
    for (int k = 0; k < size; k++) h = (h ^ (c[3 + k] >> 1)) * 0x85EBCA6Bu;
    return h >> 26;
}

// counts of original clauses of sizes 3..8 on both lists, 8 bits each (saturating;
// the XOR thresholds up to size 8 are at most 128), the per-sequence bucket counts;
// *molten: a clause of size > 2 is molten
__device__ __forceinline__ uint64_t xor_size_counts(const VeArgs &A, uint32_t dx, bool &molten, SeqCounts &seq)
{
    uint64_t packed = 0;
    molten = false;
    for (int s = 0; s < 2; s++) {
        const uint32_t l = dx ^ (uint32_t)s, sz = A.ot[l].sz;
        for (uint32_t i = 0; i < sz; i++) {
            const uint32_t *c = cl_at(A, l, i);
            const uint32_t cs = (uint32_t)c_size(c);
            if (c_orig(c) && cs >= 3u) {
                seq.add(seq_bucket(c, (int)cs));
                if (cs <= 8u) {
                    const uint32_t sh = (cs - 3u) * 8u;
                    if (((packed >> sh) & 0xFFu) != 0xFFu) packed += 1ull << sh;
                }
            }
            if (cs > 2u && c_molten(c)) molten = true;
        }
    }
    return packed;
}

// The size and sequence counts cover both lists, so the two XOR searches of a
// variable (positive, then negative side) share them; pre.have says they are made.
struct XorPre {
    uint64_t counts = 0;
    SeqCounts seq;
    bool have = false;
};

__device__ __forceinline__ bool any_molten(const VeArgs &A, uint32_t dx)
{
    for (int s = 0; s < 2; s++) {
        const uint32_t l = dx ^ (uint32_t)s, sz = A.ot[l].sz;
        for (uint32_t i = 0; i < sz; i++) {
            const uint32_t *c = cl_at(A, l, i);
            if (c_size(c) > 2 && c_molten(c)) return true;
        }
    }
    return false;
}

__device__ __forceinline__ bool find_xor_gate(const VeArgs &A, uint32_t dx, int nOrgCls, int &nAdded, uint64_t &words,
                                              uint32_t &units, XorPre &pre)
{
    const uint32_t fx = dx ^ 1u, v = dx >> 1;
    const uint32_t is = A.ot[dx].sz;
    bool molten = false;
    if (!pre.have) {
        pre.counts = xor_size_counts(A, dx, molten, pre.seq);
        pre.have = true;
    }
    else molten = any_molten(A, dx);
    const uint64_t counts = pre.counts;
    const SeqCounts &seq = pre.seq;
    // A clause passes the bucket test below only if its bucket holds at least 4 (size
    // 3, which also needs 4 clauses of size 3) or 7 (larger sizes): with no such bucket
    // every attempt fails, and the loop's only effect is the first failing attempt's
    // cleanup.
    if (!seq.b2 || (!(seq.b0 & seq.b1 & seq.b2) && (counts & 0xFFu) < 4u)) {
        if (!molten) return false;
        for (uint32_t i = 0; i < is; i++) {
            const uint32_t *ci = cl_at(A, dx, i);
            const int size = c_size(ci);
            if (c_orig(ci) && size >= 3 && size - 1 <= A.xor_max_arity) {
                freeze_arities(A, dx, fx);
                break;
            }
        }
        return false;
    }
    uint32_t lits[21];
    for (uint32_t i = 0; i < is; i++) {
        uint32_t *ci = cl_at(A, dx, i);
        if (!c_orig(ci)) continue;
        const int size = c_size(ci), arity = size - 1;
        if (size < 3 || arity > A.xor_max_arity) continue;
        bool possible;
        if (size <= 8) possible = ((counts >> ((size - 3) * 8)) & 0xFFu) >= (1u << (size - 1));
        else possible = xor_possible(A, dx, fx, size);
        if (possible) {   // the bucket of ci's sequence bounds its variants' count from above
            const uint32_t bc = seq.get(seq_bucket(ci, size));
            if (bc < 7u && bc < (1u << (size - 1))) possible = false;
        }
        if (possible) possible = xor_variants_possible(A, dx, ci, size);
        if (!possible) {
            if (molten) freeze_arities(A, dx, fx);   // the failing attempt's only lasting effect
            molten = false;
            continue;
        }
        molten = false;
        for (int k = 0; k < size; k++) lits[k] = ci[3 + k];
        uint32_t parity = 0;
        int itargets = 1 << arity;
        while (--itargets && make_arity(A, parity, lits, size));
        if (itargets) freeze_arities(A, dx, fx);
        else {
            melt(A, ci);
            nAdded = 0;
            if (count_subst(A, v, nOrgCls, dx, fx, R_SUBST, nAdded, words, units)) { freeze_arities(A, dx, fx); break; }
            return true;
        }
    }
    return false;
}

// model words written by toblivion / save_BN_gate for the smaller original side
__device__ __forceinline__ uint64_t model_words(const VeArgs &A, uint32_t p, int pOrgs, int nOrgs)
{
    const uint32_t side = pOrgs > nOrgs ? (p ^ 1u) : p;
    const uint32_t sz = A.ot[side].sz;
    uint64_t w = 2;
    for (uint32_t i = 0; i < sz; i++) {
        const uint32_t *c = cl_at(A, side, i);
        if (c_orig(c)) w += (uint64_t)c_size(c) + 1;
    }
    return w;
}

__device__ __forceinline__ uint32_t bn_units(const VeArgs &A, uint32_t p, uint32_t def)
{   // substitute_single (equivalence.h:42-86): clauses that collapse to a unit
    const uint32_t n = p ^ 1u, def_f = def ^ 1u;
    uint32_t u = 0;
    for (int s = 0; s < 2; s++) {
        const uint32_t lit = s ? p : n;          // negatives first
        const uint32_t kill = s ? def_f : def;   // literal whose presence deletes the clause
        const uint32_t into = s ? def : def_f;   // literal substituted for lit
        const uint32_t sz = A.ot[lit].sz;
        for (uint32_t i = 0; i < sz; i++) {
            const uint32_t *c = cl_at(A, lit, i);
            if (c_learnt(c) || c_molten(c) || has_lit(c, kill)) continue;
            if (c_orig(c) && c_size(c) - (has_lit(c, into) ? 1 : 0) == 1) u++;
        }
    }
    return u;
}


struct VeDetectOut {
    uint32_t *type, *aux, *nres, *nunits;
    uint64_t *nwords, *nmodel;
};

// detection of one elected variable (pvs[q] = v); A carries the lane's stage. With
// function-table gates on (fun_en), a variable that no other gate defines is left
// pending for k_ve_fun, its aux word saying whether a gate search counted resolvents.
__device__ __forceinline__ void ve_detect_var(const VeArgs &A, const VeDetectOut &O, uint64_t q, uint32_t v, bool fun_en)
{
        const uint32_t p = 2 * v, n = p + 1;
        const int pOrgs = count_orgs(A, p), nOrgs = count_orgs(A, n);
        uint32_t t = T_NONE, a = 0, nr = 0, nu = 0;
        uint64_t nw = 0, nm = 0;
        if (!pOrgs || !nOrgs) {
            t = T_PURE;
            nm = model_words(A, p, pOrgs, nOrgs);
            a = (uint32_t)(pOrgs > nOrgs);   // model side
        }
        else {
            const uint32_t def = find_bn_gate(A, p);
            if (def) {
                t = T_BN;
                a = def;
                nm = model_words(A, p, pOrgs, nOrgs);
                nu = bn_units(A, p, def);
            }
            else {
                const int nOrgCls = pOrgs + nOrgs;
                int nAdded = 0;
                uint64_t words = 0;
                uint32_t units = 0;
                bool gate = false;
                if (nOrgCls > 2) {
                    if (find_ao_gate(A, n, nOrgCls, nAdded, words, units)) gate = true;
                    else if (!nAdded && find_ao_gate(A, p, nOrgCls, nAdded, words, units)) gate = true;
                }
                if (!gate && nOrgCls > 3) {
                    XorPre xpre;
                    if (find_ite_gate(A, p, nOrgCls, nAdded, words, units)) gate = true;
                    else if (!nAdded && find_ite_gate(A, n, nOrgCls, nAdded, words, units)) gate = true;
                    else if (find_xor_gate(A, p, nOrgCls, nAdded, words, units, xpre)) gate = true;
                    else if (!nAdded && find_xor_gate(A, n, nOrgCls, nAdded, words, units, xpre)) gate = true;
                }
                if (gate) t = T_SUBST;
                else if (fun_en && nOrgCls > 2) {
                    t = T_FUNPEND;
                    a = nAdded != 0;
                }
                else if (!nAdded) {
                    if ((uint64_t)A.ot[p].sz * A.ot[n].sz >= RES_WAVE_MIN) t = T_RESPEND;
                    else if (!count_subst(A, v, nOrgCls, p, n, R_ALL, nAdded, words, units)) t = T_RES;
                }
                if (t == T_SUBST || t == T_RES) {
                    nr = (uint32_t)nAdded;
                    nw = words;
                    nu = units;
                    nm = model_words(A, p, pOrgs, nOrgs);
                    a = (uint32_t)(pOrgs > nOrgs);   // model side
                }
            }
        }
        O.type[q] = t | ((t == T_BN ? (uint32_t)(pOrgs > nOrgs) : 0u) << 8);
        O.aux[q] = a;
        O.nres[q] = nr;
        O.nwords[q] = nw;
        O.nunits[q] = nu;
        O.nmodel[q] = nm;
}

// One 64-lane block per 64 elected variables. The lanes' variables are staged in
// LDS in rounds: each round takes the pending lanes (in lane order) whose clauses
// fit in the block's arena together; a variable larger than the whole arena is
// detected from the pool instead. Detection of one variable touches only its own
// clauses, so the rounds do not change results.
template <uint32_t ARENA>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_ve_detect(VeArgs G, VeDetectOut O, const uint32_t *__restrict__ pvs, uint64_t npvs, uint32_t arena_words, int fun_en)
{
    static_assert(ARENA <= VE_ARENA_WORDS, "arena larger than the staged-pointer check");
    __shared__ uint32_t arena[ARENA];
    if (arena_words > ARENA) arena_words = ARENA;
    const StageSrc S{G.mem, G.ot, G.occ};
    // elected variables come in increasing score order: the heaviest blocks go first
    const uint64_t nblk = (npvs + 63) / 64;
    for (uint64_t it = blockIdx.x; it < nblk; it += gridDim.x) {
        const uint64_t q = (nblk - 1 - it) * 64 + threadIdx.x;
        const bool active = q < npvs;
        const uint32_t v = active ? pvs[q] : 0u;
        stage_rounds(S, arena, arena_words, active, v, 0, [&](const uint32_t *tab) {
            if (tab) {
                const VeArgs A{G.mem, G.ot, G.occ, G.rlimit, G.lbound, G.xor_max_arity, v, G.ot[2 * v].sz, tab, arena};
                ve_detect_var(A, O, q, v, fun_en != 0);
            }
            else ve_detect_var(G, O, q, v, fun_en != 0);   // larger than the arena: from the pool
        });
    }
}

// ------------------------------------------------------------------ function tables
// Function-table gates (option ve_fun_en; the GPU reference's find_fun_gate,
// src/gpu/function.cuh:263-311), for the variables that detection left pending. The
// variables of x's original clauses other than x, at most FUN_VARS of them, are
// numbered in order of first appearance; a clause without x's literal is then a
// Boolean function of them, a truth table of 2^FUN_VARS bits = FUN_WORDS words. One
// wave takes one variable and lane w holds word w of every table, so building a
// clause's table costs each lane one OR per literal.
//   collapse: x is defined when (positive side) AND (negative side) is false;
//   core: a positive clause whose earlier positive clauses already contradict the
//     negative side, and a negative clause whose earlier negative clauses are false
//     by themselves, leave the core (molten), as the reference's backward loops do
//     (their decisions do not depend on one another, so one forward pass makes them);
//   resolvents: every pair but those of two clauses outside the core, under the
//     clause bound; past it the clauses are frozen again and plain resolution is
//     not tried.
// Otherwise, plain resolution decides as it does without the option.
constexpr int FUN_VARS = 12, FUN_WORDS = 64;

// word w of the table of core variable k's literal (neg: of its negation)
__device__ __forceinline__ uint64_t fun_lit_word(int k, uint32_t neg, uint32_t w)
{
    uint64_t m;
    switch (k) {
        case 0: m = 0xaaaaaaaaaaaaaaaaull; break;
        case 1: m = 0xccccccccccccccccull; break;
        case 2: m = 0xf0f0f0f0f0f0f0f0ull; break;
        case 3: m = 0xff00ff00ff00ff00ull; break;
        case 4: m = 0xffff0000ffff0000ull; break;
        case 5: m = 0xffffffff00000000ull; break;
        default: m = ((w >> (k - 6)) & 1u) ? ~0ull : 0ull;
    }
    return neg ? ~m : m;
}

// word w of the table of clause c without the literal `lit`
__device__ __forceinline__ uint64_t fun_clause_word(const uint32_t *c, uint32_t lit, const uint32_t *cv, int nc, uint32_t w)
{
    uint64_t f = 0;
    const int sz = c_size(c);
    for (int k = 0; k < sz; k++) {
        const uint32_t o = c[3 + k];
        if (o == lit) continue;
        int m = 0;
        while (m < nc && cv[m] != (o >> 1)) m++;
        f |= fun_lit_word(m, o & 1u, w);
    }
    return f;
}

__device__ __forceinline__ uint64_t fun_and_list(const VeArgs &A, uint32_t lit, const uint32_t *cv, int nc, uint32_t w)
{
    uint64_t f = ~0ull;
    const uint32_t sz = A.ot[lit].sz;
    for (uint32_t i = 0; i < sz; i++) {
        const uint32_t *c = cl_at(A, lit, i);
        if (c_orig(c)) f &= fun_clause_word(c, lit, cv, nc, w);
    }
    return f;
}

// one wave per group of 64 elected variables; its pending variables one after another
__global__ void __launch_bounds__(64) k_ve_fun(VeArgs A, VeDetectOut O, const uint32_t *__restrict__ pvs, uint64_t npvs)
{
    __shared__ uint32_t cv[FUN_VARS];
    __shared__ int nc_s;
    const uint32_t lane = threadIdx.x;
    for (uint64_t base = (uint64_t)blockIdx.x * 64; base < npvs; base += (uint64_t)gridDim.x * 64) {
        const bool pend = base + lane < npvs && (O.type[base + lane] & 0xFFu) == T_FUNPEND;
        uint64_t todo = __ballot(pend);
        while (todo) {
            const uint64_t q = base + (uint64_t)__builtin_ctzll(todo);
            todo &= todo - 1;
            const uint32_t v = pvs[q], p = 2 * v, n = p + 1;
            const bool counted = O.aux[q] != 0;
            __syncthreads();   // the previous variable's readers of cv are done
            if (lane == 0) {
                int nc = 0;
                for (uint32_t s = 0; s < 2 && nc >= 0; s++) {
                    const uint32_t lit = p | s, sz = A.ot[lit].sz;
                    for (uint32_t i = 0; i < sz && nc >= 0; i++) {
                        const uint32_t *c = cl_at(A, lit, i);
                        if (!c_orig(c)) continue;
                        const int cs = c_size(c);
                        for (int k = 0; k < cs; k++) {
                            const uint32_t u = c[3 + k] >> 1;
                            if (u == v) continue;
                            int m = 0;
                            while (m < nc && cv[m] != u) m++;
                            if (m < nc) continue;
                            if (nc == FUN_VARS) { nc = -1; break; }
                            cv[nc++] = u;
                        }
                    }
                }
                nc_s = nc;
            }
            __syncthreads();
            const int nc = nc_s;
            bool collapsed = false, core = false;
            if (nc >= 0) {
                const uint64_t neg = fun_and_list(A, n, cv, nc, lane);
                const uint64_t pos = fun_and_list(A, p, cv, nc, lane);
                collapsed = __ballot((pos & neg) != 0) == 0;
                if (collapsed) {
                    for (uint32_t s = 0; s < 2; s++) {
                        const uint32_t lit = p | s, sz = A.ot[lit].sz;
                        uint64_t pre = ~0ull;   // the earlier original clauses of the list
                        for (uint32_t i = 0; i < sz; i++) {
                            uint32_t *c = A.mem + A.occ[A.ot[lit].off + i];
                            if (!c_orig(c)) continue;
                            const uint64_t f = s ? pre : (neg & pre);
                            if (__ballot(f != 0) == 0) {
                                core = true;
                                if (lane == 0) melt(A, c);
                            }
                            pre &= fun_clause_word(c, lit, cv, nc, lane);
                        }
                    }
                }
            }
            if (lane == 0) {
                const int pOrgs = count_orgs(A, p), nOrgs = count_orgs(A, n), nOrgCls = pOrgs + nOrgs;
                uint32_t t = T_NONE;
                int nAdded = counted ? 1 : 0;
                uint64_t words = 0;
                uint32_t units = 0;
                if (collapsed) {
                    nAdded = 0;
                    if (!count_subst(A, v, nOrgCls, p, n, R_CORE, nAdded, words, units)) t = T_CORE;
                    else if (core) {   // freezeClauses (src/gpu/elimination.cuh:106-120)
                        for (uint32_t s = 0; s < 2; s++) {
                            const uint32_t lit = p | s, sz = A.ot[lit].sz;
                            for (uint32_t i = 0; i < sz; i++) {
                                uint32_t *c = A.mem + A.occ[A.ot[lit].off + i];
                                if (c_orig(c)) freeze(A, c);
                            }
                        }
                    }
                }
                if (t == T_NONE && !nAdded && !count_subst(A, v, nOrgCls, p, n, R_ALL, nAdded, words, units)) t = T_RES;
                O.type[q] = t;
                O.aux[q] = (uint32_t)(pOrgs > nOrgs);   // model side
                O.nres[q] = t != T_NONE ? (uint32_t)nAdded : 0u;
                O.nwords[q] = t != T_NONE ? words : 0ull;
                O.nunits[q] = t != T_NONE ? units : 0u;
                O.nmodel[q] = t != T_NONE ? model_words(A, p, pOrgs, nOrgs) : 0ull;
            }
        }
    }
}

// ------------------------------------------------------------------ plain resolution by lane groups
// countResolvents (src/cpu/simplify.h:255-287) for the variables detection left
// pending (at least RES_WAVE_MIN resolvent pairs): a group of RES_G lanes takes one
// variable, RES_G pairs of |occ(x)| x |occ(~x)| at a time; a wave runs 64 / RES_G
// variables at once. The sequential count fails at the first resolvent that makes the
// count exceed the clauses before or is longer than the clause limit; that happens iff
// the total count exceeds the bound or some counted resolvent is too long, so the group
// checks both after every step and the decision and the counts on success are the same.
constexpr int RES_G = 16;

template <typename T>
__device__ __forceinline__ T group_total(T v)
{
#pragma unroll
    for (int o = RES_G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, RES_G);
    return v;
}

// LDS accesses of the wave's lanes ordered across the point (the wave runs in lockstep)
__device__ __forceinline__ void res_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS per group: the variable's clause images [ref_lo, ref_hi, word0, sig, size, lits..]
// (as staged by stage.hpp) and, per side, each entry's image offset; a variable that
// does not fit is read from the pool
constexpr uint32_t RES_WORDS = 1024, RES_ENT = 128;

__global__ void __launch_bounds__(64) k_ve_res(VeArgs A, VeDetectOut O, const uint32_t *__restrict__ pvs, uint64_t npvs)
{
    constexpr uint32_t NG = 64 / RES_G;
    __shared__ uint32_t img_all[NG][RES_WORDS];
    __shared__ uint16_t pos_all[NG][2][RES_ENT];
    const uint32_t lane = threadIdx.x, gl = lane % RES_G, g = lane / RES_G;
    const uint64_t gmask = (RES_G == 64 ? ~0ull : ((1ull << RES_G) - 1ull)) << (g * RES_G);
    uint32_t *img = img_all[g];
    uint16_t(*pos)[RES_ENT] = pos_all[g];
    for (uint64_t base = (uint64_t)blockIdx.x * 64; base < npvs; base += (uint64_t)gridDim.x * 64) {
        const bool pend = base + lane < npvs && (O.type[base + lane] & 0xFFu) == T_RESPEND;
        uint64_t todo = __ballot(pend);
        while (todo) {
            // group g takes the g-th pending variable of this step
            uint64_t t = todo;
            for (uint32_t k = 0; k < g && t; k++) t &= t - 1;
            for (uint32_t k = 0; k < NG && todo; k++) todo &= todo - 1;
            if (!t) continue;
            const uint64_t q = base + (uint64_t)__builtin_ctzll(t);
            const uint32_t v = pvs[q], p = 2 * v, n = p + 1;
            const OL lp = A.ot[p], ln = A.ot[n];
            // stage the images (prefix sums of their sizes over the group), counting originals
            const uint32_t ne = lp.sz + ln.sz;
            bool fits = lp.sz <= RES_ENT && ln.sz <= RES_ENT;
            uint32_t po = 0, no = 0, total_w = 0;
            unsigned long long before = 0;
            for (uint32_t k0 = 0; k0 < ne; k0 += RES_G) {
                const uint32_t k = k0 + gl;
                uint64_t r = 0;
                uint32_t w = 0;
                if (k < ne) {
                    r = k < lp.sz ? A.occ[lp.off + k] : A.occ[ln.off + (k - lp.sz)];
                    const uint32_t *c = A.mem + r;
                    const uint32_t sz = c[2];
                    if (c_orig(c)) {
                        (k < lp.sz ? po : no)++;
                        before += sz;
                    }
                    w = 5u + sz;
                }
                uint32_t incl = w;
#pragma unroll
                for (int o = 1; o < RES_G; o <<= 1) {
                    const uint32_t x = __shfl_up(incl, o, RES_G);
                    if (gl >= (uint32_t)o) incl += x;
                }
                const uint32_t at = total_w + incl - w;
                if (fits && k < ne && at + w <= RES_WORDS) {
                    const uint32_t *src = A.mem + r;
                    img[at] = (uint32_t)r;
                    img[at + 1] = (uint32_t)(r >> 32);
                    for (uint32_t x = 0; x + 2 < w; x++) img[at + 2 + x] = src[x];
                    if (k < lp.sz) pos[0][k] = (uint16_t)(at + 2);
                    else pos[1][k - lp.sz] = (uint16_t)(at + 2);
                }
                total_w += __shfl(incl, RES_G - 1, RES_G);
                fits = fits && total_w <= RES_WORDS;
            }
            res_wave_sync();   // the group's images are in LDS before any lane reads them
            po = group_total(po);
            no = group_total(no);
            before = group_total(before);
            const uint32_t clsbefore = po + no;
            const uint64_t npairs = (uint64_t)lp.sz * ln.sz;
            uint32_t added = 0, unitc = 0, total = 0;
            unsigned long long lits = 0;
            bool over = false, fail = false;
            uint32_t i = gl / ln.sz, j = gl % ln.sz;   // pair b0 + gl = i |occ(~x)| + j, advanced by RES_G
            const uint32_t di = RES_G / ln.sz, dj = RES_G % ln.sz;
            for (uint64_t b0 = 0; b0 < npairs && !fail; b0 += RES_G) {
                if (b0 + gl < npairs) {
                    const uint32_t *ci = fits ? img + pos[0][i] : A.mem + A.occ[lp.off + i];
                    const uint32_t *cj = fits ? img + pos[1][j] : A.mem + A.occ[ln.off + j];
                    if (c_orig(ci) && c_orig(cj)) {
                        const int rs = merge_len(v, ci, cj);
                        if (rs > 1) {
                            added++;
                            lits += (unsigned)rs;
                            if (A.rlimit && rs > A.rlimit) over = true;
                        }
                        else if (rs == 1) unitc++;
                    }
                }
                i += di;
                j += dj;
                if (j >= ln.sz) { j -= ln.sz; i++; }
                total = group_total(added);
                fail = total > clsbefore || (__ballot(over) & gmask) != 0;
            }
            if (!fail) {
                lits = group_total(lits);
                unitc = group_total(unitc);
                if (A.lbound && lits > before) fail = true;
            }
            if (gl == 0) {
                O.type[q] = fail ? T_NONE : T_RES;
                O.aux[q] = (uint32_t)(po > no);   // model side
                O.nres[q] = fail ? 0u : total;
                O.nwords[q] = fail ? 0ull : 3ull * total + lits;
                O.nunits[q] = fail || !total ? 0u : unitc;
                O.nmodel[q] = fail ? 0ull : model_words(A, p, (int)po, (int)no);
            }
            res_wave_sync();   // reads of this variable's images end before the next staging
        }
    }
}

struct VeOut {
    uint32_t *resolved;
    uint32_t *units;
    uint64_t *refs;
    uint8_t *vstate;
    uint32_t *pvs;
    uint64_t ref_base, data_base, model_base;
    ProofLog P;   // proof events: resolvents, substituted clauses and units, decided after the pass
    const uint32_t *vorg;   // the caller's variable mapping for the model records (null: identity)
    uint32_t *rw;           // rewrite map of the occurrence-table update (substituted clauses)
    uint64_t rw_words;
    uint64_t *rmlog;        // and its removal log (the substituted literal)
    uint64_t rmcap;
    unsigned long long *cnt;
};

__device__ __forceinline__ void save_side(const VeArgs &A, const uint32_t *vorg, uint32_t *&rs, uint32_t p, bool pbig)
{   // saveClause/saveWitness for the smaller side (simplify.h:299-331, model.cpp:263-307),
    // literals written in original numbering
    const uint32_t n = p ^ 1u;
    const uint32_t side = pbig ? n : p;
    const uint32_t sz = A.ot[side].sz;
    for (uint32_t i = 0; i < sz; i++) {
        const uint32_t *c = cl_at(A, side, i);
        if (!c_orig(c)) continue;
        const int cs = c_size(c);
        int pos = -1;
        for (int k = 0; k < cs; k++) { rs[k] = org_lit(vorg, c[3 + k]); if (c[3 + k] == side) pos = k; }
        if (pos > 0) { const uint32_t t = rs[pos]; rs[pos] = rs[0]; rs[0] = t; }
        rs[cs] = (uint32_t)cs;
        rs += cs + 1;
    }
    rs[0] = org_lit(vorg, pbig ? p : n);   // witness
    rs[1] = 1;
    rs += 2;
}

__device__ __forceinline__ void delete_lists(const VeArgs &A, uint32_t p)
{   // toblivion: every clause on both lists is deleted, lists cleared
    for (int s = 0; s < 2; s++) {
        OL *o = A.ot + (p ^ (uint32_t)s);
        const uint64_t *d = A.occ + o->off;
        for (uint32_t i = 0; i < o->sz; i++) c_markdel(A.mem + d[i]);
        o->sz = 0;
    }
}

__device__ __forceinline__ void substitute_one(uint32_t dx, uint32_t *org, uint32_t def, uint32_t *&us)
{   // substitute_single(dx, org, def) (equivalence.h:24-40)
    uint32_t *lits = org + 3;
    const int sz = c_size(org);
    int n = 0;
    for (int i = 0; i < sz; i++) {
        const uint32_t lit = lits[i];
        if (lit == dx) lits[n++] = def;
        else if (lit != def) lits[n++] = lit;
    }
    org[2] = (uint32_t)n;
    if (n == 1) { *us++ = lits[0]; return; }
    sort_u32_small(lits, n);
    uint32_t sig = 0;
    for (int i = 0; i < n; i++) sig |= maphash(lits[i]);
    org[1] = sig;
}

__global__ void k_ve_apply(VeArgs A, VeOut O, const uint32_t *__restrict__ type, const uint32_t *__restrict__ aux,
                           uint32_t npvs, const uint32_t *__restrict__ nres, const uint64_t *__restrict__ roff,
                           const uint64_t *__restrict__ woff, const uint64_t *__restrict__ uoff,
                           const uint64_t *__restrict__ moff)
{
    for (uint64_t q = gtid(); q < npvs; q += gstride()) {
        const uint32_t tw = type[q], t = tw & 0xFFu;
        if (t == T_NONE) continue;
        const uint32_t v = O.pvs[q], p = 2 * v, n = p + 1;
        uint32_t *rs = O.resolved + O.model_base + moff[q];
        uint32_t *us = O.units + uoff[q];
        uint32_t seq = 0;
        if (t == T_BN) {
            const uint32_t def = aux[q], def_f = def ^ 1u;
            save_side(A, O.vorg, rs, p, (tw >> 8) & 1u);
            // negatives, then positives
            for (int s = 0; s < 2; s++) {
                const uint32_t lit = s ? p : n, kill = s ? def_f : def, into = s ? def : def_f;
                const OL o = A.ot[lit];
                for (uint32_t i = 0; i < o.sz; i++) {
                    uint32_t *c = A.mem + A.occ[o.off + i];
                    if (c_learnt(c) || c_molten(c) || has_lit(c, kill)) c_markdel(c);
                    else if (c_orig(c)) {
                        const uint32_t *u0 = us;
                        mark_rewritten(O.rw, O.rw_words, A.occ[o.off + i]);
                        log_removed(O.rmlog, &O.cnt[CNT_OTRM], O.rmcap, lit, A.occ[o.off + i]);
                        substitute_one(lit, c, into, us);
                        if (us != u0) plog(O.P, (uint32_t)q, seq, PE_CAND_UNIT, us[-1]);
                        else plog(O.P, (uint32_t)q, seq, PE_CAND_ADD, A.occ[o.off + i]);
                    }
                }
            }
        }
        else {
            const bool pbig = aux[q] != 0;
            if ((t == T_SUBST || t == T_RES || t == T_CORE) && nres[q]) {   // xsubstitute/xresolve only if nAddedCls
                uint32_t dx = p, fx = n;
                if (A.ot[dx].sz > A.ot[fx].sz) { dx = n; fx = p; }
                const OL om = A.ot[dx], oo = A.ot[fx];
                uint64_t w = O.data_base + woff[q];
                uint64_t r = O.ref_base + roff[q];
                const int rule = t == T_SUBST ? R_SUBST : t == T_CORE ? R_CORE : R_ALL;
                for (uint32_t i = 0; i < om.sz; i++) {
                    const uint32_t *ci = A.mem + A.occ[om.off + i];
                    if (!c_orig(ci)) continue;
                    const bool a = c_molten(ci);
                    for (uint32_t j = 0; j < oo.sz; j++) {
                        const uint32_t *cj = A.mem + A.occ[oo.off + j];
                        if (!c_orig(cj)) continue;
                        if (skip_pair(rule, a, c_molten(cj))) continue;
                        // the length is known first so that nothing is written past this
                        // variable's region (merge_out writes before it sees a tautology)
                        const int rl = merge_len(v, ci, cj);
                        if (rl == 0) continue;
                        int len;
                        if (rl == 1) {
                            uint32_t one[2];
                            merge_out(v, ci, cj, one, len);
                            *us++ = one[0];
                            plog(O.P, (uint32_t)q, seq, PE_CAND_UNIT, one[0]);
                            continue;
                        }
                        uint32_t *dst = A.mem + w;
                        merge_out(v, ci, cj, dst + 3, len);
                        uint32_t sig = 0;
                        for (int k = 0; k < len; k++) sig |= maphash(dst[3 + k]);
                        dst[0] = ST_ORIGINAL | 8u;   // markAdded
                        dst[1] = sig;
                        dst[2] = (uint32_t)len;
                        plog(O.P, (uint32_t)q, seq, PE_CAND_ADD, w);
                        O.refs[r++] = w;
                        w += 3 + (uint64_t)len;
                    }
                }
            }
            save_side(A, O.vorg, rs, p, pbig);
            delete_lists(A, p);
        }
        O.vstate[v] = MELTED_M;
        O.pvs[q] = 0;
        plog_end(O.P, (uint32_t)q, seq);
    }
}

// The reference's atomic BVE (ve_k, src/gpu/kernels.cu:298-373, option veatomic,
// src/gpu/options.cu:24, off by default) reserves each eliminated variable's space in
// the pool, the unit list and the model stack with atomics as it goes, so their order
// depends on the schedule. With ve_atomic_en the detection is the same and each block
// reserves the regions of its 256 elected variables with one atomic per counter
// (elected order inside a block, block order as the blocks arrive) instead of the four
// scans; cur[0..3] count references, units, words and model words.
__global__ void __launch_bounds__(256) k_ve_reserve(const uint32_t *__restrict__ nres,
                                                    const uint32_t *__restrict__ nunits,
                                                    const uint64_t *__restrict__ nwords,
                                                    const uint64_t *__restrict__ nmodel, uint32_t npvs,
                                                    uint64_t *__restrict__ roff, uint64_t *__restrict__ uoff,
                                                    uint64_t *__restrict__ woff, uint64_t *__restrict__ moff,
                                                    unsigned long long *cur)
{
    // the loop bound is the same for every thread of a block: block_reserve needs all of them
    for (uint64_t q0 = (uint64_t)blockIdx.x * blockDim.x; q0 < npvs; q0 += gstride()) {
        const uint64_t q = q0 + threadIdx.x;
        const bool in = q < npvs;
        const uint64_t r = block_reserve(in ? nres[q] : 0u, cur + 0);
        const uint64_t u = block_reserve(in ? nunits[q] : 0u, cur + 1);
        const uint64_t w = block_reserve(in ? nwords[q] : 0u, cur + 2);
        const uint64_t m = block_reserve(in ? nmodel[q] : 0u, cur + 3);
        if (in) {
            roff[q] = r;
            uoff[q] = u;
            woff[q] = w;
            moff[q] = m;
        }
    }
}

// batch: the starts of the non-empty record chunks of one pass, appended in elected
// order (= stack order) to the chunk list
__global__ void k_chunk_flags(const uint64_t *__restrict__ moff, uint32_t np, uint32_t *__restrict__ flags)
{
    for (uint64_t q = gtid(); q < np; q += gstride()) flags[q] = moff[q + 1] > moff[q];
}

__global__ void k_chunk_append(const uint64_t *__restrict__ moff, const uint64_t *__restrict__ pos, uint32_t np,
                               uint64_t model_base, uint64_t *__restrict__ out)
{
    for (uint64_t q = gtid(); q < np; q += gstride())
        if (moff[q + 1] > moff[q]) out[pos[q]] = model_base + moff[q];
}

// batch: append the record chunks of one pass (per-variable offsets moff into the
// records that start at c.nres) to the chunk list
static void append_chunks(Ctx &c, const uint64_t *moff, uint32_t np)
{
    ensure(c, c.mchunk, c.nchunk + np + 1, true);
    ensure(c, c.tmpu32c, (uint64_t)np + 1);
    ensure(c, c.tmpu64c, (uint64_t)np + 2);
    k_chunk_flags<<<grid_for(np, 256, 8192), 256, 0, c.stream>>>(moff, np, c.tmpu32c.p);
    scan_u32_to_u64(c, c.tmpu32c.p, c.tmpu64c.p, np, c.cnt.p + CNT_TOTAL);
    k_chunk_append<<<grid_for(np, 256, 8192), 256, 0, c.stream>>>(moff, c.tmpu64c.p, np, c.nres, c.mchunk.p + c.nchunk);
    read_counters(c, CNT_TOTAL + 1);
    c.nchunk += c.hcnt[CNT_TOTAL];
}

void ve(Ctx &c)
{
    if (!c.opts.ve_en || !c.npvs) return;
    const uint32_t np = c.npvs;
    VeArgs A{c.mem.p, c.ot.p, c.occ.p, c.opts.ve_clause_limit, c.opts.ve_lbound_en, c.opts.xor_max_arity};
    // per-variable outputs: type, aux (u32) | nres, nunits (u32 -> scanned) | nwords, nmodel (u64)
    ensure(c, c.tmpu32a, 2ull * np + 2);
    ensure(c, c.tmpu32b, 2ull * np + 2);
    ensure(c, c.tmpu64a, 2ull * np + 2);
    ensure(c, c.tmpu64b, 4ull * np + 12);   // + the reservation counters of ve_atomic_en
    uint32_t *type = c.tmpu32a.p, *aux = c.tmpu32a.p + np + 1;
    uint32_t *nres = c.tmpu32b.p, *nunits = c.tmpu32b.p + np + 1;
    uint64_t *nwords = c.tmpu64a.p, *nmodel = c.tmpu64a.p + np + 1;
    uint64_t *roff = c.tmpu64b.p, *uoff = roff + np + 2, *woff = uoff + np + 2, *moff = woff + np + 2;
    {
        const VeDetectOut DO{type, aux, nres, nunits, nwords, nmodel};
        KTimer kt(c, K_VE_DETECT, pv_bytes(c, 0, np, true));
        if (c.tune.ve_arena == 4096)
            k_ve_detect<4096><<<grid_for(np, 64, 16384), 64, 0, c.stream>>>(A, DO, c.pvs.p, np, c.stage_words,
                                                                            c.opts.ve_fun_en);
        else
            k_ve_detect<8192><<<grid_for(np, 64, 16384), 64, 0, c.stream>>>(A, DO, c.pvs.p, np, c.stage_words,
                                                                            c.opts.ve_fun_en);
    }
    if (c.opts.ve_fun_en) {
        const VeDetectOut DO{type, aux, nres, nunits, nwords, nmodel};
        KTimer kt(c, K_VE_FUN, pv_bytes(c, 0, np, true));
        k_ve_fun<<<grid_for(np, 64, 16384), 64, 0, c.stream>>>(A, DO, c.pvs.p, np);
    }
    {
        const VeDetectOut DO{type, aux, nres, nunits, nwords, nmodel};
        KTimer kt(c, K_VE_RES, 0.0);
        k_ve_res<<<grid_for(np, 64, 16384), 64, 0, c.stream>>>(A, DO, c.pvs.p, np);
    }
    uint64_t tot[4 * 2];
    if (c.opts.ve_atomic_en && !c.batch) {
        uint64_t *cur = c.tmpu64b.p + 4ull * np + 8;
        PFG_HIP(hipMemsetAsync(cur, 0, 32, c.stream));
        k_ve_reserve<<<grid_for(np, 256, 8192), 256, 0, c.stream>>>(nres, nunits, nwords, nmodel, np, roff, uoff,
                                                                    woff, moff, (unsigned long long *)cur);
        PFG_HIP(hipMemcpyAsync(tot, cur, 32, hipMemcpyDeviceToHost, c.stream));
    }
    else {   // exclusive scans in elected order, the totals stored one past the end
        scan_u32_to_u64(c, nres, roff, np, roff + np);
        scan_u32_to_u64(c, nunits, uoff, np, uoff + np);
        scan_u64(c, nwords, woff, np, woff + np);
        scan_u64(c, nmodel, moff, np, moff + np);
        PFG_HIP(hipMemcpyAsync(&tot[0], roff + np, 8, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipMemcpyAsync(&tot[1], uoff + np, 8, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipMemcpyAsync(&tot[2], woff + np, 8, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipMemcpyAsync(&tot[3], moff + np, 8, hipMemcpyDeviceToHost, c.stream));
    }
    PFG_HIP(hipStreamSynchronize(c.stream));
    const uint64_t add_refs = tot[0], add_units = tot[1], add_words = tot[2], add_model = tot[3];
    ensure(c, c.mem, c.msz + add_words + 1, true);
    ensure(c, c.refs, c.nrefs + add_refs + 1, true);
    ensure(c, c.resolved, c.nres + add_model + 1, true);
    ensure(c, c.units, add_units + 1);
    A.mem = c.mem.p;
    uint32_t *units = c.units.p;
    VeOut O{c.resolved.p, units, c.refs.p, c.vstate.p, c.pvs.p, c.nrefs, c.msz, c.nres,
            proof_begin(c, np, c.nrefs + add_refs + add_units + np), c.has_vorg ? c.vorg.p : nullptr,
            c.otdirty.p, c.otdirty.p ? c.otdirty_words : 0, c.otrm.p, c.otrm_cap, (unsigned long long *)c.cnt.p};
    {
        KTimer kt(c, K_VE_APPLY, 4.0 * add_words + 8.0 * add_refs + 4.0 * add_model);
        k_ve_apply<<<grid_for(np, 64, 16384), 64, 0, c.stream>>>(A, O, type, aux, np, nres, roff, woff, uoff, moff);
    }
    if (c.batch && add_model) append_chunks(c, moff, np);
    c.nrefs += add_refs;
    c.msz += add_words;
    c.nres += add_model;
    proof_end(c, np, true);   // before the units are assigned: decisions read the values before the pass
    if (add_units) merge_units(c, units, add_units);
}

// ------------------------------------------------------------------ BCE
// Blocked clause elimination after VE (src/cpu/elimination.cpp:131-166): for every
// elected variable v still in the list (VE zeroes the ones it eliminated) whose two
// lists hold at most bce_limit entries, an original clause of the negative list is
// blocked when its resolvent on v with every original clause of the positive list
// is a tautology; it is saved to the model stack with witness -v and deleted.
//
// The clauses of two elected variables are disjoint (LCVE freezes every variable of
// an elected variable's clauses, and VE's resolvents hold no elected variable), so
// the variables are independent and one wave takes one variable, a lane per
// negative clause. The stack order is the sequential one: elected order, then list
// order. A count pass sizes each variable's records; after the scan the apply pass
// decides again (only for variables with records) and writes them.

struct BceArgs {
    uint32_t *mem;
    const OL *ot;
    const uint64_t *occ;
    uint32_t limit;
    const uint32_t *vorg;   // model records in original numbering (null: identity)
};

// isTautology (src/cpu/simplify.h:119-137): the resolvent of c1 and c2 on x holds a
// complementary pair (only the common prefix of the two sorted clauses is scanned)
__device__ __forceinline__ bool resolvent_is_tautology(uint32_t x, const uint32_t *c1, const uint32_t *c2)
{
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c1 + 3, *b = c2 + 3;
    int i = 0, j = 0;
    while (i < n1 && j < n2) {
        const uint32_t l1 = a[i], l2 = b[j], v1 = l1 >> 1, v2 = l2 >> 1;
        if (v1 == x) i++;
        else if (v2 == x) j++;
        else if ((l1 ^ l2) == 1u) return true;
        else if (v1 < v2) i++;
        else if (v2 < v1) j++;
        else { i++; j++; }
    }
    return false;
}

template <bool APPLY>
__global__ void __launch_bounds__(64) k_bce(BceArgs A, const uint32_t *__restrict__ pvs, uint32_t np,
                                            uint64_t *__restrict__ nmodel, uint32_t *__restrict__ resolved,
                                            const uint64_t *__restrict__ moff)
{
    const int lane = __lane_id();
    for (uint32_t q = blockIdx.x; q < np; q += gridDim.x) {
        const uint32_t v = pvs[q];
        uint64_t words = 0;
        if (v && (!APPLY || moff[q + 1] > moff[q])) {
            const uint32_t p = 2 * v, n = p + 1;
            const OL lp = A.ot[p], ln = A.ot[n];
            if (lp.sz <= A.limit && ln.sz <= A.limit) {
                uint32_t *rs = APPLY ? resolved + moff[q] : nullptr;
                for (uint32_t j0 = 0; j0 < ln.sz; j0 += 64) {
                    const uint32_t j = j0 + (uint32_t)lane;
                    uint32_t *neg = nullptr;
                    uint32_t w = 0;
                    if (j < ln.sz) {
                        neg = A.mem + A.occ[ln.off + j];
                        if (c_orig(neg)) {
                            bool all = true;
                            for (uint32_t k = 0; k < lp.sz && all; k++) {
                                const uint32_t *pos = A.mem + A.occ[lp.off + k];
                                if (c_orig(pos) && !resolvent_is_tautology(v, neg, pos)) all = false;
                            }
                            if (all) w = (uint32_t)c_size(neg) + 1;
                        }
                    }
                    uint32_t incl = w;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        const uint32_t t = __shfl_up(incl, o, 64);
                        if (lane >= o) incl += t;
                    }
                    if (APPLY && w) {   // saveClause(neg, size, -v) (model.cpp:263-280), then markDeleted
                        uint32_t *r = rs + (incl - w);
                        const int cs = c_size(neg);
                        int at = -1;
                        for (int k = 0; k < cs; k++) {
                            r[k] = org_lit(A.vorg, neg[3 + k]);
                            if (neg[3 + k] == n) at = k;
                        }
                        if (at > 0) { const uint32_t t = r[at]; r[at] = r[0]; r[0] = t; }
                        r[cs] = (uint32_t)cs;
                        c_markdel(neg);
                    }
                    const uint32_t tot = __shfl(incl, 63, 64);
                    if (APPLY) rs += tot;
                    words += tot;
                }
            }
        }
        if (!APPLY && lane == 0) nmodel[q] = words;
    }
}

void bce(Ctx &c)
{
    if (!c.opts.bce_en || !c.npvs) return;
    const uint32_t np = c.npvs;
    const BceArgs A{c.mem.p, c.ot.p, c.occ.p, (uint32_t)std::max<int32_t>(c.opts.bce_limit, 0),
                    c.has_vorg ? c.vorg.p : nullptr};
    ensure(c, c.tmpu64a, (uint64_t)np + 2);
    ensure(c, c.tmpu64b, (uint64_t)np + 2);
    uint64_t *nmodel = c.tmpu64a.p, *moff = c.tmpu64b.p;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(np, 65536);
    {
        KTimer kt(c, K_BCE_COUNT, pv_bytes(c, 0, np, true));
        k_bce<false><<<grid, 64, 0, c.stream>>>(A, c.pvs.p, np, nmodel, nullptr, nullptr);
    }
    scan_u64(c, nmodel, moff, np, moff + np);
    uint64_t add_model = 0;
    PFG_HIP(hipMemcpyAsync(&add_model, moff + np, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    if (!add_model) return;
    ensure(c, c.resolved, c.nres + add_model + 1, true);
    {
        KTimer kt(c, K_BCE_APPLY, 4.0 * add_model);
        k_bce<true><<<grid, 64, 0, c.stream>>>(A, c.pvs.p, np, nullptr, c.resolved.p + c.nres, moff);
    }
    if (c.batch) append_chunks(c, moff, np);
    c.nres += add_model;
}

}  // namespace pfg
