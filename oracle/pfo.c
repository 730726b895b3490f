/*
 * pfo.c — CPU restatement of ParaFROST's SIGmA simplifier, src/cpu semantics.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the GPU simplifier. Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg load this library;
 * the product never links or calls it.
 *
 * Every function names the reference file:line whose behaviour it restates
 * (paths relative to the ParaFROST repository). Quirks of the reference are kept
 * on purpose because parity is bit-exact, for example:
 *   - BVE resolution with only unit resolvents adds nothing (bve.cpp:130-136:
 *     xresolve runs only when nAddedCls > 0, and countResolvents counts size > 1);
 *   - ERE picks the search list only from literals met before one antecedent is
 *     exhausted (redundancy.h:87-110);
 *   - occurrence lists longer than 20 are sorted in libstdc++ std::sort order
 *     (sort.h:78-90), which is not stable; pfo_introsort_* restates it.
 * Units found by SUB strengthening are enqueued immediately and their satisfied
 * clauses deleted before later variables are processed (elimination.cpp:214-220).
 */
#include "pfo.h"
#include <stdlib.h>
#include <string.h>
#include <limits.h>

/* ------------------------------------------------------------------ */
/* literal / clause encoding (src/cpu/const.h:52-86, sclause.h:28-35) */
/* ------------------------------------------------------------------ */
#define ST_ORIGINAL 0u
#define ST_LEARNT 1u
#define ST_DELETED 2u
#define MELTED_M 1u
#define FROZEN_M 2u
#define USAGET3 1u
#define ABSV(l) ((l) >> 1)
#define FLIP(l) ((l) ^ 1u)
#define MAPHASH(l) (1u << ((l) & 31u))
#define RESOLUTION 1
#define SUBSTITUTION 2
#define HSE_MAX_CL_SIZE 1000 /* subsume.h:25 */

typedef struct { uint64_t *d; int32_t sz, cap; } ol_t;

struct pfo {
    uint32_t nvars;
    pfo_opts opts;
    uint32_t *mem; uint64_t msz, mcap;
    uint64_t *scnf; uint64_t nscnf, scap;
    ol_t *ot;
    uint32_t *ps, *ns;
    int8_t *value;
    uint8_t *vstate, *frozen;
    uint8_t *ifrozen;          /* ifreeze/iassume marks (solve.h:841), skipped by LCVE */
    uint32_t *trail; uint32_t ntrail, propagated;
    uint32_t *resolved; uint64_t nres, rescap;
    uint32_t *pvs; uint32_t npvs;
    uint32_t *eligible, *scores, *fstack; uint32_t nfstack;
    int32_t phase; uint32_t mu_inc; int unsat;
    int64_t nclauses, nliterals;
    uint32_t *outc; int32_t outn, outcap;
    uint8_t *proof; uint64_t nproof, proofcap; int proof_closed;   /* binary DRAT (proof_en) */
    uint32_t *vorg;            /* variable -> original variable (NULL: identity), solve.h:76 */
};

static inline uint32_t *CL(const pfo *s, uint64_t r) { return s->mem + r; }
static inline uint32_t c_st(const uint32_t *c) { return c[0] & 3u; }
static inline int c_orig(const uint32_t *c) { return (c[0] & 3u) == 0; }
static inline int c_del(const uint32_t *c) { return (c[0] & ST_DELETED) != 0; }
static inline int c_learnt(const uint32_t *c) { return (c[0] & ST_LEARNT) != 0; }
static inline int c_molten(const uint32_t *c) { return (c[0] >> 2) & 1u; }
static inline void c_melt(uint32_t *c) { c[0] |= 4u; }
static inline void c_freeze(uint32_t *c) { c[0] &= ~4u; }
static inline void c_markdel(uint32_t *c) { c[0] = (c[0] & ~3u) | ST_DELETED; }
static inline void c_setst(uint32_t *c, uint32_t st) { c[0] = (c[0] & ~3u) | st; }
static inline void c_markadded(uint32_t *c) { c[0] |= 8u; }
static inline int c_size(const uint32_t *c) { return (int)c[2]; }
static inline uint32_t *c_lits(uint32_t *c) { return c + 3; }
static inline const uint32_t *c_clits(const uint32_t *c) { return c + 3; }
static inline uint32_t c_lbd(const uint32_t *c) { return c[0] >> 6; }

static void *xrealloc(void *p, size_t n) { void *q = realloc(p, n ? n : 1); if (!q) abort(); return q; }

static void ol_push(ol_t *l, uint64_t r)
{
    if (l->sz == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 4;
        l->d = (uint64_t *)xrealloc(l->d, sizeof(uint64_t) * (size_t)l->cap);
    }
    l->d[l->sz++] = r;
}

static void res_push(pfo *s, uint32_t w)
{
    if (s->nres == s->rescap) {
        s->rescap = s->rescap ? s->rescap * 2 : 1024;
        s->resolved = (uint32_t *)xrealloc(s->resolved, sizeof(uint32_t) * s->rescap);
    }
    s->resolved[s->nres++] = w;
}

static void outc_push(pfo *s, uint32_t l)
{
    if (s->outn == s->outcap) {
        s->outcap = s->outcap ? s->outcap * 2 : 64;
        s->outc = (uint32_t *)xrealloc(s->outc, sizeof(uint32_t) * (size_t)s->outcap);
    }
    s->outc[s->outn++] = l;
}

/* append a clause image to the pool and scnf; returns its reference */
static uint64_t new_clause(pfo *s, const uint32_t *lits, int n, uint32_t hdr0)
{
    const uint64_t need = s->msz + 3 + (uint64_t)n;
    if (need > s->mcap) {
        while (need > s->mcap) s->mcap = s->mcap ? s->mcap * 2 : 1024;
        s->mem = (uint32_t *)xrealloc(s->mem, sizeof(uint32_t) * s->mcap);
    }
    const uint64_t r = s->msz;
    uint32_t *c = s->mem + r;
    c[0] = hdr0; c[1] = 0; c[2] = (uint32_t)n;
    if (n) memcpy(c + 3, lits, sizeof(uint32_t) * (size_t)n);
    s->msz = need;
    if (s->nscnf == s->scap) {
        s->scap = s->scap ? s->scap * 2 : 1024;
        s->scnf = (uint64_t *)xrealloc(s->scnf, sizeof(uint64_t) * s->scap);
    }
    s->scnf[s->nscnf++] = r;
    return r;
}

static void calc_sig(uint32_t *c)
{
    uint32_t sig = 0;
    for (int i = 0; i < c_size(c); i++) sig |= MAPHASH(c[3 + i]);
    c[1] = sig;
}

static void sort_u32(uint32_t *d, int n) /* ascending, distinct values */
{
    for (int i = 1; i < n; i++) {
        uint32_t t = d[i]; int j = i;
        while (j > 0 && d[j - 1] > t) { d[j] = d[j - 1]; j--; }
        d[j] = t;
    }
}

/* ------------------------------------------------------------------ */
/* sorting: sort.h:41-90 (insertion sort <= 20, std::sort otherwise)   */
/* ------------------------------------------------------------------ */
typedef int (*lessfn)(const void *ctx, uint64_t a, uint64_t b);

static void insertion_sort_sorth(uint64_t *d, int64_t n, lessfn lt, const void *ctx)
{   /* sort.h:41-54 */
    if (n == 2 && lt(ctx, d[1], d[0])) { uint64_t t = d[1]; d[1] = d[0]; d[0] = t; }
    else if (n > 2) {
        for (int64_t i = 1; i < n; i++) {
            uint64_t tmp = d[i]; int64_t j;
            for (j = i; j > 0 && lt(ctx, tmp, d[j - 1]); j--) d[j] = d[j - 1];
            d[j] = tmp;
        }
    }
}

/* --- libstdc++ std::sort (introsort) restated: the order it produces for ties --- */
static void is_adjust_heap(uint64_t *f, int64_t hole, int64_t len, uint64_t value, lessfn lt, const void *ctx)
{
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (lt(ctx, f[child], f[child - 1])) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    int64_t parent = (hole - 1) / 2;
    while (hole > top && lt(ctx, f[parent], value)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = value;
}

static void is_heapsort(uint64_t *f, int64_t len, lessfn lt, const void *ctx)
{
    if (len >= 2) {
        int64_t parent = (len - 2) / 2;
        for (;;) {
            is_adjust_heap(f, parent, len, f[parent], lt, ctx);
            if (parent == 0) break;
            parent--;
        }
    }
    while (len > 1) {
        len--;
        uint64_t v = f[len];
        f[len] = f[0];
        is_adjust_heap(f, 0, len, v, lt, ctx);
    }
}

static void is_move_median_first(uint64_t *res, uint64_t *a, uint64_t *b, uint64_t *c, lessfn lt, const void *ctx)
{
    uint64_t *pick;
    if (lt(ctx, *a, *b)) {
        if (lt(ctx, *b, *c)) pick = b;
        else if (lt(ctx, *a, *c)) pick = c;
        else pick = a;
    }
    else if (lt(ctx, *a, *c)) pick = a;
    else if (lt(ctx, *b, *c)) pick = c;
    else pick = b;
    uint64_t t = *res; *res = *pick; *pick = t;
}

static void is_loop(uint64_t *first, uint64_t *last, int64_t depth, lessfn lt, const void *ctx)
{
    while (last - first > 16) {
        if (depth == 0) { is_heapsort(first, last - first, lt, ctx); return; }
        --depth;
        uint64_t *mid = first + (last - first) / 2;
        is_move_median_first(first, first + 1, mid, last - 1, lt, ctx);
        uint64_t *lo = first + 1, *hi = last;
        for (;;) {
            while (lt(ctx, *lo, *first)) ++lo;
            --hi;
            while (lt(ctx, *first, *hi)) --hi;
            if (!(lo < hi)) break;
            uint64_t t = *lo; *lo = *hi; *hi = t;
            ++lo;
        }
        is_loop(lo, last, depth, lt, ctx);
        last = lo;
    }
}

static void is_linear_insert(uint64_t *last, lessfn lt, const void *ctx)
{
    uint64_t v = *last;
    uint64_t *next = last - 1;
    while (lt(ctx, v, *next)) { *last = *next; last = next; --next; }
    *last = v;
}

static void is_insertion(uint64_t *first, uint64_t *last, lessfn lt, const void *ctx)
{
    if (first == last) return;
    for (uint64_t *i = first + 1; i != last; ++i) {
        if (lt(ctx, *i, *first)) {
            uint64_t v = *i;
            memmove(first + 1, first, sizeof(uint64_t) * (size_t)(i - first));
            *first = v;
        }
        else is_linear_insert(i, lt, ctx);
    }
}

static void introsort(uint64_t *d, int64_t n, lessfn lt, const void *ctx)
{
    if (n < 2) return;
    int64_t lg = 63 - __builtin_clzll((unsigned long long)n);
    is_loop(d, d + n, 2 * lg, lt, ctx);
    if (n > 16) {
        is_insertion(d, d + 16, lt, ctx);
        for (uint64_t *i = d + 16; i != d + n; ++i) is_linear_insert(i, lt, ctx);
    }
    else is_insertion(d, d + n, lt, ctx);
}

static void stable_merge_sort(uint64_t *d, int64_t n, lessfn lt, const void *ctx)
{
    if (n <= 20) { insertion_sort_sorth(d, n, lt, ctx); return; }
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = lt(ctx, d[j], d[i]) ? d[j++] : d[i++];
            while (i < mid) tmp[k++] = d[i++];
            while (j < hi) tmp[k++] = d[j++];
        }
        memcpy(d, tmp, sizeof(uint64_t) * (size_t)n);
    }
    free(tmp);
}

static int u32key_less(const void *ctx, uint64_t a, uint64_t b)
{
    const uint32_t *k = (const uint32_t *)ctx;
    return k[a] < k[b];
}

void pfo_introsort_u32(uint32_t *idx, int64_t n, const uint32_t *key)
{
    uint64_t *t = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; i++) t[i] = idx[i];
    introsort(t, n, u32key_less, key);
    for (int64_t i = 0; i < n; i++) idx[i] = (uint32_t)t[i];
    free(t);
}

/* CNF_CMP_KEY (src/cpu/simplify.h:29-43): size, first literal, last literal, signature */
static int cnf_key_less(const void *ctx, uint64_t x, uint64_t y)
{
    const uint32_t *m = (const uint32_t *)ctx;
    const uint32_t *a = m + x, *b = m + y;
    const int xs = (int)a[2], ys = (int)b[2];
    if (xs < ys) return 1;
    if (xs > ys) return 0;
    if (a[3] < b[3]) return 1;
    if (a[3] > b[3]) return 0;
    const uint32_t xb = a[3 + xs - 1], yb = b[3 + ys - 1];
    if (xb < yb) return 1;
    if (xb > yb) return 0;
    return a[1] < b[1];
}

/* Sort(Vec, CMP) (sort.h:78-90) */
static void sort_list(pfo *s, ol_t *l)
{
    if (l->sz <= 1) return;
    if (l->sz <= 20) insertion_sort_sorth(l->d, l->sz, cnf_key_less, s->mem);
    else if (s->opts.stable_sort) stable_merge_sort(l->d, l->sz, cnf_key_less, s->mem);
    else introsort(l->d, l->sz, cnf_key_less, s->mem);
}

/* ------------------------------------------------------------------ */
/* construction                                                       */
/* ------------------------------------------------------------------ */
void pfo_default_opts(pfo_opts *o)
{   /* src/cpu/options.cpp:24-46, 135, with the adjustments of options.cpp:281-311 */
    o->phases = 5; o->mu_pos = 32; o->mu_neg = 32; o->lcve_min = 1; o->lits_min = 500;
    o->shrink_rate = 2; o->ve_clause_limit = 100; o->ve_lbound_en = 0; o->xor_max_arity = 10;
    o->sub_limit = 3000; o->ere_limit = 3000; o->ere_max_resolvent = 1000;
    o->sub_en = 1; o->ve_en = 1; o->ere_en = 1; o->lbd_tier1 = 2; o->stable_sort = 0;
    o->bce_en = 0; o->bce_limit = 3000; o->proof_en = 0; o->ve_fun_en = 0;
    o->lcve_max = 0; o->lcve_clause_max = 0;
}

pfo *pfo_create(uint32_t nvars, const pfo_opts *o, const uint32_t *lits, const uint64_t *offs,
                uint64_t ncls, const uint32_t *hdr0)
{
    pfo *s = (pfo *)calloc(1, sizeof(pfo));
    s->nvars = nvars;
    if (o) s->opts = *o; else pfo_default_opts(&s->opts);
    /* the option checks of the reference (src/cpu/options.cpp:305-310; ve_en stands
       for ve_en || ve_plus_en) */
    const int any = s->opts.ve_en || s->opts.sub_en || s->opts.bce_en;
    if (!s->opts.phases && any) s->opts.phases = 1;
    if (s->opts.phases && !any) s->opts.phases = 0;
    if (s->opts.phases > 1 && !s->opts.ve_en) s->opts.phases = 1;
    const size_t nlit = 2 * ((size_t)nvars + 1);
    s->ot = (ol_t *)calloc(nlit, sizeof(ol_t));
    s->ps = (uint32_t *)calloc((size_t)nvars + 1, sizeof(uint32_t));
    s->ns = (uint32_t *)calloc((size_t)nvars + 1, sizeof(uint32_t));
    s->value = (int8_t *)malloc(nlit);
    memset(s->value, -1, nlit);
    s->vstate = (uint8_t *)calloc((size_t)nvars + 1, 1);
    s->frozen = (uint8_t *)calloc((size_t)nvars + 1, 1);
    s->ifrozen = (uint8_t *)calloc((size_t)nvars + 1, 1);
    s->trail = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nvars + 1));
    s->pvs = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nvars + 1));
    s->eligible = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nvars + 1));
    s->scores = (uint32_t *)calloc((size_t)nvars + 1, sizeof(uint32_t));
    s->fstack = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nvars + 1));
    /* extract (src/cpu/simplify.cpp:95-112): copy, sort literals, signature */
    uint32_t *tmp = NULL; int tcap = 0;
    for (uint64_t i = 0; i < ncls; i++) {
        const int n = (int)(offs[i + 1] - offs[i]);
        if (n > tcap) { tcap = n; tmp = (uint32_t *)xrealloc(tmp, sizeof(uint32_t) * (size_t)n); }
        memcpy(tmp, lits + offs[i], sizeof(uint32_t) * (size_t)n);
        sort_u32(tmp, n);
        const uint64_t r = new_clause(s, tmp, n, hdr0 ? hdr0[i] : ST_ORIGINAL);
        calc_sig(CL(s, r));
        s->nclauses++;
        s->nliterals += n;
    }
    free(tmp);
    return s;
}

void pfo_destroy(pfo *s)
{
    if (!s) return;
    const size_t nlit = 2 * ((size_t)s->nvars + 1);
    for (size_t i = 0; i < nlit; i++) free(s->ot[i].d);
    free(s->ot); free(s->ps); free(s->ns); free(s->value); free(s->vstate); free(s->frozen); free(s->ifrozen);
    free(s->trail); free(s->pvs); free(s->eligible); free(s->scores); free(s->fstack);
    free(s->mem); free(s->scnf); free(s->resolved); free(s->outc); free(s->proof); free(s->vorg);
    free(s);
}

/* ------------------------------------------------------------------ */
/* binary DRAT proof (src/cpu/proof.cpp:92-232): a line is 'a' or 'd', */
/* each literal 2v+sign in original numbering (V2DEC(vorg[v], sign),   */
/* proof.cpp:118-135) as a little-endian base-128 varint (high bit =   */
/* more), then a zero byte.                                            */
/* The simplifier adds units (learnUnit, solve.h:475-479), resolvents, */
/* strengthened and substituted clauses unless a literal is already    */
/* true (addResolvent, proof.cpp:208-222), deletes subsumed clauses    */
/* (removeClause, sclause.cpp:77-84), and ends with the empty clause   */
/* at a contradiction (learnEmpty, solve.h:250-253), after which the   */
/* reference stops (killSolver).                                       */
/* ------------------------------------------------------------------ */
static void proof_byte(pfo *s, uint8_t b)
{
    if (s->nproof == s->proofcap) {
        s->proofcap = s->proofcap ? 2 * s->proofcap : 4096;
        s->proof = (uint8_t *)xrealloc(s->proof, (size_t)s->proofcap);
    }
    s->proof[s->nproof++] = b;
}

/* V2DEC(vorg[ABS(lit)], SIGN(lit)) (model.cpp:263-280, proof.cpp:123) */
static inline uint32_t org_lit(const pfo *s, uint32_t lit)
{
    return s->vorg ? ((s->vorg[ABSV(lit)] << 1) | (lit & 1u)) : lit;
}

/* the variable mapping of the caller (identity when never set): model records and
 * proof lines are written in original numbering through it */
int pfo_set_vorg(pfo *s, const uint32_t *vorg, uint32_t n)
{
    if (n != s->nvars + 1) return -1;
    for (uint32_t v = 1; v <= s->nvars; v++)
        if (!vorg[v]) return -1;
    s->vorg = (uint32_t *)xrealloc(s->vorg, sizeof(uint32_t) * n);
    memcpy(s->vorg, vorg, sizeof(uint32_t) * n);
    return 0;
}

static void proof_line(pfo *s, uint8_t tag, const uint32_t *lits, int n)
{
    if (!s->opts.proof_en || s->proof_closed) return;
    proof_byte(s, tag);
    for (int i = 0; i < n; i++) {
        uint32_t r = org_lit(s, lits[i]);
        while (r > 127u) { proof_byte(s, (uint8_t)((r & 127u) | 128u)); r >>= 7; }
        proof_byte(s, (uint8_t)r);
    }
    proof_byte(s, 0);
}

static void proof_resolvent(pfo *s, const uint32_t *c)
{   /* PROOF::addResolvent: nothing for a clause with a true literal */
    for (int i = 0; i < c_size(c); i++)
        if (s->value[c[3 + i]] > 0) return;
    proof_line(s, 'a', c + 3, c_size(c));
}

static void learn_empty(pfo *s)
{   /* learnEmpty (solve.h:250-253); nothing is logged after it */
    proof_line(s, 'a', NULL, 0);
    s->proof_closed = 1;
    s->unsat = 1;
}

/* enqueueUnit (src/cpu/solve.h:430-441): marks the variable frozen */
static void assign_unit(pfo *s, uint32_t lit)
{
    s->vstate[ABSV(lit)] = FROZEN_M; /* learnUnit -> markFrozen (solve.h:475-479, 529-533) */
    s->value[lit] = 1;
    s->value[FLIP(lit)] = 0;
    s->trail[s->ntrail++] = lit;
}

static void enqueue_unit(pfo *s, uint32_t lit)
{
    proof_line(s, 'a', &lit, 1);   /* learnUnit -> addUnit */
    assign_unit(s, lit);
}

/* ifreeze / iunfreeze (src/cpu/assume.cpp:45-61); iassume marks its variables the
 * same way (assume.cpp:68-84). In incremental mode iassumed(v) (solve.h:864-867) makes
 * LCVE pass over the variable (lcve.cpp:57): it is never elected and never freezes. */
int pfo_ifreeze(pfo *s, uint32_t v, int on)
{
    if (v == 0 || v > s->nvars) return -1;
    s->ifrozen[v] = on ? 1 : 0;
    return 0;
}

/* root units handed over with the formula: the caller's, not logged by the simplifier */
int pfo_enqueue(pfo *s, uint32_t lit)
{
    if (s->value[lit] >= 0) return s->value[lit] ? 0 : -1;
    assign_unit(s, lit);
    return 0;
}

/* ------------------------------------------------------------------ */
/* model stack (src/cpu/model.cpp:263-307): literals in original       */
/* numbering through vorg (saveWitness / saveLiteral)                  */
/* ------------------------------------------------------------------ */
static void save_witness(pfo *s, uint32_t w) { res_push(s, org_lit(s, w)); res_push(s, 1); }

static void save_clause(pfo *s, const uint32_t *c, uint32_t witlit)
{
    const int size = c_size(c);
    const uint64_t last = s->nres;
    int pos = -1;
    for (int i = 0; i < size; i++) {
        const uint32_t lit = c[3 + i];
        res_push(s, org_lit(s, lit));
        if (lit == witlit) pos = i;
    }
    if (pos > 0) {
        uint32_t t = s->resolved[last + (uint64_t)pos];
        s->resolved[last + (uint64_t)pos] = s->resolved[last];
        s->resolved[last] = t;
    }
    res_push(s, (uint32_t)size);
}

void pfo_extend_model(const pfo *s, uint8_t *value)
{   /* MODEL::extend replay loop, src/cpu/model.cpp:122-139 */
    if (!s->nres) return;
    const uint32_t *base = s->resolved;
    const uint32_t *x = base + s->nres - 1;
    while (x > base) {
        int unsat = 1;
        uint32_t k;
        for (k = *x--; k > 1; k--, x--) {
            const uint32_t l = *x;
            if (value[ABSV(l)] == (uint8_t)!(l & 1u)) { unsat = 0; break; }
        }
        if (unsat) value[ABSV(*x)] = (uint8_t)!(*x & 1u);
        x -= k;
    }
}

/* ------------------------------------------------------------------ */
/* occurrence table                                                   */
/* ------------------------------------------------------------------ */
void pfo_create_ot(pfo *s)
{   /* createOT (src/cpu/simplify.cpp:34-53) */
    for (uint32_t v = 1; v <= s->nvars; v++) { s->ot[2 * v].sz = 0; s->ot[2 * v + 1].sz = 0; }
    for (uint64_t i = 0; i < s->nscnf; i++) {
        const uint64_t r = s->scnf[i];
        const uint32_t *c = CL(s, r);
        if (c_learnt(c) || c_orig(c)) {
            for (int k = 0; k < c_size(c); k++) ol_push(&s->ot[c[3 + k]], r);
        }
    }
}

static void reduce_ol(pfo *s, ol_t *ol)
{   /* reduceOL (src/cpu/simplify.cpp:55-65) */
    int32_t j = 0;
    for (int32_t i = 0; i < ol->sz; i++) if (!c_del(CL(s, ol->d[i]))) ol->d[j++] = ol->d[i];
    ol->sz = j;
}

static void toblivion_list(pfo *s, ol_t *l)
{   /* toblivion(OL&) (src/cpu/simplify.h:291-297) */
    for (int32_t i = 0; i < l->sz; i++) c_markdel(CL(s, l->d[i]));
    l->sz = 0;
}

static int prop_clause(pfo *s, uint32_t lit, uint32_t *c)
{   /* propClause (src/cpu/elimination.cpp:41-62) */
    uint32_t sig = 0;
    uint32_t *lits = c_lits(c);
    int j = 0;
    const int n = c_size(c);
    for (int i = 0; i < n; i++) {
        const uint32_t other = lits[i];
        if (other != lit) {
            if (s->value[other] > 0) return 1;
            lits[j++] = other;
            sig |= MAPHASH(other);
        }
    }
    c[1] = sig;
    c[2] = (uint32_t)(n - 1);
    return 0;
}

int pfo_prop(pfo *s)
{   /* prop (src/cpu/elimination.cpp:64-97) */
    const uint32_t before = s->propagated;
    while (s->propagated < s->ntrail) {
        const uint32_t assign = s->trail[s->propagated++], f = FLIP(assign);
        ol_t *ol = &s->ot[f];
        for (int32_t i = 0; i < ol->sz; i++) {
            uint32_t *c = CL(s, ol->d[i]);
            if (c_del(c)) continue;
            if (prop_clause(s, f, c)) c_markdel(c);
            else {
                const int size = c_size(c);
                if (!size) { learn_empty(s); return 0; }
                if (size == 1) {
                    const uint32_t unit = c[3];
                    if (s->value[unit] < 0) enqueue_unit(s, unit);
                    else { learn_empty(s); return 0; }
                }
            }
        }
        ol->sz = 0;
        toblivion_list(s, &s->ot[assign]);
    }
    if (s->propagated != before && !s->opts.sub_en) {
        for (uint32_t v = 1; v <= s->nvars; v++) { reduce_ol(s, &s->ot[2 * v]); reduce_ol(s, &s->ot[2 * v + 1]); }
    }
    return 1;
}

void pfo_hist(pfo *s, uint32_t *hist)
{   /* histSimp (src/cpu/histogram.cpp:80-95), per literal */
    memset(hist, 0, sizeof(uint32_t) * 2 * ((size_t)s->nvars + 1));
    for (uint64_t i = 0; i < s->nscnf; i++) {
        const uint32_t *c = CL(s, s->scnf[i]);
        if (c_del(c)) continue;
        for (int k = 0; k < c_size(c); k++) hist[c[3 + k]]++;
    }
}

/* ------------------------------------------------------------------ */
/* LCVE (src/cpu/lcve.cpp:24-99)                                       */
/* ------------------------------------------------------------------ */
static void var_reorder(pfo *s)
{   /* varReorder (lcve.cpp:24-44): scores = ps*ns (uint32), order (score, var) */
    const uint32_t nv = s->nvars;
    memset(s->ps, 0, sizeof(uint32_t) * ((size_t)nv + 1));
    memset(s->ns, 0, sizeof(uint32_t) * ((size_t)nv + 1));
    for (uint64_t i = 0; i < s->nscnf; i++) {
        const uint32_t *c = CL(s, s->scnf[i]);
        if (c_del(c)) continue;
        for (int k = 0; k < c_size(c); k++) {
            const uint32_t l = c[3 + k];
            if (l & 1u) s->ns[ABSV(l)]++; else s->ps[ABSV(l)]++;
        }
    }
    for (uint32_t v = 1; v <= nv; v++) { s->eligible[v - 1] = v; s->scores[v] = s->ps[v] * s->ns[v]; }
    /* stable LSD radix by score == total order (score, var) (radixsort.h:38-88, key.h:57-66) */
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nv + 1));
    uint32_t *a = s->eligible, *b = tmp;
    for (int sh = 0; sh < 32; sh += 8) {
        size_t cnt[257] = {0};
        for (uint32_t i = 0; i < nv; i++) cnt[((s->scores[a[i]] >> sh) & 255u) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < nv; i++) b[cnt[(s->scores[a[i]] >> sh) & 255u]++] = a[i];
        uint32_t *t = a; a = b; b = t;
    }
    /* 4 passes: result back in s->eligible */
    free(tmp);
}

static void dep_freeze(pfo *s, ol_t *ol, uint32_t cand, uint32_t pt, uint32_t nt)
{   /* depFreeze (lcve.cpp:83-99) */
    for (int32_t i = 0; i < ol->sz; i++) {
        const uint32_t *c = CL(s, ol->d[i]);
        if (c_del(c)) continue;
        for (int k = 0; k < c_size(c); k++) {
            const uint32_t v = ABSV(c[3 + k]);
            if (!s->frozen[v] && v != cand && (s->ps[v] < pt || s->ns[v] < nt)) {
                s->frozen[v] = 1;
                s->fstack[s->nfstack++] = v;
            }
        }
    }
}

static int has_long_clause(const pfo *s, const ol_t *ol, uint32_t limit)
{   /* the size test of the GPU reference's depFreeze (src/gpu/lcve.cu:105-120) */
    for (int32_t i = 0; i < ol->sz; i++) {
        const uint32_t *c = CL(s, ol->d[i]);
        if (!c_del(c) && (uint32_t)c_size(c) > limit) return 1;
    }
    return 0;
}

int pfo_lcve(pfo *s)
{   /* LCVE (lcve.cpp:46-81), with the GPU reference's two optional limits
       (src/gpu/lcve.cu:68, 105-127; both off by default) */
    var_reorder(s);
    s->npvs = 0;
    s->nfstack = 0;
    const uint32_t pt = s->opts.mu_pos << s->mu_inc, nt = s->opts.mu_neg << s->mu_inc;
    for (uint32_t i = 0; i < s->nvars; i++) {
        const uint32_t cand = s->eligible[i];
        if (s->ifrozen[cand]) continue;
        if (s->vstate[cand]) continue;
        if (s->frozen[cand]) continue;
        const uint32_t p = 2 * cand, n = p + 1;
        const uint32_t psz = (uint32_t)s->ot[p].sz, nsz = (uint32_t)s->ot[n].sz;
        if (s->ps[cand] == 0 && s->ns[cand] == 0) continue;
        if (s->opts.lcve_max && (s->ps[cand] > s->opts.lcve_max || s->ns[cand] > s->opts.lcve_max)) break;
        if (psz >= pt && nsz >= nt) break;
        if (s->opts.lcve_clause_max) {
            /* depFreeze(cand, ot[p]) && depFreeze(cand, ot[n]): a failing list's own
               freezes are undone, the positive list's stay when only the negative fails */
            if (has_long_clause(s, &s->ot[p], s->opts.lcve_clause_max)) continue;
            if (has_long_clause(s, &s->ot[n], s->opts.lcve_clause_max)) {
                dep_freeze(s, &s->ot[p], cand, pt, nt);
                continue;
            }
        }
        s->pvs[s->npvs++] = cand;
        dep_freeze(s, &s->ot[p], cand, pt, nt);
        dep_freeze(s, &s->ot[n], cand, pt, nt);
    }
    for (uint32_t i = 0; i < s->nfstack; i++) s->frozen[s->fstack[i]] = 0;
    s->nfstack = 0;
    return s->npvs >= s->opts.lcve_min;
}

void pfo_sort_ot(pfo *s)
{   /* sortOT (src/cpu/simplify.cpp:79-93) */
    for (uint32_t i = 0; i < s->npvs; i++) {
        const uint32_t v = s->pvs[i];
        sort_list(s, &s->ot[2 * v]);
        sort_list(s, &s->ot[2 * v + 1]);
    }
}

/* ------------------------------------------------------------------ */
/* SUB with self-subsuming strengthening (src/cpu/subsume.h)           */
/* ------------------------------------------------------------------ */
static inline int sig_sub(uint32_t A, uint32_t B) { return !(A & ~B); }            /* simplify.h:81 */
static inline int sig_selfsub(uint32_t A, uint32_t B)                               /* simplify.h:83-87 */
{
    const uint32_t t = B | ((B & 0xAAAAAAAAu) >> 1) | ((B & 0x55555555u) << 1);
    return !(A & ~t);
}

static int sub_cl(const uint32_t *sm, const uint32_t *lr)
{   /* sub(SCLAUSE&, SCLAUSE&) (subsume.h:39-55) */
    int it1 = 0, it2 = 0, sub = 0;
    const int n1 = c_size(sm), n2 = c_size(lr);
    const uint32_t *a = c_clits(sm), *b = c_clits(lr);
    while (it1 < n1 && it2 < n2) {
        if (a[it1] < b[it2]) it1++;
        else if (b[it2] < a[it1]) it2++;
        else { sub++; it1++; it2++; }
    }
    return sub == n1;
}

static int selfsub_cl(uint32_t x, uint32_t fx, const uint32_t *sm, const uint32_t *lr)
{   /* selfsub (subsume.h:72-99) */
    int it1 = 0, it2 = 0, sub = 0, self = 0;
    const int n1 = c_size(sm), n2 = c_size(lr);
    const uint32_t *a = c_clits(sm), *b = c_clits(lr);
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2];
        if (l1 == fx) it1++;
        else if (l2 == x) { self = 1; it2++; }
        else if (l1 < l2) it1++;
        else if (l2 < l1) it2++;
        else { sub++; it1++; it2++; }
    }
    if (sub + 1 == n1) {
        if (self) return 1;
        while (it2 < n2) { if (b[it2] == x) return 1; it2++; }
    }
    return 0;
}

static void bump_shrunken(pfo *s, uint32_t *c)
{   /* bumpShrunken (src/cpu/elimination.cpp:27-39) */
    const uint32_t old = c_lbd(c);
    if ((int)old <= s->opts.lbd_tier1) return;
    uint32_t nl = (uint32_t)(c_size(c) - 1) < old ? (uint32_t)(c_size(c) - 1) : old;
    if (nl >= old) return;
    c[0] = (c[0] & 0x3Fu) | (nl << 6);
    c[0] = (c[0] & ~0x30u) | (USAGET3 << 4);
}

static void strengthen(pfo *s, uint32_t *c, uint32_t me)
{   /* strengthen (src/cpu/elimination.cpp:199-231) */
    uint32_t sig = 0;
    int n = 0;
    uint32_t *lits = c_lits(c);
    const int sz = c_size(c);
    for (int k = 0; k < sz; k++) {
        const uint32_t lit = lits[k];
        if (lit != me) { lits[n++] = lit; sig |= MAPHASH(lit); }
    }
    c[1] = sig;
    c[2] = (uint32_t)(sz - 1);
    if (n == 1) {
        const uint32_t unit = lits[0];
        const int8_t val = s->value[unit];
        if (val < 0) {
            enqueue_unit(s, unit);
            toblivion_list(s, &s->ot[unit]);
        }
        else if (!val) learn_empty(s);
    }
    else {
        proof_resolvent(s, c);
        if (c_learnt(c)) bump_shrunken(s, c);
    }
}

static void update_ol(pfo *s, ol_t *ol)
{   /* updateOL (subsume.h:27-37) */
    int32_t j = 0;
    for (int32_t i = 0; i < ol->sz; i++) {
        uint32_t *c = CL(s, ol->d[i]);
        if (c_molten(c)) c_freeze(c);
        else if (!c_del(c)) ol->d[j++] = ol->d[i];
    }
    ol->sz = j;
}

static int subsume_l(pfo *s, ol_t *list, int32_t end, uint32_t *cand)
{   /* subsume (subsume.h:101-117) */
    const int candsz = c_size(cand);
    for (int32_t j = 0; j < end; j++) {
        uint32_t *sg = CL(s, list->d[j]);
        if (c_del(sg)) continue;
        if (c_molten(cand) && c_size(sg) > candsz) continue;
        if (c_size(sg) > 1 && sig_sub(sg[1], cand[1]) && sub_cl(sg, cand)) {
            if (c_learnt(sg) && c_orig(cand)) c_setst(sg, ST_ORIGINAL);
            if (!c_del(cand)) {   /* removeClause (sclause.cpp:77-84) */
                c_markdel(cand);
                proof_line(s, 'd', cand + 3, c_size(cand));
            }
            return 1;
        }
    }
    return 0;
}

static int selfsubsume_l(pfo *s, uint32_t x, uint32_t fx, ol_t *list, uint32_t *cand)
{   /* selfsubsume (subsume.h:119-138) */
    const int candsz = c_size(cand);
    const uint32_t candsig = cand[1];
    for (int32_t j = 0; j < list->sz; j++) {
        uint32_t *sg = CL(s, list->d[j]);
        const int subsize = c_size(sg);
        if (subsize > candsz) break;
        if (c_del(sg)) continue;
        if (subsize > 1 && sig_selfsub(sg[1], candsig) && selfsub_cl(x, fx, sg, cand)) {
            strengthen(s, cand, x);
            c_melt(cand);
            return 1;
        }
    }
    return 0;
}

static void self_sub_x(pfo *s, uint32_t p)
{   /* self_sub_x (subsume.h:140-176) */
    const uint32_t n = FLIP(p);
    ol_t *poss = &s->ot[p], *negs = &s->ot[n];
    for (int32_t i = 0; i < poss->sz; i++) {
        uint32_t *pos = CL(s, poss->d[i]);
        if (c_size(pos) > HSE_MAX_CL_SIZE) break;
        if (c_del(pos)) continue;
        selfsubsume_l(s, p, n, negs, pos);
        subsume_l(s, poss, i, pos);
    }
    update_ol(s, poss);
    for (int32_t i = 0; i < negs->sz; i++) {
        uint32_t *neg = CL(s, negs->d[i]);
        if (c_size(neg) > HSE_MAX_CL_SIZE) break;
        if (c_del(neg)) continue;
        selfsubsume_l(s, n, p, poss, neg);
        subsume_l(s, negs, i, neg);
    }
    update_ol(s, negs);
}

void pfo_sub(pfo *s)
{   /* SUB (src/cpu/elimination.cpp:108-130); sub_en || ve_plus_en */
    for (uint32_t i = 0; i < s->npvs; i++) {
        const uint32_t v = s->pvs[i];
        const uint32_t p = 2 * v, n = p + 1;
        if (s->ot[p].sz <= s->opts.sub_limit && s->ot[n].sz <= s->opts.sub_limit) self_sub_x(s, p);
    }
}

/* ------------------------------------------------------------------ */
/* BVE (src/cpu/bve.cpp, simplify.h, and.h, equivalence.h, ifthenelse.h, xor.h) */
/* ------------------------------------------------------------------ */
static int merge_len(uint32_t x, const uint32_t *c1, const uint32_t *c2)
{   /* merge(x, c1, c2) -> length, 0 on tautology (simplify.h:168-193) */
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c_clits(c1), *b = c_clits(c2);
    int it1 = 0, it2 = 0, len = n1 + n2 - 2;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = ABSV(l1), v2 = ABSV(l2);
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if ((l1 ^ l2) == 1u) return 0;
        else if (v1 < v2) it1++;
        else if (v2 < v1) it2++;
        else { it1++; it2++; len--; }
    }
    return len;
}

static int merge_out(pfo *s, uint32_t x, const uint32_t *c1, const uint32_t *c2)
{   /* merge(x, c1, c2, out_c) (simplify.h:131-166) */
    s->outn = 0;
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c_clits(c1), *b = c_clits(c2);
    int it1 = 0, it2 = 0;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = ABSV(l1), v2 = ABSV(l2);
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if ((l1 ^ l2) == 1u) return 0;
        else if (v1 < v2) { it1++; outc_push(s, l1); }
        else if (v2 < v1) { it2++; outc_push(s, l2); }
        else { it1++; it2++; outc_push(s, l1); }
    }
    while (it1 < n1) { const uint32_t l1 = a[it1++]; if (ABSV(l1) != x) outc_push(s, l1); }
    while (it2 < n2) { const uint32_t l2 = b[it2++]; if (ABSV(l2) != x) outc_push(s, l2); }
    return 1;
}

static int count_subst(pfo *s, uint32_t x, int clsbefore, ol_t *me, ol_t *other, int *nAdded, int molten_rule)
{   /* countSubstituted / countResolvents (simplify.h:227-287) */
    int nAddedLits = 0;
    const int rlimit = s->opts.ve_clause_limit;
    for (int32_t i = 0; i < me->sz; i++) {
        const uint32_t *ci = CL(s, me->d[i]);
        if (!c_orig(ci)) continue;
        const int a = c_molten(ci);
        for (int32_t j = 0; j < other->sz; j++) {
            const uint32_t *cj = CL(s, other->d[j]);
            if (!c_orig(cj)) continue;
            if (molten_rule == 1 && a == c_molten(cj)) continue;
            if (molten_rule == 2 && a && c_molten(cj)) continue;
            int rsize;
            if ((rsize = merge_len(x, ci, cj)) > 1) {
                if (++(*nAdded) > clsbefore || (rlimit && rsize > rlimit)) return 1;
                nAddedLits += rsize;
            }
        }
    }
    if (s->opts.ve_lbound_en) {
        int before = 0;
        for (int32_t i = 0; i < me->sz; i++) { const uint32_t *c = CL(s, me->d[i]); if (c_orig(c)) before += c_size(c); }
        for (int32_t i = 0; i < other->sz; i++) { const uint32_t *c = CL(s, other->d[i]); if (c_orig(c)) before += c_size(c); }
        if (nAddedLits > before) return 1;
    }
    return 0;
}

static void freeze_binaries(pfo *s, ol_t *l)
{   /* simplify.h:195-202 */
    for (int32_t i = 0; i < l->sz; i++) { uint32_t *c = CL(s, l->d[i]); if (c_orig(c) && c_size(c) == 2) c_freeze(c); }
}

static int count_orgs(pfo *s, ol_t *l)
{
    int n = 0;
    for (int32_t i = 0; i < l->sz; i++) if (c_orig(CL(s, l->d[i]))) n++;
    return n;
}

static void toblivion_var(pfo *s, uint32_t p, int pOrgs, int nOrgs)
{   /* toblivion(p, ...) (simplify.h:299-331) */
    const uint32_t n = FLIP(p);
    ol_t *poss = &s->ot[p], *negs = &s->ot[n];
    if (pOrgs > nOrgs) {
        for (int32_t i = 0; i < negs->sz; i++) {
            uint32_t *c = CL(s, negs->d[i]);
            if (c_orig(c)) save_clause(s, c, n);
            c_markdel(c);
        }
        save_witness(s, p);
        negs->sz = 0;
        toblivion_list(s, poss);
    }
    else {
        for (int32_t i = 0; i < poss->sz; i++) {
            uint32_t *c = CL(s, poss->d[i]);
            if (c_orig(c)) save_clause(s, c, p);
            c_markdel(c);
        }
        save_witness(s, n);
        poss->sz = 0;
        toblivion_list(s, negs);
    }
}

/* ---- equivalence / inverter gates (equivalence.h) ---- */
static uint32_t find_sfanin(pfo *s, uint32_t gate_out, ol_t *list)
{   /* find_sfanin (equivalence.h:88-106) */
    uint32_t imp = 0;
    int nImps = 0;
    for (int32_t i = 0; i < list->sz; i++) {
        uint32_t *c = CL(s, list->d[i]);
        if (c_orig(c) && c_size(c) == 2) {
            imp = FLIP(c[3] ^ c[4] ^ gate_out);
            c_melt(c);
            nImps++;
        }
        if (nImps > 1) return 0;
    }
    return imp;
}

static uint32_t find_bn_gate(pfo *s, uint32_t p)
{   /* find_BN_gate (equivalence.h:108-130) */
    const uint32_t n = FLIP(p);
    ol_t *poss = &s->ot[p], *negs = &s->ot[n];
    uint32_t first = find_sfanin(s, p, poss);
    if (first) {
        uint32_t second = n, def = first;
        if (second < first) { first = second; second = def; }
        for (int32_t i = 0; i < negs->sz; i++) {
            uint32_t *c = CL(s, negs->d[i]);
            if (c_orig(c) && c_size(c) == 2 && c[3] == first && c[4] == second) {
                c_melt(c);
                return def;
            }
        }
    }
    freeze_binaries(s, poss);
    return 0;
}

static void save_bn_gate(pfo *s, uint32_t p, int pOrgs, int nOrgs)
{   /* save_BN_gate (equivalence.h:132-153) */
    const uint32_t n = FLIP(p);
    ol_t *poss = &s->ot[p], *negs = &s->ot[n];
    if (pOrgs > nOrgs) {
        for (int32_t i = 0; i < negs->sz; i++) { uint32_t *c = CL(s, negs->d[i]); if (c_orig(c)) save_clause(s, c, n); }
        save_witness(s, p);
    }
    else {
        for (int32_t i = 0; i < poss->sz; i++) { uint32_t *c = CL(s, poss->d[i]); if (c_orig(c)) save_clause(s, c, p); }
        save_witness(s, n);
    }
}

static int c_has(const uint32_t *c, uint32_t lit)
{   /* SCLAUSE::has (sclause.h:139-159) on sorted literals */
    for (int i = 0; i < c_size(c); i++) if (c[3 + i] == lit) return 1;
    return 0;
}

static uint32_t substitute_one(uint32_t dx, uint32_t *org, uint32_t def)
{   /* substitute_single(dx, org, def) (equivalence.h:24-40) */
    uint32_t *lits = c_lits(org);
    const int sz = c_size(org);
    int n = 0;
    for (int i = 0; i < sz; i++) {
        const uint32_t lit = lits[i];
        if (lit == dx) lits[n++] = def;
        else if (lit != def) lits[n++] = lit;
    }
    org[2] = (uint32_t)n;
    if (n == 1) return lits[0];
    sort_u32(lits, n);
    calc_sig(org);
    return 0;
}

static int substitute_single(pfo *s, uint32_t p, uint32_t def)
{   /* substitute_single(p, def, ot) (equivalence.h:42-86); returns 1 on conflict */
    const uint32_t n = FLIP(p), def_f = FLIP(def);
    ol_t *poss = &s->ot[p], *negs = &s->ot[n];
    for (int32_t i = 0; i < negs->sz; i++) {
        uint32_t *neg = CL(s, negs->d[i]);
        if (c_learnt(neg) || c_molten(neg) || c_has(neg, def)) c_markdel(neg);
        else if (c_orig(neg)) {
            const uint32_t unit = substitute_one(n, neg, def_f);
            if (unit) {
                const int8_t val = s->value[unit];
                if (val < 0) enqueue_unit(s, unit);
                else if (!val) return 1;
            }
            else proof_resolvent(s, neg);
        }
    }
    for (int32_t i = 0; i < poss->sz; i++) {
        uint32_t *pos = CL(s, poss->d[i]);
        if (c_learnt(pos) || c_molten(pos) || c_has(pos, def_f)) c_markdel(pos);
        else if (c_orig(pos)) {
            const uint32_t unit = substitute_one(p, pos, def);
            if (unit) {
                const int8_t val = s->value[unit];
                if (val < 0) enqueue_unit(s, unit);
                else if (!val) return 1;
            }
            else proof_resolvent(s, pos);
        }
    }
    return 0;
}

/* ---- AND/OR gates (and.h) ---- */
static int find_ao_gate(pfo *s, uint32_t dx, int nOrgCls, int *nAdded)
{   /* find_AO_gate (and.h:43-85) */
    s->outn = 0;
    const uint32_t x = ABSV(dx);
    ol_t *itarget = &s->ot[dx];
    uint32_t sig = 0;
    for (int32_t i = 0; i < itarget->sz; i++) {   /* find_fanin (and.h:25-41) */
        uint32_t *c = CL(s, itarget->d[i]);
        if (c_orig(c) && c_size(c) == 2) {
            const uint32_t imp = FLIP(c[3] ^ c[4] ^ dx);
            outc_push(s, imp);
            sig |= MAPHASH(imp);
            c_melt(c);
        }
    }
    if (s->outn > 1) {
        const uint32_t f_dx = FLIP(dx);
        outc_push(s, f_dx);
        sig |= MAPHASH(f_dx);
        sort_u32(s->outc, s->outn);
        ol_t *otarget = &s->ot[f_dx];
        for (int32_t i = 0; i < otarget->sz; i++) {
            uint32_t *c = CL(s, otarget->d[i]);
            if (c_orig(c) && c_size(c) == s->outn && sig_sub(c[1], sig)) {
                int eq = 1;
                for (int k = 0; k < s->outn; k++) if (c[3 + k] != s->outc[k]) { eq = 0; break; }
                if (!eq) continue;
                c_melt(c);
                *nAdded = 0;
                if (count_subst(s, x, nOrgCls, itarget, otarget, nAdded, 1)) { c_freeze(c); break; }
                return 1;
            }
        }
    }
    freeze_binaries(s, itarget);
    return 0;
}

/* ---- ITE gates (ifthenelse.h) ---- */
static uint32_t *fast_equality_check(pfo *s, uint32_t x, uint32_t y, uint32_t z)
{   /* fast_equality_check (ifthenelse.h:24-40) */
    uint32_t t;
    if (s->ot[y].sz > s->ot[z].sz) { t = y; y = z; z = t; }
    if (s->ot[x].sz > s->ot[y].sz) { t = x; x = y; y = t; }
    ol_t *list = &s->ot[x];
    /* sort3 (simplify.h:107-119) */
    if (y > z) { t = y; y = z; z = t; }
    if (x > z) { t = x; x = z; z = t; }
    if (x > y) { t = x; x = y; y = t; }
    for (int32_t i = 0; i < list->sz; i++) {
        uint32_t *c = CL(s, list->d[i]);
        if (c_molten(c)) continue;
        if (c_orig(c) && c_size(c) == 3 && c[3] == x && c[4] == y && c[5] == z) return c;
    }
    return NULL;
}

static int find_ite_gate(pfo *s, uint32_t dx, int nOrgCls, int *nAdded)
{   /* find_ITE_gate (ifthenelse.h:42-95) */
    const uint32_t fx = FLIP(dx);
    ol_t *itarget = &s->ot[dx];
    for (int32_t i = 0; i < itarget->sz; i++) {
        uint32_t *ci = CL(s, itarget->d[i]);
        if (!(c_orig(ci) && c_size(ci) == 3)) continue;
        uint32_t xi = ci[3], yi = ci[4], zi = ci[5], t;
        if (yi == dx) { t = xi; xi = yi; yi = t; }
        if (zi == dx) { t = xi; xi = zi; zi = t; }
        for (int32_t j = i + 1; j < itarget->sz; j++) {
            uint32_t *cj = CL(s, itarget->d[j]);
            if (!(c_orig(cj) && c_size(cj) == 3)) continue;
            uint32_t xj = cj[3], yj = cj[4], zj = cj[5];
            if (yj == dx) { t = xj; xj = yj; yj = t; }
            if (zj == dx) { t = xj; xj = zj; zj = t; }
            if (ABSV(yi) == ABSV(zj)) { t = yj; yj = zj; zj = t; }
            if (ABSV(zi) == ABSV(zj)) continue;
            if (yi != FLIP(yj)) continue;
            uint32_t *d1 = fast_equality_check(s, fx, yi, FLIP(zi));
            if (!d1) continue;
            uint32_t *d2 = fast_equality_check(s, fx, yj, FLIP(zj));
            if (!d2) continue;
            c_melt(ci); c_melt(cj); c_melt(d1); c_melt(d2);
            *nAdded = 0;
            if (count_subst(s, ABSV(dx), nOrgCls, itarget, &s->ot[fx], nAdded, 1)) {
                c_freeze(ci); c_freeze(cj); c_freeze(d1); c_freeze(d2);
                return 0;
            }
            return 1;
        }
    }
    return 0;
}

/* ---- XOR gates (xor.h) ---- */
static void freeze_arities(pfo *s, ol_t *me, ol_t *other)
{   /* xor.h:30-44 */
    for (int32_t i = 0; i < me->sz; i++) { uint32_t *c = CL(s, me->d[i]); if (c_size(c) > 2 && c_molten(c)) c_freeze(c); }
    for (int32_t i = 0; i < other->sz; i++) { uint32_t *c = CL(s, other->d[i]); if (c_size(c) > 2 && c_molten(c)) c_freeze(c); }
}

static int check_arity(const uint32_t *c, const uint32_t *lits, int size)
{   /* xor.h:54-68 */
    for (int k = 0; k < c_size(c); k++) {
        const uint32_t lit = c[3 + k];
        int j;
        for (j = 0; j < size; j++) if (lit == lits[j]) break;
        if (j == size) return 0;
    }
    return 1;
}

static int make_arity(pfo *s, uint32_t *parity, uint32_t *lits, int size)
{   /* makeArity (xor.h:70-99) */
    const uint32_t old = *parity;
    while (__builtin_parity(++(*parity)));
    for (int k = 0; k < size; k++) {
        const uint32_t bit = 1u << k;
        if ((*parity & bit) != (old & bit)) lits[k] = FLIP(lits[k]);
    }
    uint32_t best = lits[0];
    int32_t minsize = s->ot[best].sz;
    for (int k = 1; k < size; k++) {
        const int32_t lsize = s->ot[lits[k]].sz;
        if (lsize < minsize) { minsize = lsize; best = lits[k]; }
    }
    ol_t *l = &s->ot[best];
    for (int32_t i = 0; i < l->sz; i++) {
        uint32_t *c = CL(s, l->d[i]);
        if (c_orig(c) && c_size(c) == size && check_arity(c, lits, size)) { c_melt(c); return 1; }
    }
    return 0;
}

static int find_xor_gate(pfo *s, uint32_t dx, int nOrgCls, int *nAdded)
{   /* find_XOR_gate (xor.h:101-147) */
    const uint32_t fx = FLIP(dx), v = ABSV(dx);
    ol_t *itarget = &s->ot[dx], *otarget = &s->ot[fx];
    for (int32_t i = 0; i < itarget->sz; i++) {
        uint32_t *ci = CL(s, itarget->d[i]);
        if (!c_orig(ci)) continue;
        const int size = c_size(ci), arity = size - 1;
        if (size < 3 || arity > s->opts.xor_max_arity) continue;
        s->outn = 0;
        for (int k = 0; k < size; k++) outc_push(s, ci[3 + k]);
        uint32_t parity = 0;
        int itargets = 1 << arity;
        while (--itargets && make_arity(s, &parity, s->outc, size));
        if (itargets) freeze_arities(s, itarget, otarget);
        else {
            c_melt(ci);
            *nAdded = 0;
            if (count_subst(s, v, nOrgCls, itarget, otarget, nAdded, 1)) { freeze_arities(s, itarget, otarget); break; }
            return 1;
        }
    }
    return 0;
}

static void new_resolvent(pfo *s)
{   /* newResolvent (src/cpu/bve.cpp:192-220) */
    const int size = s->outn;
    if (size == 1) {
        const uint32_t unit = s->outc[0];
        const int8_t val = s->value[unit];
        if (val < 0) enqueue_unit(s, unit);
        else if (!val) learn_empty(s);
    }
    else {
        const uint64_t r = new_clause(s, s->outc, size, ST_ORIGINAL);
        uint32_t *c = CL(s, r);
        proof_resolvent(s, c);
        calc_sig(c);
        c_markadded(c);
    }
}

/* rule: 0 every pair, 1 gate clauses with non-gate ones only (xsubstitute), 2 not two
   non-core clauses (function-table gates) */
static void x_resolve(pfo *s, uint32_t x, int substitute)
{   /* xsubstitute / xresolve (src/cpu/bve.cpp:153-190) */
    uint32_t dx = 2 * x, fx = dx + 1;
    if (s->ot[dx].sz > s->ot[fx].sz) { uint32_t t = dx; dx = fx; fx = t; }
    ol_t *me = &s->ot[dx], *other = &s->ot[fx];
    for (int32_t i = 0; i < me->sz; i++) {
        const uint64_t ri = me->d[i];
        if (!c_orig(CL(s, ri))) continue;
        const int a = c_molten(CL(s, ri));
        for (int32_t j = 0; j < other->sz; j++) {
            const uint64_t rj = other->d[j];
            if (!c_orig(CL(s, rj))) continue;
            if (substitute == 1 && a == c_molten(CL(s, rj))) continue;
            if (substitute == 2 && a && c_molten(CL(s, rj))) continue;
            /* the pool may move when a resolvent is appended: re-read pointers */
            if (merge_out(s, x, CL(s, ri), CL(s, rj))) new_resolvent(s);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Function-table gates (option ve_fun_en; the GPU reference's         */
/* find_fun_gate, src/gpu/function.cuh:263-311, "inspired by the       */
/* function table reasoning in Lingeling"). src/cpu has no counterpart, */
/* and the reference numbers the variables by a global map; here the   */
/* distinct variables other than x of x's original clauses are         */
/* numbered in order of first appearance over the positive list, then  */
/* the negative list, at most FUN_VARS of them (no decision depends on */
/* the numbering: tables are only tested for falsity). A clause        */
/* without x's literal is a function of them, kept as a truth table of */
/* 2^FUN_VARS bits: 64 words, word w bit a is the assignment 64 w + a, */
/* variable k is bit k of it. Parity pinned by brute force (equisat,   */
/* model extension, the skipped resolvents) in tests/test_model.py.    */
/* ------------------------------------------------------------------ */
#define FUN_VARS 12
#define FUN_WORDS 64
static const uint64_t fun_magic[6] = {0xaaaaaaaaaaaaaaaaULL, 0xccccccccccccccccULL, 0xf0f0f0f0f0f0f0f0ULL,
                                      0xff00ff00ff00ff00ULL, 0xffff0000ffff0000ULL, 0xffffffff00000000ULL};

/* the word w of the table of variable k's literal (neg: its negation) */
static uint64_t fun_lit_word(int k, int neg, int w)
{
    uint64_t v = k < 6 ? fun_magic[k] : (((uint64_t)w >> (k - 6)) & 1u) ? ~0ull : 0ull;
    return neg ? ~v : v;
}

static int fun_index(const uint32_t *cv, int nc, uint32_t v)
{
    for (int k = 0; k < nc; k++) if (cv[k] == v) return k;
    return -1;
}

/* f &= every original clause of l[0, tail), without `lit`, as a function */
static void fun_and_list(pfo *s, const ol_t *l, int32_t tail, uint32_t lit, const uint32_t *cv, int nc, uint64_t *f)
{
    for (int32_t i = 0; i < tail; i++) {
        const uint32_t *c = CL(s, l->d[i]);
        if (!c_orig(c)) continue;
        uint64_t cls[FUN_WORDS] = {0};
        for (int k = 0; k < c_size(c); k++) {
            const uint32_t o = c[3 + k];
            if (o == lit) continue;
            const int m = fun_index(cv, nc, ABSV(o));
            for (int w = 0; w < FUN_WORDS; w++) cls[w] |= fun_lit_word(m, (int)(o & 1u), w);
        }
        for (int w = 0; w < FUN_WORDS; w++) f[w] &= cls[w];
    }
}

static int fun_false(const uint64_t *f)
{
    for (int w = 0; w < FUN_WORDS; w++) if (f[w]) return 0;
    return 1;
}

/* 1: x is defined by its clauses as a function of at most FUN_VARS variables (the
   positive and the negative side cannot both hold), with the resolvents under the
   bound; the clauses outside the minimized core are molten and *nAdded counts the
   resolvents of pairs that are not both outside it. On a count past the bound the
   clauses are frozen again and *nAdded keeps the count. */
static int find_fun_gate(pfo *s, uint32_t p, int nOrgCls, int *nAdded)
{
    const uint32_t n = p ^ 1u, x = ABSV(p);
    ol_t *poss = &s->ot[p], *negs = &s->ot[n];
    uint32_t cv[FUN_VARS];
    int nc = 0;
    for (int side = 0; side < 2; side++) {
        const ol_t *l = side ? negs : poss;
        for (int32_t i = 0; i < l->sz; i++) {
            const uint32_t *c = CL(s, l->d[i]);
            if (!c_orig(c)) continue;
            for (int k = 0; k < c_size(c); k++) {
                const uint32_t v = ABSV(c[3 + k]);
                if (v == x || fun_index(cv, nc, v) >= 0) continue;
                if (nc == FUN_VARS) return 0;
                cv[nc++] = v;
            }
        }
    }
    uint64_t pos[FUN_WORDS], neg[FUN_WORDS], fun[FUN_WORDS];
    for (int w = 0; w < FUN_WORDS; w++) pos[w] = neg[w] = ~0ull;
    fun_and_list(s, poss, poss->sz, p, cv, nc, pos);
    fun_and_list(s, negs, negs->sz, n, cv, nc, neg);
    for (int w = 0; w < FUN_WORDS; w++) if (pos[w] & neg[w]) return 0;   /* collapsefun */
    /* core minimization (function.cuh:288-299): a positive clause that the earlier
       positive clauses and the negative side already contradict without it leaves the
       core, as does a negative clause that the earlier negative clauses contradict */
    int core = 0;
    for (int32_t i = poss->sz - 1; i >= 0; i--) {
        memcpy(fun, neg, sizeof(fun));
        fun_and_list(s, poss, i, p, cv, nc, fun);
        uint32_t *c = CL(s, poss->d[i]);
        if (fun_false(fun) && c_orig(c)) { c_melt(c); core = 1; }
    }
    for (int32_t i = negs->sz - 1; i >= 0; i--) {
        for (int w = 0; w < FUN_WORDS; w++) fun[w] = ~0ull;
        fun_and_list(s, negs, i, n, cv, nc, fun);
        uint32_t *c = CL(s, negs->d[i]);
        if (fun_false(fun) && c_orig(c)) { c_melt(c); core = 1; }
    }
    *nAdded = 0;
    if (count_subst(s, x, nOrgCls, poss, negs, nAdded, 2)) {
        if (core) {   /* freezeClauses (elimination.cuh:106-120) */
            for (int32_t i = 0; i < poss->sz; i++) c_freeze(CL(s, poss->d[i]));
            for (int32_t i = 0; i < negs->sz; i++) c_freeze(CL(s, negs->d[i]));
        }
        return 0;
    }
    return 1;
}

#define FUNCTION 4

void pfo_ve(pfo *s)
{   /* bve (src/cpu/bve.cpp:28-151) */
    if (!s->opts.ve_en) return;
    for (uint32_t i = 0; i < s->npvs; i++) {
        uint32_t v = s->pvs[i];
        const uint32_t p = 2 * v, n = p + 1;
        ol_t *poss = &s->ot[p], *negs = &s->ot[n];
        const int pOrgs = count_orgs(s, poss), nOrgs = count_orgs(s, negs);
        if (!pOrgs || !nOrgs) {
            toblivion_var(s, p, pOrgs, nOrgs);
            v = 0;
        }
        else {
            uint32_t def = find_bn_gate(s, p);
            if (def) {
                save_bn_gate(s, p, pOrgs, nOrgs);
                if (substitute_single(s, p, def)) learn_empty(s);
                v = 0;
            }
            else {
                const int nOrgCls = pOrgs + nOrgs;
                int nAdded = 0, type = 0;
                s->outn = 0;
                if (nOrgCls > 2) {
                    if (find_ao_gate(s, n, nOrgCls, &nAdded)) type = SUBSTITUTION;
                    else if (!nAdded && find_ao_gate(s, p, nOrgCls, &nAdded)) type = SUBSTITUTION;
                }
                if (!type && nOrgCls > 3) {
                    if (find_ite_gate(s, p, nOrgCls, &nAdded)) type = SUBSTITUTION;
                    else if (!nAdded && find_ite_gate(s, n, nOrgCls, &nAdded)) type = SUBSTITUTION;
                    else if (find_xor_gate(s, p, nOrgCls, &nAdded)) type = SUBSTITUTION;
                    else if (!nAdded && find_xor_gate(s, n, nOrgCls, &nAdded)) type = SUBSTITUTION;
                }
                /* function-table gates after the other gates (kernels.cu:543-545) */
                if (!type && s->opts.ve_fun_en && nOrgCls > 2 && find_fun_gate(s, p, nOrgCls, &nAdded)) type = FUNCTION;
                if (!type && !nAdded && !count_subst(s, v, nOrgCls, poss, negs, &nAdded, 0)) type = RESOLUTION;
                if (type & (SUBSTITUTION | FUNCTION)) {
                    if (nAdded) x_resolve(s, v, type == FUNCTION ? 2 : 1);
                    toblivion_var(s, p, pOrgs, nOrgs);
                    v = 0;
                }
                else if (type & RESOLUTION) {
                    if (nAdded) x_resolve(s, v, 0);
                    toblivion_var(s, p, pOrgs, nOrgs);
                    v = 0;
                }
            }
        }
        if (!v) {
            s->vstate[s->pvs[i]] = MELTED_M; /* markEliminated (solve.h:535-541) */
            s->pvs[i] = 0;
        }
    }
}

/* ------------------------------------------------------------------ */
/* BCE (src/cpu/elimination.cpp:131-166)                               */
/* ------------------------------------------------------------------ */
static int is_tautology(uint32_t x, const uint32_t *c1, const uint32_t *c2)
{   /* isTautology (src/cpu/simplify.h:119-137) */
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c1 + 3, *b = c2 + 3;
    int it1 = 0, it2 = 0;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = l1 >> 1, v2 = l2 >> 1;
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if ((l1 ^ l2) == 1u) return 1;
        else if (v1 < v2) it1++;
        else if (v2 < v1) it2++;
        else { it1++; it2++; }
    }
    return 0;
}

/* an original clause of the negative list whose resolvents with every original clause
 * of the positive list are tautologies is saved (witness: the negative literal) and
 * deleted; elected variables zeroed by VE are passed over */
void pfo_bce(pfo *s)
{
    if (!s->opts.bce_en) return;
    for (uint32_t i = 0; i < s->npvs; i++) {
        const uint32_t v = s->pvs[i];
        if (!v) continue;
        const uint32_t p = 2 * v, n = p + 1;
        const ol_t *poss = &s->ot[p], *negs = &s->ot[n];
        if (poss->sz > s->opts.bce_limit || negs->sz > s->opts.bce_limit) continue;
        for (int32_t j = 0; j < negs->sz; j++) {
            uint32_t *neg = CL(s, negs->d[j]);
            if (!c_orig(neg)) continue;
            int all = 1;
            for (int32_t k = 0; k < poss->sz && all; k++) {
                const uint32_t *pos = CL(s, poss->d[k]);
                if (c_orig(pos) && !is_tautology(v, neg, pos)) all = 0;
            }
            if (all) {
                save_clause(s, neg, n);
                c_markdel(neg);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* ERE (src/cpu/redundancy.h, elimination.cpp:169-197)                 */
/* ------------------------------------------------------------------ */
static int merge_ere(uint32_t x, const uint32_t *c1, const uint32_t *c2, const uint32_t *target)
{   /* merge_ere (redundancy.h:37-85) */
    const int n1 = c_size(c1), n2 = c_size(c2), n = c_size(target);
    const uint32_t *a = c_clits(c1), *b = c_clits(c2), *t = c_clits(target);
    int it1 = 0, it2 = 0, it = 0;
    while (it < n && it1 < n1 && it2 < n2) {
        const uint32_t lit = t[it];
        if (ABSV(lit) == x) return 0;
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = ABSV(l1), v2 = ABSV(l2);
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if (v1 < v2) { if (l1 != lit) return 0; it1++; it++; }
        else if (v2 < v1) { if (l2 != lit) return 0; it2++; it++; }
        else { if (l1 != lit) return 0; it1++; it2++; it++; }
    }
    while (it < n && it1 < n1) { const uint32_t l1 = a[it1]; if (ABSV(l1) != x && l1 != t[it++]) return 0; it1++; }
    while (it < n && it2 < n2) { const uint32_t l2 = b[it2]; if (ABSV(l2) != x && l2 != t[it++]) return 0; it2++; }
    return 1;
}

static void forward_equ(pfo *s, uint32_t x, const uint32_t *c1, const uint32_t *c2, int maxsize)
{   /* forward_equ (redundancy.h:87-143) */
    const int n1 = c_size(c1), n2 = c_size(c2);
    const uint32_t *a = c_clits(c1), *b = c_clits(c2);
    int len = n1 + n2 - 2, it1 = 0, it2 = 0;
    int32_t minsize = INT_MAX, lsize;
    uint32_t msig = 0, best = 0;
    while (it1 < n1 && it2 < n2) {
        const uint32_t l1 = a[it1], l2 = b[it2], v1 = ABSV(l1), v2 = ABSV(l2);
        if (v1 == x) it1++;
        else if (v2 == x) it2++;
        else if ((l1 ^ l2) == 1u) return;
        else if (v1 < v2) { it1++; lsize = s->ot[l1].sz; if (lsize < minsize) { minsize = lsize; best = l1; } msig |= MAPHASH(l1); }
        else if (v2 < v1) { it2++; lsize = s->ot[l2].sz; if (lsize < minsize) { minsize = lsize; best = l2; } msig |= MAPHASH(l2); }
        else { it1++; it2++; len--; lsize = s->ot[l1].sz; if (lsize < minsize) { minsize = lsize; best = l1; } msig |= MAPHASH(l1); }
    }
    if (len > 1 && (!maxsize || len <= maxsize)) {
        const uint32_t type = (c_learnt(c1) || c_learnt(c2)) ? ST_LEARNT : ST_ORIGINAL;
        ol_t *minlist = &s->ot[best];
        for (int32_t i = 0; i < minsize; i++) {
            uint32_t *c = CL(s, minlist->d[i]);
            if (len == c_size(c) && (c_learnt(c) || c_st(c) == type) && sig_sub(msig, c[1]) && merge_ere(x, c1, c2, c)) {
                if (!c_del(c)) {   /* removeClause (sclause.cpp:67-75) */
                    c_markdel(c);
                    proof_line(s, 'd', c + 3, c_size(c));
                }
                break;
            }
        }
    }
}

void pfo_ere(pfo *s)
{   /* ERE (src/cpu/elimination.cpp:169-197) */
    if (!s->opts.ere_en) return;
    const int maxsize = s->opts.ere_max_resolvent;
    for (uint32_t q = 0; q < s->npvs; q++) {
        const uint32_t x = s->pvs[q];
        uint32_t dx = 2 * x, fx = dx + 1;
        if (s->ot[dx].sz > s->ot[fx].sz) { uint32_t t = dx; dx = fx; fx = t; }
        ol_t *me = &s->ot[dx], *other = &s->ot[fx];
        if (me->sz <= s->opts.ere_limit && other->sz <= s->opts.ere_limit) {
            for (int32_t i = 0; i < me->sz; i++) {
                const uint32_t *ci = CL(s, me->d[i]);
                if (c_del(ci)) continue;
                for (int32_t j = 0; j < other->sz; j++) {
                    const uint32_t *cj = CL(s, other->d[j]);
                    if (c_del(cj)) continue;
                    forward_equ(s, x, ci, cj, maxsize);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* counting, PV filter, GC, round driver                               */
/* ------------------------------------------------------------------ */
void pfo_count_all(pfo *s, int64_t *cls, int64_t *lits)
{   /* countAll (src/cpu/solve.h:724-734) */
    int64_t nc = 0, nl = 0;
    for (uint64_t i = 0; i < s->nscnf; i++) {
        const uint32_t *c = CL(s, s->scnf[i]);
        if (c_orig(c) || c_learnt(c)) { nc++; nl += c_size(c); }
    }
    *cls = nc; *lits = nl;
}

void pfo_filter_pvs(pfo *s)
{   /* filterPVs (src/cpu/solve.h:697-707) */
    uint32_t n = 0;
    for (uint32_t i = 0; i < s->npvs; i++) {
        const uint32_t x = s->pvs[i];
        if (x) { if (s->vstate[x]) continue; s->pvs[n++] = x; }
    }
    s->npvs = n;
}

void pfo_shrink(pfo *s)
{   /* shrinkSimp (src/cpu/simplify.cpp:216-227): stable removal of deleted clauses;
       the pool is compacted in the same order */
    uint64_t j = 0, w = 0;
    for (uint64_t i = 0; i < s->nscnf; i++) {
        const uint64_t r = s->scnf[i];
        const uint32_t *c = CL(s, r);
        if (c_del(c)) continue;
        const uint64_t words = 3 + (uint64_t)c_size(c);
        memmove(s->mem + w, c, sizeof(uint32_t) * words);
        s->scnf[j++] = w;
        w += words;
    }
    s->nscnf = j;
    s->msz = w;
}

void pfo_resize_cnf(pfo *s)
{   /* resizeCNF (src/cpu/solve.h:684-688) */
    const int times = s->phase + 1;
    if (times > 1 && times != s->opts.phases && s->opts.shrink_rate && (times % s->opts.shrink_rate) == 0)
        pfo_shrink(s);
}

int pfo_stop(pfo *s, int64_t lr)
{   /* stop (src/cpu/solve.h:784-786) */
    return (s->phase == s->opts.phases) || (lr <= s->opts.lits_min && s->phase > 2);
}

void pfo_next_phase(pfo *s)
{   /* simplify.cpp:197-199 */
    s->phase++;
    s->mu_inc++;
    s->mu_inc += (s->phase == s->opts.phases);
}

int pfo_sigmify(pfo *s)
{   /* sigmifying round loop and write-back (src/cpu/simplify.cpp:150-214) */
    s->phase = 0; s->mu_inc = 0;
    int64_t litsbefore = s->nliterals, diff = INT64_MAX;
    while (litsbefore) {
        pfo_resize_cnf(s);
        pfo_create_ot(s);
        if (!pfo_prop(s)) return -1;
        if (!pfo_lcve(s)) break;
        pfo_sort_ot(s);
        if (pfo_stop(s, diff)) { pfo_ere(s); break; }
        if (s->opts.sub_en || s->opts.ve_en) pfo_sub(s);
        if (s->unsat) return -1;
        pfo_ve(s);
        if (s->unsat) return -1;
        pfo_bce(s);
        int64_t nc, nl;
        pfo_count_all(s, &nc, &nl);
        pfo_filter_pvs(s);
        s->nclauses = nc; s->nliterals = nl;
        diff = litsbefore - s->nliterals; litsbefore = s->nliterals;
        pfo_next_phase(s);
    }
    if (!pfo_prop(s)) return -1;
    int64_t nc, nl;
    pfo_count_all(s, &nc, &nl);
    s->nclauses = nc; s->nliterals = nl;
    pfo_shrink(s);
    return 0;
}

/* ------------------------------------------------------------------ */
/* state access                                                       */
/* ------------------------------------------------------------------ */
int pfo_unsat(const pfo *s) { return s->unsat; }
uint64_t pfo_dump_proof(const pfo *s, uint8_t *out)
{
    if (out && s->nproof) memcpy(out, s->proof, (size_t)s->nproof);
    return s->nproof;
}
int32_t pfo_phase(const pfo *s) { return s->phase; }
uint32_t pfo_mu_inc(const pfo *s) { return s->mu_inc; }
uint64_t pfo_num_refs(const pfo *s) { return s->nscnf; }

uint64_t pfo_dump_cnf(const pfo *s, uint32_t *out)
{
    uint64_t w = 0;
    for (uint64_t i = 0; i < s->nscnf; i++) {
        const uint32_t *c = CL(s, s->scnf[i]);
        const uint64_t words = 3 + (uint64_t)c_size(c);
        if (out) memcpy(out + w, c, sizeof(uint32_t) * words);
        w += words;
    }
    return w;
}

/* Variable mapping with write-back after the simplifier: VMAP::initiate / mapLit /
 * mapShrinkLits / mapShrinkVars (src/cpu/vmap.h:44-170) as ParaFROST::map(true) applies
 * them (vmap.cpp:64-131; clauses through newBeginning -> newClause -> mapClause,
 * simplify.cpp:247-265, sclause.cpp:22-33). Variables without a state are kept in
 * order; the first variable that has a state and a value (a level-0 unit) is kept as
 * well and stands for every assigned literal. The state itself is not changed.
 * Outputs: live clauses in clause order (lits, offs[ncls + 1]), vorg[nvars_new + 1]
 * (new variable -> variable), the mapped trail, vstate[nvars_new + 1].
 * Returns the clause count. */
uint64_t pfo_vmap(const pfo *s, uint32_t *lits, uint64_t *offs, uint32_t *vorg, uint32_t *nvars_new, uint32_t *trail,
                  uint32_t *ntrail, uint8_t *vstate)
{
    uint32_t *mapped = (uint32_t *)calloc((size_t)s->nvars + 1, sizeof(uint32_t));
    uint32_t nv = 0, first = 0, mfirst = 0;
    int8_t vfirst = -1;
    for (uint32_t v = 1; v <= s->nvars; v++) {   /* initiate (vmap.h:61-72) */
        if (!s->vstate[v]) mapped[v] = ++nv;
        else if (!first && s->value[2 * v] >= 0) {
            first = v;
            vfirst = s->value[2 * v];
            mapped[v] = ++nv;
            mfirst = nv;
        }
    }
    uint64_t n = 0, ncls = 0;
    offs[0] = 0;
    for (uint64_t i = 0; i < s->nscnf; i++) {   /* mapClause over the written-back clauses */
        const uint32_t *c = CL(s, s->scnf[i]);
        if (c_del(c)) continue;
        for (int k = 0; k < c_size(c); k++) {
            const uint32_t lit = c[3 + k], v = lit >> 1;
            uint32_t out;
            if (mapped[v] && v != first) out = 2 * mapped[v] + (lit & 1u);   /* mapLit (vmap.h:44-60) */
            else if (s->value[lit] < 0) out = 0;
            else {
                out = 2 * mfirst;
                if (vfirst != s->value[lit]) out ^= 1u;
            }
            lits[n++] = out;
        }
        offs[++ncls] = n;
    }
    vorg[0] = 0;   /* mapShrinkVars(vorg) with vorg the identity */
    for (uint32_t v = 1; v <= s->nvars; v++)
        if (mapped[v]) vorg[mapped[v]] = v;
    uint32_t nt = 0;   /* mapShrinkLits(trail) (vmap.h:162-171) */
    for (uint32_t t = 0; t < s->ntrail; t++) {
        const uint32_t lit = s->trail[t], m = mapped[lit >> 1];
        if (m) trail[nt++] = 2 * m + (lit & 1u);
    }
    vstate[0] = 0;   /* mapSP: vstate of the kept variables */
    for (uint32_t v = 1; v <= s->nvars; v++)
        if (mapped[v]) vstate[mapped[v]] = s->vstate[v];
    free(mapped);
    *nvars_new = nv;
    *ntrail = nt;
    return ncls;
}

uint32_t pfo_trail(const pfo *s, uint32_t *out)
{
    if (out) memcpy(out, s->trail, sizeof(uint32_t) * s->ntrail);
    return s->ntrail;
}

uint64_t pfo_resolved(const pfo *s, uint32_t *out)
{
    if (out && s->nres) memcpy(out, s->resolved, sizeof(uint32_t) * s->nres);
    return s->nres;
}

void pfo_vstate(const pfo *s, uint8_t *out) { memcpy(out, s->vstate, (size_t)s->nvars + 1); }

uint32_t pfo_pvs(const pfo *s, uint32_t *out)
{
    if (out) memcpy(out, s->pvs, sizeof(uint32_t) * s->npvs);
    return s->npvs;
}

uint32_t pfo_ot_size(const pfo *s, uint32_t lit) { return (uint32_t)s->ot[lit].sz; }

int32_t pfo_ot_list(const pfo *s, uint32_t lit, int64_t *out_idx)
{
    const ol_t *l = &s->ot[lit];
    for (int32_t i = 0; i < l->sz; i++) {
        const uint64_t r = l->d[i];
        uint64_t lo = 0, hi = s->nscnf;
        int64_t found = -1;
        while (lo < hi) {
            uint64_t mid = (lo + hi) >> 1;
            if (s->scnf[mid] < r) lo = mid + 1; else hi = mid;
        }
        if (lo < s->nscnf && s->scnf[lo] == r) found = (int64_t)lo;
        if (out_idx) out_idx[i] = found;
    }
    return l->sz;
}
