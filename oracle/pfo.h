/*
 * pfo — CPU restatement of ParaFROST's SIGmA simplifier (reference src/cpu semantics).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the GPU
 * simplifier: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it. The product path never links or calls it.
 *
 * State mirrors the reference layouts: a uint32 clause pool holding SCLAUSE
 * images (3-word header + literals, src/cpu/sclause.h:28-35), a vector of
 * uint64 clause references in "scnf" order, occurrence lists per literal,
 * the per-literal value array, per-variable state, the unit trail and the
 * model-reconstruction ("resolved") stack.
 */
#ifndef PFO_H
#define PFO_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfo_opts {
    int32_t phases;            /* opt_phases            (src/cpu/options.cpp:43) = 5     */
    uint32_t mu_pos, mu_neg;   /* opt_mu_pos / mu_neg   (options.cpp:41-42)      = 32    */
    uint32_t lcve_min;         /* opt_lcve_min          (options.cpp:40)         = 1     */
    int64_t lits_min;          /* opt_ve_phase_min      (options.cpp:44)         = 500   */
    int32_t shrink_rate;       /* opt_cnf_free (gcfreq) (options.cpp:36)         = 2     */
    int32_t ve_clause_limit;   /* opt_ve_clause_max     (options.cpp:45)         = 100   */
    int32_t ve_lbound_en;      /* opt_ve_lbound_en      (options.cpp:32)         = 0     */
    int32_t xor_max_arity;     /* opt_xor_max_arity     (options.cpp:46)         = 10    */
    int32_t sub_limit;         /* opt_sub_max_occurs    (options.cpp:39)         = 3000  */
    int32_t ere_limit;         /* opt_ere_max_occurs    (options.cpp:37)         = 3000  */
    int32_t ere_max_resolvent; /* opt_ere_clause_max    (options.cpp:38)         = 1000  */
    int32_t sub_en, ve_en, ere_en;                 /* options.cpp:27-33, 281-311 */
    int32_t lbd_tier1;         /* opt_lbd_tier1         (options.cpp:135)        = 2     */
    int32_t stable_sort;       /* 0: lists > 20 sorted by libstdc++ std::sort order (reference);
                                  1: stable order (key, creation) for every list */
    int32_t bce_en;            /* opt_bce_en            (options.cpp:26)         = 0     */
    int32_t bce_limit;         /* opt_bce_max_occurs    (options.cpp:35)         = 3000  */
    int32_t proof_en;          /* opt_proof_en          (options.cpp:64)         = 0     (binary DRAT) */
    int32_t ve_fun_en;         /* GPU reference's vefun (src/gpu/options.cpp:31): function-table gates
                                  (no src/cpu counterpart: the variable numbering is chosen
                                  here; no result depends on it) = 0 */
    uint32_t lcve_max;         /* GPU reference's lcvemax (src/gpu/options.cpp:36, lcve.cu:68): the walk
                                  stops at a candidate with more live occurrences of either literal
                                  (0: no limit; src/cpu has none) = 0 */
    uint32_t lcve_clause_max;  /* GPU reference's lcveclausemax (options.cpp:37, lcve.cu:105-120): a
                                  candidate whose list holds a live clause longer than this is not
                                  elected (0: no limit) = 0 */
} pfo_opts;

typedef struct pfo pfo;

void pfo_default_opts(pfo_opts *o);

/* create a simplifier over nvars variables and load ncls clauses (CSR form, any
 * literal order; duplicates/tautologies are not allowed). hdr0 (nullable) gives the
 * packed status word of each clause (ORIGINAL when NULL). */
pfo *pfo_create(uint32_t nvars, const pfo_opts *o, const uint32_t *lits, const uint64_t *offs,
                uint64_t ncls, const uint32_t *hdr0);
void pfo_destroy(pfo *s);

/* freeze (on = 1) or melt (on = 0) variable v for LCVE across later sigmify calls:
 * the incremental ifreeze/iunfreeze (src/cpu/assume.cpp:45-61); 0 ok, -1 bad variable */
int pfo_ifreeze(pfo *s, uint32_t v, int on);

/* the caller's variable mapping vorg[nvars + 1] (vorg[v] = original variable of v;
 * ParaFROST's vorg, src/cpu/solve.h:76, vmap.cpp:103): model records and proof lines
 * are written in original numbering (model.cpp:263-280, proof.cpp:118-135). 0 ok,
 * -1 bad size or a zero entry. Without it the mapping is the identity. */
int pfo_set_vorg(pfo *s, const uint32_t *vorg, uint32_t n);

/* assign units before simplification (they are propagated by the first prop()). */
int pfo_enqueue(pfo *s, uint32_t lit);

/* the whole simplifier: src/cpu/simplify.cpp:129-214 (sigmifying) up to shrinkSimp */
int pfo_sigmify(pfo *s);

/* individual passes (same order as the reference round loop) */
void pfo_resize_cnf(pfo *s);          /* resizeCNF  (src/cpu/solve.h:684-688)        */
void pfo_create_ot(pfo *s);           /* createOT   (src/cpu/simplify.cpp:34-53)     */
int  pfo_prop(pfo *s);                /* prop       (src/cpu/elimination.cpp:64-97)  */
int  pfo_lcve(pfo *s);                /* LCVE       (src/cpu/lcve.cpp:46-81)         */
void pfo_sort_ot(pfo *s);             /* sortOT     (src/cpu/simplify.cpp:79-93)     */
int  pfo_stop(pfo *s, int64_t lr);    /* stop       (src/cpu/solve.h:784-786)        */
void pfo_ere(pfo *s);                 /* ERE        (src/cpu/elimination.cpp:169-197)*/
void pfo_sub(pfo *s);                 /* SUB        (src/cpu/elimination.cpp:108-130)*/
void pfo_ve(pfo *s);                  /* VE/bve     (src/cpu/bve.cpp:28-151)         */
void pfo_bce(pfo *s);                 /* BCE        (src/cpu/elimination.cpp:131-166)*/
void pfo_count_all(pfo *s, int64_t *cls, int64_t *lits); /* countAll (solve.h:724-734)  */
void pfo_filter_pvs(pfo *s);          /* filterPVs  (src/cpu/solve.h:697-707)        */
void pfo_shrink(pfo *s);              /* shrinkSimp (src/cpu/simplify.cpp:216-227)   */
void pfo_hist(pfo *s, uint32_t *hist);/* histSimp   (src/cpu/histogram.cpp:80-95) per literal */
void pfo_next_phase(pfo *s);          /* phase++, mu_inc update (simplify.cpp:197-199) */

/* variable mapping with write-back after the simplifier (ParaFROST::map(true),
 * src/cpu/vmap.cpp:64-131, vmap.h:44-171); the state is not changed. Buffers: lits
 * (live literals), offs (live clauses + 1), vorg and vstate (nvars + 1), trail (ntrail).
 * Returns the mapped clause count. */
uint64_t pfo_vmap(const pfo *s, uint32_t *lits, uint64_t *offs, uint32_t *vorg, uint32_t *nvars_new, uint32_t *trail,
                  uint32_t *ntrail, uint8_t *vstate);

/* state access */
int      pfo_unsat(const pfo *s);
uint64_t pfo_dump_proof(const pfo *s, uint8_t *out);   /* the simplifier's binary DRAT lines */
int32_t  pfo_phase(const pfo *s);
uint32_t pfo_mu_inc(const pfo *s);
uint64_t pfo_num_refs(const pfo *s);
/* dump scnf: per clause [word0, sig, size, lits...]; returns words written (out may be NULL) */
uint64_t pfo_dump_cnf(const pfo *s, uint32_t *out);
uint32_t pfo_trail(const pfo *s, uint32_t *out);
uint64_t pfo_resolved(const pfo *s, uint32_t *out);
void     pfo_vstate(const pfo *s, uint8_t *out);          /* nvars+1 bytes */
uint32_t pfo_pvs(const pfo *s, uint32_t *out);
/* occurrence list of a literal as clause indices in scnf order (-1 if ref unknown) */
int32_t  pfo_ot_list(const pfo *s, uint32_t lit, int64_t *out_idx);
uint32_t pfo_ot_size(const pfo *s, uint32_t lit);

/* libstdc++ std::sort order restatement, exposed for its own pin test:
 * sorts idx[0..n) by key[idx] ascending with the introsort steps of libstdc++. */
void pfo_introsort_u32(uint32_t *idx, int64_t n, const uint32_t *key);

/* model replay (src/cpu/model.cpp:104-142): value[v] in {0,1} for v in 1..nvars is
 * extended in place from the resolved stack. */
void pfo_extend_model(const pfo *s, uint8_t *value);

#ifdef __cplusplus
}
#endif
#endif
