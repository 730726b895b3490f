/*
 * pfgpu.h — C ABI of the MI355X SIGmA simplifier (libpfgpu.so).
 *
 * One opaque context per device-resident CNF instance. A context owns its HBM
 * pools (clause pool, clause references, occurrence table, variable state,
 * unit trail, model-reconstruction stack) and one HIP stream. Every call returns
 * a status code and never exits the process. Calls on one context are not
 * re-entrant; distinct contexts may be driven from distinct host threads.
 *
 * Data contract (kept from ParaFROST):
 *   literal      = 2*var + sign (sign 1 = negative), variables 1..nvars
 *                  (src/gpu/const.cuh, src/cpu/const.h:72-78)
 *   clause image = [word0 {status:2, molten:1, added:1, usage:2, lbd:26}, sig, size, lits...]
 *                  (SCLAUSE, src/gpu/sclause.h:37-186); references are uint64 word offsets
 *                  into the clause pool (S_REF, src/gpu/sclause.h)
 *   occurrence list per literal = {uint64 offset, uint32 size, uint32 capacity} over a
 *                  uint64 reference array (OT/OL, src/gpu/simptypes.cuh:279-318)
 *   resolved stack = model-reconstruction records (src/gpu/model.cuh:28-53)
 * Semantics of every pass are those of the reference simplifier ParaFROST src/cpu
 * (bit-exact, including clause order); see the per-function notes.
 */
#ifndef PFGPU_H
#define PFGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PFG_OK = 0,
    PFG_EINVAL = 1,     /* bad argument or call order                         */
    PFG_ENOMEM = 2,     /* HBM budget exceeded (reference CNFALLOC/OTALLOC_FAIL) */
    PFG_EDEVICE = 3,    /* HIP runtime error                                  */
    PFG_EUNSAT = 4,     /* formula proven unsatisfiable (reference learnEmpty)   */
    PFG_EINTERNAL = 5,  /* an internal invariant failed                      */
    PFG_EINTR = 6       /* stopped by pfg_interrupt (reference interrupted()) */
};

typedef struct pfg_ctx pfg_ctx;

/* Simplifier options; defaults equal the reference (src/cpu/options.cpp:24-46). */
typedef struct pfg_opts {
    int32_t phases;            /* 5    */
    uint32_t mu_pos, mu_neg;   /* 32   */
    uint32_t lcve_min;         /* 1    */
    int64_t lits_min;          /* 500  (vephasemin) */
    int32_t shrink_rate;       /* 2    (gcfreq) */
    int32_t ve_clause_limit;   /* 100  */
    int32_t ve_lbound_en;      /* 0    */
    int32_t xor_max_arity;     /* 10   */
    int32_t sub_limit;         /* 3000 */
    int32_t ere_limit;         /* 3000 */
    int32_t ere_max_resolvent; /* 1000 */
    int32_t sub_en, ve_en, ere_en;  /* 1, 1, 1 */
    int32_t lbd_tier1;         /* 2    */
    int32_t stable_sort;       /* 0: lists > 20 ordered as libstdc++ std::sort (reference) */
    int32_t bce_en;            /* 0    (opt_bce_en, src/cpu/options.cpp:26) */
    int32_t bce_limit;         /* 3000 (bcemax, src/cpu/options.cpp:35) */
    int32_t proof_en;          /* 0    (proof, src/cpu/options.cpp:64): binary DRAT lines of the simplifier,
                                  read with pfg_download_proof; single-instance contexts only */
    int32_t ve_fun_en;         /* 0    function-table gates in BVE (the GPU reference's vefun,
                                  src/gpu/options.cpp:31, function.cuh:263-311; src/cpu has none: the core
                                  variables are numbered by first appearance, which no result
                                  depends on); off by default, as in the reference */
    int32_t ve_atomic_en;      /* 0    BVE places each eliminated variable's resolvents, units and model
                                  records by atomic reservation instead of in elected order (the GPU
                                  reference's veatomic, src/gpu/options.cu:24, ve_k kernels.cu:298-373,
                                  off by default): the same clauses and records, their order in the pool
                                  and on the stack set by the schedule; single-instance contexts only */
    uint32_t lcve_max;         /* 0    the GPU reference's lcvemax (src/gpu/options.cpp:36, lcve.cu:68; 3000
                                  there, no such limit in src/cpu): the election stops at a candidate with
                                  more live occurrences of either literal; 0: no limit */
    uint32_t lcve_clause_max;  /* 0    the GPU reference's lcveclausemax (options.cpp:37, lcve.cu:105-127;
                                  30000 there): a candidate whose positive list holds a live clause longer
                                  than this is passed over; one whose negative list does freezes its positive
                                  list's neighbours and is not elected; 0: no limit */
} pfg_opts;

/* ---- Batches (no counterpart call in the reference, which simplifies one formula per
 * solver): many independent instances in one context, so that each kernel launch of a
 * round serves all of them. Instance i owns its own variables 1..nvars[i]; inside the
 * context they occupy consecutive ranges, each starting at a multiple of 16 (so that
 * clause signatures, which hash lit & 31, do not change). Every instance gets exactly the result that
 * pfg_sigmify gives it alone (each keeps its own round count, temperatures, LCVE
 * breaks and stop decision, and a contradiction refutes only its own instance). */
int pfg_create_batch(pfg_ctx **out, int device, uint64_t hbm_budget, uint32_t ninst, const uint32_t *nvars,
                     const pfg_opts *opts);
int pfg_batch_ninst(pfg_ctx *ctx, uint32_t *ninst);
/* lay the batch context out for another set of instances (whose aligned variable total
 * fits the context's), forgetting the clause database and all variable state; the
 * device allocations are kept, so a stream of batches is admitted without reallocating */
int pfg_batch_relayout(pfg_ctx *ctx, uint32_t ninst, const uint32_t *nvars);
/* clauses in CSR form, each instance in its own numbering; instance i has clauses
 * [inst_first[i], inst_first[i + 1]) (inst_first has ninst + 1 entries). pfg_upload_cnf
 * and pfg_upload_cnf_device take the context's numbering (instance literals shifted by
 * twice the instance's base: base[0] = 0, base[i + 1] = round_up_16(base[i] + nvars[i])). */
int pfg_upload_batch(pfg_ctx *ctx, const uint32_t *lits, const uint64_t *offs, const uint64_t *inst_first);
/* sigmify every instance; status[i] = PFG_OK or PFG_EUNSAT */
int pfg_sigmify_batch(pfg_ctx *ctx, int32_t *status);
/* one instance's clauses, model records, variable states (nvars[i] + 1 bytes) and units,
 * in its own numbering; NULL buffers query the sizes */
int pfg_batch_download(pfg_ctx *ctx, uint32_t inst, uint32_t *words, uint64_t *nwords, uint64_t *ncls, uint32_t *resolved,
                       uint64_t *nres, uint8_t *vstate, uint32_t *trail, uint32_t *ntrail);
/* No counterpart in the reference (its write-back copies, transfer.cu:104-179): the clause
 * images and model records of pfg_batch_download without the copy -- pointers into the
 * context's pinned host memory, valid until the next relayout, upload, sigmify, reset or
 * destroy of the context. For serving loops that hand results on or copy only what they keep. */
int pfg_batch_view(pfg_ctx *ctx, uint32_t inst, const uint32_t **words, uint64_t *nwords, uint64_t *ncls,
                   const uint32_t **resolved, uint64_t *nres);
/* Write-back for incremental calls, kept in HBM (cacheCNF/writeBack, src/gpu/transfer.cu:44-179,
 * without the host copy): one instance's clauses after pfg_sigmify_batch as CSR in its own
 * numbering, into caller-owned device memory: d_offs gets ncls + 1 offsets (from 0), d_lits
 * nlits literals, clause order as in pfg_batch_download. Valid until the next upload or
 * sigmify on the context; NULL buffers query the sizes. The clauses can be given back to
 * pfg_upload_cnf_device / pfg_upload_batch together with new clauses. */
int pfg_batch_export(pfg_ctx *ctx, uint32_t inst, uint32_t *d_lits, uint64_t *d_offs, uint64_t *nlits, uint64_t *ncls);
/* No counterpart in the reference: pfg_batch_export for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops). Same arguments and
 * the same output (clause order, literals, offsets), but it allocates nothing: it walks
 * the clause pool directly and keeps its per-chunk counts in a fixed device array of
 * 128 KB per GPU, held until the process ends. Limit: a pool of at most 2^33 clauses
 * (PFG_EINVAL beyond). Calls that overlap on one GPU share that array and race; run
 * them one after another. */
int pfg_batch_export_noalloc(pfg_ctx *ctx, uint32_t inst, uint32_t *d_lits, uint64_t *d_offs, uint64_t *nlits,
                             uint64_t *ncls);

/* Launch and algorithm choices of the kernels (no counterpart in the reference, whose
 * OPTIMIZEBLOCKS* macros halve block sizes by a fixed rule, src/gpu/grid.cuh:54-85).
 * Results never depend on them; parafrost_amd/tune.py measures candidates on a
 * workload and records the best. */
typedef struct pfg_tuning {
    int32_t ot_build;      /* occurrence table: 0 LSD radix sort of (literal, reference) keys, and
                              from the second round of a call on an update of the last table (the
                              keys of clauses deleted or rewritten since are dropped, those of
                              rewritten and appended clauses sorted and merged in); 1 counting
                              (histogram, atomic fill, every list sorted by reference); 2 the radix
                              sort every round; 3 as 0, updating however much changed, every update
                              checked against a full sort (PFG_EINTERNAL on a difference; a test
                              setting) */
    uint32_t clause_grid;  /* 256-thread blocks of the clause-streaming passes (LCVE, histogram, fill),
                              rounded up to a multiple of 8 (one window per XCD) */
    uint32_t sub_group;    /* lanes per elected variable in subsumption: 8, 16, 32 or 64 */
    uint32_t stage_words;  /* LDS words per block of the staged kernels (pfg_set_stage_words) */
    int32_t lcve_graph;    /* 1: the LCVE election rounds are captured once per call as a HIP graph
                              (four rounds, round number read on the device) and replayed */
    uint32_t ve_arena;     /* LDS words of a 64-lane BVE detection block: 8192 or 4096 (twice the
                              blocks per CU; variables that do not fit together wait a round) */
    uint32_t init_grid;    /* 256-thread blocks of LCVE's record-building pass (0: clause_grid): fewer
                              blocks keep each XCD's window of the rank table in its L2 */
} pfg_tuning;
int pfg_set_tuning(pfg_ctx *ctx, const pfg_tuning *t);
int pfg_get_tuning(pfg_ctx *ctx, pfg_tuning *t);

/* per-pass device time accumulated over the context's life, in milliseconds */
typedef struct pfg_times {
    double upload, hist, ot, sort_ot, prop, lcve, sub, ve, ere, gc, count, total, bce;
} pfg_times;

void pfg_default_opts(pfg_opts *o);

/* ParaFROST: cuMM::init + allocations of awaken() (src/gpu/simplify.cu:91-153, memory.cu:82-206).
 * hbm_budget = 0 means no limit beyond the device. */
int pfg_create(pfg_ctx **out, int device, uint64_t hbm_budget, uint32_t nvars, const pfg_opts *opts);
int pfg_destroy(pfg_ctx *ctx);

/* ParaFROST: extract + reflectCNF + prepareCNFAsync (transfer.cu:24-42, kernels.cu:61-70):
 * clauses in CSR form (host memory), literal order free; hdr0 (nullable) = packed status word. */
int pfg_upload_cnf(pfg_ctx *ctx, const uint32_t *lits, const uint64_t *offs, uint64_t ncls,
                   const uint32_t *hdr0);
/* the same with lits/offs already in device memory (HBM-resident input) */
int pfg_upload_cnf_device(pfg_ctx *ctx, const uint32_t *d_lits, const uint64_t *d_offs, uint64_t ncls,
                          uint64_t nlits);
/* root-level units to propagate (ParaFROST trail entries before sigmify) */
int pfg_enqueue_units(pfg_ctx *ctx, const uint32_t *lits, uint32_t n);
/* ParaFROST::ifreeze / iunfreeze (src/cpu/assume.cpp:45-61; iassume marks the same way,
 * assume.cpp:68-84): on = 1 freezes, on = 0 melts the listed variables. A frozen variable is
 * never elected by LCVE (iassumed, src/cpu/lcve.cpp:57; src/gpu/lcve.cu:62), so it is never
 * eliminated. The marks persist across sigmify calls and uploads; pfg_reset and
 * pfg_batch_relayout clear them. In a batch context variables are in the global numbering
 * (instance i's variable v is hibase[i] + v). */
int pfg_ifreeze(pfg_ctx *ctx, const uint32_t *vars, uint32_t n, int on);

/* ParaFROST::sigmify (src/gpu/simplify.cu:75-89 / src/cpu/simplify.cpp:115-214): the whole
 * round loop on the device, ending with GC of deleted clauses. */
int pfg_sigmify(pfg_ctx *ctx);

/* Individual passes, in the reference round order (src/cpu/simplify.cpp:182-200). */
int pfg_resize_cnf(pfg_ctx *ctx);                 /* resizeCNF/reallocCNF  (simplify.cu:30-46)        */
int pfg_create_ot(pfg_ctx *ctx);                  /* histSimp+resizeOTAsync+createOTAsync              */
int pfg_reduce_ot(pfg_ctx *ctx);                  /* reduceOTAsync         (kernels.cu:98-105)          */
int pfg_prop(pfg_ctx *ctx);                       /* prop                  (elimbcp.cu:103-193)         */
int pfg_lcve(pfg_ctx *ctx, uint32_t *npvs);       /* varReorder + LCVE     (lcve.cu:23-130)             */
int pfg_sort_ot(pfg_ctx *ctx);                    /* sortOT                (segsort.cu:32-41)           */
int pfg_sub(pfg_ctx *ctx);                        /* subAsync              (kernels.cu:1133-1154)       */
int pfg_ve(pfg_ctx *ctx);                         /* veAsync + postVE      (kernels.cu:1085-1120)       */
int pfg_bce(pfg_ctx *ctx);                        /* BCE (elimination.cpp:131-166; bceAsync kernels.cu:1156-1168), if bce_en */
int pfg_ere(pfg_ctx *ctx);                        /* ereAsync              (kernels.cu:1171-1195)       */
int pfg_count(pfg_ctx *ctx, int64_t *cls, int64_t *lits); /* parcountAll (kernels.cu:855-925)           */
int pfg_filter_pvs(pfg_ctx *ctx);                 /* filterPVs / DeviceSelect::If (elimination.cu:37-50) */
int pfg_gc(pfg_ctx *ctx);                         /* compactCNF            (recycle.cu:71-97)           */
int pfg_stop(pfg_ctx *ctx, int64_t lr, int *stop);/* stop                  (solve.h:773-778)            */
int pfg_next_phase(pfg_ctx *ctx);                 /* phase++, mu_inc       (simplify.cu:200-201)        */

/* histogram of live literals (histSimp, histogram.cu:34-51) into host hist[2*(nvars+1)] */
int pfg_hist(pfg_ctx *ctx, uint32_t *hist);

/* state download (cacheCNF / cacheUnits / cacheResolved / cacheEliminated, transfer.cu:104-179).
 * Sizes are returned through the size pointers; a NULL buffer only queries the size. */
int pfg_download_cnf(pfg_ctx *ctx, uint32_t *words, uint64_t *nwords, uint64_t *ncls);
/* ParaFROST::map(true) after sigmify (src/cpu/vmap.cpp:64-131, vmap.h:44-171; write-back
 * through newBeginning, simplify.cpp:247-265): the variable mapping computed on the device.
 * Variables without a state keep their order; the first variable that has a state and a
 * value (a level-0 unit) is kept too and stands for every assigned literal; the rest are
 * dropped. words: the clauses as SCLAUSE images in clause order with renumbered literals
 * and their signatures; vorg[nvars_new + 1]: new variable -> context variable; trail: the
 * mapped units; vstate[nvars_new + 1]. The context itself is not changed. NULL buffers
 * query the sizes. Single-instance contexts only. */
int pfg_download_mapped(pfg_ctx *ctx, uint32_t *words, uint64_t *nwords, uint64_t *ncls, uint32_t *vorg,
                        uint32_t *nvars_new, uint32_t *trail, uint32_t *ntrail, uint8_t *vstate);
int pfg_download_trail(pfg_ctx *ctx, uint32_t *lits, uint32_t *n);
int pfg_download_resolved(pfg_ctx *ctx, uint32_t *words, uint64_t *n);
/* The binary DRAT proof lines the simplifier wrote since pfg_create/pfg_reset (PROOF::addUnit,
 * addResolvent, deleteClause, addEmpty; src/cpu/proof.cpp:92-232 in binary mode, variables in the
 * context's numbering). bytes NULL: *n = the size. Succeeds after a refutation too. */
int pfg_download_proof(pfg_ctx *ctx, uint8_t *bytes, uint64_t *n);
int pfg_download_vstate(pfg_ctx *ctx, uint8_t *vstate);          /* nvars + 1 bytes */
int pfg_download_pvs(pfg_ctx *ctx, uint32_t *pvs, uint32_t *n);
/* occurrence list of one literal as clause positions (index in clause order) */
int pfg_download_ot_list(pfg_ctx *ctx, uint32_t lit, int64_t *idx, uint32_t *n);

int pfg_times_get(pfg_ctx *ctx, pfg_times *t);

/* ParaFROST's interrupted() (src/gpu/solve.h:181; polled before every round,
 * src/gpu/simplify.cu:182, and before SUB/VE/BCE/ERE, src/gpu/elimination.cu:28,55,65,76,
 * where the reference then ends the solver): asks the running or next pfg_sigmify /
 * pfg_sigmify_batch on the context to stop at its next poll. That call leaves the round
 * loop, propagates pending units, compacts the clause database as at a normal end and
 * returns PFG_EINTR (per instance: PFG_EINTR for instances that were still running);
 * the context then holds a valid, partly simplified formula. The request is consumed
 * by the call it stops. Safe to call from another host thread. */
int pfg_interrupt(pfg_ctx *ctx);

/* The solver's variable mapping (vorg, src/cpu/solve.h:76 / src/gpu/simplify.cu:141-146
 * with initDevVorg, src/gpu/kernels.cu:739): vorg[v] = the original variable of context
 * variable v, n = nvars + 1 entries (vorg[0] unused). After a map(true) the solver's
 * variables are renumbered, and model records (saveWitness/saveLiteral,
 * src/cpu/model.cpp:263-280; ORIGINIZELIT, src/gpu/primitives.cuh:36-46) and proof lines
 * (proof.cpp:118-135) are written in original numbering through it. vorg NULL restores
 * the identity. Single-instance contexts; pfg_reset clears it. */
int pfg_set_vorg(pfg_ctx *ctx, const uint32_t *vorg, uint32_t n);

/* How often the less common code paths ran since pfg_create / pfg_reset_stats (no
 * counterpart in the reference): [0] extra ERE claiming rounds (run when a target was
 * claimed by two elected variables), [1] unit-capable SUB variables run alone on one
 * wave, [2] of those, the ones whose lists exceed the wave's LDS arrays (sub_limit
 * above 3000), [3] LCVE calls whose election rounds ran from a captured graph,
 * [4] elected occurrence lists sorted by the exact introsort (past 20 entries with two
 * equal keys, or past 256 entries) rather than by stable ranks, [5] occurrence tables
 * made by updating the previous round's table instead of sorting every key again.
 * *n = the number of counters; at most max are written. */
int pfg_path_counts(pfg_ctx *ctx, uint64_t *counts, int max, int *n);

/* per-kernel device time (HIP events around each launch) and algorithmic HBM bytes.
 * kernel = the one HIP kernel the timer brackets, or NULL when it spans several launches.
 * Bytes of the per-variable kernels (sort_pv_lists, sub_range, ve_detect, ere_claim) are
 * only computed while accounting is on (it adds a reduction before each of them). */
typedef struct pfg_kstat {
    const char *name;
    const char *kernel;
    double ms;
    uint64_t launches;
    double bytes;
} pfg_kstat;
/* Statistics and settings (pfg_kernel_stats, pfg_set_accounting, pfg_set_stage_words,
 * pfg_reset_stats) return PFG_OK also after the formula was refuted; every other call
 * reports PFG_EUNSAT from then on. */
int pfg_kernel_stats(pfg_ctx *ctx, pfg_kstat *out, int max, int *n);
int pfg_set_accounting(pfg_ctx *ctx, int on);
/* Tuning: cap the LDS words per block that the staged per-variable kernels (SUB,
 * BVE detection) use for a variable's clauses; a variable that does not fit is
 * processed from the pool. Results do not depend on it. Default: the whole arena. */
int pfg_set_stage_words(pfg_ctx *ctx, uint32_t words);
int pfg_reset_stats(pfg_ctx *ctx);
/* forget the clause database and all variable state (allocations are kept); the
   context can then take a new upload of the same variable range */
int pfg_reset(pfg_ctx *ctx);
int pfg_is_unsat(pfg_ctx *ctx);
const char *pfg_last_error(pfg_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
