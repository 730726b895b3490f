// Internal context of one device-resident simplifier instance.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <stdexcept>
#include <vector>
#include <array>
#include <atomic>
#include <algorithm>
#include "../../include/pfgpu.h"

namespace pfg {

// occurrence list header (OL, src/gpu/simptypes.cuh:66-67): offset into the
// occurrence array, current size, capacity (= histogram count at build time)
struct OL {
    uint64_t off;
    uint32_t sz, cap;
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define PFG_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::pfg::Error(PFG_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Growable pinned host buffer of T (download staging). A grown buffer's old block is
// freed only at release(): hipHostFree waits for the whole device, which would stall
// the other contexts' streams (serving slots) at every growth. Growth is by half, so
// the retired blocks add up to at most twice the current one.
template <typename T>
struct PinBuf {
    T *p = nullptr;
    uint64_t cap = 0;
    std::vector<T *> retired;
    void ensure(uint64_t n)
    {
        if (n <= cap) return;
        if (p) retired.push_back(p);
        p = nullptr;
        const uint64_t nc = std::max<uint64_t>(n + n / 4 + 1024, cap + cap / 2);
        cap = 0;
        PFG_HIP(hipHostMalloc((void **)&p, nc * sizeof(T), hipHostMallocDefault));
        cap = nc;
    }
    void release()
    {
        for (T *q : retired) (void)hipHostFree(q);
        retired.clear();
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Growable device buffer of T.
template <typename T>
struct DBuf {
    T *p = nullptr;
    uint64_t cap = 0;
};

// device counter slots (Ctx::cnt) and flags. Slot CNT_CONF counts contradictions;
// the variables involved follow it (at most CONF_CAP), so that a batch can tell
// which instances were refuted.
enum { CNT_TOTAL = 0, CNT_A = 1, CNT_B = 2, CNT_C = 3, CNT_D = 4, CNT_SUBFROM = 5, CNT_SORTX = 7, CNT_NTRAIL = 8, CNT_FLAGS = 9, CNT_ACCT = 10,
       CNT_ROUND = 11, CNT_OTRM = 12, CNT_CONF = 16 };
enum { FLAG_UNSAT = 1, FLAG_INTERNAL = 2 };
constexpr uint64_t CONF_CAP = 1ull << 20;

// event timers: most wrap exactly one kernel (kernel_symbols names it), the
// rest a short sequence of launches (kernel_symbols entry null)
enum KId {
    K_UPLOAD, K_HIST, K_OT_FILL, K_OT_SORT, K_PROP, K_LCVE_RADIX, K_LCVE_INIT, K_LCVE_PASS_A, K_LCVE_PASS_B,
    K_SORT_PV, K_SUB, K_VE_DETECT, K_VE_APPLY, K_BCE_COUNT, K_BCE_APPLY, K_ERE, K_GC, K_COUNT, K_SUB_REST,
    K_COUNT_INST, K_LCVE_GRAPH, K_VE_FUN, K_VE_RES, K_OT_UPDATE, K_NUM
};
extern const char *const kernel_names[K_NUM];
extern const char *const kernel_symbols[K_NUM];

// how often the less common code paths ran (pfg_path_counts): the tests check
// that their cases reach them
enum PathId {
    PATH_ERE_ROUNDS,     // extra ERE claiming rounds (a target claimed by two elected variables)
    PATH_SUB_WAVE,       // unit-capable variables run alone on one wave
    PATH_SUB_WAVE_BIG,   // ... of them with lists longer than the LDS arrays
    PATH_LCVE_GRAPH,     // LCVE calls whose election rounds ran as a captured graph
    PATH_SORT_EXACT,     // elected lists sorted by the exact introsort (equal keys past 20 entries, or long)
    PATH_OT_UPDATE,      // occurrence tables made by updating the last one
    PATH_NUM
};

// launch and algorithm choices (pfg_tuning); results never depend on them
struct Tuning {
    int32_t ot_build = 0;          // 0 radix (updated between rounds), 1 counting, 2 radix every round,
                                   // 3 as 0 with every update checked against a full build
    uint32_t clause_grid = 2048;   // blocks of 256 threads for the XCD-windowed clause passes
    uint32_t sub_group = 16;       // lanes per variable in the first subsumption configuration
    int32_t lcve_graph = 0;        // LCVE election rounds replayed from a captured graph
    uint32_t ve_arena = 8192;      // LDS words of a BVE detection block (4096: twice the blocks per CU)
    uint32_t init_grid = 0;        // blocks of the LCVE record-building pass (0: clause_grid)
};

struct KStat {
    double ms = 0, bytes = 0;
    uint64_t launches = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, allocated = 0;
    pfg_opts opts{};
    uint32_t nvars = 0, nvars_cap = 0;   // variables in use; per-variable buffers are sized for nvars_cap
    uint64_t nlits2 = 0;  // 2 * (nvars + 1)

    // clause pool (SCLAUSE images) and references in clause order
    DBuf<uint32_t> mem, mem2;   // mem2: destination of compaction
    uint64_t msz = 0;           // words in use
    DBuf<uint64_t> refs, refs2;
    uint64_t nrefs = 0;

    // occurrence table
    DBuf<OL> ot;
    DBuf<uint64_t> occ;
    DBuf<uint32_t> hist;        // per literal
    bool ot_valid = false;
    bool hist_live = false;     // hist holds the live-literal counts of the current database
    // what the next radix build updates from (tuning ot_build 0 and 3): the sorted
    // (literal << 32 | reference) keys of the last build, every clause's live size at
    // that build, a bitmap over reference >> 2 of the clauses rewritten in place since
    // then (equivalence substitution in BVE keeps the size), and a log of the keys of
    // literals removed from clauses in place since then (count in cnt[CNT_OTRM])
    DBuf<uint64_t> otk, otk2;
    uint64_t otk_n = 0, otk_nrefs = 0;
    DBuf<uint32_t> otsz;
    DBuf<uint32_t> otdirty;
    uint64_t otdirty_words = 0;
    DBuf<uint64_t> otrm;
    uint64_t otrm_cap = 0;
    DBuf<uint32_t> otpos;       // the update: bit per old key position, set for the dropped keys
    DBuf<uint64_t> otcnt;       // ... and per tile of old keys, the dropped ones
    DBuf<uint64_t> otchk;       // ot_build 3: the lists and bounds of a full build, for the check
    bool otk_valid = false;
    bool trace = false;         // PFG_TRACE set: per-round progress on stderr
    bool account = false;       // compute algorithmic bytes of the per-variable kernels (extra reductions)
    uint32_t stage_words = 0xFFFFFFFFu;   // cap on the LDS words per block the staged kernels use

    // variables
    DBuf<int8_t> value;         // per literal: -1 unassigned, 0 false, 1 true
    DBuf<uint8_t> vstate;       // per variable: MELTED 1, FROZEN 2
    DBuf<uint8_t> ifrozen;      // per variable: frozen for LCVE by the caller (incremental ifreeze)
    DBuf<uint32_t> trail;       // units (nvars + 1)
    uint32_t ntrail = 0, propagated = 0;
    DBuf<uint32_t> resolved;    // model stack
    uint64_t nres = 0;
    DBuf<uint64_t> mchunk;      // batch: first word of each eliminated variable's records, in stack order
    uint64_t nchunk = 0;
    DBuf<uint32_t> vorg;        // the caller's variable -> original variable map (pfg_set_vorg)
    bool has_vorg = false;      // false: identity
    DBuf<uint32_t> pvs;         // elected variables
    DBuf<uint32_t> units;       // units produced by one BVE pass, in elected order
    uint32_t npvs = 0;

    // scratch
    DBuf<uint32_t> scores, eligible, rank, tmpu32a, tmpu32b, tmpu32c, tmpu32d;
    DBuf<uint8_t> tmpu8;
    DBuf<uint64_t> tmpu64a, tmpu64b, tmpu64c;
    DBuf<uint32_t> owner;       // per (ref >> 1): ERE target claims (all-ones when idle)
    DBuf<uint32_t> claim;       // bitmap over (ref >> 1): prop claims
    DBuf<uint32_t> plong;       // prop: the long lists of a batch of units
    DBuf<uint32_t> erec;        // LCVE: per-clause candidate-rank records
    DBuf<uint32_t> ufirst;      // per variable: first unit occurrence (all-ones when idle)
    DBuf<uint64_t> cnt;         // small device counters
    // binary DRAT output (opts.proof_en, single-instance contexts)
    DBuf<uint8_t> proof;        // the bytes written so far
    uint64_t nproof = 0;
    bool proof_closed = false;  // the empty clause has been written
    DBuf<uint8_t> pev, pord;    // one pass's events (PEv), as logged and in loop order
    DBuf<uint32_t> pnev, pbytes;
    DBuf<uint64_t> pcnt, poff, pboff, pfirst;   // counters, offsets; per literal: first unit event
    uint64_t pcap = 0;
    uint64_t *hcnt = nullptr;   // pinned mirror
    DBuf<uint8_t> scan_tmp;     // scan partial sums
    DBuf<uint8_t> radix_tmp;    // radix sort temporaries

    // batch of independent instances (ninst == 1: one instance). Instance i owns the
    // global variables (hibase[i], hibase[i + 1]]; its literals are shifted by 2 hibase[i].
    uint32_t ninst = 1;
    bool batch = false;                   // created as a batch (per-instance round control)
    std::vector<uint32_t> hibase, inst_nvars;
    DBuf<uint32_t> ibase;                 // device copy of hibase
    std::vector<uint8_t> imu_host;        // per instance: mu_inc of this round, 0xFF = not electing
    DBuf<uint8_t> imu;
    DBuf<uint64_t> ib0;                   // per instance: static break, then dynamic break (2 ninst)
    DBuf<uint64_t> inel;                  // per instance: elected variables
    DBuf<uint64_t> icnt;                  // per instance: live clauses, live literals (2 ninst)
    uint64_t batch_gen = 0, cache_gen = ~0ull;   // host copy of the batch state for per-instance downloads
    PinBuf<uint32_t> cache_mem, pin_res, pin_trail;   // pinned staging of the downloads
    PinBuf<uint64_t> pin_refs;
    PinBuf<uint8_t> pin_vstate;
    std::vector<uint64_t> cache_start;              // per instance: first word of its clause images
    std::vector<uint64_t> cache_cstart;             // per instance: its first clause image (grouped order)
    DBuf<uint64_t> img_off;                         // per grouped image: word offset in mem2
    std::vector<uint64_t> cache_rstart;             // per instance: first word of its model records (pin_res)
    std::vector<std::vector<uint32_t>> cache_trail;

    // round state
    int32_t phase = 0;
    uint32_t mu_inc = 0;
    int64_t nclauses = 0, nliterals = 0;
    bool unsat = false;
    bool sorted_lists = false;

    pfg_times times{};
    uint64_t path[PATH_NUM] = {};
    Tuning tune;
    KStat kstat[K_NUM];
    std::vector<std::array<hipEvent_t, 2>> kev;     // event pairs of pending kernel timers
    std::vector<std::pair<int, double>> kpend;      // (kernel, algorithmic bytes) per pending pair
    int kopen = 0;                                  // timers whose end event is not recorded yet
    hipEvent_t ev[8][2] = {};   // per nesting depth of pass timers
    int ev_depth = 0;
    std::string last_error;
    std::atomic<int> interrupt{0};   // set by pfg_interrupt, polled between passes and rounds
};

// allocation helpers (budget-checked)
void dalloc_raw(Ctx &c, void **p, uint64_t bytes);
void dfree_raw(Ctx &c, void *p, uint64_t bytes);

template <typename T>
void ensure(Ctx &c, DBuf<T> &b, uint64_t n, bool keep = false)
{
    if (n <= b.cap) return;
    uint64_t ncap = b.cap ? b.cap : 1024;
    while (ncap < n) ncap += ncap / 2 + 1024;
    void *np = nullptr;
    dalloc_raw(c, &np, ncap * sizeof(T));
    if (keep && b.p && b.cap)
        PFG_HIP(hipMemcpyAsync(np, b.p, b.cap * sizeof(T), hipMemcpyDeviceToDevice, c.stream));
    if (b.p) dfree_raw(c, b.p, b.cap * sizeof(T));   // queued after the copy on the same stream
    b.p = (T *)np;
    b.cap = ncap;
}

template <typename T>
void release(Ctx &c, DBuf<T> &b)
{
    if (b.p) dfree_raw(c, b.p, b.cap * sizeof(T));
    b.p = nullptr;
    b.cap = 0;
}

// read a few device counters into the pinned mirror (synchronizes the stream)
void read_counters(Ctx &c, int n);

// event pair around one kernel launch; resolved lazily (no stream sync per launch)
struct KTimer {
    Ctx &c;
    size_t slot;
    KTimer(Ctx &cc, int kid, double bytes);
    ~KTimer();
};
void resolve_ktimers(Ctx &c);

struct PassTimer {
    Ctx &c;
    double *acc;
    int d;
    PassTimer(Ctx &cc, double *a);
    ~PassTimer();
};

}  // namespace pfg
