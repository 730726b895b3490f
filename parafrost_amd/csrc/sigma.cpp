// Host side of the MI355X simplifier: the reference's simplification entry
// points (ParaFROST::sigmify / sigmifying and the pass wrappers) driving the
// device passes, and the C ABI declared in include/pfgpu.h.
#include "ctx.hpp"
#include "device.hpp"
#include "passes.hpp"
#include <chrono>
#include <cstring>
#include <vector>
#include <algorithm>
#include <new>
#include <cstdlib>
#include <mutex>
#include <set>

namespace pfg {

// Device buffers come from the device's stream-ordered pool on the context's stream:
// a free waits for the work queued before it on that stream only, where hipFree would
// wait for the whole device (other contexts' streams included: serving slots, a
// caller's own work). The pool keeps freed blocks for reuse instead of returning them
// to the driver at every synchronization.
static void pool_setup(int device)
{
    static std::mutex mu;
    static std::set<int> done;
    std::lock_guard<std::mutex> lock(mu);
    if (done.count(device)) return;
    hipMemPool_t pool = nullptr;
    PFG_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t keep = UINT64_MAX;
    PFG_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    done.insert(device);
}

void dalloc_raw(Ctx &c, void **p, uint64_t bytes)
{
    if (c.budget && c.allocated + bytes > c.budget)
        throw Error(PFG_ENOMEM, "HBM budget exceeded (" + std::to_string(c.allocated + bytes) + " > " +
                                    std::to_string(c.budget) + " bytes)");
    hipError_t e = hipMallocAsync(p, bytes ? bytes : 1, c.stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(e == hipErrorOutOfMemory ? PFG_ENOMEM : PFG_EDEVICE,
                    std::string("hipMallocAsync: ") + hipGetErrorString(e));
    }
    c.allocated += bytes;
}

void dfree_raw(Ctx &c, void *p, uint64_t bytes)
{
    (void)hipFreeAsync(p, c.stream);
    c.allocated -= std::min(bytes, c.allocated);
}

void read_counters(Ctx &c, int n)
{
    PFG_HIP(hipMemcpyAsync(c.hcnt, c.cnt.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
}

PassTimer::PassTimer(Ctx &cc, double *a) : c(cc), acc(a), d(cc.ev_depth < 8 ? cc.ev_depth++ : -1)
{
    if (d >= 0) (void)hipEventRecord(c.ev[d][0], c.stream);
}

PassTimer::~PassTimer()
{
    if (d < 0) return;
    (void)hipEventRecord(c.ev[d][1], c.stream);
    if (hipEventSynchronize(c.ev[d][1]) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c.ev[d][0], c.ev[d][1]) == hipSuccess) *acc += ms;
    }
    c.ev_depth--;
}

const char *const kernel_names[K_NUM] = {
    "upload_prep", "hist", "ot_fill", "ot_sort", "prop", "lcve_radix", "lcve_init", "lcve_pass_a", "lcve_pass_b",
    "sort_pv_lists", "sub", "ve_detect", "ve_apply", "bce_count", "bce_apply", "ere_probe", "gc_move", "count",
    "sub_rest", "count_inst", "lcve_graph", "ve_fun", "ve_res", "ot_update"};
const char *const kernel_symbols[K_NUM] = {
    "k_upload", nullptr, nullptr, nullptr, nullptr, nullptr, "k_elect_init", "k_elect_pass_a", "k_elect_pass_b",
    "k_sort_pv_wave", "k_sub_grp", "k_ve_detect", "k_ve_apply", nullptr, nullptr, "k_ere_probe", "k_gc_move", "k_count",
    nullptr, "k_count_inst", nullptr, "k_ve_fun", "k_ve_res", "k_otu_merge"};

// Pending event pairs are resolved in batches, never while a timer is open (its end
// event is not recorded yet).
KTimer::KTimer(Ctx &cc, int kid, double bytes) : c(cc)
{
    if (c.kpend.size() >= 2048 && !c.kopen) resolve_ktimers(c);
    c.kopen++;
    slot = c.kpend.size();
    if (c.kev.size() <= slot) {
        std::array<hipEvent_t, 2> e{};
        PFG_HIP(hipEventCreate(&e[0]));
        PFG_HIP(hipEventCreate(&e[1]));
        c.kev.push_back(e);
    }
    c.kpend.emplace_back(kid, bytes);
    (void)hipEventRecord(c.kev[slot][0], c.stream);
}

KTimer::~KTimer()
{
    (void)hipEventRecord(c.kev[slot][1], c.stream);
    c.kopen--;
}

void resolve_ktimers(Ctx &c)
{
    if (c.kpend.empty() || c.kopen) return;
    (void)hipStreamSynchronize(c.stream);
    for (size_t i = 0; i < c.kpend.size(); i++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c.kev[i][0], c.kev[i][1]) == hipSuccess) {
            KStat &k = c.kstat[c.kpend[i].first];
            k.ms += ms;
            k.bytes += c.kpend[i].second;
            k.launches++;
        }
    }
    // a failed query must not stay behind as the calling thread's last HIP error
    (void)hipGetLastError();
    c.kpend.clear();
}

// instance layout of a context: consecutive variable ranges of the given sizes, each
// starting at a multiple of 16 so that shifted literals keep their signature bits
// (lit & 31); the padding variables have no clauses
uint64_t instance_bases(uint32_t ninst, const uint32_t *nvars, std::vector<uint32_t> &base)
{
    base.assign(ninst + 1, 0);
    uint64_t at = 0;
    for (uint32_t i = 0; i < ninst; i++) {
        base[i] = (uint32_t)at;
        at += nvars[i];
        if (i + 1 < ninst) at = (at + 15) & ~15ull;
        if (at >= (1ull << 31)) return at;
    }
    base[ninst] = (uint32_t)at;
    return at;
}

void set_instances(Ctx &c, uint32_t ninst, const uint32_t *nvars)
{
    c.ninst = ninst;
    instance_bases(ninst, nvars, c.hibase);
    c.inst_nvars.assign(nvars, nvars + ninst);
    c.imu_host.assign(ninst, 0);
    ensure(c, c.ibase, ninst + 1);
    ensure(c, c.imu, ninst + 8);
    ensure(c, c.ib0, 2ull * ninst);
    ensure(c, c.inel, ninst);
    ensure(c, c.icnt, 2ull * ninst);
    PFG_HIP(hipMemcpyAsync(c.ibase.p, c.hibase.data(), 4ull * (ninst + 1), hipMemcpyHostToDevice, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
}

// The reference simplifier's round structure (src/cpu/simplify.cpp:150-214,
// src/gpu/simplify.cu:155-234), one method per reference entry point.
struct Sigma {
    Ctx &c;
    explicit Sigma(Ctx &cc) : c(cc) {}

    void resizeCNF()
    {   // solve.h:684-688 / simplify.cu:30-46: GC every shrink_rate rounds
        const int times = c.phase + 1;
        if (times > 1 && times != c.opts.phases && c.opts.shrink_rate && (times % c.opts.shrink_rate) == 0) shrinkSimp();
    }
    void shrinkSimp() { PassTimer t(c, &c.times.gc); shrink(c); }   // live counts do not change
    void createOT() { PassTimer t(c, &c.times.ot); create_ot(c); }
    bool propagate()
    {
        PassTimer t(c, &c.times.prop);
        const uint32_t t0 = c.ntrail;
        const bool was_unsat = c.unsat;
        const bool ok = prop(c);
        proof_units(c, t0, c.ntrail);   // learnUnit of each unit the propagation enqueued
        if (!ok && !was_unsat) proof_empty(c);
        return ok;
    }
    bool LCVE()
    {
        PassTimer t(c, &c.times.lcve);
        return lcve(c) >= c.opts.lcve_min;
    }
    void sortOT() { PassTimer t(c, &c.times.sort_ot); sort_ot(c); }
    bool stop(int64_t lr) const { return (c.phase == c.opts.phases) || (lr <= c.opts.lits_min && c.phase > 2); }
    void ERE() { PassTimer t(c, &c.times.ere); c.hist_live = false; ere(c); }
    void SUB() { c.hist_live = false; if (c.opts.sub_en || c.opts.ve_en) { PassTimer t(c, &c.times.sub); sub(c); } }
    void VE() { c.hist_live = false; if (c.opts.ve_en) { PassTimer t(c, &c.times.ve); ve(c); } }
    void BCE() { c.hist_live = false; if (c.opts.bce_en) { PassTimer t(c, &c.times.bce); bce(c); } }
    void countAll(int64_t &cls, int64_t &lits, bool with_hist = false)
    {
        PassTimer t(c, &c.times.count);
        count_all(c, cls, lits, with_hist);
    }
    void filterPVs() { filter_pvs(c); }
    void nextPhase()
    {
        c.phase++;
        c.mu_inc++;
        c.mu_inc += (c.phase == c.opts.phases);
    }

    // interrupted() (src/gpu/solve.h:181): pfg_interrupt asked this call to stop. The
    // reference polls it before every round (simplify.cu:182) and before SUB, VE, BCE
    // and ERE (elimination.cu:28,55,65,76); here the call then leaves the round loop
    // and writes back what the finished passes produced.
    bool interrupted() { return c.interrupt.exchange(0) != 0; }

    int sigmifying()
    {
        c.phase = 0;
        c.mu_inc = 0;
        int64_t litsbefore = c.nliterals, diff = INT64_MAX;
        bool intr = false;
        while (litsbefore) {
            if ((intr = interrupted())) break;
            resizeCNF();
            createOT();
            if (!propagate()) return PFG_EUNSAT;
            if (!LCVE()) break;
            sortOT();
            if (stop(diff)) {
                if ((intr = interrupted())) break;
                ERE();
                break;
            }
            if ((intr = interrupted())) break;
            SUB();
            if (c.unsat) return PFG_EUNSAT;
            if ((intr = interrupted())) break;
            VE();
            if (c.unsat) return PFG_EUNSAT;
            if ((intr = interrupted())) break;
            BCE();
            int64_t nc, nl;
            countAll(nc, nl);
            filterPVs();
            c.nclauses = nc;
            c.nliterals = nl;
            diff = litsbefore - c.nliterals;
            litsbefore = c.nliterals;
            nextPhase();
        }
        if (c.ntrail > c.propagated) {
            if (!c.ot_valid) createOT();
            if (!propagate()) return PFG_EUNSAT;
        }
        int64_t nc, nl;
        countAll(nc, nl);
        c.nclauses = nc;
        c.nliterals = nl;
        shrinkSimp();
        return intr ? PFG_EINTR : PFG_OK;
    }

    // ---- batches: the same round structure, decided per instance
    // instance of global variable v (host side)
    uint32_t inst_of(uint64_t v) const
    {
        const auto it = std::lower_bound(c.hibase.begin(), c.hibase.end(), (uint32_t)v);
        const uint64_t i = (uint64_t)(it - c.hibase.begin());
        return (uint32_t)(i == 0 ? 0 : std::min<uint64_t>(i - 1, c.ninst - 1));
    }

    // contradictions reported since the last call refute their instances
    void collect_conflicts(std::vector<uint8_t> &st)
    {
        // the flags and the contradiction count in one copy (slots CNT_FLAGS .. CNT_CONF)
        uint64_t hc[CNT_CONF - CNT_FLAGS + 1];
        PFG_HIP(hipMemcpyAsync(hc, c.cnt.p + CNT_FLAGS, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
        const uint64_t h[2] = {hc[0], hc[CNT_CONF - CNT_FLAGS]};
        if (!(h[0] & FLAG_UNSAT)) return;
        if (h[1] > CONF_CAP) throw Error(PFG_EINTERNAL, "more contradictions in one pass than the conflict list holds");
        std::vector<uint64_t> vars(h[1]);
        if (h[1]) PFG_HIP(hipMemcpyAsync(vars.data(), c.cnt.p + CNT_CONF + 1, 8 * h[1], hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_FLAGS, 0, 8, c.stream));
        PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_CONF, 0, 8, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
        for (const uint64_t v : vars) st[inst_of(v)] = B_REFUTED;
        c.unsat = false;
    }

    enum { B_RUN = 0, B_DONE = 1, B_REFUTED = 2 };

    // sigmifying (src/cpu/simplify.cpp:150-214) for every instance of the batch at
    // once: an instance leaves the loop where the reference's loop would (no
    // literals, fewer than lcve_min elected, or stop() after which it runs ERE) and
    // is refuted where the reference would return UNSAT; the others go on.
    int sigmifying_batch(int32_t *status)
    {
        const uint32_t ni = c.ninst;
        std::vector<uint8_t> st(ni, B_RUN), cls(ni, 0);
        std::vector<int32_t> phase(ni, 0);
        std::vector<uint32_t> mu(ni, 0);
        std::vector<int64_t> nc, nl;
        {
            PassTimer t(c, &c.times.count);
            count_inst(c, nc, nl);
        }
        std::vector<int64_t> before = nl, diff(ni, INT64_MAX);
        std::vector<uint64_t> nel(ni);
        int round = 0;
        bool intr = false;
        for (;;) {
            if ((intr = interrupted())) break;
            bool any = false;
            for (uint32_t i = 0; i < ni; i++) {
                if (st[i] == B_RUN && before[i] == 0) st[i] = B_DONE;
                any |= st[i] == B_RUN;
            }
            if (!any) break;
            if (c.trace) {
                uint32_t run = 0;
                int64_t lits_run = 0, lits_all = 0;
                for (uint32_t i = 0; i < ni; i++) {
                    lits_all += before[i];
                    if (st[i] == B_RUN) { run++; lits_run += before[i]; }
                }
                fprintf(stderr, "[pfg] batch round %d: %u of %u instances running, literals %lld of %lld\n", round, run, ni,
                        (long long)lits_run, (long long)lits_all);
            }
            // Compaction keeps clause order, so its schedule is free: between rounds it runs
            // only when deleted images take a tenth of the pool or more (the passes skip
            // deleted clauses either way); the call ends with one as always.
            const int times = round + 1;
            if (times > 1 && c.opts.shrink_rate && (times % c.opts.shrink_rate) == 0) {
                uint64_t live = 0;
                for (uint32_t i = 0; i < ni; i++) live += 3ull * (uint64_t)nc[i] + (uint64_t)nl[i];
                if (10 * live < 9 * c.msz) shrinkSimp();
            }
            createOT();
            propagate();
            collect_conflicts(st);
            for (uint32_t i = 0; i < ni; i++) c.imu_host[i] = st[i] == B_RUN ? (uint8_t)std::min<uint32_t>(mu[i], 254u) : 0xFFu;
            {
                PassTimer t(c, &c.times.lcve);
                lcve(c);
            }
            PFG_HIP(hipMemcpyAsync(nel.data(), c.inel.p, 8ull * ni, hipMemcpyDeviceToHost, c.stream));
            PFG_HIP(hipStreamSynchronize(c.stream));
            for (uint32_t i = 0; i < ni; i++) {
                cls[i] = 0;
                if (st[i] != B_RUN) continue;
                if (nel[i] < c.opts.lcve_min) { st[i] = B_DONE; continue; }
                const bool stop = (phase[i] == c.opts.phases) || (diff[i] <= c.opts.lits_min && phase[i] > 2);
                cls[i] = stop ? 2 : 1;
            }
            uint32_t na = 0, nb = 0;
            partition_pvs(c, cls, na, nb);
            sortOT();
            if (nb) {
                uint32_t *base = c.pvs.p;
                c.pvs.p = base + na;
                c.npvs = nb;
                ERE();
                c.pvs.p = base;
            }
            c.npvs = na;
            for (uint32_t i = 0; i < ni; i++) if (cls[i] == 2) st[i] = B_DONE;
            if (na) {
                SUB();
                collect_conflicts(st);
                VE();
                collect_conflicts(st);
                BCE();
            }
            {
                PassTimer t(c, &c.times.count);
                count_inst(c, nc, nl);
            }
            for (uint32_t i = 0; i < ni; i++) {
                if (cls[i] != 1 || st[i] != B_RUN) continue;
                diff[i] = before[i] - nl[i];
                before[i] = nl[i];
                phase[i]++;
                mu[i]++;
                mu[i] += (phase[i] == c.opts.phases);
            }
            if (++round > (1 << 20)) throw Error(PFG_EINTERNAL, "batch did not finish");
        }
        if (c.ntrail > c.propagated) {
            if (!c.ot_valid) createOT();
            propagate();
            collect_conflicts(st);
        }
        int64_t tc, tl;
        countAll(tc, tl);
        c.nclauses = tc;
        c.nliterals = tl;
        shrinkSimp();
        for (uint32_t i = 0; i < ni; i++) status[i] = st[i] == B_REFUTED ? PFG_EUNSAT : st[i] == B_RUN && intr ? PFG_EINTR : PFG_OK;
        c.batch_gen++;
        return intr ? PFG_EINTR : PFG_OK;
    }
};

}  // namespace pfg

using namespace pfg;

struct pfg_ctx {
    Ctx c;
};

#define API_BEGIN                  \
    if (!ctx) return PFG_EINVAL;   \
    Ctx &c = ctx->c;               \
    try {                          \
        PFG_HIP(hipSetDevice(c.device));
// a kernel launch that failed in the call is reported here, not left behind as the
// calling thread's last HIP error
#define API_LAUNCH_CHECK                                                                            \
    {                                                                                               \
        const hipError_t le_ = hipGetLastError();                                                   \
        if (le_ != hipSuccess) throw ::pfg::Error(PFG_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(le_)); \
    }
#define API_END                                  \
    API_LAUNCH_CHECK                             \
    }                                            \
    catch (const Error &e) {                     \
        c.last_error = e.what();                 \
        return e.code;                           \
    }                                            \
    catch (const std::bad_alloc &) {             \
        c.last_error = "host allocation failed"; \
        return PFG_ENOMEM;                       \
    }                                            \
    return c.unsat ? PFG_EUNSAT : PFG_OK;
// bookkeeping calls (statistics, settings) succeed whatever the formula's state
#define API_END_OK                               \
    API_LAUNCH_CHECK                             \
    }                                            \
    catch (const Error &e) {                     \
        c.last_error = e.what();                 \
        return e.code;                           \
    }                                            \
    catch (const std::bad_alloc &) {             \
        c.last_error = "host allocation failed"; \
        return PFG_ENOMEM;                       \
    }                                            \
    return PFG_OK;

extern "C" {

void pfg_default_opts(pfg_opts *o)
{
    o->phases = 5; o->mu_pos = 32; o->mu_neg = 32; o->lcve_min = 1; o->lits_min = 500;
    o->shrink_rate = 2; o->ve_clause_limit = 100; o->ve_lbound_en = 0; o->xor_max_arity = 10;
    o->sub_limit = 3000; o->ere_limit = 3000; o->ere_max_resolvent = 1000;
    o->sub_en = 1; o->ve_en = 1; o->ere_en = 1; o->lbd_tier1 = 2; o->stable_sort = 0;
    o->bce_en = 0; o->bce_limit = 3000; o->proof_en = 0; o->ve_fun_en = 0;
    o->ve_atomic_en = 0; o->lcve_max = 0; o->lcve_clause_max = 0;
}


// the option checks of the reference simplifier (src/cpu/options.cpp:305-310; ve_en
// stands for its ve_en || ve_plus_en)
static void normalize_opts(pfg_opts &o)
{
    const bool any = o.ve_en || o.sub_en || o.bce_en;
    if (!o.phases && any) o.phases = 1;
    if (o.phases && !any) o.phases = 0;
    if (o.phases > 1 && !o.ve_en) o.phases = 1;
}

int pfg_create(pfg_ctx **out, int device, uint64_t hbm_budget, uint32_t nvars, const pfg_opts *opts)
{
    if (!out || !nvars) return PFG_EINVAL;
    *out = nullptr;
    pfg_ctx *ctx = new (std::nothrow) pfg_ctx();
    if (!ctx) return PFG_ENOMEM;
    Ctx &c = ctx->c;
    c.device = device;
    c.budget = hbm_budget;
    c.trace = getenv("PFG_TRACE") != nullptr;
    c.nvars = nvars;
    c.nvars_cap = nvars;
    c.nlits2 = 2ull * ((uint64_t)nvars + 1);
    if (opts) c.opts = *opts; else pfg_default_opts(&c.opts);
    normalize_opts(c.opts);
    try {
        PFG_HIP(hipSetDevice(device));
        pool_setup(device);
        PFG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        for (auto &e : c.ev) { PFG_HIP(hipEventCreate(&e[0])); PFG_HIP(hipEventCreate(&e[1])); }
        PFG_HIP(hipHostMalloc((void **)&c.hcnt, (CNT_CONF + 1) * sizeof(uint64_t), hipHostMallocDefault));
        ensure(c, c.cnt, CNT_CONF + 1 + CONF_CAP);
        PFG_HIP(hipMemsetAsync(c.cnt.p, 0, (CNT_CONF + 1) * 8, c.stream));
        ensure(c, c.value, c.nlits2 + 8);
        PFG_HIP(hipMemsetAsync(c.value.p, 0xFF, c.value.cap, c.stream));
        ensure(c, c.vstate, (uint64_t)nvars + 1);
        PFG_HIP(hipMemsetAsync(c.vstate.p, 0, c.vstate.cap, c.stream));
        ensure(c, c.ifrozen, (uint64_t)nvars + 1);
        PFG_HIP(hipMemsetAsync(c.ifrozen.p, 0, c.ifrozen.cap, c.stream));
        ensure(c, c.trail, (uint64_t)nvars + 1);
        ensure(c, c.ufirst, (uint64_t)nvars + 1);
        PFG_HIP(hipMemsetAsync(c.ufirst.p, 0xFF, c.ufirst.cap * 4, c.stream));
        ensure(c, c.pvs, (uint64_t)nvars + 1);
        ensure(c, c.resolved, 1024);
        set_instances(c, 1, &nvars);
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    catch (const Error &e) {
        c.last_error = e.what();
        int code = e.code;
        pfg_destroy(ctx);
        return code;
    }
    *out = ctx;
    return PFG_OK;
}

int pfg_destroy(pfg_ctx *ctx)
{
    if (!ctx) return PFG_EINVAL;
    Ctx &c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    release(c, c.mem); release(c, c.mem2); release(c, c.refs); release(c, c.refs2);
    release(c, c.ot); release(c, c.occ); release(c, c.hist); release(c, c.value); release(c, c.vstate); release(c, c.ifrozen);
    release(c, c.trail); release(c, c.resolved); release(c, c.mchunk); release(c, c.img_off); release(c, c.pvs); release(c, c.units);
    release(c, c.scores); release(c, c.eligible); release(c, c.rank); release(c, c.tmpu32a); release(c, c.tmpu32b); release(c, c.tmpu32c); release(c, c.tmpu32d);
    release(c, c.tmpu8); release(c, c.tmpu64a); release(c, c.tmpu64b); release(c, c.tmpu64c);
    release(c, c.owner); release(c, c.claim); release(c, c.plong); release(c, c.erec); release(c, c.ufirst); release(c, c.cnt);
    release(c, c.scan_tmp); release(c, c.radix_tmp);
    release(c, c.otk); release(c, c.otk2); release(c, c.otsz); release(c, c.otdirty);
    release(c, c.otrm); release(c, c.otpos); release(c, c.otcnt); release(c, c.otchk);
    release(c, c.proof); release(c, c.pev); release(c, c.pord); release(c, c.pnev); release(c, c.pbytes);
    release(c, c.pcnt); release(c, c.poff); release(c, c.pboff); release(c, c.pfirst);
    release(c, c.ibase); release(c, c.imu); release(c, c.ib0); release(c, c.inel); release(c, c.icnt); release(c, c.vorg);
    if (c.stream) (void)hipStreamSynchronize(c.stream);   // the frees are queued on the stream
    c.cache_mem.release(); c.pin_res.release(); c.pin_trail.release(); c.pin_refs.release(); c.pin_vstate.release();
    if (c.hcnt) (void)hipHostFree(c.hcnt);
    for (auto &e : c.kev) { (void)hipEventDestroy(e[0]); (void)hipEventDestroy(e[1]); }
    for (auto &e : c.ev) {
        if (e[0]) (void)hipEventDestroy(e[0]);
        if (e[1]) (void)hipEventDestroy(e[1]);
    }
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
    return PFG_OK;
}

int pfg_upload_cnf(pfg_ctx *ctx, const uint32_t *lits, const uint64_t *offs, uint64_t ncls, const uint32_t *hdr0)
{
    if (!ctx || (!lits && ncls) || !offs) return PFG_EINVAL;
    API_BEGIN
    const uint64_t nlits = offs[ncls];
    for (uint64_t i = 0; i < nlits; i++)
        if ((lits[i] >> 1) == 0 || (lits[i] >> 1) > c.nvars) throw Error(PFG_EINVAL, "literal out of range");
    PassTimer t(c, &c.times.upload);
    uint32_t *dl = nullptr, *dh = nullptr;
    uint64_t *doff = nullptr;
    dalloc_raw(c, (void **)&dl, nlits * 4 + 4);
    dalloc_raw(c, (void **)&doff, (ncls + 1) * 8);
    PFG_HIP(hipMemcpyAsync(dl, lits, nlits * 4, hipMemcpyHostToDevice, c.stream));
    PFG_HIP(hipMemcpyAsync(doff, offs, (ncls + 1) * 8, hipMemcpyHostToDevice, c.stream));
    if (hdr0) {
        dalloc_raw(c, (void **)&dh, ncls * 4 + 4);
        PFG_HIP(hipMemcpyAsync(dh, hdr0, ncls * 4, hipMemcpyHostToDevice, c.stream));
    }
    upload(c, dl, doff, ncls, nlits, dh);
    c.batch_gen++;
    PFG_HIP(hipStreamSynchronize(c.stream));
    dfree_raw(c, dl, nlits * 4 + 4);
    dfree_raw(c, doff, (ncls + 1) * 8);
    if (dh) dfree_raw(c, dh, ncls * 4 + 4);
    API_END
}

int pfg_upload_cnf_device(pfg_ctx *ctx, const uint32_t *d_lits, const uint64_t *d_offs, uint64_t ncls, uint64_t nlits)
{
    if (!ctx || !d_offs) return PFG_EINVAL;
    API_BEGIN
    PassTimer t(c, &c.times.upload);
    upload(c, d_lits, d_offs, ncls, nlits, nullptr);
    c.batch_gen++;
    PFG_HIP(hipStreamSynchronize(c.stream));
    API_END
}

int pfg_enqueue_units(pfg_ctx *ctx, const uint32_t *lits, uint32_t n)
{
    if (!ctx || (n && !lits)) return PFG_EINVAL;
    API_BEGIN
    for (uint32_t i = 0; i < n; i++)
        if ((lits[i] >> 1) == 0 || (lits[i] >> 1) > c.nvars) throw Error(PFG_EINVAL, "literal out of range");
    enqueue_units_host(c, lits, n);
    API_END
}

int pfg_ifreeze(pfg_ctx *ctx, const uint32_t *vars, uint32_t n, int on)
{
    if (!ctx || (n && !vars)) return PFG_EINVAL;
    API_BEGIN
    for (uint32_t i = 0; i < n; i++)
        if (vars[i] == 0 || vars[i] > c.nvars) throw Error(PFG_EINVAL, "variable out of range");
    set_ifrozen(c, vars, n, on != 0);
    API_END
}

int pfg_sigmify(pfg_ctx *ctx)
{
    API_BEGIN
    double t0 = 0;
    int rc;
    {
        PassTimer t(c, &t0);
        Sigma s(c);
        rc = s.sigmifying();
    }
    resolve_ktimers(c);
    c.times.total += t0;
    if (rc != PFG_OK) return rc;
    API_END
}

int pfg_resize_cnf(pfg_ctx *ctx) { API_BEGIN Sigma(c).resizeCNF(); API_END }
int pfg_create_ot(pfg_ctx *ctx) { API_BEGIN Sigma(c).createOT(); API_END }
int pfg_reduce_ot(pfg_ctx *ctx) { API_BEGIN if (!c.ot_valid) throw Error(PFG_EINVAL, "no occurrence table"); reduce_ot(c); API_END }
int pfg_prop(pfg_ctx *ctx) { API_BEGIN if (!c.ot_valid && c.ntrail > c.propagated) throw Error(PFG_EINVAL, "no occurrence table"); Sigma(c).propagate(); API_END }

int pfg_lcve(pfg_ctx *ctx, uint32_t *npvs)
{
    API_BEGIN
    if (!c.ot_valid) throw Error(PFG_EINVAL, "no occurrence table");
    Sigma(c).LCVE();
    if (npvs) *npvs = c.npvs;
    API_END
}

int pfg_sort_ot(pfg_ctx *ctx) { API_BEGIN Sigma(c).sortOT(); API_END }
int pfg_sub(pfg_ctx *ctx) { API_BEGIN Sigma(c).SUB(); API_END }
int pfg_ve(pfg_ctx *ctx) { API_BEGIN Sigma(c).VE(); API_END }
int pfg_bce(pfg_ctx *ctx) { API_BEGIN Sigma(c).BCE(); API_END }
int pfg_ere(pfg_ctx *ctx) { API_BEGIN Sigma(c).ERE(); API_END }

int pfg_count(pfg_ctx *ctx, int64_t *cls, int64_t *lits)
{
    API_BEGIN
    int64_t a, b;
    Sigma(c).countAll(a, b);
    c.nclauses = a;
    c.nliterals = b;
    if (cls) *cls = a;
    if (lits) *lits = b;
    API_END
}

int pfg_filter_pvs(pfg_ctx *ctx) { API_BEGIN Sigma(c).filterPVs(); API_END }
int pfg_gc(pfg_ctx *ctx) { API_BEGIN Sigma(c).shrinkSimp(); API_END }
int pfg_next_phase(pfg_ctx *ctx) { API_BEGIN Sigma(c).nextPhase(); API_END }

int pfg_stop(pfg_ctx *ctx, int64_t lr, int *stop)
{
    API_BEGIN
    if (stop) *stop = Sigma(c).stop(lr) ? 1 : 0;
    API_END
}

int pfg_hist(pfg_ctx *ctx, uint32_t *h)
{
    API_BEGIN
    PassTimer t(c, &c.times.hist);
    hist(c);
    if (h) {
        PFG_HIP(hipMemcpyAsync(h, c.hist.p, c.nlits2 * 4, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    API_END
}

int pfg_create_batch(pfg_ctx **out, int device, uint64_t hbm_budget, uint32_t ninst, const uint32_t *nvars,
                     const pfg_opts *opts)
{
    if (!out || !ninst || !nvars) return PFG_EINVAL;
    if (opts && opts->proof_en) return PFG_EINVAL;   // the proof is one instance's log
    // a batch groups each instance's model records from the elected order
    if (opts && opts->ve_atomic_en) return PFG_EINVAL;
    for (uint32_t i = 0; i < ninst; i++)
        if (!nvars[i]) return PFG_EINVAL;
    std::vector<uint32_t> base;
    const uint64_t total = instance_bases(ninst, nvars, base);
    if (total >= (1ull << 31)) return PFG_EINVAL;
    int rc = pfg_create(out, device, hbm_budget, (uint32_t)total, opts);
    if (rc != PFG_OK) return rc;
    pfg_ctx *ctx = *out;
    API_BEGIN
    set_instances(c, ninst, nvars);
    c.batch = true;
    API_END
}

int pfg_batch_relayout(pfg_ctx *ctx, uint32_t ninst, const uint32_t *nvars)
{
    if (!ctx || !ninst || !nvars) return PFG_EINVAL;
    API_BEGIN
    if (!c.batch) return PFG_EINVAL;
    for (uint32_t i = 0; i < ninst; i++)
        if (!nvars[i]) return PFG_EINVAL;
    std::vector<uint32_t> base;
    const uint64_t total = instance_bases(ninst, nvars, base);
    if (total > c.nvars_cap) return PFG_EINVAL;
    c.nvars = (uint32_t)total;
    c.nlits2 = 2ull * (total + 1);
    set_instances(c, ninst, nvars);
    // forget the previous clause database and variable state
    PFG_HIP(hipMemsetAsync(c.value.p, 0xFF, c.value.cap, c.stream));
    PFG_HIP(hipMemsetAsync(c.vstate.p, 0, c.vstate.cap, c.stream));
    PFG_HIP(hipMemsetAsync(c.ifrozen.p, 0, c.ifrozen.cap, c.stream));
    // every counter but CNT_SORTX (a path count, kept until pfg_reset_stats)
    PFG_HIP(hipMemsetAsync(c.cnt.p, 0, CNT_SORTX * 8, c.stream));
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_SORTX + 1, 0, (CNT_CONF - CNT_SORTX) * 8, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    c.ntrail = c.propagated = 0;
    c.nres = c.nchunk = 0;
    c.npvs = 0;
    c.msz = c.nrefs = 0;
    c.phase = 0;
    c.mu_inc = 0;
    c.nclauses = c.nliterals = 0;
    c.unsat = false;
    c.ot_valid = c.hist_live = false;
    c.otk_valid = false;
    c.batch_gen++;
    API_END
}

int pfg_batch_ninst(pfg_ctx *ctx, uint32_t *ninst)
{
    if (!ctx || !ninst) return PFG_EINVAL;
    *ninst = ctx->c.ninst;
    return PFG_OK;
}

int pfg_upload_batch(pfg_ctx *ctx, const uint32_t *lits, const uint64_t *offs, const uint64_t *inst_first)
{
    if (!ctx || !offs || !inst_first) return PFG_EINVAL;
    Ctx &c0 = ctx->c;
    const uint64_t ncls = inst_first[c0.ninst];
    std::vector<uint32_t> g(offs[ncls]);
    for (uint32_t i = 0; i < c0.ninst; i++) {
        if (inst_first[i + 1] < inst_first[i]) return PFG_EINVAL;
        const uint32_t nv = c0.inst_nvars[i], shift = 2u * c0.hibase[i];
        for (uint64_t k = offs[inst_first[i]]; k < offs[inst_first[i + 1]]; k++) {
            const uint32_t l = lits[k];
            if ((l >> 1) == 0 || (l >> 1) > nv) return PFG_EINVAL;
            g[k] = l + shift;
        }
    }
    return pfg_upload_cnf(ctx, g.data(), offs, ncls, nullptr);
}

int pfg_sigmify_batch(pfg_ctx *ctx, int32_t *status)
{
    if (!ctx || !status) return PFG_EINVAL;
    API_BEGIN
    double t0 = 0;
    int rc;
    {
        PassTimer t(c, &t0);
        Sigma s(c);
        rc = s.sigmifying_batch(status);
    }
    c.times.total += t0;
    if (rc != PFG_OK) return rc;
    API_END
}

// Host copy of the batch state for per-instance downloads, rebuilt once per batch:
// clause images and model records grouped by instance on the device, then pinned
// downloads; units are bucketed on the host. The grouped images stay in mem2 for exports.
static void batch_cache(Ctx &c)
{
    if (c.cache_gen == c.batch_gen) return;
    // clause images and model records grouped by instance on the device, then pinned
    // downloads; units are bucketed on the host
    const auto tq0 = std::chrono::steady_clock::now();
    ensure(c, c.mem2, c.msz + 1);
    batch_images(c, c.mem2.p, c.cache_start, c.cache_cstart);
    const uint64_t nw = c.cache_start[c.ninst];
    if (c.batch) {
        ensure(c, c.tmpu32d, c.nres + 1);
        batch_records(c, c.tmpu32d.p, c.cache_rstart);
    }
    else   // one instance: the stack as it is
        c.cache_rstart.assign({0, c.nres});
    const uint64_t nr = c.cache_rstart[c.ninst];
    const auto tq1 = std::chrono::steady_clock::now();
    c.cache_mem.ensure(nw + 1);
    c.pin_res.ensure(nr + 1);
    c.pin_trail.ensure((uint64_t)c.ntrail + 1);
    c.pin_vstate.ensure((uint64_t)c.nvars + 1);
    const auto tq2 = std::chrono::steady_clock::now();
    if (nw) PFG_HIP(hipMemcpyAsync(c.cache_mem.p, c.mem2.p, nw * 4, hipMemcpyDeviceToHost, c.stream));
    if (nr)
        PFG_HIP(hipMemcpyAsync(c.pin_res.p, c.batch ? c.tmpu32d.p : c.resolved.p, nr * 4, hipMemcpyDeviceToHost,
                               c.stream));
    PFG_HIP(hipMemcpyAsync(c.pin_vstate.p, c.vstate.p, (uint64_t)c.nvars + 1, hipMemcpyDeviceToHost, c.stream));
    if (c.ntrail) PFG_HIP(hipMemcpyAsync(c.pin_trail.p, c.trail.p, 4ull * c.ntrail, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    const auto tq3 = std::chrono::steady_clock::now();
    const uint32_t *trail = c.pin_trail.p;
    const uint64_t ntr = c.ntrail;
    Sigma S(c);
    c.cache_trail.assign(c.ninst, {});
    for (uint64_t k = 0; k < ntr; k++) c.cache_trail[S.inst_of(trail[k] >> 1)].push_back(trail[k]);
    c.cache_gen = c.batch_gen;
    if (c.trace) {
        const auto tq4 = std::chrono::steady_clock::now();
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        fprintf(stderr, "[pfg] batch results: group %.2f ms, pin %.2f ms, copy %.2f ms (%llu words, %llu records), "
                        "units %.2f ms\n", ms(tq0, tq1), ms(tq1, tq2), ms(tq2, tq3), (unsigned long long)nw,
                (unsigned long long)nr, ms(tq3, tq4));
    }
}

// one instance's part of the batch state, literals in the instance's own numbering:
// its clauses (SCLAUSE images, batch order), model records, variable states (nvars_i + 1
// bytes) and units; a NULL buffer only queries the size
int pfg_batch_download(pfg_ctx *ctx, uint32_t inst, uint32_t *words, uint64_t *nwords, uint64_t *ncls, uint32_t *resolved,
                       uint64_t *nres, uint8_t *vstate, uint32_t *trail, uint32_t *ntrail)
{
    API_BEGIN
    if (inst >= c.ninst) return PFG_EINVAL;
    batch_cache(c);
    const uint32_t lo = c.hibase[inst], shift = 2u * lo;
    const uint64_t w0 = c.cache_start[inst], w1 = c.cache_start[inst + 1];
    if (words && w1 > w0) std::memcpy(words, c.cache_mem.p + w0, (w1 - w0) * 4);
    if (nwords) *nwords = w1 - w0;
    if (ncls) *ncls = c.cache_cstart[inst + 1] - c.cache_cstart[inst];
    const uint64_t r0 = c.cache_rstart[inst], r1 = c.cache_rstart[inst + 1];
    if (resolved && r1 > r0) std::memcpy(resolved, c.pin_res.p + r0, (r1 - r0) * 4);
    if (nres) *nres = r1 - r0;
    if (vstate) {
        vstate[0] = 0;
        std::memcpy(vstate + 1, c.pin_vstate.p + lo + 1, c.inst_nvars[inst]);
    }
    if (trail) for (size_t k = 0; k < c.cache_trail[inst].size(); k++) trail[k] = c.cache_trail[inst][k] - shift;
    if (ntrail) *ntrail = (uint32_t)c.cache_trail[inst].size();
    API_END
}

// pfg_batch_download's clause images and model records of one instance, in place in the
// pinned cache (valid until the batch changes)
int pfg_batch_view(pfg_ctx *ctx, uint32_t inst, const uint32_t **words, uint64_t *nwords, uint64_t *ncls,
                   const uint32_t **resolved, uint64_t *nres)
{
    API_BEGIN
    if (inst >= c.ninst) return PFG_EINVAL;
    batch_cache(c);
    const uint64_t w0 = c.cache_start[inst], w1 = c.cache_start[inst + 1];
    const uint64_t r0 = c.cache_rstart[inst], r1 = c.cache_rstart[inst + 1];
    if (words) *words = c.cache_mem.p + w0;
    if (nwords) *nwords = w1 - w0;
    if (ncls) *ncls = c.cache_cstart[inst + 1] - c.cache_cstart[inst];
    if (resolved) *resolved = c.pin_res.p + r0;
    if (nres) *nres = r1 - r0;
    API_END
}

// one instance's clauses as CSR in caller-owned device memory, in the instance's
// numbering (clause order of the download); NULL buffers only query the sizes
int pfg_batch_export(pfg_ctx *ctx, uint32_t inst, uint32_t *d_lits, uint64_t *d_offs, uint64_t *nlits, uint64_t *ncls)
{
    API_BEGIN
    if (inst >= c.ninst) return PFG_EINVAL;
    batch_cache(c);
    const uint64_t c0 = c.cache_cstart[inst], nc = c.cache_cstart[inst + 1] - c0;
    const uint64_t nl = (c.cache_start[inst + 1] - c.cache_start[inst]) - 3 * nc;
    if (nlits) *nlits = nl;
    if (ncls) *ncls = nc;
    if (d_offs) batch_export(c, c.mem2.p, c0, nc, nl, d_lits, d_offs);
    API_END
}

int pfg_batch_export_noalloc(pfg_ctx *ctx, uint32_t inst, uint32_t *d_lits, uint64_t *d_offs, uint64_t *nlits,
                             uint64_t *ncls)
{
    API_BEGIN
    batch_export_noalloc(c, inst, d_lits, d_offs, nlits, ncls);
    API_END
}

int pfg_download_mapped(pfg_ctx *ctx, uint32_t *words, uint64_t *nwords, uint64_t *ncls, uint32_t *vorg,
                        uint32_t *nvars_new, uint32_t *trail, uint32_t *ntrail, uint8_t *vstate)
{
    API_BEGIN
    if (c.batch) throw Error(PFG_EINVAL, "variable mapping is for single-instance contexts");
    if (c.unsat) return PFG_EUNSAT;
    std::vector<uint32_t> w, vo, tr;
    std::vector<uint8_t> vs;
    uint64_t nc = 0;
    uint32_t nv = 0;
    vmap_download(c, w, nc, vo, nv, tr, vs);
    if (words && !w.empty()) std::memcpy(words, w.data(), w.size() * 4);
    if (nwords) *nwords = w.size();
    if (ncls) *ncls = nc;
    if (vorg) std::memcpy(vorg, vo.data(), vo.size() * 4);
    if (nvars_new) *nvars_new = nv;
    if (trail && !tr.empty()) std::memcpy(trail, tr.data(), tr.size() * 4);
    if (ntrail) *ntrail = (uint32_t)tr.size();
    if (vstate) std::memcpy(vstate, vs.data(), vs.size());
    API_END
}

int pfg_download_cnf(pfg_ctx *ctx, uint32_t *words, uint64_t *nwords, uint64_t *ncls)
{
    API_BEGIN
    std::vector<uint64_t> refs(c.nrefs);
    std::vector<uint32_t> mem(c.msz);
    if (c.nrefs) PFG_HIP(hipMemcpyAsync(refs.data(), c.refs.p, c.nrefs * 8, hipMemcpyDeviceToHost, c.stream));
    if (c.msz) PFG_HIP(hipMemcpyAsync(mem.data(), c.mem.p, c.msz * 4, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    uint64_t w = 0;
    for (uint64_t i = 0; i < c.nrefs; i++) {
        const uint32_t *cl = mem.data() + refs[i];
        const uint64_t n = 3 + (uint64_t)cl[2];
        if (words) std::memcpy(words + w, cl, n * 4);
        w += n;
    }
    if (nwords) *nwords = w;
    if (ncls) *ncls = c.nrefs;
    API_END
}

int pfg_download_trail(pfg_ctx *ctx, uint32_t *lits, uint32_t *n)
{
    API_BEGIN
    if (lits && c.ntrail) {
        PFG_HIP(hipMemcpyAsync(lits, c.trail.p, c.ntrail * 4ull, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    if (n) *n = c.ntrail;
    API_END
}

int pfg_download_proof(pfg_ctx *ctx, uint8_t *bytes, uint64_t *n)
{
    API_BEGIN
    if (bytes && c.nproof) {
        PFG_HIP(hipMemcpyAsync(bytes, c.proof.p, c.nproof, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    if (n) *n = c.nproof;
    API_END_OK
}

int pfg_download_resolved(pfg_ctx *ctx, uint32_t *words, uint64_t *n)
{
    API_BEGIN
    if (words && c.nres) {
        PFG_HIP(hipMemcpyAsync(words, c.resolved.p, c.nres * 4, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    if (n) *n = c.nres;
    API_END
}

int pfg_download_vstate(pfg_ctx *ctx, uint8_t *vstate)
{
    API_BEGIN
    if (!vstate) throw Error(PFG_EINVAL, "null buffer");
    PFG_HIP(hipMemcpyAsync(vstate, c.vstate.p, (uint64_t)c.nvars + 1, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    API_END
}

int pfg_download_pvs(pfg_ctx *ctx, uint32_t *pvs, uint32_t *n)
{
    API_BEGIN
    if (pvs && c.npvs) {
        PFG_HIP(hipMemcpyAsync(pvs, c.pvs.p, c.npvs * 4ull, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
    }
    if (n) *n = c.npvs;
    API_END
}

int pfg_download_ot_list(pfg_ctx *ctx, uint32_t lit, int64_t *idx, uint32_t *n)
{
    API_BEGIN
    if (!c.ot_valid || lit >= c.nlits2) throw Error(PFG_EINVAL, "no occurrence table / bad literal");
    OL o;
    PFG_HIP(hipMemcpyAsync(&o, c.ot.p + lit, sizeof(OL), hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    if (n) *n = o.sz;
    if (idx && o.sz) {
        std::vector<uint64_t> refs(c.nrefs), ent(o.sz);
        PFG_HIP(hipMemcpyAsync(refs.data(), c.refs.p, c.nrefs * 8, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipMemcpyAsync(ent.data(), c.occ.p + o.off, o.sz * 8ull, hipMemcpyDeviceToHost, c.stream));
        PFG_HIP(hipStreamSynchronize(c.stream));
        for (uint32_t k = 0; k < o.sz; k++) {
            auto it = std::lower_bound(refs.begin(), refs.end(), ent[k]);
            idx[k] = (it != refs.end() && *it == ent[k]) ? (int64_t)(it - refs.begin()) : -1;
        }
    }
    API_END
}

int pfg_kernel_stats(pfg_ctx *ctx, pfg_kstat *out, int max, int *n)
{
    API_BEGIN
    resolve_ktimers(c);
    int k = 0;
    for (int i = 0; i < K_NUM && k < max; i++) {
        if (!c.kstat[i].launches) continue;
        if (out) {
            out[k].name = kernel_names[i];
            out[k].kernel = kernel_symbols[i];
            out[k].ms = c.kstat[i].ms;
            out[k].launches = c.kstat[i].launches;
            out[k].bytes = c.kstat[i].bytes;
        }
        k++;
    }
    if (n) *n = k;
    API_END_OK
}

int pfg_set_stage_words(pfg_ctx *ctx, uint32_t words)
{
    API_BEGIN
    c.stage_words = words;
    API_END_OK
}

int pfg_set_accounting(pfg_ctx *ctx, int on)
{
    API_BEGIN
    c.account = on != 0;
    API_END_OK
}

int pfg_reset_stats(pfg_ctx *ctx)
{
    API_BEGIN
    resolve_ktimers(c);
    for (auto &k : c.kstat) k = KStat{};
    c.times = pfg_times{};
    for (auto &p : c.path) p = 0;
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_SORTX, 0, 8, c.stream));
    API_END_OK
}

int pfg_reset(pfg_ctx *ctx)
{
    API_BEGIN
    PFG_HIP(hipMemsetAsync(c.value.p, 0xFF, c.value.cap, c.stream));
    PFG_HIP(hipMemsetAsync(c.vstate.p, 0, c.vstate.cap, c.stream));
    PFG_HIP(hipMemsetAsync(c.ifrozen.p, 0, c.ifrozen.cap, c.stream));
    // every counter but CNT_SORTX (a path count, kept until pfg_reset_stats)
    PFG_HIP(hipMemsetAsync(c.cnt.p, 0, CNT_SORTX * 8, c.stream));
    PFG_HIP(hipMemsetAsync(c.cnt.p + CNT_SORTX + 1, 0, (CNT_CONF - CNT_SORTX) * 8, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    c.ntrail = c.propagated = 0;
    c.nres = c.nchunk = 0;
    c.npvs = 0;
    c.msz = 0;
    c.nrefs = 0;
    c.phase = 0;
    c.mu_inc = 0;
    c.nclauses = c.nliterals = 0;
    c.unsat = false;
    c.ot_valid = false;
    c.hist_live = false;
    c.otk_valid = false;
    c.nproof = 0;
    c.proof_closed = false;
    c.has_vorg = false;
    API_END
}

int pfg_times_get(pfg_ctx *ctx, pfg_times *t)
{
    if (!ctx || !t) return PFG_EINVAL;
    *t = ctx->c.times;
    return PFG_OK;
}

int pfg_is_unsat(pfg_ctx *ctx) { return ctx && ctx->c.unsat ? 1 : 0; }

int pfg_set_tuning(pfg_ctx *ctx, const pfg_tuning *t)
{
    if (!ctx || !t) return PFG_EINVAL;
    Ctx &c = ctx->c;
    if (t->ot_build < 0 || t->ot_build > 3 || t->lcve_graph < 0 || t->lcve_graph > 1 || !t->clause_grid ||
        t->clause_grid > (1u << 20) || t->init_grid > (1u << 20) || !(t->ve_arena == 4096 || t->ve_arena == 8192) ||
        !(t->sub_group == 8 || t->sub_group == 16 || t->sub_group == 32 || t->sub_group == 64)) {
        c.last_error = "tuning value out of range";
        return PFG_EINVAL;
    }
    c.tune.ot_build = t->ot_build;
    c.tune.clause_grid = (t->clause_grid + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    c.tune.sub_group = t->sub_group;
    c.tune.lcve_graph = t->lcve_graph;
    c.tune.ve_arena = t->ve_arena;
    c.tune.init_grid = (t->init_grid + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    c.stage_words = t->stage_words;
    return PFG_OK;
}

int pfg_get_tuning(pfg_ctx *ctx, pfg_tuning *t)
{
    if (!ctx || !t) return PFG_EINVAL;
    const Ctx &c = ctx->c;
    t->ot_build = c.tune.ot_build;
    t->clause_grid = c.tune.clause_grid;
    t->sub_group = c.tune.sub_group;
    t->stage_words = c.stage_words;
    t->lcve_graph = c.tune.lcve_graph;
    t->ve_arena = c.tune.ve_arena;
    t->init_grid = c.tune.init_grid;
    return PFG_OK;
}

int pfg_interrupt(pfg_ctx *ctx)
{
    if (!ctx) return PFG_EINVAL;
    ctx->c.interrupt.store(1);
    return PFG_OK;
}

int pfg_set_vorg(pfg_ctx *ctx, const uint32_t *vorg, uint32_t n)
{
    if (!ctx) return PFG_EINVAL;
    API_BEGIN
    if (c.batch) throw Error(PFG_EINVAL, "a variable mapping applies to single-instance contexts");
    if (!vorg) { c.has_vorg = false; return c.unsat ? PFG_EUNSAT : PFG_OK; }
    if ((uint64_t)n != (uint64_t)c.nvars + 1) throw Error(PFG_EINVAL, "vorg needs nvars + 1 entries");
    for (uint32_t v = 1; v < n; v++)
        if (vorg[v] == 0 || vorg[v] >= (1u << 31)) throw Error(PFG_EINVAL, "vorg entry out of range");
    ensure(c, c.vorg, n);
    PFG_HIP(hipMemcpyAsync(c.vorg.p, vorg, 4ull * n, hipMemcpyHostToDevice, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    c.has_vorg = true;
    API_END_OK
}

int pfg_path_counts(pfg_ctx *ctx, uint64_t *counts, int max, int *n)
{
    if (!ctx || (max > 0 && !counts)) return PFG_EINVAL;
    API_BEGIN
    // counted on the device (no host read in the pass itself)
    uint64_t sx = 0;
    PFG_HIP(hipMemcpyAsync(&sx, c.cnt.p + CNT_SORTX, 8, hipMemcpyDeviceToHost, c.stream));
    PFG_HIP(hipStreamSynchronize(c.stream));
    c.path[PATH_SORT_EXACT] = sx;
    const int k = std::min<int>(max, PATH_NUM);
    for (int i = 0; i < k; i++) counts[i] = c.path[i];
    if (n) *n = PATH_NUM;
    API_END_OK
}

const char *pfg_last_error(pfg_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

}  // extern "C"
