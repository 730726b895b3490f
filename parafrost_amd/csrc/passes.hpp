// Host launchers of the simplifier passes (one translation unit per pass family).
#pragma once
#include "ctx.hpp"
#include <vector>

namespace pfg {

// cnf.hip
void upload(Ctx &c, const uint32_t *d_lits, const uint64_t *d_offs, uint64_t ncls, uint64_t nlits,
            const uint32_t *d_hdr0);
void hist(Ctx &c);                       // c.hist = live literal counts
void create_ot(Ctx &c);                  // histogram, list layout, fill in clause order
void reduce_ot(Ctx &c);
bool prop(Ctx &c);                       // false on conflict
void count_all(Ctx &c, int64_t &cls, int64_t &lits, bool with_hist = false);
void shrink(Ctx &c);                     // stable compaction of live clauses
void filter_pvs(Ctx &c);
void sort_ot(Ctx &c);                    // PV lists by the reference clause key
uint64_t compact_u32(Ctx &c, const uint32_t *in, const uint32_t *flags, uint64_t n, uint32_t *out);
void enqueue_units_host(Ctx &c, const uint32_t *lits, uint32_t n);
// batches
void count_inst(Ctx &c, std::vector<int64_t> &cls, std::vector<int64_t> &lits);
void partition_pvs(Ctx &c, const std::vector<uint8_t> &cls, uint32_t &na, uint32_t &nb);
void batch_images(Ctx &c, uint32_t *out, std::vector<uint64_t> &start, std::vector<uint64_t> &cstart);
void batch_export(Ctx &c, const uint32_t *img, uint64_t c0, uint64_t ncls, uint64_t nlits, uint32_t *d_lits,
                  uint64_t *d_offs);
void batch_records(Ctx &c, uint32_t *out, std::vector<uint64_t> &start);
void set_ifrozen(Ctx &c, const uint32_t *vars, uint32_t n, bool on);
void vmap_download(Ctx &c, std::vector<uint32_t> &words, uint64_t &ncls, std::vector<uint32_t> &vorg, uint32_t &nvars_new,
                   std::vector<uint32_t> &trail, std::vector<uint8_t> &vstate);
void merge_units(Ctx &c, const uint32_t *d_units, uint64_t n);  // ordered dedupe into the trail
// algorithmic bytes of reading the occurrence lists of elected variables pvs[b, e):
// 16 per list header + 8 per entry (+ 12 + 4|c| per listed clause when with_clauses);
// 0 unless accounting is on (synchronizes the stream)
double pv_bytes(Ctx &c, uint32_t b, uint32_t e, bool with_clauses);

// sigma.cpp
void set_instances(Ctx &c, uint32_t ninst, const uint32_t *nvars);
uint64_t instance_bases(uint32_t ninst, const uint32_t *nvars, std::vector<uint32_t> &base);

// lcve.hip
uint32_t lcve(Ctx &c);                   // returns #elected
// sub.hip
void sub(Ctx &c);
// ve.hip
void ve(Ctx &c);
void bce(Ctx &c);                        // blocked clause elimination (bce_en)
// ere.hip
void ere(Ctx &c);
// proof.hip (binary DRAT, opts.proof_en): a pass logs events through the ProofLog
// of proof_begin (capacity in events; null log when the proof is off), and
// proof_end orders, decides (decide: units are assigned after the pass) and
// appends them; proof_units appends the units trail[from, to) of a propagation
struct ProofLog;
ProofLog proof_begin(Ctx &c, uint32_t np, uint64_t cap);
void proof_end(Ctx &c, uint32_t np, bool decide);
void proof_units(Ctx &c, uint32_t from, uint32_t to);
void proof_empty(Ctx &c);                  // the empty clause; nothing is written after it

// pfg_batch_export without allocation (noalloc.hip)
void batch_export_noalloc(Ctx &c, uint32_t inst, uint32_t *d_lits, uint64_t *d_offs, uint64_t *nlits, uint64_t *ncls);

}  // namespace pfg
