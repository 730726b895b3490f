/*
 * ref_harness.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * Drives the reference CPU simplifier (ParaFROST src/cpu, compiled from its own
 * sources by oracle/Makefile into oracle/_ref/) through the exact step sequence of
 * ParaFROST::sigmifying() (src/cpu/simplify.cpp:150-214) and dumps the simplifier
 * state so the C restatement (oracle/pfo.c) can be compared with it word for word.
 *
 * usage: pfref <input.cnf> <out.bin> [reference options, e.g. --phases=3]
 * env PFREF_TRACE=<prefix>: also dump the state after every pass to
 *   <prefix>.<round>.<pass>.bin  (pass in prop, lcve, sort, ere, sub, ve)
 * env PFREF_MAP=1: after the simplifier, run the reference's variable mapping with
 *   write-back (ParaFROST::map(true), src/cpu/vmap.cpp:64-131; newBeginning,
 *   simplify.cpp:247-265) and write a mapped dump instead:
 *   u32 'PFRM', u32 nvars_new, u64 nclauses, per clause (u32 size, lits[]),
 *   u32 nvorg, vorg[] (new variable -> original), u32 ntrail, trail[],
 *   u8 vstate[nvars_new + 1]
 * env PFREF_SECOND=1 (with PFREF_MAP): after the mapped dump, finish the write-back
 *   as sigmifying() does (rebuildWT, retrail; simplify.cpp:206-208), lower the phases
 *   as sigmify() does after a call (simplify.cpp:125-128) and run the simplifier a
 *   second time on the mapped formula; its state goes to <out.bin>.2 in the dump
 *   format below, followed by u64 = the model-stack size before the second call
 *   (its records are the stack's suffix, written through the mapped vorg), and the
 *   proof length at its start goes to PFREF_PROOF_MARK's file with ".2" appended
 * env PFREF_PROOF_MARK=<file>: with -proof --proofout=<path>, write the proof's
 *   length in bytes (u64) when the simplifier starts, so that the simplifier's part
 *   of the proof can be told from the parser's and shrinkTop's lines
 * env PFREF_TIME=1: print "PFREF_TIME <seconds>" to stderr, the wall time of the
 *   simplifier (rootify .. shrinkSimp, or .. the refutation; not the parse, not the
 *   dump of a simplified run; a refuted run's time includes its dump)
 * env PFREF_FREEZE=<file>: u32 variables (little endian) frozen as by ifreeze()
 *   (src/cpu/assume.cpp:45-52) with the solver in incremental mode, so that
 *   iassumed() (src/cpu/solve.h:864-867) holds for them during LCVE
 *
 * Dump format (little endian):
 *   u32 'PFRD', u32 nvars, u64 nclauses, u64 words, words[] (per clause: hdr0, sig, size, lits)
 *   u32 ntrail, trail[]; u64 nresolved, resolved[]; u8 vstate[nvars+1]; u32 npvs, pvs[]
 *   u32 nlists, then per PV literal (2v, 2v+1 for each PV in order): u32 n, i64 scnf index[]
 *   u32 status (0 ok, 1 unsat)
 */
#include "solve.h"
#include "control.h"
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <sys/stat.h>
#include <string>
#include <vector>

using namespace pFROST;

bool quiet_en = false;
bool competition_en = false;
int verbose = -1;

static void put32(FILE* f, uint32_t v) { fwrite(&v, 4, 1, f); }
static void put64(FILE* f, uint64_t v) { fwrite(&v, 8, 1, f); }

static uint32_t hdr0(SCLAUSE& c)
{
	return (uint32_t(c.status()) & 3u) | (uint32_t(c.molten()) << 2) | (uint32_t(c.added()) << 3)
		| ((uint32_t(c.usage()) & 3u) << 4) | (uint32_t(c.lbd()) << 6);
}

static void dump(ParaFROST* s, const std::string& path, bool with_lists, uint32_t status)
{
	FILE* f = fopen(path.c_str(), "wb");
	if (!f) { perror(path.c_str()); exit(3); }
	put32(f, 0x44524650u);
	put32(f, inf.maxVar);
	SCNF& scnf = s->scnf;
	uint64_t words = 0;
	for (uint32_t i = 0; i < scnf.size(); i++) words += 3 + uint64_t(scnf[i]->size());
	put64(f, scnf.size());
	put64(f, words);
	std::unordered_map<const void*, int64_t> index;
	for (uint32_t i = 0; i < scnf.size(); i++) {
		SCLAUSE& c = *scnf[i];
		index[scnf[i]] = i;
		put32(f, hdr0(c));
		put32(f, c.sig());
		put32(f, uint32_t(c.size()));
		for (int k = 0; k < c.size(); k++) put32(f, c[k]);
	}
	put32(f, s->trail.size());
	for (uint32_t i = 0; i < s->trail.size(); i++) put32(f, s->trail[i]);
	put64(f, s->model.resolved.size());
	for (uint32_t i = 0; i < s->model.resolved.size(); i++) put32(f, s->model.resolved[i]);
	for (uint32_t v = 0; v <= inf.maxVar; v++) {
		uint8_t st = v ? uint8_t(s->sp->vstate[v].state) : 0;
		fwrite(&st, 1, 1, f);
	}
	put32(f, s->PVs.size());
	for (uint32_t i = 0; i < s->PVs.size(); i++) put32(f, s->PVs[i]);
	if (with_lists && s->ot.size()) {
		put32(f, 2 * s->PVs.size());
		for (uint32_t i = 0; i < s->PVs.size(); i++) {
			for (uint32_t sgn = 0; sgn < 2; sgn++) {
				OL& ol = s->ot[2 * s->PVs[i] + sgn];
				put32(f, uint32_t(ol.size()));
				for (int k = 0; k < ol.size(); k++) {
					auto it = index.find(ol[k]);
					put64(f, it == index.end() ? uint64_t(-1) : uint64_t(it->second));
				}
			}
		}
	}
	else put32(f, 0);
	put32(f, status);
	fclose(f);
}

static void dump_mapped(ParaFROST* s, const std::string& path)
{
	FILE* f = fopen(path.c_str(), "wb");
	if (!f) { perror(path.c_str()); exit(3); }
	put32(f, 0x4d524650u);
	put32(f, inf.maxVar);
	put64(f, s->orgs.size());
	for (uint32_t i = 0; i < s->orgs.size(); i++) {
		CLAUSE& c = s->cm[s->orgs[i]];
		put32(f, uint32_t(c.size()));
		for (int k = 0; k < c.size(); k++) put32(f, c[k]);
	}
	put32(f, s->vorg.size());
	for (uint32_t i = 0; i < s->vorg.size(); i++) put32(f, s->vorg[i]);
	put32(f, s->trail.size());
	for (uint32_t i = 0; i < s->trail.size(); i++) put32(f, s->trail[i]);
	for (uint32_t v = 0; v <= inf.maxVar; v++) {
		uint8_t st = v ? uint8_t(s->sp->vstate[v].state) : 0;
		fwrite(&st, 1, 1, f);
	}
	fclose(f);
}

int main(int argc, char** argv)
{
	if (argc < 3) { fprintf(stderr, "usage: pfref <input.cnf> <out.bin> [options]\n"); return 2; }
	const char* trace = getenv("PFREF_TRACE");
	std::vector<char*> args;
	args.push_back(argv[0]);
	args.push_back(argv[1]);
	for (int i = 3; i < argc; i++) args.push_back(argv[i]);
	int nargs = int(args.size());
	args.push_back(nullptr);
	parseArguments(nargs, args.data());
	quiet_en = true;
	verbose = 0;
	ParaFROST* s = new ParaFROST(argv[1]);
	pfrost = s;
	if (const char* fz = getenv("PFREF_FREEZE")) {
		FILE* f = fopen(fz, "rb");
		if (!f) { perror(fz); return 3; }
		s->incremental = true;
		s->ifrozen.expand(inf.maxVar + 1, 0);
		uint32_t v;
		while (fread(&v, 4, 1, f) == 1) {
			if (v == 0 || v > inf.maxVar) { fprintf(stderr, "bad frozen variable %u\n", v); return 3; }
			s->ifrozen[v] = 1;
		}
		fclose(f);
	}
	const std::string out = argv[2];
	int round = 0;
	auto tdump = [&](const char* pass) {
		if (trace) dump(s, std::string(trace) + "." + std::to_string(round) + "." + pass + ".bin", true, 0);
	};
	auto mark_proof = [&](const std::string& suffix) {
		if (const char* mk = getenv("PFREF_PROOF_MARK")) {
			fflush(NULL);
			struct stat sb;
			uint64_t at = (s->opts.proof_en && stat(s->opts.proof_path, &sb) == 0) ? uint64_t(sb.st_size) : 0;
			FILE* f = fopen((std::string(mk) + suffix).c_str(), "wb");
			if (!f) { perror(mk); exit(3); }
			put64(f, at);
			fclose(f);
		}
	};
	/* ParaFROST::sigmifying(), src/cpu/simplify.cpp:150-214, without the timers,
	   sleep/interrupt hooks and the write-back to the CDCL database.
	   Returns 0 simplified, 1 refuted (dumped), 2 nothing to simplify (dumped). */
	auto sigmifying = [&](const std::string& dst, const std::string& mark_suffix) -> int {
		s->rootify();
		s->shrinkTop(false);
		mark_proof(mark_suffix);
		if (s->orgs.empty()) { dump(s, dst, false, 0); return 2; }
		s->awaken();
		if (s->simpstate == AWAKEN_FAIL) { fprintf(stderr, "awaken failed\n"); exit(4); }
		int64 litsbefore = inf.nLiterals, diff = INT64_MAX;
		while (litsbefore) {
			s->resizeCNF();
			s->createOT();
			if (!s->prop()) { dump(s, dst, false, 1); return 1; }
			tdump("prop");
			if (!s->LCVE()) break;
			tdump("lcve");
			s->sortOT();
			tdump("sort");
			if (s->stop(diff)) { s->ERE(); tdump("ere"); break; }
			s->SUB();
			tdump("sub");
			s->VE();
			tdump("ve");
			s->BCE();
			s->countAll(), s->filterPVs();
			inf.nClauses = inf.n_cls_after, inf.nLiterals = inf.n_lits_after;
			diff = litsbefore - inf.nLiterals, litsbefore = inf.nLiterals;
			s->phase++, s->mu_inc++;
			s->mu_inc += s->phase == s->opts.phases;
			round++;
		}
		if (!s->prop()) { dump(s, dst, false, 1); return 1; }
		s->countFinal();
		s->shrinkSimp();
		return 0;
	};
	const auto t_start = std::chrono::steady_clock::now();
	const int outcome = sigmifying(out, "");
	if (getenv("PFREF_TIME"))   /* a refuted run includes the dump of its final state */
		fprintf(stderr, "PFREF_TIME %.6f\n",
			std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count());
	if (outcome) return 0;
	if (getenv("PFREF_MAP")) {
		s->map(true);
		dump_mapped(s, out);
		if (!getenv("PFREF_SECOND")) return 0;
		/* the rest of sigmifying's write-back, then a second sigmify() call */
		s->rebuildWT(s->opts.sigma_priorbins);
		if (s->retrail()) { fprintf(stderr, "refuted after the mapping\n"); return 5; }
		if (s->opts.phases > 2) s->opts.phases--;
		const uint64_t nres_first = s->model.resolved.size();
		round = 0;
		const std::string out2 = out + ".2";
		if (sigmifying(out2, ".2") == 0) dump(s, out2, false, 0);
		FILE* f = fopen(out2.c_str(), "ab");
		if (!f) { perror(out2.c_str()); return 3; }
		put64(f, nres_first);
		fclose(f);
		return 0;
	}
	dump(s, out, false, 0);
	return 0;
}
