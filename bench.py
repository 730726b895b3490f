"""Benchmark of the MI355X SIGmA simplifier.

Default workload (the headline metric, one GPU): batching -- 4,096 independent
random 3-SAT / 4-SAT instances (variables log-uniform in [1,000, 200,000], ratios
U[4.0, 4.3] / U[9.5, 10.0], seed 4, 4.06e9 literals), simplified in batches of
instances that share one context's kernel launches, against the plain path (one
context per instance); then serving -- a Poisson stream of incremental BMC-like
requests (random AIGs of 10k-100k AND gates, 1-64 frames, seed 5) with continuous
batching over HBM-resident instances, p50/p99 latency per offered load, against the
same request chain served one at a time.

A step = one pass of the batched path over every instance (inputs resident in HBM;
each batch's clause image is rebuilt from them, then the whole round loop runs on the
device). value = input literals / second of the batched path.

Prints one JSON line. Extra objects:
  plain, vs_plain -- the plain per-instance path on a spread sample of the instances
  served          -- the serving workload: literals / s, p50 / p99 latency per offered
                     load, and its plain chain
  writeback_ms    -- every instance's results (clauses, model records, variable states,
                     units) brought to host memory once, outside the timed steps
  roofline        -- HBM roofline of the dominant kernel: algorithmic bytes per launch
                     (computed on the device by one extra, untimed step) over the
                     kernel's average HIP-event duration, against 8 TB/s
  cpu_baseline    -- the reference CPU simplifier built from its sources (or, where
                     that build is absent, the oracle, its C restatement), single
                     thread, on a bounded sample of the same instances.
Other workloads: --config c1 | c2 | c3 (one instance), c4 or c5 alone.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="headline", choices=["headline", "c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--c5-requests", type=int, default=64, help="c5: requests in the stream")
    ap.add_argument("--c5-rates", default="50,200,500,2000", help="c5: offered loads swept (requests / s)")
    ap.add_argument("--c5-slots", type=int, default=3, help="c5: batches in flight (one context each)")
    ap.add_argument("--c5-batch-lits", type=float, default=2.5e8, help="c5: literal budget of one batch")
    ap.add_argument("--batch-lits", type=float, default=1.0e9,
                    help="c4: literal budget of one batch (instances are admitted in order until it is reached)")
    ap.add_argument("--c4-instances", type=int, default=4096, help="c4: number of instances")
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="c4: instances timed on the plain path (0: none, for profiler runs of the batched path)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="c2-c4: measure candidate tunings on the workload first and record the best "
                         "(parafrost_amd/tuning/<config>.json); by default the recorded tuning is applied")
    ap.add_argument("--no-tuning", action="store_true", help="run with the library's default tuning")
    ap.add_argument("--bce", action="store_true", help="c1-c3: blocked clause elimination on (off by default, as in the reference)")
    ap.add_argument("--proof", action="store_true", help="c1-c3: binary DRAT proof output on (reports its size)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 256 MiB (calibrates FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--summarize-trace", metavar="KERNEL_STATS_CSV",
                    help="print a per-step kernel table from a rocprofv3 --stats kernel_stats.csv and exit")
    ap.add_argument("--trace-steps", type=int, default=1, help="sigmify calls covered by the traced run")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="print per-kernel HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                         "counter_collection.csv files, scaled by the copy calibration, and exit")
    return ap.parse_args()


COPY_BYTES = 256 << 20


def _kname(n: str) -> str:
    n = n.split("(")[0].replace("void ", "").strip()
    return n.split("::")[-1].split("<")[0]


def summarize_trace(path: str, steps: int):
    import csv
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print(f"total kernel time per sigmify: {tot / steps / 1e6:.3f} ms ({steps} sigmify calls traced)")
    print()
    print("| kernel | launches / call | ms / call | avg us / launch | share |")
    print("|---|---|---|---|---|")
    for r in rows:
        t = float(r["TotalDurationNs"])
        if t / tot < 0.002:
            continue
        print(f"| {_kname(r['Name'])} | {int(r['Calls']) / steps:.1f} | {t / steps / 1e6:.3f} | "
              f"{float(r['AverageNs']) / 1e3:.1f} | {100 * t / tot:.1f}% |")


def _pmc(path: str):
    import csv, collections
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[_kname(r["Kernel_Name"])].append(float(r["Counter_Value"]) * 1024.0)   # KiB per dispatch
    return d


def summarize_pmc(fetch: str, write: str, cfetch: str, cwrite: str):
    f, w, cf, cw = _pmc(fetch), _pmc(write), _pmc(cfetch), _pmc(cwrite)
    cf_b = [v for k, vs in cf.items() if "copyBuffer" in k for v in vs]
    cw_b = [v for k, vs in cw.items() if "copyBuffer" in k for v in vs]
    fs = COPY_BYTES / (sum(cf_b) / len(cf_b))
    ws = COPY_BYTES / (sum(cw_b) / len(cw_b))
    print(f"# copy calibration: 256 MiB device copy reads FETCH_SIZE x{1 / fs:.3f}, WRITE_SIZE x{1 / ws:.3f} "
          f"of the true bytes; factors applied: fetch {fs:.4f}, write {ws:.4f}")
    print("# kernel launches fetch_bytes_per_launch write_bytes_per_launch")
    for k in sorted(f, key=lambda k: -sum(f[k])):
        fb = sum(f[k]) / len(f[k]) * fs
        wb = (sum(w[k]) / len(w[k]) * ws) if k in w else 0.0
        print(f"{k} {len(f[k])} {fb:.0f} {wb:.0f}")


def traffic_for(cfg: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary of this workload."""
    p = os.path.join(ROOT, "profiles", f"traffic_{cfg}.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        t = line.split()
        if len(t) == 4 and not line.startswith("#") and t[0] == kernel:
            return int(float(t[2]) + float(t[3]))
    return None


def copy_calibration():
    """One plain device-to-device copy of 256 MiB, three times: the known-size kernel
    against which HBM byte counters are scaled."""
    import torch
    n = 256 << 20
    a = torch.ones(n // 4, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}))


REF_BIN = os.path.join(ROOT, "oracle", "_ref", "pfref")


def cpu_baseline_reference(sample, scale: str):
    """The reference CPU simplifier itself (ParaFROST src/cpu, compiled from its
    sources into oracle/_ref/pfref), default options, on the sample written as DIMACS;
    its own timer covers the simplifier only (not the parse): about 10 s in total."""
    import subprocess
    import tempfile
    reps, t_total, lits = 0, 0.0, 0
    with tempfile.TemporaryDirectory() as td:
        path, out = os.path.join(td, "sample.cnf"), os.path.join(td, "out.bin")
        sample.write_dimacs(path)
        env = dict(os.environ, PFREF_TIME="1")
        while t_total < 10.0 and reps < 20:
            r = subprocess.run([REF_BIN, path, out], env=env, capture_output=True, text=True, timeout=600)
            t = [float(x.split()[1]) for x in r.stderr.splitlines() if x.startswith("PFREF_TIME ")]
            if r.returncode != 0 or not t:
                return None
            t_total += t[0]
            lits += sample.nlits
            reps += 1
    return dict({"value": lits / t_total, "unit": "literals/s", "cores": 1, "kind": "reference",
                 "sample": f"{reps} x full simplification by the reference CPU simplifier (src/cpu, built from its "
                           f"sources; parse excluded) of the same generator at {scale} ({sample.nvars} vars, "
                           f"{sample.ncls} clauses, {sample.nlits} literals), single thread, {t_total:.1f} s"},
                **host_info())


def cpu_baseline(cfg: str):
    """The reference CPU simplifier (oracle/_ref/pfref) timed on a bounded sample: the
    same generator at 1/10 scale, full simplification, repeated until about 10 s;
    literals/s. Where that build is absent, the oracle (its single-thread C
    restatement) is timed the same way instead."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from parafrost_amd import gen
    sample = {"c1": lambda: gen.config("c1"),
              "c2": lambda: gen.geometric(200_000, 840_000, 2, 8, seed=2),
              "c3": lambda: gen.scalefree(500_000, 2_000_000, 2.5, 0.2, seed=3)}[cfg]()
    scale = "full scale" if cfg == "c1" else "1/10 scale"
    if os.path.exists(REF_BIN):
        ref = cpu_baseline_reference(sample, scale)
        if ref is not None:
            return ref
    reps, t_total, lits = 0, 0.0, 0
    while t_total < 10.0 and reps < 20:
        o = po.Oracle(sample, po.default_opts())
        t0 = time.perf_counter()
        o.sigmify()
        t_total += time.perf_counter() - t0
        lits += sample.nlits
        reps += 1
        del o
    return {"value": lits / t_total, "unit": "literals/s", "cores": 1, "kind": "port",
            "sample": f"{reps} x full sigmify of the same generator at {scale} "
                      f"({sample.nvars} vars, {sample.ncls} clauses, {sample.nlits} literals), "
                      f"single thread, {t_total:.1f} s"}


def roofline(kt: dict, ka: dict):
    """Per-kernel HBM roofline from the event timers: achieved = algorithmic bytes per
    launch / average launch duration. Only timers that bracket exactly one kernel and
    have defined algorithmic bytes qualify; the dominant one is the largest total time."""
    rows = {}
    for name, k in kt.items():
        a = ka.get(name)
        if not k["kernel"] or not a or a["bytes"] <= 0 or not k["launches"]:
            continue
        avg_ms = k["ms"] / k["launches"]
        bpl = a["bytes"] / a["launches"]
        gbs = bpl / (avg_ms * 1e-3) / 1e9
        rows[name] = {"kernel": k["kernel"], "ms_total": round(k["ms"], 3), "launches": k["launches"],
                      "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": int(bpl),
                      "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if not rows:
        return None, rows
    dom = max(rows, key=lambda n: rows[n]["ms_total"])
    r = rows[dom]
    roof = {"bound": "hbm", "achieved": r["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["frac"],
            "traffic": None, "kernel": r["kernel"], "avg_launch_ms": r["avg_launch_ms"],
            "bytes_per_launch": r["bytes_per_launch"], "launches": r["launches"]}
    return roof, rows


def run_c4(args):
    """Batching workload: 4,096 independent random 3-SAT / 4-SAT instances (variables
    log-uniform in [1,000, 200,000]; ratios U[4.0, 4.3] / U[9.5, 10.0]; seed 4). The
    instances are admitted in order into batches of at most --batch-lits literals; each
    batch is one context whose kernel launches serve all of its instances. A step = every
    batch simplified once (inputs resident in HBM). The plain path (one context per
    instance) is timed on an evenly spread sample of instances. Returns the JSON record."""
    import torch
    import parafrost_amd as pf
    from parafrost_amd import gen
    t0 = time.perf_counter()
    B = gen.c4(ninst=args.c4_instances)
    gen_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    # batches by literal budget, in instance order
    lit_first = B.offs[B.cls_first.astype(np.int64)]
    groups, a = [], 0
    while a < B.ninst:
        b = a + 1
        while b < B.ninst and int(lit_first[b + 1] - lit_first[a]) <= args.batch_lits:
            b += 1
        groups.append((a, b))
        a = b
    # group inputs in batch numbering relative to the group's first instance, resident in HBM
    inputs = []
    for (a, b) in groups:
        ca, cb = int(B.cls_first[a]), int(B.cls_first[b])
        la, lb = int(B.offs[ca]), int(B.offs[cb])
        lits = torch.from_numpy((B.lits[la:lb] - np.uint32(2 * int(B.base[a]))).view(np.int32)).to(dev)
        offs = torch.from_numpy((B.offs[ca:cb + 1] - np.uint64(la)).view(np.int64)).to(dev)
        inputs.append((B.inst_nvars[a:b].copy(), lits, offs, cb - ca, lb - la))
    torch.cuda.synchronize()
    cap = max(int(gen.batch_bases(nv)[-1]) for nv, *_ in inputs)
    # one context sized for the largest group, laid out again for every group
    s = pf.BatchSimplifier(np.array([cap], np.uint32), pf.default_opts())

    def step(account=False, download=None):
        st = []
        s.set_accounting(account)
        for (nv, lits, offs, ncls, nl) in inputs:
            s.relayout(nv)
            s.upload_device(lits.data_ptr(), offs.data_ptr(), ncls, nl)
            st.append(s.sigmify_batch())
            if download is not None:
                download(len(nv))
        s.set_accounting(False)
        return st

    tuning = setup_tuning(args, s, step, "c4", "batched c4 workload")
    for _ in range(args.warmup):
        step()
    s.reset_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    kt = s.kernel_times()
    pass_ms = {k: round(v / args.steps, 2) for k, v in s.times().items()}
    s.reset_times()
    step(account=True)
    ka = s.kernel_times()
    roof, kernels = roofline(kt, ka)
    if roof:
        roof["traffic"] = traffic_for("c4", roof["kernel"])
        if roof["traffic"] is not None:
            roof["traffic_source"] = "profiles/traffic_c4.txt (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE)"
    unsat = int(sum(int((x != 0).sum()) for x in st))
    # write-back: every instance's results to host memory, outside the timed steps: the batch's
    # clause images and model records grouped by instance and copied once into the context's
    # pinned memory (each instance's arrays viewed there), units and variable states copied
    wb = {"s": 0.0, "words": 0}

    def download(n):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(n):
            w, t, r, vs = s.instance_view(i)
            wb["words"] += len(w) + len(t) + len(r) + len(vs) // 4
        wb["s"] += time.perf_counter() - t1

    step(download=download)
    s.close()
    del inputs
    # plain path on a spread sample (none with --plain-sample 0: profiler runs of the batched path)
    pick = np.unique(np.linspace(0, B.ninst - 1, min(args.plain_sample, B.ninst)).round().astype(np.int64))
    if args.plain_sample <= 0:
        pick = pick[:0]
    plain_t, plain_l = 0.0, 0
    for i in pick:
        c = B.instance(int(i))
        dl = torch.from_numpy(c.lits.view(np.int32)).to(dev)
        do = torch.from_numpy(c.offs.view(np.int64)).to(dev)
        s = pf.Simplifier(c.nvars, pf.default_opts())
        for rep in range(2):   # the second run is timed
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            s.reset_for_upload()
            s.upload_device(dl.data_ptr(), do.data_ptr(), c.ncls, c.nlits)
            try:
                s.sigmify()
            except pf.Unsat:
                pass
            torch.cuda.synchronize()
            if rep:
                plain_t += time.perf_counter() - t1
                plain_l += c.nlits
        s.close()
    plain = plain_l / plain_t if plain_t else None
    value = B.nlits / dt
    out = {
        "metric": "simplified literals/sec on 1 MI355X (batched/served vs plain); p50/p99 latency",
        "value": value, "unit": "literals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"path": "batched", "workload": f"{B.ninst} random 3-SAT/4-SAT instances, vars log-uniform "
                   "[1000, 200000], ratios U[4.0,4.3]/U[9.5,10.0], seed 4",
                   "instances": B.ninst, "variables": B.nvars, "clauses": B.ncls, "input_literals": B.nlits,
                   "batches": len(groups), "batch_literal_budget": args.batch_lits, "refuted_instances": unsat,
                   "generation_s": round(gen_s, 1), "pass_ms_per_step": pass_ms},
        "plain": {"value": plain, "unit": "literals/s",
                  "sample": f"{len(pick)} instances spread over the batch, {plain_l} literals, one context each, "
                            "second of two runs timed"},
        "vs_plain": value / plain if plain else None,
        "tuning": tuning,
        "writeback_ms": round(wb["s"] * 1e3, 1),
        "writeback": f"every instance's clauses, model records, variable states and units in host memory after its "
                     f"batch ({wb['words']} words): one grouped copy per batch into pinned memory, instances viewed "
                     f"there; once, outside the timed steps",
        "roofline": roof, "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_instances(B)
    return out


def run_c5(args):
    """Serving workload: a stream of incremental BMC-like requests (random AIGs of
    10k-100k 2-input ANDs, 30% inverted edges, 64 latches, Tseitin-encoded, unrolled
    1-64 frames; half of the requests extend an instance served before by 1-8 frames;
    seed 5), replayed as a Poisson stream at each offered load. Each request simplifies
    its instance's resident simplified clauses plus its new frames (units enqueued,
    interface variables frozen). Three slots of continuous batching (--c5-slots), 2.5e8 literals per
    batch. Latency = results on the host - arrival (the served path's in its slot
    contexts' pinned memory, the plain chain's in host memory it reuses). The
    tuning is the record for this workload (parafrost_amd/tuning/c5.json; --autotune
    measures it: each candidate is a served run at the last offered load). The plain path runs the same
    request chain one request at a time in a single-instance context with the state
    on the host; both count the literals of the formulas they simplify (equal)."""
    import parafrost_amd as pf  # noqa: F401  (loads torch and the library first)
    from parafrost_amd import serve
    t0 = time.perf_counter()
    reqs = serve.make_requests(args.c5_requests, seed=5)
    gen_s = time.perf_counter() - t0
    rates = [float(x) for x in args.c5_rates.split(",")]
    # warm-up: load the code objects
    w = serve.make_requests(3, seed=99, gates=(2000, 4000), max_frames=4)
    serve.Server(slots=2).run(w, rate=1000.0)

    def reset():
        for r in reqs:
            r.done = r.arrival = 0.0
            r.status, r.result, r.nlits = -1, None, 0

    tuning = serve_tuning(args, reqs, rates[-1], reset)
    knobs = {k: v for k, v in tuning.items() if k != "source"}
    sweep = []
    for rate in rates:
        for r in reqs:
            r.done = r.arrival = 0.0
            r.status, r.result, r.nlits = -1, None, 0
        st = serve.Server(slots=args.c5_slots, batch_lits=args.c5_batch_lits, tuning=knobs).run(reqs, rate=rate)
        release_device_memory()
        sweep.append({k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()})
        print(f"# c5 rate {rate}: p50 {st['p50_ms']:.1f} ms p99 {st['p99_ms']:.1f} ms "
              f"{st['literals_per_s'] / 1e6:.1f} M literals/s ({st['literals']} literals)", file=sys.stderr, flush=True)
    served_lits = sweep[-1]["literals"]
    # plain: the same chain, one request at a time, state on the host
    p = serve.PlainChain(max(r.nvars for r in reqs), reuse_host=True)
    p.run(serve.make_requests(1, seed=98, gates=(2000, 4000), max_frames=2)[0])   # warm-up
    p.state.clear()
    t1 = time.perf_counter()
    plain_l = sum(p.run(r)[2] for r in reqs)
    plain_t = time.perf_counter() - t1
    p.close()
    plain = plain_l / plain_t
    print(f"# c5 plain: {plain / 1e6:.1f} M literals/s ({plain_l} literals, {plain_t:.2f} s)", file=sys.stderr, flush=True)
    best = max(sweep, key=lambda x: x["literals_per_s"])
    out = {
        "metric": "simplified literals/sec on 1 MI355X (batched/served vs plain); p50/p99 latency",
        "value": best["literals_per_s"], "unit": "literals/s", "n_gpus": 1, "steps": len(rates), "warmup": 1,
        "ms_per_step": best["span_s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"path": f"served (incremental, HBM-resident instances, continuous batching, {args.c5_slots} slots, "
                           f"{args.c5_batch_lits:.3g} literals per batch)",
                   "workload": f"{len(reqs)} incremental BMC-like requests: random AIGs 10k-100k ANDs, 30% inverted "
                               "edges, 64 latches, Tseitin, 1-64 frames, half extend an instance served before; seed 5",
                   "requests": len(reqs), "literals": served_lits, "generation_s": round(gen_s, 1)},
        "served": sweep, "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"],
        "plain": {"value": plain, "unit": "literals/s", "literals": plain_l, "seconds": round(plain_t, 3),
                  "path": "the same request chain one at a time: host formula uploaded, single-instance "
                          "context, results downloaded, clauses written back on the host"},
        "vs_plain": best["literals_per_s"] / plain,
        "tuning": tuning,
    }
    return out


def release_device_memory():
    """drop a finished server's resident instances (device tensors) before the next run"""
    import gc
    import torch
    gc.collect()
    torch.cuda.empty_cache()


def serve_tuning(args, reqs, rate, reset):
    """the slots' tuning for the serving workload: --no-tuning the library defaults,
    --autotune measured now (coordinate descent, each candidate a served run at `rate`),
    otherwise the record parafrost_amd/tuning/c5.json when it matches the device"""
    import parafrost_amd as pf
    from parafrost_amd import serve, tune
    probe = pf.BatchSimplifier([16], pf.default_opts())
    defaults = probe.tuning()
    probe.close()
    if args.no_tuning:
        return {"source": "library defaults", **defaults}

    class Slots:   # the tuning every slot context of the next run gets
        def __init__(self):
            self.t = dict(defaults)

        def tuning(self):
            return dict(self.t)

        def set_tuning(self, **kw):
            self.t.update(kw)

    slots = Slots()
    if args.autotune:
        def step():
            reset()
            serve.Server(slots=args.c5_slots, batch_lits=args.c5_batch_lits, tuning=slots.t).run(reqs, rate=rate)
            release_device_memory()
        rec = tune.tune(slots, step, lambda: None, "c5", reps=1, workload="served c5 workload",
                        log=lambda m: print("# " + m, file=sys.stderr, flush=True))
        return {"source": f"measured now ({len(rec['trials'])} candidates)", **slots.tuning()}
    rec = tune.apply(slots, "c5")
    src = "parafrost_amd/tuning/c5.json" if rec else f"library defaults (no record of c5 for {tune._device_arch()})"
    return {"source": src, **slots.tuning()}


def setup_tuning(args, s, step, name, workload):
    """--autotune: measure and record; otherwise apply the recorded tuning (if any)."""
    import torch
    from parafrost_amd import tune
    if args.no_tuning:
        return {"source": "library defaults", **s.tuning()}
    if args.autotune:
        rec = tune.tune(s, step, torch.cuda.synchronize, name, reps=1, workload=workload,
                        log=lambda m: print("# " + m, file=sys.stderr, flush=True))
        return {"source": f"measured now ({len(rec['trials'])} candidates)", **s.tuning()}
    rec = tune.apply(s, name)
    src = (f"parafrost_amd/tuning/{name}.json" if rec else
           f"library defaults (no record of {name} for {tune._device_arch()}: {tune._device_name()})")
    return {"source": src, **s.tuning()}


def host_info():
    """the GPU box host's online cores and CPU model"""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"host_cores_online": os.cpu_count(), "cpu_model": model}


def cpu_baseline_instances(B):
    """The reference CPU simplifier (oracle/_ref/pfref, default options, its own timer
    around the simplifier) on instances of the batch spread over all sizes, in a seeded
    random order, until about 15 s of simplifier time; the oracle (its C restatement)
    where that build is absent."""
    order = np.argsort(B.inst_nvars, kind="stable")
    pick = order[np.linspace(0, len(order) - 1, 200).round().astype(np.int64)]
    pick = pick[np.random.default_rng(4).permutation(len(pick))]   # spread over the sizes until the time is up
    t_total, lits, n = 0.0, 0, 0
    kind = "reference" if os.path.exists(REF_BIN) else "port"
    if kind == "reference":
        import subprocess
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path, outp = os.path.join(td, "i.cnf"), os.path.join(td, "o.bin")
            for i in pick:
                c = B.instance(int(i))
                c.write_dimacs(path)
                r = subprocess.run([REF_BIN, path, outp], env=dict(os.environ, PFREF_TIME="1"), capture_output=True,
                                   text=True, timeout=600)
                t = [float(x.split()[1]) for x in r.stderr.splitlines() if x.startswith("PFREF_TIME ")]
                if r.returncode != 0:
                    kind = "port"
                    break
                if not t:   # refuted: the simplifier ends the process before its timer prints
                    continue
                t_total += t[0]
                lits += c.nlits
                n += 1
                if t_total > 15.0:
                    break
    if kind == "port":
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        t_total, lits, n = 0.0, 0, 0
        for i in pick:
            c = B.instance(int(i))
            o = po.Oracle(c, po.default_opts())
            t0 = time.perf_counter()
            o.sigmify()
            t_total += time.perf_counter() - t0
            lits += c.nlits
            n += 1
            del o
            if t_total > 15.0:
                break
    what = ("the reference CPU simplifier (src/cpu, built from its sources; parse excluded)" if kind == "reference"
            else "the oracle (single-thread C restatement of the reference CPU simplifier)")
    return dict({"value": lits / t_total, "unit": "literals/s", "cores": 1, "kind": kind,
                 "sample": f"{n} instances of the batch (spread over the sizes, seeded random order), full "
                           f"simplification each by {what}, single thread, {lits} literals in {t_total:.1f} s"},
                **host_info())


def run_headline(args):
    """The metric's workloads: the batched path (c4) as the line, the served path (c5)
    as its `served` object."""
    out = run_c4(args)
    print(f"# c4 batched {out['value'] / 1e6:.1f} M literals/s, plain {(out['plain']['value'] or 0) / 1e6:.1f}, "
          f"x{out['vs_plain'] or 0:.2f}", file=sys.stderr, flush=True)
    c5 = run_c5(args)
    out["served"] = {"workload": c5["config"]["workload"], "path": c5["config"]["path"],
                     "value": c5["value"], "unit": "literals/s", "p50_ms": c5["p50_ms"], "p99_ms": c5["p99_ms"],
                     "sweep": c5["served"], "plain": c5["plain"], "vs_plain": c5["vs_plain"], "tuning": c5["tuning"]}
    out["config"]["served_workload"] = c5["config"]["workload"]
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU (sharding across GPUs is designed, not built); "
                                   "--gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.trace_steps)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.config == "headline":
        print(json.dumps(run_headline(args)))
        return
    if args.config == "c4":
        print(json.dumps(run_c4(args)))
        return
    if args.config == "c5":
        print(json.dumps(run_c5(args)))
        return
    import torch
    import parafrost_amd as pf
    from parafrost_amd import gen

    dev = torch.device("cuda:0")
    cnf = gen.config(args.config)
    L = int(cnf.nlits)
    d_lits = torch.from_numpy(cnf.lits.view(np.int32)).to(dev)
    d_offs = torch.from_numpy(cnf.offs.view(np.int64)).to(dev)
    torch.cuda.synchronize()

    s = pf.Simplifier(cnf.nvars, pf.default_opts(bce_en=int(args.bce), proof_en=int(args.proof)))

    outcome = {"refuted": False}

    def step():
        s.reset_for_upload()
        s.upload_device(d_lits.data_ptr(), d_offs.data_ptr(), cnf.ncls, L)
        try:
            s.sigmify()
        except pf.Unsat:   # c3: the simplifier proves the formula unsatisfiable mid-way
            outcome["refuted"] = True

    tuning = setup_tuning(args, s, step, args.config, f"single instance {args.config}")
    for _ in range(args.warmup):
        step()
    s.reset_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    times = s.times()
    kt = s.kernel_times()
    try:
        stats = s.stats()
    except pf.Unsat:
        stats = {"clauses": None, "literals": None, "eliminated": None}

    # algorithmic bytes per launch: one more, untimed step with accounting on
    # (the step is deterministic, so every launch repeats with the same sizes)
    s.reset_times()
    s.set_accounting(True)
    step()
    s.set_accounting(False)
    ka = s.kernel_times()
    torch.cuda.synchronize()
    roof, kernels = roofline(kt, ka)
    if roof:
        roof["traffic"] = traffic_for(args.config, roof["kernel"])
        if roof["traffic"] is not None:
            roof["traffic_source"] = f"profiles/traffic_{args.config}.txt (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE)"
    out = {
        "metric": "simplified literals/sec on 1 MI355X (batched/served vs plain); p50/p99 latency",
        "value": L * args.steps / dt,
        "unit": "literals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"path": "plain (one instance per context)", "workload": {"c1": "uniform 3-SAT 20k vars / 85.2k clauses, seed 1",
                                "c2": "mixed-width CNF 2M vars / 8.4M clauses, widths 2-8 geometric, seed 2",
                                "c3": "scale-free CNF 5M vars / 20M clauses, 20% Tseitin gates, seed 3"}[args.config],
                   "input_literals": L, "variables": cnf.nvars, "clauses": cnf.ncls,
                   "passes": "OT+prop+LCVE+sortOT+SUB+BVE(gates)" + ("+BCE" if args.bce else "") + "+ERE+GC, 5 phases"
                             + (", DRAT proof output" if args.proof else ""),
                   "clauses_out": stats["clauses"], "literals_out": stats["literals"],
                   "eliminated": stats["eliminated"], "pass_ms_per_step": {k: round(v / args.steps, 3) for k, v in times.items()},
                   "outcome": ("refuted: the formula is unsatisfiable; a step is the rounds up to the refutation"
                               if outcome["refuted"] else "simplified")},
        "tuning": tuning,
        "roofline": roof,
        "kernels": kernels,
    }
    if args.proof:
        out["config"]["proof_bytes_per_step"] = int(s.proof().size)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
