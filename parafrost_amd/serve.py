"""Serving: a request loop over the batch simplifier with continuous batching.

A request asks for the simplification of one BMC-like instance: a random AIG
unrolled over time frames (``gen.aig_bmc``). A *new* request brings frames
[0, t); an *extend* request brings frames [t, t') of an instance seen before,
whose clauses stay resident in HBM, so only the new frames cross PCIe. Every
request is simplified from its instance's whole clause set (frames [0, t')),
so its result does not depend on what else is in flight.

Worker slots each own a batch context (its own HIP stream). A free slot admits
the oldest waiting requests until the batch's literal budget is reached, lays its
context out for them, and simplifies them in one pass; requests that arrive
meanwhile wait for the next free slot, so the work of two batches overlaps on
the device. Latency of a request = results on the host - arrival.
"""
from __future__ import annotations

import dataclasses
import threading
import time
from collections import deque

import numpy as np

from . import BatchSimplifier, default_opts
from . import gen


@dataclasses.dataclass
class Request:
    rid: int
    key: int             # instance
    seed: int
    ngates: int
    t0: int              # frames [t0, t1) carried by this request
    t1: int
    cnf: gen.Cnf         # those frames' clauses (host)
    nvars: int           # variables of frames [0, t1)
    arrival: float = 0.0
    done: float = 0.0
    nlits: int = 0       # literals simplified (frames [0, t1))
    status: int = -1
    result: tuple | None = None   # (clause words, units, model records, variable states)


def make_requests(n: int, seed: int = 5, gates=(10_000, 100_000), max_frames: int = 64, extend_frac: float = 0.5,
                  latches: int = 64):
    """A stream of n requests: new instances (1..max_frames frames) and, with
    probability extend_frac, extensions of an earlier instance by 1..8 frames."""
    rng = np.random.default_rng(seed)
    frames, meta, out = {}, {}, []
    for rid in range(n):
        ext = [k for k, t in frames.items() if t < max_frames]
        if ext and rng.random() < extend_frac:
            key = int(rng.choice(ext))
            t0 = frames[key]
            t1 = min(max_frames, t0 + int(rng.integers(1, 9)))
            s, ng = meta[key]
        else:
            key = len(meta)
            s, ng = int(rng.integers(1, 1 << 62)), int(rng.integers(gates[0], gates[1] + 1))
            meta[key] = (s, ng)
            t0, t1 = 0, int(rng.integers(1, max_frames + 1))
        frames[key] = t1
        cnf, nv = gen.aig_bmc(s, ng, t0, t1, nlatches=latches)
        out.append(Request(rid, key, s, ng, t0, t1, cnf, nv))
    return out


class Server:
    """slots: concurrent batch contexts; batch_lits: literal budget of one batch;
    store_bytes: HBM budget of the resident instances (least recently used first out;
    an evicted instance is regenerated from its seed when extended)."""

    def __init__(self, slots: int = 2, batch_lits: float = 2.5e8, store_bytes: float = 64e9, opts=None,
                 keep_results: bool = False):
        import torch
        self.torch = torch
        self.dev = torch.device("cuda:0")
        torch.cuda.init()   # in this thread: the workers must not be the first to touch the device
        torch.zeros(1, device=self.dev)
        self.slots = slots
        self.batch_lits = batch_lits
        self.store_bytes = store_bytes
        self.opts = opts or default_opts()
        self.keep = keep_results
        self.store = {}          # key -> (lits (device, own numbering), offs (device), nvars, frames, last use)
        self.store_used = 0
        self.pending = deque()
        self.cv = threading.Condition()
        self.closed = False
        self.finished = []
        self.phase_s = {"intake": 0.0, "assemble": 0.0, "upload": 0.0, "simplify": 0.0, "results": 0.0}
        self.batches = 0

    # ---- resident instances
    def _resident(self, r: Request):
        """Device clauses of r's instance, frames [0, r.t1), appending r's frames."""
        torch = self.torch
        ent = self.store.get(r.key)
        if ent is not None and ent[3] == r.t0:
            lits, offs = ent[0], ent[1]
            add_l = torch.from_numpy(r.cnf.lits.view(np.int32)).to(self.dev)
            add_o = torch.from_numpy((r.cnf.offs[1:] + np.uint64(lits.numel())).view(np.int64)).to(self.dev)
            lits = torch.cat([lits, add_l])
            offs = torch.cat([offs, add_o])
            self.store_used -= ent[0].numel() * 4 + ent[1].numel() * 8
        else:   # new, or evicted: the request's frames, or the whole instance again
            cnf = r.cnf if r.t0 == 0 else gen.aig_bmc(r.seed, r.ngates, 0, r.t1)[0]
            lits = torch.from_numpy(cnf.lits.view(np.int32)).to(self.dev)
            offs = torch.from_numpy(cnf.offs.view(np.int64)).to(self.dev)
        self.store[r.key] = (lits, offs, r.nvars, r.t1, time.perf_counter())
        self.store_used += lits.numel() * 4 + offs.numel() * 8
        while self.store_used > self.store_bytes and len(self.store) > 1:
            old = min((k for k in self.store if k != r.key), key=lambda k: self.store[k][4])
            e = self.store.pop(old)
            self.store_used -= e[0].numel() * 4 + e[1].numel() * 8
        return lits, offs

    # ---- request intake
    def submit(self, r: Request):
        with self.cv:
            r.arrival = time.perf_counter()
            lits, offs = self._resident(r)
            self.phase_s["intake"] += time.perf_counter() - r.arrival
            r.nlits = int(lits.numel())
            self.pending.append((r, lits, offs))
            self.cv.notify()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()

    def _admit(self):
        with self.cv:
            while not self.pending and not self.closed:
                self.cv.wait()
            if not self.pending:
                return None
            batch, lits = [], 0
            while self.pending and (not batch or lits + self.pending[0][0].nlits <= self.batch_lits):
                e = self.pending.popleft()
                batch.append(e)
                lits += e[0].nlits
            return batch

    # ---- worker
    def _worker(self, cap_vars: int):
        torch = self.torch
        ctx = BatchSimplifier([cap_vars], self.opts)
        torch.cuda.set_device(self.dev)
        try:
            while True:
                batch = self._admit()
                if batch is None:
                    break
                ta = time.perf_counter()
                nv = np.array([e[0].nvars for e in batch], np.uint32)
                base = gen.batch_bases(nv)
                if int(base[-1]) > ctx.nvars_cap:
                    ctx.close()
                    ctx = BatchSimplifier([int(base[-1])], self.opts)
                parts_l, parts_o, at = [], [torch.zeros(1, dtype=torch.int64, device=self.dev)], 0
                for (r, l, o), b in zip(batch, base[:-1]):
                    parts_l.append(l + 2 * int(b))
                    parts_o.append(o[1:] + at)
                    at += l.numel()
                d_l = torch.cat(parts_l)
                d_o = torch.cat(parts_o)
                torch.cuda.current_stream().synchronize()   # not the device: the other slot keeps running
                tb = time.perf_counter()
                ctx.relayout(nv)
                ctx.upload_device(d_l.data_ptr(), d_o.data_ptr(), d_o.numel() - 1, d_l.numel())
                tc = time.perf_counter()
                st = ctx.sigmify_batch()
                td = time.perf_counter()
                for i, (r, _, _) in enumerate(batch):
                    res = ctx.instance(i)
                    r.status = int(st[i])
                    if self.keep:
                        r.result = res
                    r.done = time.perf_counter()
                te = time.perf_counter()
                with self.cv:
                    self.finished.extend(e[0] for e in batch)
                    self.batches += 1
                    for k, dt in (("assemble", tb - ta), ("upload", tc - tb), ("simplify", td - tc), ("results", te - td)):
                        self.phase_s[k] += dt
        finally:
            ctx.close()

    def run(self, requests, rate: float, cap_vars: int | None = None):
        """Replay requests as a Poisson stream of the given rate (requests / s)."""
        rng = np.random.default_rng(12345)
        gaps = rng.exponential(1.0 / rate, len(requests))
        cap = cap_vars or int(gen.batch_bases(np.array([r.nvars for r in requests], np.uint32))[-1])
        threads = [threading.Thread(target=self._worker, args=(cap,)) for _ in range(self.slots)]
        for t in threads:
            t.start()
        t0 = time.perf_counter()
        due = t0
        for r, g in zip(requests, gaps):
            due += g
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            self.submit(r)
        self.close()
        for t in threads:
            t.join()
        lat = np.array([r.done - r.arrival for r in requests])
        span = max(r.done for r in requests) - min(r.arrival for r in requests)
        return {"requests": len(requests), "offered_rate": rate, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
                "literals": int(sum(r.nlits for r in requests)),
                "literals_per_s": float(sum(r.nlits for r in requests) / span), "span_s": float(span),
                "batches": self.batches, "phase_s": {k: round(v, 3) for k, v in self.phase_s.items()}}
