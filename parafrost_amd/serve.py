"""Serving: incremental simplification requests with continuous batching.

A request asks for the simplification of one BMC-like instance: a random AIG
unrolled over time frames (``gen.aig_bmc``). A *new* request brings frames
[0, t); an *extend* request brings frames [t, t') of an instance served before.
Requests follow the reference's incremental use of the simplifier (isolve:
clauses added with iadd/itoClause, interface variables kept by ifreeze, then
sigmify; src/cpu/solveinc.cpp, addinc.cpp, assume.cpp):

* the instance's state after its last request stays resident in HBM: its
  simplified clauses (written back on the device by ``pfg_batch_export``) and
  its units;
* a request simplifies that resident formula together with its new frames'
  clauses, the resident units enqueued on the root trail, and the variables
  its next frames will refer to (``gen.aig_bmc_interface``) frozen, unless the
  instance has reached its last frame;
* the result of a request is its simplified clauses, its units, the model
  records it added (the model stack grows by them) and its variable states.

So only new frames cross PCIe, and the work per request is the simplified
formula plus the new frames, not the whole unrolled instance. Requests on one
instance are served in arrival order, one at a time; a request's result depends
only on its instance's earlier requests, never on the other requests it was
batched with (``PlainChain`` runs the same chain without batching or residency
and must give the same results).

Worker slots each own a batch context (its own HIP stream). A free slot admits
the oldest waiting requests whose instance is idle until the batch's literal
budget is reached, lays its context out for them and simplifies them in one
pass; requests that arrive meanwhile wait for the next free slot, so the work of
the slots' batches overlaps on the device. The resident store has an HBM budget: the
least recently used idle instances beyond it are moved to host memory and
brought back by their next request. Latency of a request = results on the host
- arrival. Results (clauses, units, model records, variable states) are host arrays:
the clause words and model records view the slot context's pinned host memory,
where the batch's results are downloaded in one copy, valid until the slot's next
batch; keep_results=True gives every request arrays of its own instead.
"""
from __future__ import annotations

import dataclasses
import threading
import time
from collections import deque

import numpy as np

from . import BatchSimplifier, HostPool, Simplifier, Unsat, default_opts
from . import gen


@dataclasses.dataclass
class Request:
    rid: int
    key: int             # instance
    seed: int
    ngates: int
    t0: int              # frames [t0, t1) carried by this request
    t1: int
    max_frames: int      # the instance is never extended past this frame
    cnf: gen.Cnf         # those frames' clauses (host)
    nvars: int           # variables of frames [0, t1)
    latches: int = 64
    arrival: float = 0.0
    done: float = 0.0
    nlits: int = 0       # literals of the formula simplified (resident clauses + new frames)
    status: int = -1
    result: tuple | None = None   # (clause words, units, model records, variable states)
    d_new: tuple | None = None    # new frames' clauses on the device (lits, offs)

    def frozen(self) -> np.ndarray:
        """variables the instance's later frames refer to (none after its last frame)"""
        if self.t1 >= self.max_frames:
            return np.zeros(0, np.uint32)
        return gen.aig_bmc_interface(self.seed, self.ngates, self.t1, nlatches=self.latches)


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
        out.append(Request(rid, key, s, ng, t0, t1, max_frames, cnf, nv, latches))
    return out


@dataclasses.dataclass
class Resident:
    """An instance's state after its last request, in its own numbering."""
    lits: object          # device int32 tensor, or None when moved to host
    offs: object          # device int64 tensor (ncls + 1)
    hlits: np.ndarray | None
    hoffs: np.ndarray | None
    units: np.ndarray
    frames: int
    refuted: bool = False
    last: float = 0.0
    busy: bool = False

    def nbytes(self) -> int:
        return 0 if self.lits is None else self.lits.numel() * 4 + self.offs.numel() * 8

    def nlits(self) -> int:
        if self.lits is not None:
            return int(self.lits.numel())
        return 0 if self.hlits is None else len(self.hlits)   # a placeholder holds nothing yet


def pick(pending, store, budget):
    """Admission: scan the waiting requests oldest first; take each one whose instance
    is idle (no request of it in flight or waiting ahead of it) while the literal
    estimate (resident clauses + new frames) fits the budget; the first one always
    fits. Taken instances are marked busy (a new instance gets a busy placeholder).
    Returns (batch, still waiting)."""
    batch, lits, seen, keep = [], 0, set(), deque()
    for r in pending:
        res = store.get(r.key)
        idle = r.key not in seen and not (res is not None and res.busy)
        seen.add(r.key)
        if not idle:
            keep.append(r)
            continue
        est = int(r.cnf.nlits) + (res.nlits() if res is not None and r.t0 > 0 and not res.refuted else 0)
        if not batch or lits + est <= budget:
            batch.append(r)
            lits += est
            if res is None:
                store[r.key] = Resident(None, None, None, None, np.zeros(0, np.uint32), 0, busy=True)
            else:
                res.busy = True
        else:
            keep.append(r)
    return batch, keep


class Server:
    """slots: concurrent batch contexts; batch_lits: literal budget of one batch;
    store_bytes: HBM budget of the resident instances."""

    def __init__(self, slots: int = 2, batch_lits: float = 2.5e8, store_bytes: float = 64e9, opts=None,
                 keep_results: bool = False, device: str = "cuda:0", context=None, tuning=None):
        """context: the batch context class (BatchSimplifier; tests of the request loop
        itself pass a host stand-in together with device="cpu"); tuning: launch choices
        (pfg_tuning fields) set on every slot context."""
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.ctx_cls = context or BatchSimplifier
        if self.dev.type == "cuda":
            torch.cuda.init()   # in this thread: the workers must not be the first to touch the device
            torch.zeros(1, device=self.dev)
        self.slots = slots
        self.batch_lits = batch_lits
        self.store_bytes = store_bytes
        self.opts = opts or default_opts()
        self.tuning = dict(tuning or {})
        self.keep = keep_results
        self.store: dict[int, Resident] = {}
        self.store_used = 0
        self.spills = 0
        self.pending = deque()
        self.cv = threading.Condition()
        self.closed = False
        self.finished = []
        self.phase_s = {"intake": 0.0, "assemble": 0.0, "upload": 0.0, "simplify": 0.0, "results": 0.0,
                        "download": 0.0, "export": 0.0}
        self.batches = 0
        self.error = None   # the first exception of a worker: stops the server, raised by run()

    # ---- request intake
    def submit(self, r: Request):
        torch = self.torch
        t = time.perf_counter()
        r.arrival = t
        # only the new frames cross PCIe
        r.d_new = (torch.from_numpy(r.cnf.lits.view(np.int32)).to(self.dev),
                   torch.from_numpy(r.cnf.offs.view(np.int64)).to(self.dev))
        with self.cv:
            self.phase_s["intake"] += time.perf_counter() - t
            self.pending.append(r)
            self.cv.notify()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()

    def _admit(self):
        """The oldest waiting requests whose instance is idle, up to the literal budget
        (requests on one instance keep their order)."""
        with self.cv:
            while True:
                if self.error is not None:
                    return None
                batch, self.pending = pick(self.pending, self.store, self.batch_lits)
                if batch:
                    return batch
                if self.closed and not self.pending:
                    return None
                self.cv.wait()

    def _formula(self, r: Request, res: Resident | None):
        """device (lits, offs) of the resident clauses followed by the new frames'"""
        torch = self.torch
        nl, no = r.d_new
        if res is None or res.frames == 0:
            return nl, no
        if res.lits is None:   # moved to host: bring it back
            res.lits = torch.from_numpy(res.hlits.view(np.int32)).to(self.dev)
            res.offs = torch.from_numpy(res.hoffs.view(np.int64)).to(self.dev)
            res.hlits = res.hoffs = None
            with self.cv:
                self.store_used += res.nbytes()
        return torch.cat([res.lits, nl]), torch.cat([res.offs, no[1:] + res.lits.numel()])

    def _worker(self, cap_vars: int):
        try:
            self._serve(cap_vars)
        except BaseException as e:   # noqa: BLE001  (handed to run())
            with self.cv:
                if self.error is None:
                    self.error = e
                self.closed = True
                self.cv.notify_all()

    def _sync(self):
        """wait for this thread's queued tensor work (not the device: the other slot keeps running)"""
        if self.dev.type == "cuda":
            self.torch.cuda.current_stream().synchronize()

    def _context(self, cap_vars: int):
        ctx = self.ctx_cls([cap_vars], self.opts)
        if self.tuning:
            ctx.set_tuning(**self.tuning)
        return ctx

    def _serve(self, cap_vars: int):
        torch = self.torch
        ctx = self._context(cap_vars)
        if self.dev.type == "cuda":
            torch.cuda.set_device(self.dev)
        try:
            while True:
                batch = self._admit()
                if batch is None:
                    break
                ta = time.perf_counter()
                work = []
                for r in batch:
                    res = self.store[r.key]
                    if res.refuted:   # an unsatisfiable instance stays unsatisfiable
                        r.status, r.result = 4, None
                        continue
                    if res.frames != r.t0:
                        raise RuntimeError(f"request {r.rid}: instance {r.key} is at frame {res.frames}, "
                                           f"the request starts at {r.t0}")
                    work.append((r, res, self._formula(r, res)))
                nv = np.array([r.nvars for r, _, _ in work], np.uint32)
                if len(work):
                    base = gen.batch_bases(nv)
                    if int(base[-1]) > ctx.nvars_cap:
                        ctx.close()
                        ctx = self._context(int(base[-1]))
                    parts_l, parts_o, at = [], [torch.zeros(1, dtype=torch.int64, device=self.dev)], 0
                    for (r, _, (l, o)), b in zip(work, base[:-1]):
                        r.nlits = int(l.numel())
                        parts_l.append(l + 2 * int(b))
                        parts_o.append(o[1:] + at)
                        at += l.numel()
                    d_l = torch.cat(parts_l)
                    d_o = torch.cat(parts_o)
                    self._sync()
                tb = time.perf_counter()
                if len(work):
                    ctx.relayout(nv)
                    ctx.upload_device(d_l.data_ptr(), d_o.data_ptr(), d_o.numel() - 1, d_l.numel())
                    for i, (r, res, _) in enumerate(work):
                        fz = r.frozen()
                        if len(fz):
                            ctx.ifreeze_instance(i, fz)
                        if len(res.units):
                            ctx.enqueue_instance(i, res.units)
                tc = time.perf_counter()
                st = ctx.sigmify_batch() if len(work) else []
                td = time.perf_counter()
                new = []
                t_dl = t_ex = 0.0
                for i, (r, res, _) in enumerate(work):
                    t0 = time.perf_counter()
                    # results not kept: clause words and model records viewed in the slot context's
                    # pinned memory, valid until its next batch (as the reused host memory was)
                    w, t, rec, vs = ctx.instance(i) if self.keep else ctx.instance_view(i)
                    t1 = time.perf_counter()
                    t_dl += t1 - t0
                    r.status = int(st[i])
                    if r.status == 0:
                        lits, offs = ctx.export(i, self.dev)
                        t_ex += time.perf_counter() - t1
                        # the units stay with the instance: its own copy, not the slot's reused memory
                        new.append((r, Resident(lits, offs, None, None, t.copy(), r.t1)))
                    else:
                        new.append((r, Resident(None, None, None, None, np.zeros(0, np.uint32), r.t1, refuted=True)))
                    if self.keep:
                        r.result = (w, t, rec, vs)
                now = time.perf_counter()
                for r in batch:
                    r.done = now
                te = time.perf_counter()
                with self.cv:
                    for r, nres in new:
                        self.store_used -= self.store[r.key].nbytes()
                        nres.last = te
                        self.store[r.key] = nres
                        self.store_used += nres.nbytes()
                    for r in batch:
                        self.store[r.key].busy = False
                        r.d_new = None
                    self._spill()
                    self.finished.extend(batch)
                    self.batches += 1
                    for k, dt in (("assemble", tb - ta), ("upload", tc - tb), ("simplify", td - tc), ("results", te - td),
                                  ("download", t_dl), ("export", t_ex)):
                        self.phase_s[k] += dt
                    self.cv.notify_all()
        finally:
            ctx.close()

    def _spill(self):
        """move the least recently used idle instances to host memory while over budget"""
        while self.store_used > self.store_bytes:
            idle = [k for k, e in self.store.items() if not e.busy and e.lits is not None]
            if not idle:
                return
            k = min(idle, key=lambda q: self.store[q].last)
            e = self.store[k]
            self.store_used -= e.nbytes()
            e.hlits = e.lits.cpu().numpy().view(np.uint32)
            e.hoffs = e.offs.cpu().numpy().view(np.uint64)
            e.lits = e.offs = None
            self.spills += 1

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
            if self.error is not None:
                break
            due += g
            d = due - time.perf_counter()
            if d > 0:
                time.sleep(d)
            self.submit(r)
        self.close()
        for t in threads:
            t.join()
        if self.error is not None:
            raise RuntimeError(f"serving worker failed: {self.error!r}") from self.error
        lat = np.array([r.done - r.arrival for r in requests])
        span = max(r.done for r in requests) - min(r.arrival for r in requests)
        lits = sum(r.nlits for r in requests)
        return {"requests": len(requests), "offered_rate": rate, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
                "literals": int(lits), "literals_per_s": float(lits / span), "span_s": float(span),
                "batches": self.batches, "spills": self.spills,
                "phase_s": {k: round(v, 3) for k, v in self.phase_s.items()}}


class PlainChain:
    """The same incremental requests on the plain per-instance path: one request at a
    time in a single-instance context, the instance's state kept on the host between
    its requests (clauses written back from the downloaded database), the whole
    formula uploaded from host memory for every request."""

    def __init__(self, cap_vars: int, opts=None, reuse_host: bool = False):
        """reuse_host: results are downloaded into reused host memory and stay valid
        until the next run() (as a served batch's results until the slot's next batch)"""
        self.g = Simplifier(cap_vars, opts or default_opts())
        self.pool = HostPool() if reuse_host else None
        self.state = {}   # key -> (Cnf of the simplified clauses, units, refuted)

    def close(self):
        self.g.close()

    def run(self, r: Request):
        """Simplify r; returns (status, (clause words, units, model records, variable states), literals)."""
        prev = self.state.get(r.key)
        if prev is not None and prev[2]:
            return 4, None, 0
        if prev is None:
            lits, offs, units = r.cnf.lits, r.cnf.offs, np.zeros(0, np.uint32)
        else:
            c, units, _ = prev
            lits = np.concatenate([c.lits, r.cnf.lits])
            offs = np.concatenate([c.offs, r.cnf.offs[1:] + np.uint64(c.nlits)])
        cnf = gen.Cnf(r.nvars, lits, offs)
        g = self.g
        g.reset_for_upload()
        fz = r.frozen()
        if len(fz):
            g.ifreeze(fz)
        g.upload(cnf)
        if len(units):
            g.enqueue_units(units)
        try:
            g.sigmify()
        except Unsat:
            self.state[r.key] = (None, None, True)
            return 4, None, int(cnf.nlits)
        if self.pool is not None:
            self.pool.reset()
        w, t, rec = g.cnf_words(self.pool), g.trail(self.pool), g.resolved(self.pool)
        vs = g.vstate()[:r.nvars + 1]
        self.state[r.key] = (gen.images_to_csr(w, r.nvars), t.copy(), False)
        return 0, (w, t, rec, vs), int(cnf.nlits)
