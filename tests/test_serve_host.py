"""Host-side pieces of serving (no GPU): the request stream, the generator's
frame-by-frame growth and frame interface, and the host write-back of clause images."""
import numpy as np
import pytest

from parafrost_amd import gen


def test_aig_frames_concatenate():
    for seed in (1, 2, 3):
        a, _ = gen.aig_bmc(seed, 700, 0, 3)
        b, nvb = gen.aig_bmc(seed, 700, 3, 7)
        w, nvw = gen.aig_bmc(seed, 700, 0, 7)
        assert nvb == nvw
        assert np.array_equal(np.concatenate([a.lits, b.lits]), w.lits)
        assert a.ncls + b.ncls == w.ncls


def test_aig_clauses_well_formed():
    c, nv = gen.aig_bmc(9, 1200, 0, 5)
    assert c.lits.min() >= 2 and (c.lits >> 1).max() <= nv
    for k in range(c.ncls):
        cl = c.lits[c.offs[k]:c.offs[k + 1]]
        assert 2 <= len(cl) <= 3
        assert len(set((cl >> 1).tolist())) == len(cl)   # distinct variables, no tautology


def test_request_stream_shape():
    from parafrost_amd import serve
    reqs = serve.make_requests(40, seed=3, gates=(200, 400), max_frames=16)
    frames = {}
    for r in reqs:
        assert r.t1 > r.t0
        assert frames.get(r.key, 0) == r.t0   # extensions continue where the instance stopped
        frames[r.key] = r.t1
    assert any(r.t0 > 0 for r in reqs) and any(r.t0 == 0 for r in reqs)


def test_interface_covers_later_frames():
    """Clauses of frames t, t + 1, .. refer to variables of frames [0, t) only through
    the interface of frame t (the variables an incremental instance keeps frozen)."""
    for seed, t in ((4, 1), (5, 3), (6, 6)):
        first_new = gen.aig_bmc(seed, 900, 0, t)[1] + 1
        later, _ = gen.aig_bmc(seed, 900, t, t + 4)
        old = np.unique(later.lits >> 1)
        old = old[old < first_new]
        iface = gen.aig_bmc_interface(seed, 900, t)
        assert len(old) > 0
        assert np.all(np.isin(old, iface)), (seed, t)
        assert iface.min() >= 1 and iface.max() < first_new


def test_images_to_csr_roundtrip():
    c, nv = gen.aig_bmc(12, 400, 0, 3)
    words = []
    for k in range(c.ncls):
        cl = c.lits[c.offs[k]:c.offs[k + 1]]
        words += [0, 0, len(cl)] + cl.tolist()
    back = gen.images_to_csr(np.array(words, np.uint32), nv)
    assert np.array_equal(back.lits, c.lits) and np.array_equal(back.offs, c.offs)
    with pytest.raises(ValueError):
        gen.images_to_csr(np.array([0, 0, 5, 2], np.uint32), nv)


def test_request_frozen_sets():
    from parafrost_amd import serve
    reqs = serve.make_requests(30, seed=7, gates=(200, 300), max_frames=6)
    for r in reqs:
        fz = r.frozen()
        if r.t1 >= r.max_frames:
            assert len(fz) == 0
        else:
            assert len(fz) > 0 and fz.max() <= r.nvars


# ---- admission and the request loop, with a host stand-in for the batch context

def _reqs():
    from parafrost_amd import serve
    return serve.make_requests(28, seed=31, gates=(1500, 5000), max_frames=10)


def test_pick_keeps_instance_order_and_budget():
    from collections import deque
    from parafrost_amd import serve
    reqs = _reqs()
    store = {}
    batch, keep = serve.pick(deque(reqs), store, budget=1e12)
    keys = [r.key for r in batch]
    assert len(keys) == len(set(keys)), "one request per instance in a batch"
    for r in keep:   # every request left behind waits for an earlier one of its instance
        assert any(b.key == r.key and b.rid < r.rid for b in batch)
    assert all(store[k].busy for k in keys)
    # the extension of a busy new instance waits and does not look at its (empty) state
    batch2, keep2 = serve.pick(keep, store, budget=1e12)
    assert not batch2 and len(keep2) == len(keep)
    # a tiny budget still admits the oldest request
    b3, k3 = serve.pick(deque(_reqs()), {}, budget=1)
    assert len(b3) == 1 and b3[0].rid == 0 and len(k3) == 27


class _HostContext:
    """Stand-in for BatchSimplifier: echoes the uploaded formula as the result."""
    fail = False

    def __init__(self, nvars, opts=None):
        self.nvars_cap = int(nvars[0])

    def relayout(self, nv):
        self.nv = np.asarray(nv)

    def upload_device(self, d_lits, d_offs, ncls, nlits):
        pass

    def ifreeze_instance(self, i, v):
        pass

    def enqueue_instance(self, i, u):
        pass

    def sigmify_batch(self):
        if _HostContext.fail:
            raise RuntimeError("device failure")
        return np.zeros(len(self.nv), np.int32)

    def instance(self, i, pool=None):
        return (np.zeros(3, np.uint32), np.array([2 * i + 2], np.uint32), np.zeros(0, np.uint32),
                np.zeros(int(self.nv[i]) + 1, np.uint8))

    def export(self, i, device=None):
        import torch
        return torch.arange(5, dtype=torch.int32), torch.tensor([0, 2, 5], dtype=torch.int64)

    def close(self):
        pass


def test_request_loop_serves_every_request_in_order():
    from parafrost_amd import serve
    reqs = _reqs()
    srv = serve.Server(slots=2, batch_lits=1.5e5, store_bytes=1, keep_results=True, device="cpu",
                       context=_HostContext)
    st = srv.run(reqs, rate=5000.0)
    assert st["requests"] == len(reqs) and st["spills"] > 0
    assert all(r.status == 0 and r.result is not None for r in reqs)
    done = {}
    for r in sorted(reqs, key=lambda q: q.done):   # an instance's requests finish in order
        assert done.get(r.key, -1) < r.rid
        done[r.key] = r.rid
    # extensions simplified the resident clauses (5 literals) plus their new frames
    assert all(r.nlits == (5 if r.t0 else 0) + r.cnf.nlits for r in reqs)


def test_request_loop_reports_worker_failure():
    """A failing worker stops the loop and run() raises instead of waiting forever."""
    from parafrost_amd import serve
    _HostContext.fail = True
    try:
        srv = serve.Server(slots=2, batch_lits=1.5e5, device="cpu", context=_HostContext)
        with pytest.raises(RuntimeError, match="device failure"):
            srv.run(_reqs(), rate=5000.0)
    finally:
        _HostContext.fail = False


def test_host_pool_slices_stay_valid_until_reset():
    from parafrost_amd import HostPool
    p = HostPool()
    a = p.take("w", 10, np.uint32)
    a[:] = 7
    b = p.take("w", 1 << 17, np.uint32)   # grows: a new array, a keeps its memory
    b[:] = 9
    assert (a == 7).all() and (b == 9).all() and not np.shares_memory(a, b)
    c = p.take("w", 5, np.uint32)
    assert not np.shares_memory(b, c)
    p.reset()
    d = p.take("w", 5, np.uint32)   # after reset the newest array is reused from its start
    assert np.shares_memory(c, d)
