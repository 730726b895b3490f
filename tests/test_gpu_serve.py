"""Serving with continuous batching over incremental requests: every request's
result equals the same request chain run on the plain per-instance path (one
request at a time, no batching, no residency) and the oracle's chain, bit for
bit, whatever it was batched with and whether its instance stayed in HBM or was
moved to host memory in between.
"""
import numpy as np
import pytest

import pyoracle as po
from parafrost_amd import gen

pytestmark = pytest.mark.gpu


def _plain(reqs):
    from parafrost_amd import serve
    cap = max(r.nvars for r in reqs)
    p = serve.PlainChain(cap)
    try:
        return {r.rid: p.run(r) for r in reqs}
    finally:
        p.close()


def _same(a, b, rid):
    for x, y, what in zip(a, b, ("clauses", "units", "model records", "variable states")):
        if what == "units":
            x, y = np.sort(x), np.sort(y)
        assert np.array_equal(x, y), f"request {rid}: {what} differ"


def _check_against_plain(reqs):
    plain = _plain(reqs)
    for r in reqs:
        st, res, nl = plain[r.rid]
        assert st == r.status, f"request {r.rid}: status plain {st}, served {r.status}"
        if st:
            continue
        assert nl == r.nlits, f"request {r.rid}: formula size plain {nl}, served {r.nlits}"
        _same(res, r.result, r.rid)


def test_serving_equals_plain_chain():
    from parafrost_amd import serve
    reqs = serve.make_requests(28, seed=31, gates=(1500, 5000), max_frames=10)
    assert sum(r.t0 > 0 for r in reqs) >= 5, "the stream should hold extension requests"
    srv = serve.Server(slots=2, batch_lits=1.5e5, keep_results=True)
    stats = srv.run(reqs, rate=400.0)
    print(stats, flush=True)
    assert stats["requests"] == len(reqs) and stats["batches"] > 1
    _check_against_plain(reqs)
    # residency pays: an extension simplifies less than its whole unrolled instance
    ext = [r for r in reqs if r.t0 > 0 and r.status == 0]
    assert ext and all(r.nlits < gen.aig_bmc(r.seed, r.ngates, 0, r.t1)[0].nlits for r in ext)


@pytest.mark.parametrize("case", ["small", "c5_gates"])
def test_serving_matches_oracle_chain(case):
    """Every request against the oracle's own request chain. c5_gates: requests in the
    serving workload's size range (10k-100k AND gates), one of them an extension."""
    from parafrost_amd import serve
    if case == "small":
        reqs = serve.make_requests(10, seed=32, gates=(300, 900), max_frames=6, extend_frac=0.6)
        srv = serve.Server(slots=2, batch_lits=4e3, keep_results=True)
    else:
        reqs = serve.make_requests(4, seed=36, gates=(10_000, 100_000), max_frames=5, extend_frac=0.6)
        assert any(r.t0 > 0 for r in reqs) and min(r.ngates for r in reqs) >= 10_000
        srv = serve.Server(slots=2, batch_lits=2e6, keep_results=True)
    print(f"serving vs oracle chain: {case}, gates {[r.ngates for r in reqs]}", flush=True)
    srv.run(reqs, rate=1000.0)
    state = {}
    for r in reqs:
        prev = state.get(r.key)
        if prev is not None and prev[2]:
            assert r.status != 0
            continue
        if prev is None:
            cnf, units = r.cnf, np.zeros(0, np.uint32)
        else:
            c, units, _ = prev
            cnf = gen.Cnf(r.nvars, np.concatenate([c.lits, r.cnf.lits]),
                          np.concatenate([c.offs, r.cnf.offs[1:] + np.uint64(c.nlits)]))
        o = po.Oracle(gen.Cnf(r.nvars, cnf.lits, cnf.offs), po.default_opts())
        o.ifreeze(r.frozen())
        assert o.enqueue(units)
        ro = o.sigmify()
        assert (ro != 0) == (r.status != 0), f"request {r.rid}"
        if ro:
            state[r.key] = (None, None, True)
            continue
        so = o.state()
        _same((so.cnf, so.trail, so.resolved, so.vstate), r.result, r.rid)
        fz = r.frozen()
        assert not np.any(so.vstate[fz] == 1), "a frozen interface variable was eliminated"
        state[r.key] = (gen.images_to_csr(so.cnf, r.nvars), so.trail, False)


def test_serving_store_spills_to_host():
    """With a tiny HBM budget idle instances move to host memory between their
    requests; results are unchanged."""
    from parafrost_amd import serve
    reqs = serve.make_requests(14, seed=33, gates=(800, 2000), max_frames=8, extend_frac=0.7)
    srv = serve.Server(slots=1, batch_lits=5e4, store_bytes=1, keep_results=True)
    stats = srv.run(reqs, rate=2000.0)
    assert stats["spills"] > 0
    _check_against_plain(reqs)


def test_serving_with_reused_host_memory():
    """Results not kept: every slot downloads into host memory it reuses from batch to
    batch (so does the plain chain). The instances' state must not live there: the
    statuses and formula sizes of a long extension chain equal the plain chain's."""
    from parafrost_amd import serve
    reqs = serve.make_requests(24, seed=34, gates=(800, 3000), max_frames=8, extend_frac=0.8)
    serve.Server(slots=2, batch_lits=3e4).run(reqs, rate=2000.0)
    p = serve.PlainChain(max(r.nvars for r in reqs), reuse_host=True)
    try:
        for r in reqs:
            st, _, nl = p.run(r)
            assert st == r.status, f"request {r.rid}: status plain {st}, served {r.status}"
            if st == 0:
                assert nl == r.nlits, f"request {r.rid}: formula size plain {nl}, served {r.nlits}"
    finally:
        p.close()
