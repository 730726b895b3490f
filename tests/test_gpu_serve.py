"""Serving with continuous batching: every request's result is the same as when
the request is simplified alone (bit-exact), whatever it was batched with, and
resident instances grown by extension requests equal the instance generated whole.
"""
import numpy as np
import pytest

import pyoracle as po
from parafrost_amd import gen

pytestmark = pytest.mark.gpu


def _alone(r):
    import parafrost_amd as pf
    cnf, nv = gen.aig_bmc(r.seed, r.ngates, 0, r.t1)
    assert nv == r.nvars
    g = pf.BatchSimplifier([nv], pf.default_opts())
    try:
        g.upload_instances([cnf])
        st = g.sigmify_batch()
        return int(st[0]), g.instance(0), cnf
    finally:
        g.close()


def test_serving_results_equal_alone():
    from parafrost_amd import serve
    reqs = serve.make_requests(28, seed=31, gates=(1500, 5000), max_frames=10)
    assert any(r.t0 > 0 for r in reqs), "the stream should hold extension requests"
    srv = serve.Server(slots=2, batch_lits=1.5e5, keep_results=True)
    stats = srv.run(reqs, rate=400.0)
    print(stats, flush=True)
    assert stats["requests"] == len(reqs)
    for r in reqs:
        st, res, cnf = _alone(r)
        assert r.nlits == cnf.nlits
        assert st == r.status, f"request {r.rid}: status alone {st}, in the mix {r.status}"
        if st:
            continue
        for a, b, what in zip(res, r.result, ("clauses", "units", "model records", "variable states")):
            if what == "units":
                a, b = np.sort(a), np.sort(b)
            assert np.array_equal(a, b), f"request {r.rid}: {what} differ alone vs in the mix"


def test_serving_matches_oracle_on_small_requests():
    from parafrost_amd import serve
    reqs = serve.make_requests(6, seed=32, gates=(300, 900), max_frames=4)
    srv = serve.Server(slots=2, batch_lits=4e3, keep_results=True)
    srv.run(reqs, rate=1000.0)
    for r in reqs:
        cnf, _ = gen.aig_bmc(r.seed, r.ngates, 0, r.t1)
        o = po.Oracle(cnf, po.default_opts())
        ro = o.sigmify()
        assert (ro != 0) == (r.status != 0)
        if ro:
            continue
        so = o.state()
        w, t, res, vs = r.result
        assert np.array_equal(so.cnf, w)
        assert np.array_equal(np.sort(so.trail), np.sort(t))
        assert np.array_equal(so.resolved, res)
        assert np.array_equal(so.vstate, vs)


def test_serving_store_eviction_regenerates():
    """With a tiny resident budget instances are evicted and rebuilt from their seed;
    results still equal the alone runs."""
    from parafrost_amd import serve
    reqs = serve.make_requests(10, seed=33, gates=(800, 2000), max_frames=8, extend_frac=0.7)
    srv = serve.Server(slots=1, batch_lits=5e4, store_bytes=1, keep_results=True)
    srv.run(reqs, rate=2000.0)
    for r in reqs:
        st, res, _ = _alone(r)
        assert st == r.status
        if not st:
            assert np.array_equal(res[0], r.result[0]) and np.array_equal(res[2], r.result[2])
