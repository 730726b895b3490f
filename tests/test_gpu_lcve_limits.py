"""The GPU reference's optional LCVE limits (lcve_max, lcve_clause_max; off by default)
pass by pass against the oracle: candidates over the occurrence limit end the walk
where it reaches them unfrozen (dynamic breaks that are not hot by list size, so many
of them), candidates with a long clause on the positive list are passed over, and
those with one on the negative list freeze through their positive list only. The
election rounds then hold candidates that do not freeze through a clause ranked
before an elected one that does, which must leave them undecided."""
import numpy as np
import pytest

import lockstep as ls
from parafrost_amd import gen
from test_oracle_lcve_limits import HAND, _cnf

pytestmark = pytest.mark.gpu


def test_hand_built():
    for kw in (dict(), dict(lcve_clause_max=2), dict(lcve_clause_max=3)):
        print(f"hand-built {kw}", flush=True)
        ls.run_lockstep(_cnf(5, HAND), phases=1, **kw)


CASES = [("uniform", dict(lcve_max=5)), ("uniform", dict(lcve_max=8)), ("geometric", dict(lcve_max=6)),
         ("scalefree", dict(lcve_max=10)), ("geometric", dict(lcve_clause_max=3)),
         ("geometric", dict(lcve_clause_max=5)), ("scalefree", dict(lcve_clause_max=4)),
         ("uniform", dict(lcve_clause_max=2)), ("geometric", dict(lcve_max=7, lcve_clause_max=4)),
         ("scalefree", dict(lcve_max=12, lcve_clause_max=6, lcve_graph=1))]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_limits_lockstep(k):
    kind, kw = CASES[k]
    kw = dict(kw)
    graph = kw.pop("lcve_graph", 0)
    for seed in range(2):
        rng = np.random.default_rng(9700 + 10 * k + seed)
        nv = int(rng.integers(2000, 8000))
        if kind == "uniform":
            cnf = gen.uniform_ksat(nv, int(nv * 4.2), 3, 9700 + seed)
        elif kind == "geometric":
            cnf = gen.geometric(nv, int(nv * 3.6), 2, 8, 9700 + seed)
        else:
            cnf = gen.scalefree(nv, int(nv * 3.5), 2.5, 0.2, 9700 + seed)
        print(f"{kind} {kw} graph={graph} seed {seed}: {nv} vars", flush=True)
        ls.run_lockstep(cnf, setup=(lambda g: g.set_tuning(lcve_graph=1)) if graph else None, phases=4, **kw)


def test_limits_in_a_batch():
    """A batch with both limits equals its instances alone."""
    import parafrost_amd as pf
    import pyoracle as po
    rng = np.random.default_rng(9800)
    cnfs = [gen.geometric(int(n), int(n * 3.6), 2, 8, 9800 + i) for i, n in enumerate(rng.integers(300, 4000, 10))]
    kw = dict(lcve_max=7, lcve_clause_max=4)
    g = pf.BatchSimplifier([c.nvars for c in cnfs], pf.default_opts(**kw))
    try:
        g.upload_instances(cnfs)
        st = g.sigmify_batch()
        for i, c in enumerate(cnfs):
            o = po.Oracle(c, po.default_opts(**kw))
            ro = o.sigmify()
            assert (ro != 0) == (st[i] != 0)
            if ro == 0:
                w, t, r, vs = g.instance(i)
                so = o.state()
                assert np.array_equal(so.cnf, w) and np.array_equal(so.resolved, r) and np.array_equal(so.vstate, vs)
    finally:
        g.close()
