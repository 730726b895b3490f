"""Every tuning gives the same results: the occurrence-table build (radix sort or
counting), the XCD-window grid of the clause passes, the subsumption lane group, the
BVE detection arena and the LCVE record-pass grid are compared with the oracle pass by
pass, single instances and batches."""
import numpy as np
import pytest

import lockstep as ls
from parafrost_amd import gen

pytestmark = pytest.mark.gpu

TUNINGS = [dict(ot_build=1), dict(ot_build=1, clause_grid=8), dict(sub_group=8), dict(sub_group=32),
           dict(sub_group=64), dict(ot_build=1, sub_group=8, clause_grid=8192), dict(lcve_graph=1),
           dict(lcve_graph=1, clause_grid=8), dict(ve_arena=4096), dict(ve_arena=4096, init_grid=8),
           dict(init_grid=512, ve_arena=4096, stage_words=1500)]


@pytest.mark.parametrize("k", range(len(TUNINGS)))
def test_tunings_lockstep(k):
    t = TUNINGS[k]
    print(f"tuning {t}", flush=True)
    graph_calls = 0
    for seed in range(3):
        rng = np.random.default_rng(9300 + seed)
        nv = int(rng.integers(200, 3000))
        if seed == 0:
            cnf = gen.uniform_ksat(nv, int(nv * 4.2), 3, 9300 + seed)
        elif seed == 1:
            cnf = gen.geometric(nv, int(nv * 3.0), 2, 8, 9300 + seed)
        else:   # long lists: staged-size limits of the lane groups, std::sort-ordered lists
            cnf = gen.scalefree(max(nv, 2000), int(max(nv, 2000) * 4.0), 2.5, 0.2, 9300 + seed)
        seen = {}
        ls.run_lockstep(cnf, setup=lambda g: g.set_tuning(**t), phases=4, mu_pos=64, mu_neg=64,
                        inspect=lambda g: seen.update(g.path_counts()))
        graph_calls += seen.get("lcve_graph_calls", 0)   # (a refuted formula ends before the last comparison)
    if t.get("lcve_graph"):
        assert graph_calls >= 1   # the election rounds ran from the captured graph


def test_tuning_batch_counting_build():
    """A batch with the counting build and small groups equals its instances alone."""
    import parafrost_amd as pf
    import pyoracle as po
    rng = np.random.default_rng(9400)
    cnfs = [gen.uniform_ksat(int(n), int(n * rng.uniform(3.8, 4.3)), 3, 9400 + i)
            for i, n in enumerate(rng.integers(100, 3000, 12))]
    g = pf.BatchSimplifier([c.nvars for c in cnfs], pf.default_opts())
    try:
        g.set_tuning(ot_build=1, sub_group=8, clause_grid=64, lcve_graph=1)
        g.upload_instances(cnfs)
        st = g.sigmify_batch()
        for i, c in enumerate(cnfs):
            o = po.Oracle(c, po.default_opts())
            ro = o.sigmify()
            assert (ro != 0) == (st[i] != 0)
            if ro == 0:
                w, t, r, vs = g.instance(i)
                so = o.state()
                assert np.array_equal(so.cnf, w) and np.array_equal(so.resolved, r) and np.array_equal(so.vstate, vs)
    finally:
        g.close()


def test_tuning_rejects_bad_values():
    import parafrost_amd as pf
    g = pf.Simplifier(10, pf.default_opts())
    try:
        with pytest.raises(pf.SimplifierError):
            g.set_tuning(sub_group=12)
        with pytest.raises(pf.SimplifierError):
            g.set_tuning(ot_build=2)
        with pytest.raises(pf.SimplifierError):
            g.set_tuning(ve_arena=2048)
        assert g.tuning()["sub_group"] in (8, 16, 32, 64)
    finally:
        g.close()
