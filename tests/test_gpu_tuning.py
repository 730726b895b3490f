"""Every tuning gives the same results: the occurrence-table build (radix sort updated
between rounds, radix sort every round, counting), the XCD-window grid of the clause
passes, the subsumption lane group, the BVE detection arena and the LCVE record-pass
grid are compared with the oracle pass by pass, single instances and batches. With
ot_build 3 every occurrence-table update is also checked against a full sort (the
library raises on a difference)."""
import numpy as np
import pytest

import adversarial as adv
import lockstep as ls
from parafrost_amd import gen

pytestmark = pytest.mark.gpu

TUNINGS = [dict(ot_build=1), dict(ot_build=1, clause_grid=8), dict(sub_group=8), dict(sub_group=32),
           dict(sub_group=64), dict(ot_build=1, sub_group=8, clause_grid=8192), dict(lcve_graph=1),
           dict(lcve_graph=1, clause_grid=8), dict(ve_arena=4096), dict(ve_arena=4096, init_grid=8),
           dict(init_grid=512, ve_arena=4096, stage_words=1500), dict(ot_build=2), dict(ot_build=3),
           dict(ot_build=3, lcve_graph=1, ve_arena=4096)]


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
            g.set_tuning(ot_build=4)
        with pytest.raises(pf.SimplifierError):
            g.set_tuning(ve_arena=2048)
        assert g.tuning()["sub_group"] in (8, 16, 32, 64)
    finally:
        g.close()


def _formula(kind, seed):
    rng = np.random.default_rng(9500 + seed)
    nv = int(rng.integers(8000, 20000))
    if kind == "uniform":
        return gen.uniform_ksat(nv, int(nv * 4.2), 3, 9500 + seed)
    if kind == "ksat4":
        return gen.uniform_ksat(nv, int(nv * 9.7), 4, 9500 + seed)
    if kind == "geometric":
        return gen.geometric(nv, int(nv * 4.2), 2, 8, 9500 + seed)
    if kind == "scalefree":
        return gen.scalefree(nv, int(nv * 4.0), 2.5, 0.2, 9500 + seed)
    if kind == "equivalences":
        return adv.equivalences(seed)
    return adv.adversarial(seed)


@pytest.mark.parametrize("kind", ["uniform", "ksat4", "geometric", "scalefree", "equivalences", "adversarial"])
def test_ot_update_equals_full_sort(kind):
    """Whole sigmify calls with every occurrence-table update checked against a full
    sort, the results compared with the oracle."""
    import parafrost_amd as pf
    import pyoracle as po
    updates = 0
    for seed in range(2):
        cnf = _formula(kind, seed)
        print(f"{kind} seed {seed}: {cnf.nvars} vars, {cnf.ncls} clauses", flush=True)
        oo, go = ls._opts_pair(phases=5)
        o = po.Oracle(cnf, oo)
        ro = o.sigmify()
        g = pf.Simplifier(cnf.nvars, go)
        try:
            g.set_tuning(ot_build=3)
            g.upload(cnf)
            try:
                g.sigmify()
                rg = 0
            except pf.Unsat:
                rg = -1
            assert (ro != 0) == (rg != 0)
            if ro == 0:
                ls.compare(o.state(), ls.state_gpu(g), "sigmify")
            updates += g.path_counts()["ot_updates"]
        finally:
            g.close()
    # later rounds' tables came from updates; scale-free formulas change more than a
    # quarter of their keys per round, so their tables are sorted in full
    assert updates >= 1 or kind == "scalefree"


def test_ot_update_batch_checked():
    """A batch whose every occurrence-table update is checked against a full sort
    equals its instances alone (3-SAT, 4-SAT, planted equivalences)."""
    import parafrost_amd as pf
    import pyoracle as po
    rng = np.random.default_rng(9600)
    cnfs = []
    for i, n in enumerate(rng.integers(500, 6000, 14)):
        k = i % 3
        cnfs.append(gen.uniform_ksat(int(n), int(n * rng.uniform(4.0, 4.3)), 3, 9600 + i) if k == 0 else
                    gen.uniform_ksat(int(n), int(n * rng.uniform(9.5, 10.0)), 4, 9600 + i) if k == 1 else
                    adv.equivalences(100 + i))
    g = pf.BatchSimplifier([c.nvars for c in cnfs], pf.default_opts())
    try:
        g.set_tuning(ot_build=3)
        g.upload_instances(cnfs)
        st = g.sigmify_batch()
        assert g.path_counts()["ot_updates"] >= 1
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
