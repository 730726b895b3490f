"""Batches of independent instances in one context: each instance must come out
exactly as the oracle (and the single-instance GPU path) simplifies it alone --
clause database, model stack, variable states, units (as a set) and the
unsatisfiable verdict -- whatever the other instances in the batch do.
"""
import numpy as np
import pytest

import pyoracle as po
from parafrost_amd import gen

pytestmark = pytest.mark.gpu


def _instance(rng, k):
    nv = int(rng.integers(20, 600))
    kind = k % 4
    seed = int(rng.integers(1, 1 << 30))
    if kind == 0:
        return gen.uniform_ksat(nv, int(nv * rng.uniform(2.0, 4.6)), 3, seed)
    if kind == 1:
        return gen.geometric(nv, int(nv * rng.uniform(1.5, 4.0)), 2, 8, seed)
    if kind == 2:
        nv = max(nv, 64)
        return gen.scalefree(nv, int(nv * rng.uniform(2.0, 5.0)), 2.5, 0.3, seed)
    # dense binary-heavy scale-free instances: often refuted by the simplifier
    nv = max(nv, 64)
    return gen.scalefree(nv, int(nv * rng.uniform(4.0, 8.0)), 2.5, 0.1, seed)


def _full_3cnf():
    """All eight clauses over three variables: unsatisfiable."""
    lits, offs = [], [0]
    for m in range(8):
        for v in range(3):
            lits.append(2 * (v + 1) + ((m >> v) & 1))
        offs.append(len(lits))
    return gen.Cnf(3, np.array(lits, np.uint32), np.array(offs, np.uint64))


def _check(cnfs, g=None, freezes=None, units=None, **kw):
    """freezes / units: per instance (own numbering), or None"""
    import parafrost_amd as pf
    own = g is None
    if own:
        g = pf.BatchSimplifier([c.nvars for c in cnfs], pf.default_opts(**kw))
    else:
        g.relayout([c.nvars for c in cnfs])
    try:
        g.upload_instances(cnfs)
        for i in range(len(cnfs)):
            if freezes is not None and len(freezes[i]):
                g.ifreeze_instance(i, freezes[i])
            if units is not None and len(units[i]):
                g.enqueue_instance(i, units[i])
        status = g.sigmify_batch()
        n_unsat = 0
        for i, cnf in enumerate(cnfs):
            o = po.Oracle(cnf, po.default_opts(**kw))
            if freezes is not None:
                o.ifreeze(freezes[i])
            if units is not None and len(units[i]):
                assert o.enqueue(units[i])
            ro = o.sigmify()
            assert (ro != 0) == (status[i] != 0), f"instance {i}: unsat oracle {ro} batch {status[i]}"
            if ro != 0:
                n_unsat += 1
                continue
            so = o.state()
            w, t, r, vs = g.instance(i)
            assert np.array_equal(so.cnf, w), f"instance {i}: clause database differs"
            assert np.array_equal(np.sort(so.trail), np.sort(t)), f"instance {i}: units differ"
            assert np.array_equal(so.resolved, r), f"instance {i}: model stack differs"
            assert np.array_equal(so.vstate, vs), f"instance {i}: variable states differ"
        return n_unsat
    finally:
        if own:
            g.close()


@pytest.mark.parametrize("seed", range(6))
def test_batch_matches_each_instance_alone(seed):
    print(f"batch seed={seed}", flush=True)
    rng = np.random.default_rng(4400 + seed)
    cnfs = [_instance(rng, k) for k in range(int(rng.integers(8, 24)))]
    if seed % 2 == 0:
        cnfs.insert(int(rng.integers(0, len(cnfs))), _full_3cnf())
    kw = dict(phases=int(rng.integers(1, 7)), mu_pos=int(rng.integers(10, 60)), mu_neg=int(rng.integers(10, 60)))
    _check(cnfs, **kw)


@pytest.mark.parametrize("seed", range(3))
def test_batch_frozen_and_units(seed):
    """Incremental mode in a batch: per-instance frozen variables and root units."""
    print(f"batch frozen seed={seed}", flush=True)
    rng = np.random.default_rng(4600 + seed)
    cnfs = [_instance(rng, k) for k in range(int(rng.integers(6, 16)))]
    freezes, units = [], []
    for c in cnfs:
        freezes.append(np.unique(rng.integers(1, c.nvars + 1, max(1, int(c.nvars * rng.uniform(0.0, 0.4))))).astype(np.uint32))
        if rng.random() < 0.5:
            vs = rng.choice(np.arange(1, c.nvars + 1), size=max(1, c.nvars // 50), replace=False)
            units.append((2 * vs + rng.integers(0, 2, len(vs))).astype(np.uint32))
        else:
            units.append(np.zeros(0, np.uint32))
    _check(cnfs, freezes=freezes, units=units, phases=int(rng.integers(2, 6)))


def test_batch_with_refuted_instances():
    rng = np.random.default_rng(4499)
    cnfs = [_instance(rng, 3) for _ in range(10)] + [_full_3cnf(), _instance(rng, 0)]
    assert _check(cnfs) >= 1


def test_batch_of_one_equals_single():
    rng = np.random.default_rng(4500)
    _check([_instance(rng, 1)])


def test_batch_option_variants():
    rng = np.random.default_rng(4501)
    cnfs = [_instance(rng, k) for k in range(9)]
    for kw in (dict(ere_en=0), dict(sub_en=0), dict(lcve_min=30), dict(phases=8, lits_min=1 << 40)):
        print(f"batch options {kw}", flush=True)
        _check(cnfs, **kw)


def test_batch_bce():
    """Blocked clause elimination in a batch: its model records join each
    instance's record chunks in stack order."""
    from blocked import with_blocked_gadgets
    rng = np.random.default_rng(4503)
    cnfs = [with_blocked_gadgets(_instance(rng, k), int(rng.integers(1, 12)), 4503 + k) for k in range(12)]
    cnfs.insert(5, _full_3cnf())
    _check(cnfs, bce_en=1, phases=4)


def test_batch_adversarial():
    """Duplicate clauses, clauses past the subsumption size break, a hub variable, pure
    literals and implication chains, batched with ordinary and refuted instances."""
    from adversarial import adversarial
    rng = np.random.default_rng(4504)
    cnfs = [adversarial(s) for s in range(4)] + [_instance(rng, k) for k in range(6)] + [_full_3cnf()]
    _check(cnfs, phases=4)
    _check(cnfs[:4], bce_en=1, phases=3, mu_pos=60, mu_neg=60)


def test_batch_relayout_reuses_context():
    """One context laid out again for another set of instances gives the same results."""
    import parafrost_amd as pf
    from parafrost_amd import gen as G
    rng = np.random.default_rng(4502)
    g1 = [_instance(rng, k) for k in range(7)]
    g2 = [_instance(rng, k + 1) for k in range(11)] + [_full_3cnf()]
    cap = max(int(G.batch_bases([c.nvars for c in grp])[-1]) for grp in (g1, g2))
    g = pf.BatchSimplifier([cap], pf.default_opts())
    try:
        _check(g1, g)
        _check(g2, g)
        _check(g1, g)
    finally:
        g.close()


def test_batch_generator_instances_roundtrip():
    """The batch generator's instance view is the instance in its own numbering."""
    from parafrost_amd import gen as G
    b = G.c4(ninst=12, vmax=3000, threads=4)
    for i in range(b.ninst):
        c = b.instance(i)
        assert c.nvars == int(b.inst_nvars[i])
        assert c.lits.min() >= 2 and c.lits.max() <= 2 * c.nvars + 1
        assert c.nlits == c.ncls * int(b.kinds[i])


def _csr(words):
    """clause images -> (literals, offsets)"""
    lits, offs, i = [], [0], 0
    while i < len(words):
        n = int(words[i + 2])
        lits.append(words[i + 3:i + 3 + n])
        offs.append(offs[-1] + n)
        i += 3 + n
    return (np.concatenate(lits) if lits else np.zeros(0, np.uint32)), np.array(offs, np.uint64)


def test_batch_view_matches_download():
    """pfg_batch_view: the clause words and model records in place in the pinned cache equal
    the downloaded copies, for every instance (refuted ones included)."""
    import parafrost_amd as pf
    rng = np.random.default_rng(4750)
    cnfs = [_instance(rng, k) for k in range(11)] + [_full_3cnf()]
    g = pf.BatchSimplifier([c.nvars for c in cnfs], pf.default_opts())
    try:
        g.upload_instances(cnfs)
        g.sigmify_batch()
        for i in range(len(cnfs)):
            a, b = g.instance(i), g.instance_view(i)
            for x, y, what in zip(a, b, ("words", "trail", "records", "vstate")):
                assert x.dtype == y.dtype and np.array_equal(x, y), f"instance {i}: {what}"
    finally:
        g.close()


def test_batch_export_matches_download():
    """The HBM write-back of every instance is the CSR form of its downloaded clauses."""
    import torch
    import parafrost_amd as pf
    rng = np.random.default_rng(4700)
    cnfs = [_instance(rng, k) for k in range(9)] + [_full_3cnf()]
    g = pf.BatchSimplifier([c.nvars for c in cnfs], pf.default_opts())
    try:
        g.upload_instances(cnfs)
        st = g.sigmify_batch()
        for i in range(len(cnfs)):
            w, _, _, _ = g.instance(i)
            dl, do = g.export(i)
            el, eo = _csr(w)
            assert np.array_equal(dl.cpu().numpy().view(np.uint32), el), f"instance {i}: literals"
            assert np.array_equal(do.cpu().numpy().view(np.uint64), eo), f"instance {i}: offsets"
            if st[i] == 0 and len(eo) > 1:
                # what was exported uploads again as the same clauses
                h = pf.Simplifier(cnfs[i].nvars, pf.default_opts(phases=0))
                try:
                    h.upload_device(dl.data_ptr(), do.data_ptr(), len(eo) - 1, len(el))
                    assert np.array_equal(_csr(h.cnf_words())[0], el)
                finally:
                    h.close()
    finally:
        g.close()


def test_c4_workload_instances_match_oracle():
    """24 instances of the batching workload itself (the c4 generator's own specs and
    seeds: random 3-SAT at ratio 4.0-4.3 and 4-SAT at 9.5-10.0, variables log-uniform in
    [1,000, 200,000]), taken evenly over the size order -- from 1,003 to 199,129
    variables, three above 100,000, 4-SAT and 3-SAT mixed, 22.6M literals -- in one
    batch, every instance against the oracle alone."""
    from parafrost_amd import gen as G
    nv, ncls, ks, seeds = G.c4_specs(4096)
    order = np.argsort(nv, kind="stable")
    pick = np.sort(order[np.linspace(0, len(order) - 1, 24).round().astype(np.int64)])
    assert (nv[pick] > 100_000).sum() >= 2 and set(ks[pick].tolist()) == {3, 4}
    b = G.batch_ksat(nv[pick], ncls[pick], ks[pick], seeds[pick])
    cnfs = [b.instance(i) for i in range(b.ninst)]
    print(f"c4 instances: {b.ninst}, {b.nlits} literals", flush=True)
    _check(cnfs)
