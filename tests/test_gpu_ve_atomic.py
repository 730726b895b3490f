"""BVE with atomic placement (ve_atomic_en; the GPU reference's veatomic, ve_k in
src/gpu/kernels.cu:298-373): eliminated variables reserve their resolvents, units and
model records with atomics, so the order of the pool's new clauses and of the model
stack depends on the schedule. src/cpu has no such mode, so the oracle does not give
its results; they are pinned two ways:

* after one BVE pass the atomic mode holds exactly what the scanned (elected-order)
  mode holds, which the lock-step tests pin to the oracle: the same clause images
  before the pass's appended region, the same multiset of appended clauses, the same
  multiset of model records, variable states and units;
* whole calls on formulas with a planted model A: every clause left is satisfied by A
  (resolvents, substitutions and strengthened clauses are consequences of the
  formula), so is every unit; no eliminated variable is left in a clause; and the
  model extension of A restricted to the variables left, with random values for the
  eliminated ones, satisfies the input formula. The scanned mode goes through the
  same checks.
"""
import numpy as np
import pytest

import parafrost_amd as pf
from parafrost_amd import gen

pytestmark = pytest.mark.gpu


def _planted(nv, ncls, k, seed):
    """random k-SAT clauses over distinct variables, each satisfied by the assignment A"""
    rng = np.random.default_rng(seed)
    A = rng.integers(0, 2, nv + 1).astype(np.uint8)
    out = []
    need = ncls
    while need > 0:
        m = 2 * need + 64
        v = rng.integers(1, nv + 1, (m, k))
        v.sort(axis=1)
        ok = np.all(v[:, 1:] != v[:, :-1], axis=1)
        s = rng.integers(0, 2, (m, k))
        sat = np.any(A[v] == (1 - s), axis=1)   # literal 2v+s is true iff A[v] == 1 - s
        good = (v * 2 + s)[ok & sat][:need]
        out.append(good)
        need -= len(good)
    lits = np.concatenate(out).astype(np.uint32).reshape(-1)
    offs = np.arange(0, ncls * k + 1, k, dtype=np.uint64)
    return gen.Cnf(nv, lits, offs), A


def _records(res):
    """model-stack records (tuples of literals, witness first), last record first"""
    out, x = [], len(res) - 1
    while x > 0:
        k = int(res[x])
        out.append(tuple(int(t) for t in res[x - k:x]))
        x -= k + 1
    return out


def _extend(res, value):
    """the model extension replay (src/cpu/model.cpp:122-139): from the last record to
    the first, a record none of whose literals past the witness is true sets its witness"""
    for rec in _records(res):
        if not any(value[l >> 1] == 1 - (l & 1) for l in rec[1:]):
            w = rec[0]
            value[w >> 1] = 1 - (w & 1)
    return value


def _clauses(words):
    out, i = [], 0
    while i < len(words):
        n = int(words[i + 2])
        out.append((int(words[i]), tuple(int(x) for x in words[i + 3:i + 3 + n])))
        i += 3 + n
    return out


def _to_ve(g, cnf):
    g.upload(cnf)
    g.resizeCNF()
    g.createOT()
    g.prop()
    npv = g.LCVE()
    g.sortOT()
    g.SUB()
    return npv


@pytest.mark.parametrize("kind", ["uniform", "geometric", "scalefree"])
def test_one_pass_equals_scanned(kind):
    if kind == "uniform":
        cnf = gen.uniform_ksat(30000, 126000, 3, 8801)
    elif kind == "geometric":
        cnf = gen.geometric(30000, 126000, 2, 8, 8802)
    else:
        cnf = gen.scalefree(30000, 120000, 2.5, 0.2, 8803)
    print(f"ve_atomic one pass {kind}: {cnf.nvars} vars, {cnf.ncls} clauses", flush=True)
    a = pf.Simplifier(cnf.nvars, pf.default_opts(ve_atomic_en=1))
    s = pf.Simplifier(cnf.nvars, pf.default_opts())
    try:
        npv = _to_ve(a, cnf)
        assert _to_ve(s, cnf) == npv
        assert npv > 512   # several reserving blocks of 256
        w0 = s.cnf_words()
        assert np.array_equal(a.cnf_words(), w0)
        n0 = len(w0)
        a.VE()
        s.VE()
        wa, ws = a.cnf_words(), s.cnf_words()
        assert len(wa) == len(ws) > n0
        assert np.array_equal(wa[:n0], ws[:n0])
        ca, cs = _clauses(wa[n0:]), _clauses(ws[n0:])
        assert sorted(ca) == sorted(cs)
        ra, rs = a.resolved(), s.resolved()
        assert len(ra) == len(rs)
        assert sorted(_records(ra)) == sorted(_records(rs))
        assert np.array_equal(a.vstate(), s.vstate())
        assert np.array_equal(np.sort(a.trail()), np.sort(s.trail()))
        print(f"  {npv} elected, {len(cs)} resolvents, order differs: {ca != cs}", flush=True)
    finally:
        a.close()
        s.close()


@pytest.mark.parametrize("atomic", [1, 0])
@pytest.mark.parametrize("k,ratio,seed", [(3, 4.0, 1), (3, 3.6, 2), (3, 3.0, 3)])
def test_planted_model_extension(atomic, k, ratio, seed):
    nv = 20000
    cnf, A = _planted(nv, int(nv * ratio), k, 8900 + seed)
    print(f"ve_atomic={atomic} planted k={k} ratio={ratio}", flush=True)
    g = pf.Simplifier(nv, pf.default_opts(ve_atomic_en=atomic))
    try:
        g.upload(cnf)
        g.sigmify()   # satisfiable by construction: no Unsat
        vs = g.vstate()
        elim = vs[: nv + 1] == 1
        assert elim.sum() > 100
        for hdr, lits in _clauses(g.cnf_words()):
            if hdr & 2:
                continue
            assert any(A[l >> 1] == 1 - (l & 1) for l in lits)
            assert not any(elim[l >> 1] for l in lits)
        for u in g.trail():
            assert A[int(u) >> 1] == 1 - (int(u) & 1)
        rng = np.random.default_rng(seed)
        value = A.copy()
        value[elim] = rng.integers(0, 2, int(elim.sum()))
        value = _extend(g.resolved(), value)
        L = cnf.lits.astype(np.int64)
        ok = value[L >> 1] == 1 - (L & 1)
        assert np.all(np.logical_or.reduceat(ok, cnf.offs[:-1].astype(np.int64)))
    finally:
        g.close()


def test_batch_rejects_atomic():
    with pytest.raises(pf.SimplifierError):
        pf.BatchSimplifier([100, 200], pf.default_opts(ve_atomic_en=1))
