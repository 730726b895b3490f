"""Closed forms and invariants that pin the oracle's passes independently of the
reference build: literal histogram, occurrence lists, LCVE order and election."""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from parafrost_amd import gen


def _cnf(seed):
    rng = np.random.default_rng(seed)
    nv = int(rng.integers(50, 2000))
    return gen.geometric(nv, int(nv * rng.uniform(1.5, 4.0)), 2, 8, seed) if seed % 2 else \
        gen.uniform_ksat(nv, int(nv * rng.uniform(2.0, 4.3)), 3, seed)


@pytest.mark.parametrize("seed", range(6))
def test_hist_is_bincount(seed):
    cnf = _cnf(seed)
    o = po.Oracle(cnf)
    h = np.empty(2 * (cnf.nvars + 1), np.uint32)
    o.L.pfo_hist(o.h, h)
    assert np.array_equal(h, np.bincount(cnf.lits, minlength=len(h)))


@pytest.mark.parametrize("seed", range(6))
def test_ot_lists_are_clause_order_sets(seed):
    cnf = _cnf(seed)
    o = po.Oracle(cnf)
    o.L.pfo_create_ot(o.h)
    want = {}
    for i, c in enumerate(cnf.clauses()):
        for l in c:
            want.setdefault(int(l), []).append(i)
    for lit in range(2, 2 * (cnf.nvars + 1)):
        n = o.L.pfo_ot_size(o.h, lit)
        idx = np.empty(n, np.int64)
        o.L.pfo_ot_list(o.h, lit, idx.ctypes.data)
        assert idx.tolist() == want.get(lit, [])


def _live_neighbours(st):
    nb = {}
    for c in st.clauses():
        if c[0] & 2:
            continue
        vs = {l >> 1 for l in c[2]}
        for v in vs:
            nb.setdefault(v, set()).update(vs - {v})
    return nb


@pytest.mark.parametrize("seed", range(10))
def test_lcve_independent_and_greedy(seed):
    """Elected variables share no live clause; the election is the greedy prefix
    over (score, var) order: every skipped, electable candidate before the last
    elected one has an earlier elected neighbour (src/cpu/lcve.cpp:46-99)."""
    cnf = _cnf(seed)
    mu = 8 + seed * 3
    o = po.Oracle(cnf, po.default_opts(mu_pos=mu, mu_neg=mu))
    o.L.pfo_create_ot(o.h)
    o.L.pfo_prop(o.h)
    o.L.pfo_lcve(o.h)
    st = o.state()
    pvs = [int(v) for v in st.pvs]
    assert len(set(pvs)) == len(pvs)
    nb = _live_neighbours(st)
    S = set(pvs)
    for v in pvs:
        assert not (nb.get(v, set()) & S), "elected variables share a clause"
    h = np.bincount(cnf.lits, minlength=2 * (cnf.nvars + 1)).astype(np.uint64)
    ps, ns = h[0::2], h[1::2]
    score = (ps * ns) & 0xFFFFFFFF
    order = sorted(range(1, cnf.nvars + 1), key=lambda v: (int(score[v]), v))
    rank = {v: i for i, v in enumerate(order)}
    last = max(rank[v] for v in pvs) if pvs else -1
    # elected in rank order
    assert [rank[v] for v in pvs] == sorted(rank[v] for v in pvs)
    for v in order[:last + 1]:
        if v in S or (ps[v] == 0 and ns[v] == 0):
            continue
        cold = ps[v] < mu or ns[v] < mu
        earlier = [u for u in nb.get(v, set()) if u in S and rank[u] < rank[v]]
        assert cold and earlier, f"candidate {v} skipped without an earlier elected neighbour"
