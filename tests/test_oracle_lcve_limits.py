"""The GPU reference's two optional LCVE limits in the oracle (lcvemax and
lcveclausemax, src/gpu/lcve.cu:68, 105-127; src/cpu has neither, so both are off by
default): a hand-built formula whose elected variables follow from the reference's
rules, and properties the rules fix on generated formulas:

* with lcve_max = L the walk is the default walk up to the first candidate reached
  unfrozen with more than L live occurrences of a literal, so the elected variables
  are a prefix of the default ones, none of them over L, and the prefixes are nested
  as L grows;
* with lcve_clause_max = C no elected variable has a live clause longer than C on
  either list, no two elected variables share a live clause, and a C no clause
  exceeds changes nothing.
"""
import numpy as np
import pytest

import pyoracle as po
from parafrost_amd import gen


def _cnf(nvars, clauses):
    lits = np.array([l for c in clauses for l in c], np.uint32)
    offs = np.zeros(len(clauses) + 1, np.uint64)
    offs[1:] = np.cumsum([len(c) for c in clauses])
    return gen.Cnf(nvars, lits, offs)


def _elect(cnf, **kw):
    o = po.Oracle(cnf, po.default_opts(**kw))
    L = o.L
    L.pfo_resize_cnf(o.h)
    L.pfo_create_ot(o.h)
    assert L.pfo_prop(o.h)
    L.pfo_lcve(o.h)
    return o, [int(v) for v in o.state().pvs]


def P(v):
    return 2 * v


def N(v):
    return 2 * v + 1


# Scores (ps * ns): x1 1, x2 1, x3 2, x4 1, x5 1, so the walk visits x1, x2, x4, x5, x3.
# Every variable is cold (temperature 32).
HAND = [[P(1), P(2)], [N(1), P(3), P(4)], [N(2), N(3)], [N(4), P(5)], [N(5), N(3)]]


def test_hand_built_clause_limit():
    cnf = _cnf(5, HAND)
    # default: x1 elected (freezes x2, x3, x4), x5 elected
    assert _elect(cnf)[1] == [1, 5]
    # limit 2: x1's negative list holds the ternary: x1 freezes x2 through its positive
    # list and is not elected; x4's positive list holds it: x4 is passed over; x5 elected
    assert _elect(cnf, lcve_clause_max=2)[1] == [5]
    # limit 3: nothing is longer
    assert _elect(cnf, lcve_clause_max=3)[1] == [1, 5]


def _live_hist(st, nvars):
    h = np.zeros(2 * nvars + 2, np.int64)
    for hdr, _, lits in st.clauses():
        if not hdr & 2:
            for l in lits:
                h[l] += 1
    return h


def _formulas():
    yield gen.uniform_ksat(3000, 12600, 3, 41)
    yield gen.geometric(3000, 11000, 2, 8, 42)
    yield gen.scalefree(4000, 14000, 2.5, 0.2, 43)


@pytest.mark.parametrize("k", range(3))
def test_lcve_max_prefix(k):
    cnf = list(_formulas())[k]
    o, full = _elect(cnf)
    hist = _live_hist(o.state(), cnf.nvars)
    prev = None
    for L in (2, 3, 4, 6, 8, 12, 1 << 20):
        o, pv = _elect(cnf, lcve_max=L)
        assert full[: len(pv)] == pv
        if prev is None:
            assert len(pv) < len(full)   # the limit cuts the walk short
        else:
            assert prev == pv[: len(prev)]
        for v in pv:
            assert hist[2 * v] <= L and hist[2 * v + 1] <= L
        prev = pv
    assert prev == full


@pytest.mark.parametrize("k", range(3))
def test_clause_limit_properties(k):
    cnf = list(_formulas())[k]
    longest = int(np.diff(cnf.offs).max())
    full = _elect(cnf)[1]
    assert _elect(cnf, lcve_clause_max=longest)[1] == full
    seen = []
    for C in (2, 3, 4, 6):
        o, pv = _elect(cnf, lcve_clause_max=C)
        seen.append(pv)
        st = o.state()
        elected = set(pv)
        for hdr, _, lits in st.clauses():
            if hdr & 2:
                continue
            el = [l >> 1 for l in lits if (l >> 1) in elected]
            assert len(el) <= 1   # independent
            if el:
                assert len(lits) <= C
    assert any(seen) and any(pv != full for pv in seen)
