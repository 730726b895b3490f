"""The oracle against golden vectors of the reference CPU simplifier.

tests/golden/reference_vectors.npz holds, for 48 seeded instances, the state of
the reference simplifier (ParaFROST src/cpu, built from its sources) after every
pass of rounds 0 and 1 and at the end (tests/golden/make_golden.py). The oracle
must reproduce every one of them word for word: clause images in clause order,
elected variables, elected variables' occurrence lists, unit trail, model stack
and variable states.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from oracle_steps import oracle_passes
from parafrost_amd.gen import Cnf

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


def _load():
    z = np.load(GOLD, allow_pickle=False)
    return z, [str(n) for n in z["names"]]


Z, NAMES = _load()


def _case(name):
    meta = Z[f"{name}/meta"]
    cnf = Cnf(int(meta[0]), Z[f"{name}/lits"], Z[f"{name}/offs"])
    opts = po.default_opts(phases=int(meta[1]), mu_pos=int(meta[2]), mu_neg=int(meta[3]), shrink_rate=int(meta[4]))
    return cnf, opts, int(meta[5])


def _lists(name, r, npvs):
    key = f"{name}/r{r}_sort_lists"
    if key not in Z.files:
        return None
    a = Z[key]
    k = 2 * npvs
    lens, data, out, pos = a[:k], a[k:], [], 0
    for n in lens:
        out.append(data[pos:pos + int(n)])
        pos += int(n)
    return out


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_per_pass(name):
    cnf, opts, unsat = _case(name)
    o = po.Oracle(cnf, opts)
    checked = 0
    final = None
    for rnd, p, st in oracle_passes(o):
        if p == "final":
            final = st
            continue
        if rnd > 1:
            continue
        key = f"{name}/r{rnd}_{p}_cnf"
        if key not in Z.files:
            continue
        assert np.array_equal(st.cnf, Z[key]), f"{name} round {rnd} {p}: clause database"
        assert np.array_equal(st.pvs, Z[f"{name}/r{rnd}_{p}_pvs"]), f"{name} round {rnd} {p}: elected variables"
        assert np.array_equal(st.trail, Z[f"{name}/r{rnd}_{p}_trail"]), f"{name} round {rnd} {p}: trail"
        if p == "sort":
            ref_lists = _lists(name, rnd, len(st.pvs))
            assert ref_lists is not None and len(ref_lists) == len(st.lists)
            for a, b in zip(st.lists, ref_lists):
                assert np.array_equal(a, b), f"{name} round {rnd}: occurrence list order"
        checked += 1
    if unsat:
        assert o.L.pfo_unsat(o.h)
        return
    assert checked > 0
    assert final is not None
    for f in ("cnf", "trail", "resolved", "vstate"):
        assert np.array_equal(getattr(final, f), Z[f"{name}/final_{f}"]), f"{name}: final {f}"
