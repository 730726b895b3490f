"""The oracle against golden vectors of the reference CPU simplifier.

tests/golden/reference_vectors.npz holds, for 48 seeded instances, the state of
the reference simplifier (ParaFROST src/cpu, built from its sources) after every
pass of rounds 0 and 1 and at the end (tests/golden/make_golden.py). The oracle
must reproduce every one of them word for word: clause images in clause order,
elected variables, elected variables' occurrence lists, unit trail, model stack
and variable states. Cases f* run in incremental mode: variables frozen by
ifreeze() and, for some, units enqueued before the simplifier starts.
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
    opts = po.default_opts(phases=int(meta[1]), mu_pos=int(meta[2]), mu_neg=int(meta[3]), shrink_rate=int(meta[4]),
                           bce_en=int(meta[6]), bce_limit=int(meta[7]))
    return cnf, opts, int(meta[5])


def _live(words):
    """clause images without the deleted ones (status bit 1 of word 0)"""
    out, i = [], 0
    while i < len(words):
        n = 3 + int(words[i + 2])
        if not words[i] & 2:
            out.append(words[i:i + n])
        i += n
    return np.concatenate(out) if out else np.zeros(0, np.uint32)


def _incremental(name, o):
    """frozen variables (ifreeze) and units of an incremental-mode case"""
    if f"{name}/freeze" in Z.files:
        o.ifreeze(Z[f"{name}/freeze"])
    if f"{name}/units" in Z.files:
        assert o.enqueue(Z[f"{name}/units"])


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
    _incremental(name, o)
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
        if rnd == 0 and f"{name}/units" in Z.files:
            # the reference applies parsed units before extraction (shrinkTop), the oracle
            # in its first prop, which leaves the satisfied clauses as deleted entries
            # until the next resize: compare the live clauses and the units as a set
            assert np.array_equal(_live(st.cnf), _live(Z[key])), f"{name} round 0 {p}: live clauses"
            assert np.array_equal(st.pvs, Z[f"{name}/r0_{p}_pvs"]), f"{name} round 0 {p}: elected variables"
            assert np.array_equal(np.sort(st.trail), np.sort(Z[f"{name}/r0_{p}_trail"])), f"{name} round 0 {p}: units"
            checked += 1
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
        a, b = getattr(final, f), Z[f"{name}/final_{f}"]
        if f == "trail" and f"{name}/units" in Z.files:
            a, b = np.sort(a), np.sort(b)
        assert np.array_equal(a, b), f"{name}: final {f}"


def test_frozen_variables_change_the_result_and_survive():
    """The frozen cases are not vacuous: freezing changes the outcome of some of them,
    and no frozen variable is ever eliminated (vstate MELTED = 1)."""
    changed = 0
    for name in [n for n in NAMES if f"{n}/freeze" in Z.files]:
        cnf, opts, unsat = _case(name)
        if unsat:
            continue
        fz = Z[f"{name}/freeze"]
        assert not np.any(Z[f"{name}/final_vstate"][fz] == 1), name
        o = po.Oracle(cnf, opts)
        if f"{name}/units" in Z.files:
            o.enqueue(Z[f"{name}/units"])
        o.sigmify()
        changed += not np.array_equal(o.state().cnf, Z[f"{name}/final_cnf"])
    assert changed >= 4


@pytest.mark.parametrize("name", [n for n in NAMES if f"{n}/map_lits" in Z.files])
def test_oracle_vmap_matches_reference(name):
    """The variable-mapped write-back after the simplifier (ParaFROST::map(true),
    src/cpu/vmap.cpp:64-131): mapped clauses in clause order, vorg, mapped trail
    and variable states."""
    cnf, opts, unsat = _case(name)
    o = po.Oracle(cnf, opts)
    _incremental(name, o)
    assert o.sigmify() == 0
    m = o.vmap()
    for f in ("lits", "offs", "vorg", "trail", "vstate"):
        assert np.array_equal(getattr(m, f), Z[f"{name}/map_{f}"]), f"{name}: mapped {f}"
    assert m.nvars + 1 == len(Z[f"{name}/map_vorg"])


@pytest.mark.parametrize("name", [n for n in NAMES if f"{n}/proof" in Z.files])
def test_oracle_proof_matches_reference(name):
    """The binary DRAT lines of the simplifier, byte for byte (the reference run with
    -proof; its parser's lines are not part of them)."""
    cnf, opts, unsat = _case(name)
    opts.proof_en = 1
    o = po.Oracle(cnf, opts)
    if f"{name}/freeze" in Z.files:
        o.ifreeze(Z[f"{name}/freeze"])
    o.sigmify()
    assert np.array_equal(o.proof(), Z[f"{name}/proof"])


def test_oracle_proofs_pass_rup_check():
    """Independent of the reference: every line the oracle adds is implied by unit
    propagation over the formula and the lines before it (resolvents, strengthened
    and substituted clauses and units all are), and a refutation ends with the
    empty clause."""
    import drat
    checked = refuted = 0
    for name in [n for n in NAMES if f"{n}/proof" in Z.files][:40]:
        cnf, opts, unsat = _case(name)
        opts.proof_en = 1
        o = po.Oracle(cnf, opts)
        if f"{name}/freeze" in Z.files:
            o.ifreeze(Z[f"{name}/freeze"])
        rc = o.sigmify()
        lines = drat.parse(o.proof())
        checked += drat.check_rup([list(map(int, c)) for c in cnf.clauses()], lines)
        if rc != 0:
            refuted += 1
            assert lines[-1] == ("a", ())
    assert checked > 1000 and refuted >= 1


def second_call_case(name):
    """(mapped formula, options, vorg, units) of a case's second simplifier call: the
    reference's second sigmify() on the formula map(true) left after the first call"""
    cnf, opts, unsat = _case(name)
    meta = Z[f"{name}/second_meta"]
    m = Cnf(len(Z[f"{name}/map_vorg"]) - 1, Z[f"{name}/map_lits"], Z[f"{name}/map_offs"])
    opts.phases = int(meta[0])
    opts.proof_en = 1
    return m, opts, Z[f"{name}/map_vorg"], Z[f"{name}/map_trail"], int(meta[1])


SECOND = [n for n in NAMES if f"{n}/second_meta" in Z.files]


@pytest.mark.parametrize("name", SECOND)
def test_oracle_second_call_with_vorg_matches_reference(name):
    """A second simplifier call after the variable mapping: the reference renumbers its
    variables (map(true)) and writes model records and proof lines through vorg in the
    original numbering (model.cpp:263-280, proof.cpp:118-135). The oracle, given the
    mapped formula, its vorg and the kept unit, reproduces the second call's state,
    its model records (the stack's suffix) and its proof lines."""
    m, opts, vorg, units, unsat = second_call_case(name)
    o = po.Oracle(m, opts)
    o.set_vorg(vorg)
    assert o.enqueue(units)
    rc = o.sigmify()
    assert (rc != 0) == bool(unsat)
    st = o.state()
    assert np.array_equal(st.cnf, Z[f"{name}/second_cnf"])
    assert np.array_equal(np.sort(st.trail), np.sort(Z[f"{name}/second_trail"]))
    assert np.array_equal(st.vstate, Z[f"{name}/second_vstate"])
    assert np.array_equal(st.resolved, Z[f"{name}/second_resolved"])
    assert np.array_equal(o.proof(), Z[f"{name}/second_proof"])


def test_second_call_vorg_is_not_trivial():
    """The second-call cases exercise a renumbering: vorg differs from the identity,
    records and proof lines of some cases hold renumbered literals, and without the
    mapping the oracle's records differ from the reference's."""
    moved = differs = 0
    for name in SECOND:
        m, opts, vorg, units, unsat = second_call_case(name)
        moved += int((vorg[1:] != np.arange(1, len(vorg))).any())
        if not len(Z[f"{name}/second_resolved"]):
            continue
        o = po.Oracle(m, opts)
        assert o.enqueue(units)
        o.sigmify()
        differs += not np.array_equal(o.state().resolved, Z[f"{name}/second_resolved"])
    assert moved >= len(SECOND) // 2 and differs >= 10
