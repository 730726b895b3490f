"""GPU simplifier parity with the oracle (bit-exact), through the C ABI.

Lock-step cases compare the complete state after every pass of every round
(tests/lockstep.py); end-to-end cases compare the final state of a whole
sigmify. Larger workloads add size-independent properties.
"""
import functools

import numpy as np
import pytest

import lockstep as ls
import pyoracle as po
from parafrost_amd import gen

pytestmark = pytest.mark.gpu


def _small(seed):
    rng = np.random.default_rng(7000 + seed)
    nv = int(rng.integers(20, 400))
    kind = seed % 3
    if kind == 0:
        cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.0, 4.5)), 3, 7000 + seed)
    elif kind == 1:
        cnf = gen.geometric(nv, int(nv * rng.uniform(1.5, 4.0)), 2, 8, 7000 + seed)
    else:
        nv = max(nv, 64)
        cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 4.0)), 2.5, 0.3, 7000 + seed)
    kw = dict(phases=int(rng.integers(1, 7)), mu_pos=int(rng.integers(10, 60)), mu_neg=int(rng.integers(10, 60)),
              shrink_rate=int(rng.integers(1, 4)))
    return cnf, kw


@pytest.mark.parametrize("seed", range(48))
def test_lockstep_small(seed):
    print(f"lockstep small seed={seed}", flush=True)
    cnf, kw = _small(seed)
    ls.run_lockstep(cnf, **kw)


OPTION_VARIANTS = [
    dict(stable_sort=1), dict(ere_en=0), dict(ve_lbound_en=1), dict(xor_max_arity=2), dict(xor_max_arity=20),
    dict(sub_limit=20, ere_limit=20), dict(ve_clause_limit=0), dict(ve_clause_limit=3), dict(ere_max_resolvent=0),
    dict(lits_min=1 << 40), dict(phases=8, mu_pos=10, mu_neg=10), dict(sub_en=0), dict(lcve_min=50),
    dict(bce_en=1), dict(bce_en=1, bce_limit=100, mu_pos=200, mu_neg=200), dict(bce_en=1, ve_en=0),
    dict(ve_fun_en=1), dict(ve_fun_en=1, ve_lbound_en=1, ve_clause_limit=3),
]


@pytest.mark.parametrize("variant", range(len(OPTION_VARIANTS)))
def test_lockstep_options(variant):
    kw = OPTION_VARIANTS[variant]
    print(f"lockstep options {kw}", flush=True)
    for seed in (1, 2, 5):
        cnf, base = _small(seed)
        base.update(kw)
        ls.run_lockstep(cnf, **base)


@pytest.mark.parametrize("seed", range(10))
def test_lockstep_bce(seed):
    """Blocked clause elimination on formulas with planted blocked clauses that VE
    leaves in place; the state after BCE is compared in every round."""
    from blocked import with_blocked_gadgets
    print(f"lockstep bce seed={seed}", flush=True)
    cnf, kw = _small(seed)
    rng = np.random.default_rng(7500 + seed)
    cnf = with_blocked_gadgets(cnf, int(rng.integers(2, 20)), 7500 + seed)
    ls.run_lockstep(cnf, bce_en=1, bce_limit=int(rng.integers(100, 3001)), **kw)


@pytest.mark.parametrize("seed", range(3))
def test_bce_mid_size_matches_oracle(seed):
    from blocked import with_blocked_gadgets
    print(f"bce mid size seed={seed}", flush=True)
    rng = np.random.default_rng(7600 + seed)
    nv = int(rng.integers(20000, 80000))
    if seed % 2:
        cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.5, 4.0)), 3, 7600 + seed)
    else:
        cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 3.5)), 2.5, 0.3, 7600 + seed)
    cnf = with_blocked_gadgets(cnf, nv // 20, 7600 + seed)
    ls.run_full(cnf, bce_en=1)


@pytest.mark.parametrize("seed", range(16))
def test_lockstep_proof(seed):
    """Binary DRAT lines after every pass equal the oracle's (units of one batched
    propagation as a set); covers refuted instances, BCE and the ERE round."""
    from blocked import with_blocked_gadgets
    print(f"lockstep proof seed={seed}", flush=True)
    cnf, kw = _small(seed)
    if seed % 4 == 3:
        rng = np.random.default_rng(7700 + seed)
        nv = max(64, int(rng.integers(64, 400)))
        cnf = gen.scalefree(nv, int(nv * rng.uniform(4.0, 8.0)), 2.5, 0.1, 7700 + seed)   # often refuted
    if seed % 4 == 1:
        cnf = with_blocked_gadgets(cnf, 8, 7700 + seed)
        kw["bce_en"] = 1
    ls.run_lockstep(cnf, proof_en=1, **kw)


@pytest.mark.parametrize("seed", range(4))
def test_proof_mid_size_matches_oracle(seed):
    print(f"proof mid size seed={seed}", flush=True)
    rng = np.random.default_rng(7800 + seed)
    nv = int(rng.integers(20000, 80000))
    if seed % 2:
        cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.5, 4.4)), 3, 7800 + seed)
    else:
        cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 5.0)), 2.5, 0.2, 7800 + seed)
    ls.run_full(cnf, proof_en=1, phases=int(rng.integers(2, 6)))


def test_proof_off_writes_nothing():
    import parafrost_amd as pf
    cnf, kw = _small(3)
    g = pf.Simplifier(cnf.nvars, pf.default_opts(**kw))
    try:
        g.upload(cnf)
        try:
            g.sigmify()
        except pf.Unsat:
            pass
        assert g.proof().size == 0
    finally:
        g.close()


@pytest.mark.parametrize("seed", range(6))
def test_lockstep_long_lists(seed):
    """Mid-size scale-free instances with high temperatures: elected variables with
    occurrence lists longer than 20 (std::sort order) and gate structure."""
    print(f"lockstep long lists seed={seed}", flush=True)
    rng = np.random.default_rng(9000 + seed)
    nv = int(rng.integers(3000, 12000))
    cnf = gen.scalefree(nv, int(nv * rng.uniform(3.0, 5.0)), 2.5, 0.2, 9000 + seed)
    ls.run_lockstep(cnf, lists=True, phases=5, mu_pos=int(rng.integers(64, 400)), mu_neg=int(rng.integers(64, 400)))


@pytest.mark.parametrize("words", [0, 40, 400])
@pytest.mark.parametrize("seed", [0, 1, 2, 4, 5, 7])
def test_lockstep_stage_limits(seed, words):
    """SUB and BVE detection with the LDS stage capped: 0 reads every variable from
    the pool, small caps stage a block's variables over many rounds and leave the
    larger ones unstaged. Results must not change."""
    print(f"lockstep stage limit {words} seed={seed}", flush=True)
    cnf, kw = _small(seed)
    ls.run_lockstep(cnf, setup=lambda g: g.set_stage_words(words), **kw)


@pytest.mark.parametrize("seed", range(2))
def test_lockstep_long_lists_stage_limit(seed):
    """Long lists with a stage cap below their size: the unstaged fallback on
    variables with std::sort-ordered lists and gates."""
    print(f"lockstep long lists stage limit seed={seed}", flush=True)
    rng = np.random.default_rng(9000 + seed)
    nv = int(rng.integers(3000, 12000))
    cnf = gen.scalefree(nv, int(nv * rng.uniform(3.0, 5.0)), 2.5, 0.2, 9000 + seed)
    ls.run_lockstep(cnf, lists=True, setup=lambda g: g.set_stage_words(300), phases=5,
                    mu_pos=int(rng.integers(64, 400)), mu_neg=int(rng.integers(64, 400)))


@pytest.mark.parametrize("seed", range(6))
def test_lockstep_learnt_clauses(seed):
    """Learnt clauses (status LEARNT with an LBD): ERE types, subsumption promotion,
    strengthening LBD bump (subsume.h:101-117, elimination.cpp:27-39)."""
    print(f"lockstep learnt seed={seed}", flush=True)
    cnf, kw = _small(seed + 100)
    rng = np.random.default_rng(seed)
    hdr = np.where(rng.random(cnf.ncls) < 0.3, 1 | (rng.integers(1, 30, cnf.ncls) << 6) | (1 << 4), 0).astype(np.uint32)
    oo, go = ls._opts_pair(**kw)
    o = po.Oracle(cnf, oo, hdr0=hdr)
    ro = o.sigmify()
    import parafrost_amd as pf
    g = pf.Simplifier(cnf.nvars, go)
    g.upload(cnf, hdr0=hdr)
    try:
        try:
            g.sigmify()
            rg = 0
        except pf.Unsat:
            rg = -1
        assert (ro != 0) == (rg != 0), f"unsat oracle {ro} gpu {rg}"
        if ro == 0:
            ls.compare(o.state(), ls.state_gpu(g), "learnt sigmify")
    finally:
        g.close()


@pytest.mark.parametrize("seed", range(8))
def test_lockstep_root_units(seed):
    """Units on the trail before simplification: batched device propagation."""
    print(f"lockstep units seed={seed}", flush=True)
    cnf, kw = _small(seed + 200)
    rng = np.random.default_rng(seed)
    vs = rng.choice(np.arange(1, cnf.nvars + 1), size=max(1, cnf.nvars // 25), replace=False)
    units = (2 * vs + rng.integers(0, 2, len(vs))).astype(np.uint32)
    # the proof too: the caller's units are not logged, the ones propagation finds are
    ls.run_lockstep(cnf, units=units, proof_en=1, **kw)


@pytest.mark.parametrize("seed", range(8))
def test_lockstep_frozen(seed):
    """Incremental mode: variables frozen by ifreeze are passed over by LCVE (lcve.cpp:57);
    odd seeds also put units on the trail first."""
    print(f"lockstep frozen seed={seed}", flush=True)
    cnf, kw = _small(seed + 300)
    rng = np.random.default_rng(77 + seed)
    freeze = np.unique(rng.integers(1, cnf.nvars + 1, max(1, int(cnf.nvars * rng.uniform(0.05, 0.4))))).astype(np.uint32)
    units = None
    if seed % 2:
        vs = rng.choice(np.arange(1, cnf.nvars + 1), size=max(1, cnf.nvars // 40), replace=False)
        units = (2 * vs + rng.integers(0, 2, len(vs))).astype(np.uint32)
    ls.run_lockstep(cnf, units=units, freeze=freeze, proof_en=seed % 2, **kw)


def test_frozen_marks_persist_and_clear():
    """Marks set before an upload hold for the sigmify after it, melting undoes them,
    and pfg_reset clears them."""
    import parafrost_amd as pf
    cnf, kw = _small(311)
    frz = np.arange(1, cnf.nvars + 1, 2, dtype=np.uint32)
    g = pf.Simplifier(cnf.nvars, pf.default_opts(**kw))
    try:
        def run(fz=None, melt=False):
            g.reset_for_upload()
            if fz is not None:
                g.ifreeze(fz)
                if melt:
                    g.ifreeze(fz, on=False)
            g.upload(cnf)
            try:
                g.sigmify()
            except pf.Unsat:
                return None
            return g.cnf_words(), g.vstate()
        base = run()
        frozen = run(frz)
        melted = run(frz, melt=True)
        g.reset_for_upload()
        g.ifreeze(frz)
        cleared = run()
        o = po.Oracle(cnf, po.default_opts(**kw))
        o.ifreeze(frz)
        ro = o.sigmify()
        assert (ro != 0) == (frozen is None)
        if frozen is not None:
            assert np.array_equal(frozen[0], o.state().cnf)
            assert not np.any(frozen[1][frz] == 1)
        if base is not None:
            assert np.array_equal(melted[0], base[0]) and np.array_equal(cleared[0], base[0])
            assert frozen is None or not np.array_equal(frozen[0], base[0])
    finally:
        g.close()


def _csr(words):
    lits, offs, i = [], [0], 0
    while i < len(words):
        n = int(words[i + 2])
        cl = words[i + 3:i + 3 + n]
        assert int(words[i + 1]) == int(np.bitwise_or.reduce(np.left_shift(1, cl & 31).astype(np.uint32))) if n else True
        lits.append(cl)
        offs.append(offs[-1] + n)
        i += 3 + n
    return (np.concatenate(lits) if lits else np.zeros(0, np.uint32)), np.array(offs, np.uint64)


@pytest.mark.parametrize("seed", range(8))
def test_vmap_matches_oracle(seed):
    """Variable mapping with write-back after sigmify (ParaFROST::map(true)) on the
    device: renumbered clauses (and their signatures), vorg, mapped units and variable
    states equal the oracle's; the context itself is left as it was."""
    import parafrost_amd as pf
    print(f"vmap seed={seed}", flush=True)
    cnf, kw = _small(seed + 400)
    rng = np.random.default_rng(90 + seed)
    units = np.zeros(0, np.uint32)
    if seed % 2:
        vs = rng.choice(np.arange(1, cnf.nvars + 1), size=max(1, cnf.nvars // 60), replace=False)
        units = (2 * vs + rng.integers(0, 2, len(vs))).astype(np.uint32)
    o = po.Oracle(cnf, po.default_opts(**kw))
    if len(units):
        assert o.enqueue(units)
    ro = o.sigmify()
    g = pf.Simplifier(cnf.nvars, pf.default_opts(**kw))
    try:
        g.upload(cnf)
        if len(units):
            g.enqueue_units(units)
        try:
            g.sigmify()
        except pf.Unsat:
            assert ro != 0
            return
        assert ro == 0
        before = g.cnf_words()
        w, vorg, tr, vs = g.mapped()
        m = o.vmap()
        lits, offs = _csr(w)
        assert np.array_equal(lits, m.lits) and np.array_equal(offs, m.offs)
        assert np.array_equal(vorg, m.vorg) and np.array_equal(vs, m.vstate)
        assert np.array_equal(np.sort(tr), np.sort(m.trail))
        assert np.array_equal(g.cnf_words(), before)
    finally:
        g.close()


def test_vmap_c1():
    import parafrost_amd as pf
    cnf = gen.config("c1")
    o = po.Oracle(cnf, po.default_opts())
    assert o.sigmify() == 0
    g = pf.Simplifier(cnf.nvars, pf.default_opts())
    try:
        g.upload(cnf)
        g.sigmify()
        w, vorg, tr, vs = g.mapped()
        m = o.vmap()
        lits, offs = _csr(w)
        assert np.array_equal(lits, m.lits) and np.array_equal(offs, m.offs) and np.array_equal(vorg, m.vorg)
        assert m.nvars < cnf.nvars
    finally:
        g.close()


def test_unsat_by_propagation():
    import parafrost_amd as pf
    lits = np.array([2 * 1 + 1, 2 * 2, 2 * 2 + 1, 2 * 3, 2 * 3 + 1, 2 * 1], np.uint32)  # (~1|2) (~2|3) (~3|1)... closed by units
    cnf = gen.Cnf(3, lits, np.array([0, 2, 4, 6], np.uint64))
    g = pf.Simplifier(3)
    g.upload(cnf)
    g.enqueue_units(np.array([2 * 1, 2 * 3 + 1], np.uint32))   # 1, ~3  -> 2 forced, then ~3 conflicts
    with pytest.raises(pf.Unsat):
        g.sigmify()
    g.close()


@pytest.mark.parametrize("seed", range(8))
def test_lockstep_adversarial(seed):
    print(f"lockstep adversarial seed={seed}", flush=True)
    from adversarial import adversarial
    cnf = adversarial(seed)
    rng = np.random.default_rng(seed)
    ls.run_lockstep(cnf, phases=int(rng.integers(1, 6)), mu_pos=int(rng.integers(10, 80)),
                    mu_neg=int(rng.integers(10, 80)), bce_en=seed % 2, proof_en=(seed // 2) % 2)


@pytest.mark.parametrize("seed", range(6))
def test_lockstep_xor_gates(seed):
    """Planted XOR constraints: gate search successes, failures and their cleanup."""
    print(f"lockstep xor gates seed={seed}", flush=True)
    from adversarial import xor_gates
    rng = np.random.default_rng(seed)
    ls.run_lockstep(xor_gates(seed), phases=int(rng.integers(2, 6)), xor_max_arity=int(rng.integers(2, 6)))


def test_empty_and_tiny():
    import parafrost_amd as pf
    g = pf.Simplifier(5)
    g.upload(gen.Cnf(5, np.zeros(0, np.uint32), np.zeros(1, np.uint64)))
    g.sigmify()
    assert len(g.cnf_words()) == 0
    g.close()
    cnf = gen.Cnf(2, np.array([2, 4], np.uint32), np.array([0, 2], np.uint64))
    ls.run_full(cnf)


def test_c1_lockstep():
    print("c1 lockstep", flush=True)
    ls.run_lockstep(gen.config("c1"))


def test_c1_reupload_deterministic():
    import parafrost_amd as pf
    cnf = gen.config("c1")
    g = pf.Simplifier(cnf.nvars)
    outs = []
    for _ in range(3):
        g.reset_for_upload()
        g.upload(cnf)
        g.sigmify()
        outs.append((g.cnf_words(), g.resolved(), g.vstate()))
    g.close()
    for a in outs[1:]:
        for x, y in zip(outs[0], a):
            assert np.array_equal(x, y)


@pytest.mark.slow
def test_c2_full_matches_oracle():
    """The benchmark workload end to end (2M vars, 8.4M clauses)."""
    print("c2 full", flush=True)
    ls.run_full(gen.config("c2"))


@pytest.mark.slow
def test_c2_full_bce_and_proof_match_oracle():
    """The benchmark workload with both optional passes on: BCE (nothing is blocked
    after VE on c2) and the DRAT proof (280,056 lines), state and proof bit for bit."""
    print("c2 full, BCE and proof", flush=True)
    ls.run_full(gen.config("c2"), bce_en=1, proof_en=1)


@functools.lru_cache(maxsize=1)
def _c3():
    return gen.config("c3")   # about 40 s to generate; shared by the c3 tests


@pytest.mark.slow
def test_c3_lockstep_and_invariants_of_the_gpu_state():
    """Scale-free 5M vars / 20M clauses, pass by pass against the oracle up to the
    refutation (both prove the formula unsatisfiable during a later propagation), then
    size-independent checks on the GPU's own last state before it: eliminated
    variables occur in no live clause, literals sorted and unique per clause,
    signatures consistent, model stack well formed."""
    print("c3 lockstep", flush=True)
    cnf = _c3()
    keep = {}
    rounds, final = ls.run_lockstep(cnf, lists=False, keep=keep)
    st = keep["gpu"]
    print(f"c3: {rounds} rounds compared, last state after {keep['where']}", flush=True)
    assert rounds >= 1
    w = st.cnf
    elim = st.vstate == 1
    assert elim.sum() > 0
    i, n, k = 0, len(w), 0
    while i < n and k < 300_000:
        sz = int(w[i + 2])
        cl = w[i + 3:i + 3 + sz]
        if not w[i] & 2:
            assert not elim[cl >> 1].any()
        assert np.all(np.diff(cl.astype(np.int64)) > 0)
        assert int(w[i + 1]) == int(np.bitwise_or.reduce(np.left_shift(1, cl & 31).astype(np.uint32))) if sz else True
        i += 3 + sz
        k += 1
    r = st.resolved   # model stack: records end with their size, walk backwards
    j = len(r) - 1
    while j > 0:
        k = int(r[j])
        assert k >= 1
        j -= k + 1
    assert j == -1


@pytest.mark.slow
def test_c3_proof_up_to_the_refutation():
    """c3 with the DRAT proof on: the proof lines up to the empty clause equal the
    oracle's (runs of unit lines as sets; the unit run right before the empty clause,
    whose length depends on the propagation order, left out)."""
    import drat
    print("c3 proof", flush=True)
    ls.run_full(_c3(), proof_en=1, proof_cmp=drat.fast_normalized)


def test_hbm_budget_refuses_cleanly():
    """A context whose HBM budget is too small for the formula reports PFG_ENOMEM
    (the reference's CNFALLOC_FAIL: the solver goes on without simplifying). The
    budget counts what the context holds, including buffers grown before the refusal:
    the same context then either simplifies a smaller formula exactly or refuses it
    the same way, and a fresh context with that budget simplifies it exactly."""
    import parafrost_amd as pf
    big = gen.geometric(200_000, 840_000, 2, 8, seed=2)   # needs well over 100 MB
    s5, _ = _small(5)
    small = gen.Cnf(big.nvars, s5.lits, s5.offs)           # a few hundred of the variables used
    budget = 48 << 20
    o = po.Oracle(small, po.default_opts())
    ro = o.sigmify()

    def run_small(g):
        g.upload(small)
        try:
            g.sigmify()
            rg = 0
        except pf.Unsat:
            rg = -1
        assert (ro != 0) == (rg != 0)
        if ro == 0:
            ls.compare(o.state(), ls.state_gpu(g), "small formula")

    g = pf.Simplifier(big.nvars, pf.default_opts(), hbm_budget=budget)
    try:
        with pytest.raises(pf.SimplifierError) as e:
            g.upload(big)
            g.sigmify()
        assert e.value.code == pf.PFG_ENOMEM and "budget" in str(e.value)
        g.reset_for_upload()
        try:
            run_small(g)
        except pf.SimplifierError as e2:
            assert e2.code == pf.PFG_ENOMEM, str(e2)
    finally:
        g.close()
    g = pf.Simplifier(big.nvars, pf.default_opts(), hbm_budget=budget)
    try:
        run_small(g)
    finally:
        g.close()
