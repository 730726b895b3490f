"""GPU tests of the less common paths and of the boundary additions, each against
the oracle through the C ABI:

* ERE claiming rounds: targets that two elected variables (or two resolvent pairs of
  one variable) meet, where the sequential loop's second visitor goes on to the next
  equal clause (tests/adversarial.py: ere_shared_targets);
* the unit-capable subsumption wave with lists longer than its LDS arrays
  (sub_limit above 3000);
* sortOT's exact introsort for elected lists past 20 entries holding equal keys, and
  past the wave sort's 256 entries, beside the stable ranks of the other lists;
* the caller's variable mapping (vorg): a second simplifier call on the formula the
  reference's map(true) left, with model records and proof lines in original numbering;
* pfg_interrupt: a call stopped before its first round, and one stopped while it runs;
* function-table gates (ve_fun_en): variables defined as functions of 3 others and
  variables whose negative clauses hold a contradictory prefix, where the gate leaves
  out the resolvents between clauses outside the core (tests/adversarial.py:
  function_gates), alone, tiled by the hundred, and on general formulas.
"""
import threading
import time

import numpy as np
import pytest

import lockstep as ls
import pyoracle as po
import parafrost_amd as pf
from adversarial import adversarial, ere_shared_targets, function_gates, function_gates_tiled, sort_ties, sub_long_unit_capable
from parafrost_amd import gen
from test_oracle_golden import SECOND, second_call_case

pytestmark = pytest.mark.gpu


def test_ere_shared_targets_match_the_sequential_loop():
    """x and y both resolve to a clause T that has a duplicate: the reference deletes
    T (x) and then its duplicate (y, which meets T deleted); z's two resolvent pairs do
    the same with D. The GPU needs a second claiming round for y."""
    print("ere shared targets", flush=True)
    seen = {}
    for copies in (1, 6):
        cnf = ere_shared_targets(copies)
        ls.run_lockstep(cnf, phases=0, ve_en=0, sub_en=0, proof_en=1,
                        inspect=lambda g: seen.update(g.path_counts()))
        assert seen["ere_extra_rounds"] >= 1
        st = po.Oracle(cnf, po.default_opts(phases=0, ve_en=0, sub_en=0))
        st.sigmify()
        live = [c for c in st.state().clauses() if not c[0] & 2]
        assert len(live) == len(list(cnf.clauses())) - 4 * copies   # both copies of T and of D are gone


@pytest.mark.parametrize("seed", range(4))
def test_ere_rounds_on_duplicate_heavy_formulas(seed):
    """Formulas with many duplicate clauses, ERE on in the final round, proof on."""
    print(f"ere duplicates seed={seed}", flush=True)
    cnf = adversarial(40 + seed)
    ls.run_lockstep(cnf, phases=1 + seed % 2, proof_en=1)


def test_sub_wave_with_long_lists():
    """sub_limit 5000: the unit-capable variable's lists (3,600 entries each) exceed the
    wave's LDS arrays; its per-entry arrays live in HBM. The strengthened binary makes a
    unit, and the duplicates and supersets on the lists are subsumed."""
    print("sub wave long lists", flush=True)
    cnf, frozen = sub_long_unit_capable(3500)
    seen = {}
    ls.run_lockstep(cnf, sub_limit=5000, ere_limit=5000, mu_pos=100000, mu_neg=100000, phases=1, freeze=frozen,
                    proof_en=1, inspect=lambda g: seen.update(g.path_counts()))
    assert seen["sub_wave_big_vars"] >= 1


@pytest.mark.parametrize("seed", range(2))
def test_sort_ties_take_the_exact_introsort(seed):
    """Elected lists of 16 to 300 entries, duplicates among them: the lists past 20
    entries with equal keys (and the one past 256) are sorted by the exact introsort,
    the others by stable ranks; the lists after sortOT equal the oracle's."""
    print(f"sort ties seed={seed}", flush=True)
    cnf, frozen = sort_ties(seed)
    seen = {}
    ls.run_lockstep(cnf, phases=1, freeze=frozen, inspect=lambda g: seen.update(g.path_counts()))
    assert seen["sort_exact_lists"] >= 4


@pytest.mark.parametrize("name", SECOND)
def test_second_call_with_vorg_lockstep(name):
    """The reference's second sigmify() after map(true) (golden second_* cases): mapped
    formula, its vorg and kept unit; model records and proof lines in original numbering."""
    print(f"vorg second call {name}", flush=True)
    m, opts, vorg, units, unsat = second_call_case(name)
    kw = {f: getattr(opts, f) for f, _ in po.Opts._fields_}
    ls.run_lockstep(m, units=units, vorg=vorg, **kw)


@pytest.mark.parametrize("seed", range(3))
def test_vorg_permutation_mid_size(seed):
    """A random renumbering (vorg a permutation into a larger range) on mid-size formulas
    with BCE and the proof on: every record and proof line goes through it."""
    print(f"vorg permutation seed={seed}", flush=True)
    rng = np.random.default_rng(8100 + seed)
    nv = int(rng.integers(3000, 20000))
    cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.5, 4.0)), 3, 8100 + seed) if seed % 2 else \
        gen.geometric(nv, int(nv * rng.uniform(1.5, 3.0)), 2, 8, 8100 + seed)
    vorg = np.zeros(nv + 1, np.uint32)
    vorg[1:] = rng.choice(np.arange(1, 4 * nv, dtype=np.uint32), nv, replace=False)
    ls.run_lockstep(cnf, vorg=vorg, bce_en=1, proof_en=1, phases=int(rng.integers(1, 5)))


def test_interrupt_before_the_first_round():
    """An interrupt asked for before the call stops it at its first poll: the formula is
    the uploaded one, and the next call simplifies it exactly as an uninterrupted one."""
    print("interrupt before the first round", flush=True)
    cnf = gen.uniform_ksat(3000, 12000, 3, seed=31)
    o0 = po.Oracle(cnf, po.default_opts())
    g = pf.Simplifier(cnf.nvars, pf.default_opts())
    try:
        g.upload(cnf)
        g.interrupt()
        with pytest.raises(pf.Interrupted):
            g.sigmify()
        ls.compare(o0.state(), ls.state_gpu(g), "after the interrupt")
        g.sigmify()
        o0.sigmify()
        ls.compare(o0.state(), ls.state_gpu(g), "after the second call")
    finally:
        g.close()


def test_interrupt_while_running():
    """An interrupt from another thread during a call on a 2M-variable formula: the call
    ends early with EINTR (or had already finished), the formula left behind is well
    formed (sorted literals, right signatures, no eliminated variable in a live clause;
    checked on its first 300,000 clauses), and the next call runs to the end."""
    print("interrupt while running", flush=True)
    cnf = gen.geometric(2_000_000, 8_400_000, 2, 8, seed=2)
    g = pf.Simplifier(cnf.nvars, pf.default_opts())
    try:
        g.upload(cnf)
        out = {}

        def run():
            t0 = time.perf_counter()
            try:
                g.sigmify()
                out["rc"] = "ok"
            except pf.Interrupted:
                out["rc"] = "intr"
            out["s"] = time.perf_counter() - t0

        th = threading.Thread(target=run)
        th.start()
        time.sleep(0.004)
        g.interrupt()
        th.join(120)
        assert not th.is_alive()
        print(f"call ended: {out}", flush=True)
        assert out["rc"] in ("ok", "intr")
        w, vs = g.cnf_words(), g.vstate()
        i, n = 0, 0
        while i < len(w) and n < 300_000:
            k = int(w[i + 2])
            lits = w[i + 3:i + 3 + k]
            assert not w[i] & 2 and k >= 1 and np.all(np.diff(lits.astype(np.int64)) > 0)
            assert int(w[i + 1]) == int(np.bitwise_or.reduce(np.left_shift(np.uint32(1), lits & 31))) if k else True
            assert not np.any(vs[lits >> 1] == 1)
            i += 3 + k
            n += 1
        assert n == 300_000
        g.interrupt()   # consumed by the next call: it stops at once
        with pytest.raises(pf.Interrupted):
            g.sigmify()
        g.sigmify()
    finally:
        g.close()


def test_interrupt_batch():
    """A batch stopped before its first round: every running instance reports EINTR and
    the next call gives each instance its own result."""
    print("interrupt batch", flush=True)
    cnfs = [gen.uniform_ksat(500 + 100 * i, (500 + 100 * i) * 4, 3, seed=600 + i) for i in range(5)]
    b = pf.BatchSimplifier([c.nvars for c in cnfs], pf.default_opts())
    try:
        b.upload_instances(cnfs)
        b.interrupt()
        st = np.zeros(b.ninst, np.int32)
        rc = b.L.pfg_sigmify_batch(b.h, st.ctypes.data)
        assert rc == pf.PFG_EINTR and np.all(st == pf.PFG_EINTR)
        st = b.sigmify_batch()
        for i, c in enumerate(cnfs):
            o = po.Oracle(c, po.default_opts())
            ro = o.sigmify()
            w, t, r, vs = b.instance(i)
            assert (ro != 0) == (st[i] != 0)
            if ro == 0:
                s = o.state()
                assert np.array_equal(s.cnf, w) and np.array_equal(s.resolved, r) and np.array_equal(s.vstate, vs)
    finally:
        b.close()


@pytest.mark.parametrize("seed", range(6))
def test_function_gates_lockstep(seed):
    print(f"function gates seed={seed}", flush=True)
    cnf, freeze = function_gates(seed)
    ls.run_lockstep(cnf, freeze=freeze, ve_fun_en=1, phases=1 + seed % 3, mu_pos=64, mu_neg=64, proof_en=seed % 2)


@pytest.mark.parametrize("copies,seed", [(60, 0), (700, 1)])
def test_function_gates_tiled_lockstep(copies, seed):
    """Many pending variables per group of 64 elected ones, BCE and the proof on for one."""
    print(f"function gates tiled copies={copies} seed={seed}", flush=True)
    cnf, freeze = function_gates_tiled(copies, seed)
    ls.run_lockstep(cnf, freeze=freeze, ve_fun_en=1, phases=2, mu_pos=64, mu_neg=64, proof_en=seed, bce_en=seed)


@pytest.mark.parametrize("seed", range(4))
def test_function_gates_on_general_formulas(seed):
    """Formulas of the generators and the adversarial ones: most variables have more than
    12 neighbours or no collapsing table, and fall back to plain resolution."""
    print(f"function gates general seed={seed}", flush=True)
    cnf = adversarial(seed) if seed % 2 else gen.scalefree(3000, 9000, 2.5, 0.3, 70 + seed)
    ls.run_lockstep(cnf, ve_fun_en=1, phases=3, mu_pos=40, mu_neg=40, proof_en=1)


def test_many_timed_calls_leave_no_hip_error():
    """Thousands of timed launches in one context, with the sorts of full table builds
    every round (ot_build 2): the pending event pairs are resolved between timers,
    never while one is open, and no HIP error is left behind for the calling thread
    (torch launches in between and afterwards succeed)."""
    import torch
    cnf = gen.uniform_ksat(3000, 12600, 3, 77)
    g = pf.Simplifier(cnf.nvars, pf.default_opts())
    try:
        g.set_tuning(ot_build=2)
        x = None
        for k in range(60):
            g.upload(cnf)
            g.sigmify()
            x = torch.ones(4, device="cuda") * float(k)   # a torch launch on this thread
        torch.cuda.synchronize()
        assert float(x.sum()) == 4.0 * 59
        st = g.kernel_times()
        assert st["ot_sort"]["launches"] >= 60 and st["ot_fill"]["launches"] >= 60
    finally:
        g.close()
