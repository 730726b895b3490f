"""Oracle against the reference CPU simplifier run live (oracle/_ref/pfref, built
from the reference sources by `make -C oracle ref`). Skipped where that build is
absent; the committed golden vectors cover the same ground offline."""
import os
import tempfile

import numpy as np
import pytest

import pyoracle as po
from parafrost_amd import gen

pytestmark = pytest.mark.skipif(not os.path.exists(po.REF_BIN), reason="reference build absent")


def _run(cnf, opts, freeze=None):
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "f.cnf")
        cnf.write_dimacs(p)
        ref = po.run_reference(p, os.path.join(td, "f.bin"), opts, freeze=freeze)
    o = po.Oracle(cnf, opts)
    if freeze is not None:
        o.ifreeze(freeze)
    rc = o.sigmify()
    st = o.state()
    if ref.unsat:
        assert rc != 0
        return
    for f in ("cnf", "trail", "resolved", "vstate"):
        assert np.array_equal(getattr(st, f), getattr(ref, f)), f


def test_c1_full():
    _run(gen.config("c1"), po.default_opts())


def test_c2_full():
    """the 2M-variable mixed-width workload: the oracle's final state equals the reference's word for word
    (about 25 s for the oracle, 50 s for the reference)"""
    _run(gen.config("c2"), po.default_opts())


@pytest.mark.parametrize("seed", range(12))
def test_mid_size(seed):
    rng = np.random.default_rng(500 + seed)
    nv = int(rng.integers(2000, 20000))
    kind = seed % 3
    if kind == 0:
        cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(3.0, 4.3)), 3, 500 + seed)
    elif kind == 1:
        cnf = gen.geometric(nv, int(nv * rng.uniform(2.0, 4.0)), 2, 8, 500 + seed)
    else:
        cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 5.0)), 2.5, 0.2, 500 + seed)
    opts = po.default_opts(phases=int(rng.integers(2, 7)), mu_pos=int(rng.integers(10, 200)),
                           mu_neg=int(rng.integers(10, 200)), shrink_rate=int(rng.integers(1, 4)))
    _run(cnf, opts)


@pytest.mark.parametrize("seed", range(6))
def test_mid_size_frozen(seed):
    """incremental mode: a random set of variables frozen by ifreeze()"""
    rng = np.random.default_rng(700 + seed)
    nv = int(rng.integers(2000, 20000))
    cnf = (gen.uniform_ksat(nv, int(nv * rng.uniform(3.0, 4.3)), 3, 700 + seed) if seed % 2 else
           gen.scalefree(nv, int(nv * rng.uniform(2.0, 5.0)), 2.5, 0.2, 700 + seed))
    freeze = np.unique(rng.integers(1, nv + 1, int(nv * rng.uniform(0.05, 0.4)))).astype(np.uint32)
    _run(cnf, po.default_opts(phases=int(rng.integers(2, 7))), freeze)


@pytest.mark.parametrize("seed", range(6))
def test_mid_size_vmap(seed):
    """variable mapping with write-back after the simplifier, with level-0 units"""
    rng = np.random.default_rng(750 + seed)
    nv = int(rng.integers(2000, 15000))
    cnf = (gen.geometric(nv, int(nv * rng.uniform(2.0, 3.5)), 2, 8, 750 + seed) if seed % 2 else
           gen.uniform_ksat(nv, int(nv * rng.uniform(2.5, 4.0)), 3, 750 + seed))
    vs = rng.choice(np.arange(1, nv + 1), 4, replace=False)
    units = (2 * vs + rng.integers(0, 2, 4)).astype(np.uint32)
    opts = po.default_opts(phases=int(rng.integers(1, 6)))
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "f.cnf")
        lits = np.concatenate([cnf.lits, units])
        offs = np.concatenate([cnf.offs, cnf.offs[-1] + 1 + np.arange(len(units), dtype=np.uint64)])
        gen.Cnf(cnf.nvars, lits, offs).write_dimacs(p)
        ref = po.run_reference(p, os.path.join(td, "f.bin"), opts, mapped=True)
    o = po.Oracle(cnf, opts)
    assert o.enqueue(units)
    rc = o.sigmify()
    if not isinstance(ref, po.Mapped):
        assert rc != 0
        return
    m = o.vmap()
    for f in ("lits", "offs", "vorg", "trail", "vstate"):
        assert np.array_equal(getattr(m, f), getattr(ref, f)), f


@pytest.mark.parametrize("seed", range(6))
def test_mid_size_bce(seed):
    """Blocked clause elimination on, with planted blocked clauses."""
    from blocked import with_blocked_gadgets
    rng = np.random.default_rng(900 + seed)
    nv = int(rng.integers(2000, 12000))
    if seed % 2:
        cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.5, 4.0)), 3, 900 + seed)
    else:
        cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 3.5)), 2.5, 0.3, 900 + seed)
    cnf = with_blocked_gadgets(cnf, nv // 20, 900 + seed)
    _run(cnf, po.default_opts(phases=int(rng.integers(2, 6)), bce_en=1, bce_limit=int(rng.integers(100, 3001))))


@pytest.mark.parametrize("seed", range(8))
def test_mid_size_proof(seed):
    """Binary DRAT output (-proof), byte for byte; half with BCE on, some refuted."""
    rng = np.random.default_rng(950 + seed)
    nv = int(rng.integers(2000, 12000))
    kind = seed % 4
    if kind == 0:
        cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(3.0, 4.4)), 3, 950 + seed)
    elif kind == 1:
        cnf = gen.geometric(nv, int(nv * rng.uniform(1.5, 3.5)), 2, 8, 950 + seed)
    elif kind == 2:
        cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 4.0)), 2.5, 0.3, 950 + seed)
    else:
        cnf = gen.scalefree(nv, int(nv * rng.uniform(4.0, 8.0)), 2.5, 0.1, 950 + seed)
    opts = po.default_opts(phases=int(rng.integers(2, 6)), bce_en=seed % 2, proof_en=1)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "f.cnf")
        cnf.write_dimacs(p)
        ref = po.run_reference(p, os.path.join(td, "f.bin"), opts)
    o = po.Oracle(cnf, opts)
    o.sigmify()
    assert len(ref.proof) > 0
    assert np.array_equal(o.proof(), ref.proof)


def test_reference_timer_for_the_cpu_baseline():
    """PFREF_TIME makes the reference build report the simplifier's own wall time (the
    CPU baseline of bench.py); the state it dumps is unchanged."""
    import subprocess
    cnf = gen.config("c1")
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "f.cnf")
        cnf.write_dimacs(p)
        env = dict(os.environ, PFREF_TIME="1")
        r = subprocess.run([po.REF_BIN, p, os.path.join(td, "t.bin")], env=env, capture_output=True, text=True,
                           timeout=300)
        t = [float(x.split()[1]) for x in r.stderr.splitlines() if x.startswith("PFREF_TIME ")]
        assert r.returncode == 0 and len(t) == 1 and 0 < t[0] < 60
        ref = po.run_reference(p, os.path.join(td, "f.bin"), po.default_opts())
        timed = po.read_dump(os.path.join(td, "t.bin"))
        assert np.array_equal(ref.cnf, timed.cnf) and np.array_equal(ref.resolved, timed.resolved)


@pytest.mark.parametrize("seed", range(8))
def test_adversarial(seed):
    """Duplicate clauses, clauses past SUB's size break, a hub variable, pure literals,
    implication chains; half with BCE on, half with the proof compared too."""
    from adversarial import adversarial
    rng = np.random.default_rng(seed)
    opts = po.default_opts(phases=int(rng.integers(1, 6)), mu_pos=int(rng.integers(10, 80)),
                           mu_neg=int(rng.integers(10, 80)), bce_en=seed % 2, proof_en=(seed // 2) % 2)
    cnf = adversarial(seed)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "f.cnf")
        cnf.write_dimacs(p)
        ref = po.run_reference(p, os.path.join(td, "f.bin"), opts)
    o = po.Oracle(cnf, opts)
    rc = o.sigmify()
    st = o.state()
    assert (rc != 0) == bool(ref.unsat)
    if not ref.unsat:
        for f in ("cnf", "trail", "resolved", "vstate"):
            assert np.array_equal(getattr(st, f), getattr(ref, f)), f
    if opts.proof_en:
        assert np.array_equal(o.proof(), ref.proof)


@pytest.mark.parametrize("seed", range(6))
def test_xor_gates(seed):
    """Planted XOR constraints: the XOR gate search succeeds (substitution) and fails
    (cleanup of molten clauses) as in the reference."""
    from adversarial import xor_gates
    rng = np.random.default_rng(seed)
    opts = po.default_opts(phases=int(rng.integers(2, 6)), xor_max_arity=int(rng.integers(2, 6)))
    cnf = xor_gates(seed)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "f.cnf")
        cnf.write_dimacs(p)
        ref = po.run_reference(p, os.path.join(td, "f.bin"), opts)
    o = po.Oracle(cnf, opts)
    rc = o.sigmify()
    st = o.state()
    assert (rc != 0) == bool(ref.unsat)
    if not ref.unsat:
        for f in ("cnf", "trail", "resolved", "vstate"):
            assert np.array_equal(getattr(st, f), getattr(ref, f)), f


@pytest.mark.parametrize("seed", range(6))
def test_mid_size_second_call_vorg(seed):
    """A second simplifier call after map(true): the mapped formula with its vorg and
    kept unit; model records (the stack's suffix) and proof lines in original numbering."""
    rng = np.random.default_rng(780 + seed)
    nv = int(rng.integers(2000, 12000))
    cnf = (gen.geometric(nv, int(nv * rng.uniform(1.5, 3.5)), 2, 8, 780 + seed) if seed % 2 else
           gen.uniform_ksat(nv, int(nv * rng.uniform(2.5, 4.2)), 3, 780 + seed))
    opts = po.default_opts(phases=int(rng.integers(1, 6)), bce_en=seed % 3 == 2, proof_en=1)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "f.cnf")
        if seed % 2 == 0:   # level-0 units: the mapping keeps one assigned variable
            vs = rng.choice(np.arange(1, nv + 1), 3, replace=False)
            units = (2 * vs + rng.integers(0, 2, 3)).astype(np.uint32)
            lits = np.concatenate([cnf.lits, units])
            offs = np.concatenate([cnf.offs, cnf.offs[-1] + 1 + np.arange(3, dtype=np.uint64)])
            gen.Cnf(cnf.nvars, lits, offs).write_dimacs(p)
        else:
            cnf.write_dimacs(p)
        r = po.run_reference(p, os.path.join(td, "f.bin"), opts, second=True)
    if r is None:
        return
    m, ref, nfirst = r
    o2 = po.default_opts(phases=opts.phases - 1 if opts.phases > 2 else opts.phases, bce_en=opts.bce_en, proof_en=1)
    o = po.Oracle(gen.Cnf(m.nvars, m.lits, m.offs), o2)
    o.set_vorg(m.vorg)
    assert o.enqueue(m.trail)
    rc = o.sigmify()
    assert (rc != 0) == bool(ref.unsat)
    st = o.state()
    assert np.array_equal(st.cnf, ref.cnf)
    assert np.array_equal(np.sort(st.trail), np.sort(ref.trail))
    assert np.array_equal(st.vstate, ref.vstate)
    assert np.array_equal(st.resolved, ref.resolved[nfirst:])
    assert np.array_equal(o.proof(), ref.proof)
