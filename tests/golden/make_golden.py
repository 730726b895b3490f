"""Write golden vectors of the reference CPU simplifier (ParaFROST src/cpu).

Runs the reference simplifier, compiled from the reference's own sources by
oracle/Makefile (target `ref`), on small seeded instances and stores, per case,
the input clauses and the reference's final state (clause images in clause
order, unit trail, model stack, variable states) plus per-pass dumps of the
first rounds, and the variable-mapped write-back the reference makes after the
simplifier (ParaFROST::map(true)). Cases f* run in incremental mode with variables
frozen by ifreeze() and, for some, units given as unit clauses. Cases b* run with
blocked clause elimination on, over formulas with planted blocked clauses
(tests/blocked.py). Cases without units also hold the reference's binary DRAT proof
lines of the simplifier. Cases g* and b* that the first call leaves satisfiable also
hold a second simplifier call on the mapped formula (second_*: the reference's second
sigmify() after map(true), whose model records and proof lines are written in original
numbering through the mapped vorg). Output: tests/golden/reference_vectors.npz (data only).

usage: python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pyoracle as po  # noqa: E402
from parafrost_amd import gen  # noqa: E402
from blocked import with_blocked_gadgets  # noqa: E402

PASSES = ("lcve", "sort", "sub", "ve", "ere")


def cases():
    out = []
    for seed in range(48):
        rng = np.random.default_rng(1000 + seed)
        nv = int(rng.integers(30, 250))
        kind = seed % 3
        if kind == 0:
            cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.0, 4.3)), 3, 1000 + seed)
        elif kind == 1:
            cnf = gen.geometric(nv, int(nv * rng.uniform(1.5, 4.0)), 2, 8, 1000 + seed)
        else:
            nv = max(nv, 64)
            cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 4.0)), 2.5, 0.3, 1000 + seed)
        opts = dict(phases=int(rng.integers(1, 6)), mu_pos=int(rng.integers(10, 60)),
                    mu_neg=int(rng.integers(10, 60)), shrink_rate=int(rng.integers(1, 4)))
        out.append((f"g{seed}", cnf, opts, None, None))
    # incremental mode: variables frozen by ifreeze(), half of the cases with units
    # given as unit clauses (the reference parser enqueues them; the oracle gets them
    # through pfo_enqueue)
    for seed in range(16):
        rng = np.random.default_rng(3000 + seed)
        nv = int(rng.integers(30, 250))
        kind = seed % 3
        if kind == 0:
            cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.0, 4.3)), 3, 3000 + seed)
        elif kind == 1:
            cnf = gen.geometric(nv, int(nv * rng.uniform(1.5, 4.0)), 2, 8, 3000 + seed)
        else:
            nv = max(nv, 64)
            cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 4.0)), 2.5, 0.3, 3000 + seed)
        opts = dict(phases=int(rng.integers(1, 6)), mu_pos=int(rng.integers(10, 60)),
                    mu_neg=int(rng.integers(10, 60)), shrink_rate=int(rng.integers(1, 4)))
        freeze = np.unique(rng.integers(1, nv + 1, int(nv * rng.uniform(0.05, 0.4)))).astype(np.uint32)
        units = None
        if seed % 2:
            vs = rng.choice(np.arange(1, nv + 1), 3, replace=False)
            units = (2 * vs + rng.integers(0, 2, 3)).astype(np.uint32)
        out.append((f"f{seed}", cnf, opts, freeze, units))
    # blocked clause elimination on: formulas with planted blocked clauses that VE
    # leaves in place
    for seed in range(16):
        rng = np.random.default_rng(5000 + seed)
        nv = int(rng.integers(60, 400))
        kind = seed % 3
        if kind == 0:
            cnf = gen.uniform_ksat(nv, int(nv * rng.uniform(2.0, 4.0)), 3, 5000 + seed)
        elif kind == 1:
            cnf = gen.geometric(nv, int(nv * rng.uniform(1.5, 3.0)), 2, 6, 5000 + seed)
        else:
            cnf = gen.scalefree(nv, int(nv * rng.uniform(2.0, 3.5)), 2.5, 0.3, 5000 + seed)
        cnf = with_blocked_gadgets(cnf, int(rng.integers(2, 12)), 5000 + seed)
        opts = dict(phases=int(rng.integers(1, 6)), mu_pos=int(rng.integers(10, 60)),
                    mu_neg=int(rng.integers(10, 60)), shrink_rate=int(rng.integers(1, 4)),
                    bce_en=1, bce_limit=int(rng.integers(100, 3001)))
        out.append((f"b{seed}", cnf, opts, None, None))
    return out


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True, capture_output=True)
    arrays = {}
    names = []
    with tempfile.TemporaryDirectory() as td:
        for name, cnf, opts, freeze, units in cases():
            path = os.path.join(td, name + ".cnf")
            if units is not None:
                lits = np.concatenate([cnf.lits, units])
                offs = np.concatenate([cnf.offs, cnf.offs[-1] + 1 + np.arange(len(units), dtype=np.uint64)])
                gen.Cnf(cnf.nvars, lits, offs).write_dimacs(path)
                arrays[f"{name}/units"] = units
            else:
                cnf.write_dimacs(path)
            o = po.default_opts(**opts)
            st = po.run_reference(path, os.path.join(td, name + ".bin"), o, trace_prefix=os.path.join(td, name),
                                  freeze=freeze)
            if freeze is not None:
                arrays[f"{name}/freeze"] = freeze
            # the reference's variable mapping with write-back after the simplifier
            mp = po.run_reference(path, os.path.join(td, name + ".map.bin"), o, freeze=freeze, mapped=True)
            if isinstance(mp, po.Mapped):
                for f in ("lits", "offs", "vorg", "trail", "vstate"):
                    arrays[f"{name}/map_{f}"] = getattr(mp, f)
            if isinstance(mp, po.Mapped) and freeze is None:
                # a second sigmify() on the mapped formula, proof on (records and proof
                # lines through the mapped vorg)
                o2 = po.default_opts(**opts)
                o2.proof_en = 1
                r2 = po.run_reference(path, os.path.join(td, name + ".second.bin"), o2, second=True)
                if r2 is not None:
                    m2, s2, nfirst = r2
                    assert np.array_equal(m2.lits, mp.lits) and np.array_equal(m2.vorg, mp.vorg)
                    for f in ("cnf", "trail", "vstate", "proof"):
                        arrays[f"{name}/second_{f}"] = getattr(s2, f)
                    arrays[f"{name}/second_resolved"] = s2.resolved[nfirst:]
                    arrays[f"{name}/second_meta"] = np.array([o2.phases - 1 if o2.phases > 2 else o2.phases, s2.unsat],
                                                              np.int64)
            if units is None:
                # the reference's binary DRAT lines of the simplifier (-proof)
                op = po.default_opts(**opts)
                op.proof_en = 1
                pr = po.run_reference(path, os.path.join(td, name + ".proof.bin"), op, freeze=freeze)
                arrays[f"{name}/proof"] = pr.proof
            arrays[f"{name}/lits"] = cnf.lits
            arrays[f"{name}/offs"] = cnf.offs
            arrays[f"{name}/meta"] = np.array([cnf.nvars, opts["phases"], opts["mu_pos"], opts["mu_neg"],
                                               opts["shrink_rate"], st.unsat, o.bce_en, o.bce_limit], np.int64)
            for f in ("cnf", "trail", "resolved", "vstate"):
                arrays[f"{name}/final_{f}"] = getattr(st, f)
            # per-pass states of round 0 and 1
            for r in (0, 1):
                for p in PASSES:
                    tp = os.path.join(td, f"{name}.{r}.{p}.bin")
                    if not os.path.exists(tp):
                        continue
                    ts = po.read_dump(tp)
                    arrays[f"{name}/r{r}_{p}_cnf"] = ts.cnf
                    arrays[f"{name}/r{r}_{p}_pvs"] = ts.pvs
                    arrays[f"{name}/r{r}_{p}_trail"] = ts.trail
                    if p == "sort":
                        arrays[f"{name}/r{r}_{p}_lists"] = (np.concatenate([np.array([len(x)], np.int64) for x in ts.lists] +
                                                                           [x for x in ts.lists]) if ts.lists else np.zeros(0, np.int64))
            names.append(name)
    arrays["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **arrays)
    print("wrote", len(names), "cases")


if __name__ == "__main__":
    main()
