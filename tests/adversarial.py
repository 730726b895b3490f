"""Formulas with what the reference's loops treat specially, for the parity tests."""
import numpy as np

from parafrost_amd import gen


def adversarial(seed):
    """Random 3-SAT plus what the reference's loops treat specially: duplicate clauses,
    clauses past SUB's size break (1000 literals) and past the LDS literal cache,
    a hub variable in a large share of the clauses, pure literals, implication chains
    and the largest variable index."""
    rng = np.random.default_rng(8800 + seed)
    nv = int(rng.integers(1300, 2500))
    base = gen.uniform_ksat(nv, int(nv * rng.uniform(2.0, 3.5)), 3, 8800 + seed)
    cls = [list(map(int, c)) for c in base.clauses()]
    cls += [list(c) for c in (cls[i] for i in rng.integers(0, len(cls), len(cls) // 20))]   # duplicates
    for size in (150, 700, 1001, int(rng.integers(1100, 1300))):                         # long clauses
        vs = rng.choice(np.arange(1, nv + 1), size=size, replace=False)
        cls.append(sorted(int(2 * v + rng.integers(0, 2)) for v in vs))
    hub = int(rng.integers(1, nv + 1))
    for c in cls[: len(cls) // 3]:                                                        # a hub variable
        if all((l >> 1) != hub for l in c):
            c.append(2 * hub + int(rng.integers(0, 2)))
    for v in rng.choice(np.arange(1, nv + 1), size=8, replace=False):                      # pure literals
        for _ in range(3):
            o = int(rng.integers(1, nv + 1))
            if o != v:
                cls.append(sorted([2 * int(v), 2 * o + int(rng.integers(0, 2))]))
    chain = rng.choice(np.arange(1, nv + 1), size=30, replace=False)                        # implications
    cls += [sorted([2 * int(a) + 1, 2 * int(b)]) for a, b in zip(chain[:-1], chain[1:])]
    cls.append(sorted([2 * nv, 2 * int(rng.integers(1, nv)) + 1, 2 * int(rng.integers(1, nv))]))
    clean = []
    for c in cls:   # one literal per variable, no tautologies
        seen, out = {}, []
        for l in sorted(c):
            if (l >> 1) in seen:
                continue
            seen[l >> 1] = l
            out.append(l)
        clean.append(out)
    lits = np.array([l for c in clean for l in c], np.uint32)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clean])]).astype(np.uint64)
    return gen.Cnf(nv, lits, offs)
