"""Formulas that hold blocked clauses which survive VE, for the BCE tests.

A gadget on a fresh variable x and a base variable a: three clauses
(x | -a | r | s) and four (-x | u | w) make VE on x add 12 resolvents for 9
clauses (so it fails), while two clauses (-x | a | t) are blocked on x: their
resolvent with every positive clause holds a and -a. The other variables are
drawn from the base variables with the most occurrences, so that x (with few
occurrences) is elected before them.
"""
import numpy as np

from parafrost_amd import gen


def with_blocked_gadgets(cnf: "gen.Cnf", ngadgets: int, seed: int) -> "gen.Cnf":
    rng = np.random.default_rng(seed)
    nv = cnf.nvars
    clauses = [list(cnf.lits[int(cnf.offs[i]):int(cnf.offs[i + 1])]) for i in range(len(cnf.offs) - 1)]

    def lit(v, neg):
        return 2 * int(v) + int(neg)

    occ = np.bincount(cnf.lits >> 1, minlength=cnf.nvars + 1)[1:]
    busy = np.argsort(-occ, kind="stable")[:max(8, cnf.nvars // 3)] + 1

    def pick(k, exclude=()):
        out = []
        while len(out) < k:
            v = int(busy[rng.integers(0, len(busy))])
            if v not in exclude and v not in out:
                out.append(v)
        return out

    for _ in range(ngadgets):
        nv += 1
        x = nv
        a = pick(1)[0]
        for _ in range(3):
            r, s = pick(2, (a,))
            clauses.append([lit(x, 0), lit(a, 1), lit(r, rng.integers(0, 2)), lit(s, rng.integers(0, 2))])
        for _ in range(4):
            u, w = pick(2)
            clauses.append([lit(x, 1), lit(u, rng.integers(0, 2)), lit(w, rng.integers(0, 2))])
        for _ in range(2):
            t = pick(1, (a,))[0]
            clauses.append([lit(x, 1), lit(a, 0), lit(t, rng.integers(0, 2))])
    # sorted literals; a clause never repeats a variable
    lits, offs = [], [0]
    for c in clauses:
        seen, out = set(), []
        for l in sorted(c):
            if (l >> 1) in seen:
                continue
            seen.add(l >> 1)
            out.append(l)
        lits.extend(out)
        offs.append(len(lits))
    return gen.Cnf(nv, np.array(lits, np.uint32), np.array(offs, np.uint64))
