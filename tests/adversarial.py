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


def ere_shared_targets(copies: int = 4):
    """Formulas where eager redundancy elimination meets a target more than once
    (run with phases=0, so that the first round is the ERE round). Per copy, with
    fresh variables:
      x: (x a b), (~x c)      -> resolvent (a b c)
      y: (y a c), (~y b)      -> resolvent (a b c)
      T = (a b c) twice (a duplicate)
      z: (z d), (z d e), (~z e) -> resolvent (d e) from both pairs
      D = (d e) twice
    and clauses with ~a, ~b, ~c, ~d, ~e over filler variables, so that no literal is
    pure and x, y, z come first in the elimination order (smallest scores). In the
    reference's sequential loop x deletes the first T and y meets it deleted and stops;
    z's first pair deletes the first D and its second pair meets it deleted and stops:
    both duplicates stay."""
    cls = []
    nv = 0

    def fresh(k):
        nonlocal nv
        out = list(range(nv + 1, nv + 1 + k))
        nv += k
        return out

    P = lambda v: 2 * v
    N = lambda v: 2 * v + 1
    for _ in range(copies):
        x, y, z, a, b, c, d, e = fresh(8)
        cls += [[P(x), P(a), P(b)], [N(x), P(c)], [P(y), P(a), P(c)], [N(y), P(b)],
                [P(a), P(b), P(c)], [P(a), P(b), P(c)],
                [P(z), P(d)], [P(z), P(d), P(e)], [N(z), P(e)], [P(d), P(e)], [P(d), P(e)]]
        for v in (a, b, c, d, e):
            f, g = fresh(2)
            cls += [[N(v), P(f), P(g)], [N(f), N(g), P(v)], [P(f), N(g)], [N(f), P(g)]]
    clean = [sorted(c) for c in cls]
    lits = np.array([l for c in clean for l in c], np.uint32)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clean])]).astype(np.uint64)
    return gen.Cnf(nv, lits, offs)


def sub_long_unit_capable(nlong: int = 3500, seed: int = 0):
    """One variable x with more than nlong clauses on each side (longer than the
    subsumption wave's LDS arrays) and a complementary binary pair (x a), (~x a), so
    that strengthening makes a unit: run with sub_limit above nlong and every other
    variable frozen (ifreeze), so that LCVE elects x. Its long lists hold duplicates and
    subsumed clauses for the wave to find."""
    rng = np.random.default_rng(9900 + seed)
    nv = 4 * nlong + 10
    x, a = 1, 2
    cls = [[2 * x, 2 * a], [2 * x + 1, 2 * a]]
    for s in (0, 1):
        for i in range(nlong):
            o = rng.choice(np.arange(3, nv + 1), int(rng.integers(2, 4)), replace=False)   # no other binaries
            cls.append([2 * x + s] + [int(2 * v + rng.integers(0, 2)) for v in o])
            if i % 50 == 0:   # a duplicate and a superset of the clause before
                cls.append(list(cls[-1]))
                cls.append(cls[-1] + [2 * int(rng.integers(3, nv + 1)) + 1])
    clean = []
    for c in cls:
        seen, out = set(), []
        for l in sorted(c):
            if (l >> 1) not in seen:
                seen.add(l >> 1)
                out.append(l)
        clean.append(out)
    lits = np.array([l for c in clean for l in c], np.uint32)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clean])]).astype(np.uint64)
    return gen.Cnf(nv, lits, offs), np.arange(2, nv + 1, dtype=np.uint32)


def sort_ties(seed: int = 0):
    """Elected variables whose occurrence lists are long and hold clauses with equal
    CNF_CMP_KEY (same size, first and last literal, signature): duplicates. Each
    variable x (every other variable frozen by ifreeze, so that LCVE elects the x's)
    has on its positive side n clauses (x a b) over fresh a, b, some of them twice, and
    one negative clause. The lists of 21 to 256 entries with equal keys are sorted by
    the exact introsort, those without by stable ranks; one list of 300 entries is past
    the wave sort's limit. Returns the formula and the variables to freeze."""
    rng = np.random.default_rng(9950 + seed)
    cls = []
    nv = 0

    def fresh(k):
        nonlocal nv
        out = list(range(nv + 1, nv + 1 + k))
        nv += k
        return out

    xs = []
    for n, dups in ((24, 3), (40, 0), (64, 10), (120, 1), (300, 5), (16, 4)):
        x, c = fresh(2)
        xs.append(x)
        side = []
        for _ in range(n - dups):
            a, b = fresh(2)
            side.append([2 * x, 2 * a + int(rng.integers(0, 2)), 2 * b + int(rng.integers(0, 2))])
        for i in rng.choice(len(side), dups, replace=False):
            side.append(list(side[i]))
        order = rng.permutation(len(side))
        cls += [side[i] for i in order]
        cls.append([2 * x + 1, 2 * c])
    clean = [sorted(c) for c in cls]
    lits = np.array([l for c in clean for l in c], np.uint32)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in clean])]).astype(np.uint64)
    frozen = np.array([v for v in range(1, nv + 1) if v not in set(xs)], np.uint32)
    return gen.Cnf(nv, lits, offs), frozen


def function_gates(seed: int, nv_free: int = 8):
    """Small formulas for function-table gates (ve_fun_en), at most 20 variables for brute
    force, with the variables to freeze so that only the gadget variables x are elected.

    Definition gadget: x is a random Boolean function of 3 free variables, one clause per
    input assignment (inputs equal to the assignment imply x, or ~x), plus clauses of x
    longer than the definition's. The function table collapses, but the minimization
    (a clause leaves the core when the clauses before it already contradict) only melts
    positive clauses here, so the resolvents equal plain resolution's.

    Contradiction gadget: ~x joined to every clause of the pigeonhole formula of 3
    pigeons in 2 holes (6 variables, minimally unsatisfiable, no clause pair that
    self-subsumes, so subsumption leaves it as it is), then 2 longer negative clauses of x
    after it and one positive clause. The negative prefix before the extra clauses is
    contradictory, so they leave the core; the positive clause leaves it too (the negative
    side alone is false), and the 2 resolvents between them, never tautologies, are the
    ones a gate skips and plain resolution adds."""
    rng = np.random.default_rng(9600 + seed)
    cls = []
    nv = nv_free
    # definition gadget
    ins = rng.choice(np.arange(1, nv_free + 1), size=3, replace=False)
    nv += 1
    x = nv
    table = rng.integers(0, 2, 8)
    for a in range(8):
        # inputs equal to a  ->  x == table[a]:  (some input differs) | (x or ~x)
        c = [2 * int(v) + ((a >> i) & 1) for i, v in enumerate(ins)]
        c.append(2 * x + (0 if table[a] else 1))
        cls.append(sorted(c))
    half = nv_free // 2
    for sign in (0, 1):
        for _ in range(int(rng.integers(1, 3))):
            vs = np.arange(1, half + 1) + sign * half
            cls.append(sorted([2 * x + sign] + [int(2 * v + rng.integers(0, 2)) for v in vs]))
    # contradiction gadget: pigeon i in hole j is variable p0 + 2 i + j
    p0 = nv + 1
    nv += 6
    nv += 1
    x = nv
    php = [[2 * (p0 + 2 * i), 2 * (p0 + 2 * i + 1)] for i in range(3)]
    for j in range(2):
        for i in range(3):
            for k in range(i + 1, 3):
                php.append([2 * (p0 + 2 * i + j) + 1, 2 * (p0 + 2 * k + j) + 1])
    for c in php:
        cls.append(sorted(c + [2 * x + 1]))
    a, b = rng.choice(np.arange(1, half + 1), size=2, replace=False)
    cls.append(sorted([2 * x, 2 * int(a), 2 * int(b)]))
    for _ in range(2):
        vs = rng.choice(np.arange(half + 1, nv_free + 1), size=3, replace=False)
        cls.append(sorted([2 * x + 1] + [int(2 * v + rng.integers(0, 2)) for v in vs]))
    for _ in range(int(rng.integers(nv_free, 2 * nv_free))):
        vs = rng.choice(np.arange(1, nv_free + 1), size=3, replace=False)
        cls.append(sorted(int(2 * v + rng.integers(0, 2)) for v in vs))
    lits = np.array([l for c in cls for l in c], np.uint32)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in cls])]).astype(np.uint64)
    freeze = np.array([v for v in range(1, nv + 1) if v not in (nv_free + 1, nv)], np.uint32)
    return gen.Cnf(nv, lits, offs), freeze


def function_gates_tiled(copies: int, seed: int):
    """copies of function_gates(seed + i) over disjoint variables, with their frozen
    variables: many pending function-table gates in one elimination pass"""
    lits, offs, freeze, nv = [], [np.zeros(1, np.uint64)], [], 0
    for i in range(copies):
        cnf, fr = function_gates(seed * 100003 + i)
        lits.append(cnf.lits + 2 * nv)
        offs.append(cnf.offs[1:] + offs[-1][-1])
        freeze.append(fr + nv)
        nv += cnf.nvars
    return (gen.Cnf(nv, np.concatenate(lits).astype(np.uint32), np.concatenate(offs)),
            np.concatenate(freeze).astype(np.uint32))


def equivalences(seed: int, frac: float = 0.15):
    """Random 3-SAT with planted equivalences x = a (the binaries (-x|a) and (x|-a)):
    BVE substitutes a for x in x's other clauses, rewriting them in place, mostly with
    their size kept."""
    rng = np.random.default_rng(9900 + seed)
    nv = int(rng.integers(3000, 6000))
    base = gen.uniform_ksat(nv, int(nv * 3.0), 3, 9900 + seed)
    cls = [sorted(map(int, c)) for c in base.clauses()]
    xs = rng.choice(np.arange(1, nv + 1), size=int(nv * frac), replace=False)
    for x in xs:
        a = int(rng.integers(1, nv + 1))
        if a == x:
            continue
        s = int(rng.integers(0, 2))
        cls.append(sorted([2 * int(x) + 1, 2 * a + s]))
        cls.append(sorted([2 * int(x), 2 * a + (1 - s)]))
    lits = np.array([l for c in cls for l in c], np.uint32)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in cls])]).astype(np.uint64)
    return gen.Cnf(nv, lits, offs)


def xor_gates(seed: int):
    """Random 3-SAT with planted XOR constraints x + y1 + .. + yk = b (k = 2..4, every
    clause of the parity encoding: 2^k clauses of size k + 1), so that BVE's XOR gate
    search succeeds and substitutes; some gate variables appear nowhere else."""
    rng = np.random.default_rng(9950 + seed)
    nv = int(rng.integers(1500, 3000))
    base = gen.uniform_ksat(nv, int(nv * rng.uniform(1.5, 3.0)), 3, 9950 + seed)
    cls = [sorted(map(int, c)) for c in base.clauses()]
    for _ in range(int(nv * 0.08)):
        k = int(rng.integers(2, 5))
        vs = rng.choice(np.arange(1, nv + 1), size=k + 1, replace=False)
        b = int(rng.integers(0, 2))
        for a in range(1 << (k + 1)):
            if bin(a).count("1") % 2 == b:
                continue   # the clause forbidding assignment a (bit i: value of vs[i])
            cls.append(sorted(2 * int(v) + ((a >> i) & 1) for i, v in enumerate(vs)))
    lits = np.array([l for c in cls for l in c], np.uint32)
    offs = np.concatenate([[0], np.cumsum([len(c) for c in cls])]).astype(np.uint64)
    return gen.Cnf(nv, lits, offs)
