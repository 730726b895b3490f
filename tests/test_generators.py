"""Seeded workload generators: shapes, determinism, distribution."""
import hashlib

import numpy as np

from parafrost_amd import gen


def test_c1_shape_and_determinism():
    a, b = gen.config("c1"), gen.config("c1")
    assert a.nvars == 20_000 and a.ncls == 85_200 and a.nlits == 255_600
    assert np.array_equal(a.lits, b.lits)
    for c in list(a.clauses())[:2000]:
        assert len(set(int(l) >> 1 for l in c)) == 3


def test_geometric_widths():
    c = gen.geometric(200_000, 400_000, 2, 8, seed=2)
    w = np.diff(c.offs)
    assert w.min() == 2 and w.max() == 8
    frac2 = (w == 2).mean()
    assert 0.49 < frac2 < 0.51          # P(w=2) = 1/2
    assert abs(w.mean() - 2.98) < 0.02  # 2 + sum_{k>=1} 2^-k truncated at 8


def test_scalefree_gates_and_power_law():
    c = gen.scalefree(200_000, 800_000, 2.5, 0.2, seed=3)
    assert c.ncls == 800_000
    w = np.diff(c.offs)
    assert w.min() == 2 and w.max() <= 12
    occ = np.bincount(c.lits >> 1, minlength=c.nvars + 1)[1:]
    top = np.sort(occ)[::-1]
    assert top[0] > 50 * np.median(occ[occ > 0])   # heavy tail
    for cl in list(c.clauses())[:5000]:
        assert len(set(int(l) >> 1 for l in cl)) == len(cl)


def test_seed_changes_instance():
    a = gen.uniform_ksat(1000, 4000, 3, seed=1)
    b = gen.uniform_ksat(1000, 4000, 3, seed=2)
    assert not np.array_equal(a.lits, b.lits)
