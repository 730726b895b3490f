"""The oracle's restatement of libstdc++ std::sort (used for occurrence lists longer
than 20, src/cpu/sort.h:78-90) must order ties exactly as std::sort does.

Pinned against the real std::sort of this host's libstdc++ (built at test time
with g++) on tie-heavy inputs, including inputs that drive introsort into its
heap-sort fallback."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as po

SRC = r'''
#include <algorithm>
#include <cstdint>
extern "C" void stdsort_idx(uint32_t* idx, int64_t n, const uint32_t* key) {
    std::sort(idx, idx + n, [key](uint32_t a, uint32_t b) { return key[a] < key[b]; });
}
'''


@pytest.fixture(scope="module")
def stdsort():
    d = tempfile.mkdtemp()
    src, so = os.path.join(d, "s.cpp"), os.path.join(d, "s.so")
    open(src, "w").write(SRC)
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-o", so, src], check=True)
    L = ctypes.CDLL(so)
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    L.stdsort_idx.argtypes = [u32p, ctypes.c_int64, u32p]
    return L


def _both(L, keys):
    n = len(keys)
    keys = np.ascontiguousarray(keys, np.uint32)
    a = np.arange(n, dtype=np.uint32)
    b = a.copy()
    L.stdsort_idx(a, n, keys)
    po.lib().pfo_introsort_u32(b, n, keys)
    return a, b


def median_of_3_killer(n):
    # Musser's construction against median-of-3 quicksort (drives introsort to heap sort)
    k = n // 2
    a = np.zeros(n, np.uint32)
    for i in range(1, k + 1):
        if i % 2 == 1:
            a[i - 1] = i
            a[i] = k + i
        a[k + i - 1] = 2 * i
    return a


@pytest.mark.parametrize("seed", range(30))
def test_random_ties(stdsort, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(17, 4000))
    keys = rng.integers(0, max(2, n // int(rng.integers(1, 50))), n).astype(np.uint32)
    a, b = _both(stdsort, keys)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [17, 21, 64, 100, 1000, 4096])
def test_patterns(stdsort, n):
    for keys in (np.zeros(n), np.arange(n), np.arange(n)[::-1], np.arange(n) % 7,
                 np.concatenate([np.arange(n // 2), np.arange(n - n // 2)[::-1]]), median_of_3_killer(n)):
        a, b = _both(stdsort, keys)
        assert np.array_equal(a, b)
