"""Seeded synthetic CNF workloads (ctypes over libpfgen.so, built from csrc/gen.c).

A formula is returned as a ``Cnf``: ``lits`` (uint32, literal = 2*var + sign) and
``offs`` (uint64 CSR offsets, length ncls + 1). Variables are 1..nvars.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libpfgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
        lib.pfgen_uniform_ksat.restype = ctypes.c_int64
        lib.pfgen_uniform_ksat.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                           ctypes.c_uint64, u32p, u64p]
        lib.pfgen_geometric.restype = ctypes.c_int64
        lib.pfgen_geometric.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_uint64, u32p, u64p]
        lib.pfgen_scalefree.restype = ctypes.c_int64
        lib.pfgen_scalefree.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, u32p, u64p]
        lib.pfgen_write_dimacs.restype = ctypes.c_int
        lib.pfgen_write_dimacs.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint64, u32p, u64p]
        _lib = lib
    return _lib


@dataclass
class Cnf:
    nvars: int
    lits: np.ndarray  # uint32
    offs: np.ndarray  # uint64, ncls + 1

    @property
    def ncls(self) -> int:
        return len(self.offs) - 1

    @property
    def nlits(self) -> int:
        return int(self.offs[-1])

    def write_dimacs(self, path: str) -> None:
        rc = _load().pfgen_write_dimacs(path.encode(), self.nvars, self.ncls, self.lits, self.offs)
        if rc != 0:
            raise OSError(f"cannot write {path}")

    def clauses(self):
        for i in range(self.ncls):
            yield self.lits[self.offs[i]:self.offs[i + 1]]


def _finish(nvars, lits, offs, n):
    if n < 0:
        raise ValueError("generator rejected its arguments")
    return Cnf(nvars, lits[:n].copy(), offs)


def uniform_ksat(nvars: int, ncls: int, k: int = 3, seed: int = 1) -> Cnf:
    lits = np.empty(ncls * k, np.uint32)
    offs = np.empty(ncls + 1, np.uint64)
    n = _load().pfgen_uniform_ksat(nvars, ncls, k, seed, lits, offs)
    return _finish(nvars, lits, offs, n)


def geometric(nvars: int, ncls: int, wmin: int = 2, wmax: int = 8, seed: int = 2) -> Cnf:
    lits = np.empty(ncls * wmax, np.uint32)
    offs = np.empty(ncls + 1, np.uint64)
    n = _load().pfgen_geometric(nvars, ncls, wmin, wmax, seed, lits, offs)
    return _finish(nvars, lits, offs, n)


def scalefree(nvars: int, ncls: int, beta: float = 2.5, gate_frac: float = 0.2, seed: int = 3) -> Cnf:
    lits = np.empty(ncls * 12, np.uint32)
    offs = np.empty(ncls + 1, np.uint64)
    n = _load().pfgen_scalefree(nvars, ncls, beta, gate_frac, seed, lits, offs)
    return _finish(nvars, lits, offs, n)


# The workloads the simplifier is measured and checked on.
def config(name: str) -> Cnf:
    if name == "c1":   # uniform random 3-SAT, 20,000 vars, 85,200 clauses (ratio 4.26), seed 1
        return uniform_ksat(20_000, 85_200, 3, seed=1)
    if name == "c2":   # 2M vars, 8.4M clauses, widths 2..8 geometric(1/2), seed 2
        return geometric(2_000_000, 8_400_000, 2, 8, seed=2)
    if name == "c3":   # scale-free 5M vars, 20M clauses, beta 2.5, 20% Tseitin AND/OR gates, seed 3
        return scalefree(5_000_000, 20_000_000, 2.5, 0.2, seed=3)
    raise KeyError(name)
