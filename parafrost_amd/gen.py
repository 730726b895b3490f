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
        lib.pfgen_images_to_csr.restype = ctypes.c_int64
        lib.pfgen_images_to_csr.argtypes = [u32p, ctypes.c_uint64, u32p, u64p, ctypes.POINTER(ctypes.c_uint64)]
        lib.pfgen_aig_bmc.restype = ctypes.c_int64
        lib.pfgen_aig_bmc.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, u32p, u64p,
                                      ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.c_void_p]
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


# ---------------------------------------------------------------- batches
@dataclass
class Batch:
    """Independent instances in batch numbering: instance i's variable v is the batch
    variable base[i] + v (bases are multiples of 16, as the simplifier lays them out);
    its clauses are [cls_first[i], cls_first[i + 1])."""
    inst_nvars: np.ndarray  # uint32 per instance
    base: np.ndarray        # uint64, ninst + 1 (base[ninst] = batch variable count)
    cls_first: np.ndarray   # uint64, ninst + 1
    lits: np.ndarray        # uint32, batch numbering
    offs: np.ndarray        # uint64, ncls + 1
    kinds: np.ndarray       # clause width per instance

    @property
    def ninst(self) -> int:
        return len(self.inst_nvars)

    @property
    def nvars(self) -> int:
        return int(self.base[-1])

    @property
    def ncls(self) -> int:
        return len(self.offs) - 1

    @property
    def nlits(self) -> int:
        return int(self.offs[-1])

    def instance(self, i: int) -> Cnf:
        """Instance i in its own numbering."""
        a, b = int(self.cls_first[i]), int(self.cls_first[i + 1])
        la, lb = int(self.offs[a]), int(self.offs[b])
        return Cnf(int(self.inst_nvars[i]), self.lits[la:lb] - np.uint32(2 * int(self.base[i])),
                   self.offs[a:b + 1] - np.uint64(la))


def batch_bases(nvars) -> np.ndarray:
    base = np.zeros(len(nvars) + 1, np.uint64)
    at = 0
    for i, n in enumerate(nvars):
        base[i] = at
        at += int(n)
        if i + 1 < len(nvars):
            at = (at + 15) & ~15
    base[len(nvars)] = at
    return base


def batch_ksat(nvars, ncls, ks, seeds, threads: int = 16) -> Batch:
    """Uniform random k-SAT instances (k distinct variables per clause, fair signs),
    generated in parallel straight into batch numbering."""
    import concurrent.futures as cf
    nvars = np.asarray(nvars, np.uint32)
    ncls = np.asarray(ncls, np.uint64)
    ks = np.asarray(ks, np.int64)
    n = len(nvars)
    base = batch_bases(nvars)
    cls_first = np.zeros(n + 1, np.uint64)
    cls_first[1:] = np.cumsum(ncls)
    lit_first = np.zeros(n + 1, np.uint64)
    lit_first[1:] = np.cumsum(ncls * ks.astype(np.uint64))
    lits = np.empty(int(lit_first[-1]), np.uint32)
    offs = np.empty(int(cls_first[-1]) + 1, np.uint64)
    offs[0] = 0
    lib = _load()

    def one(i):
        la, lb = int(lit_first[i]), int(lit_first[i + 1])
        ca, cb = int(cls_first[i]), int(cls_first[i + 1])
        tmp = np.empty(cb - ca + 1, np.uint64)
        part = lits[la:lb]
        got = lib.pfgen_uniform_ksat(int(nvars[i]), cb - ca, int(ks[i]), int(seeds[i]), part, tmp)
        if got != lb - la:
            raise ValueError(f"instance {i}: generator returned {got} literals, expected {lb - la}")
        part += np.uint32(2 * int(base[i]))
        offs[ca + 1:cb + 1] = tmp[1:] + np.uint64(la)

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, range(n)))
    return Batch(nvars, base, cls_first, lits, offs, ks)


def c4_specs(ninst: int = 4096, vmin: int = 1000, vmax: int = 200_000, seed: int = 4):
    """The batching workload: variable counts log-uniform in [vmin, vmax]; each instance
    is random 3-SAT at clause ratio U[4.0, 4.3] or 4-SAT at U[9.5, 10.0], 50/50."""
    rng = np.random.default_rng(seed)
    nv = np.exp(rng.uniform(np.log(vmin), np.log(vmax), ninst)).round().astype(np.uint32)
    four = rng.random(ninst) < 0.5
    ratio = np.where(four, rng.uniform(9.5, 10.0, ninst), rng.uniform(4.0, 4.3, ninst))
    ncls = np.maximum(1, (ratio * nv).round()).astype(np.uint64)
    ks = np.where(four, 4, 3).astype(np.int64)
    seeds = rng.integers(1, 1 << 62, ninst, dtype=np.int64).astype(np.uint64)
    return nv, ncls, ks, seeds


def c4(ninst: int = 4096, vmax: int = 200_000, seed: int = 4, threads: int = 16) -> Batch:
    return batch_ksat(*c4_specs(ninst, 1000, vmax, seed), threads=threads)


# ---------------------------------------------------------------- BMC-like requests
def aig_bmc(seed: int, ngates: int, t0: int, t1: int, nlatches: int = 64, ninputs: int = 32, pinv: float = 0.3):
    """Frames [t0, t1) of a random AIG (see csrc/gen.c) Tseitin-encoded: returns
    (Cnf of those frames' clauses, variable count after t1 frames). The clauses of
    frames [0, t) followed by those of [t, t') are the clauses of frames [0, t')."""
    nf = max(0, t1 - t0)
    lits = np.empty(max(1, 8 * ngates * nf), np.uint32)
    offs = np.empty(3 * ngates * nf + 1, np.uint64)
    nc = ctypes.c_uint64()
    nv = ctypes.c_uint32()
    n = _load().pfgen_aig_bmc(seed, ngates, nlatches, ninputs, pinv, t0, t1, lits, offs, ctypes.byref(nc), ctypes.byref(nv),
                              None)
    if n < 0:
        raise ValueError("generator rejected its arguments")
    return Cnf(int(nv.value), lits[:n].copy(), offs[:nc.value + 1].copy()), int(nv.value)


def aig_bmc_interface(seed: int, ngates: int, t: int, nlatches: int = 64, ninputs: int = 32, pinv: float = 0.3):
    """Variables of the same AIG's latches at the start of frame t (sorted, distinct):
    the only variables of frames [0, t) that the clauses of frames t, t + 1, .. contain.
    An instance that may be extended keeps them frozen (ifreeze)."""
    out = np.zeros(max(1, nlatches), np.uint32)
    lits = np.zeros(1, np.uint32)
    offs = np.zeros(1, np.uint64)
    nc = ctypes.c_uint64()
    nv = ctypes.c_uint32()
    if _load().pfgen_aig_bmc(seed, ngates, nlatches, ninputs, pinv, t, t, lits, offs, ctypes.byref(nc), ctypes.byref(nv),
                             out.ctypes.data) < 0:
        raise ValueError("generator rejected its arguments")
    return np.unique(out[:nlatches])


def images_to_csr(words, nvars: int) -> Cnf:
    """Clause images ([word0, sig, size, lits..] per clause) as a Cnf (host write-back)."""
    w = np.ascontiguousarray(words, np.uint32)
    lits = np.empty(max(1, len(w)), np.uint32)
    offs = np.empty(len(w) // 3 + 2, np.uint64)
    nc = ctypes.c_uint64()
    n = _load().pfgen_images_to_csr(w, len(w), lits, offs, ctypes.byref(nc))
    if n < 0:
        raise ValueError("malformed clause images")
    return Cnf(nvars, lits[:n].copy(), offs[:nc.value + 1].copy())
