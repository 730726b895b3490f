"""MI355X-native SIGmA simplifier (ParaFROST's GPU inprocessing engine, rebuilt).

Python view of libpfgpu.so. ``Simplifier`` mirrors the reference entry points
(ParaFROST::sigmify and the pass wrappers of src/cpu/simplify.cpp / src/gpu/simplify.cu);
every method calls the C ABI declared in include/pfgpu.h. There is no fallback:
if the native library is missing or the device call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import gen  # noqa: F401  (workload generators)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PFG_OK, PFG_EINVAL, PFG_ENOMEM, PFG_EDEVICE, PFG_EUNSAT, PFG_EINTERNAL, PFG_EINTR = range(7)
_ERRNAMES = {1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "EUNSAT", 5: "EINTERNAL", 6: "EINTR"}
PATH_NAMES = ("ere_extra_rounds", "sub_wave_vars", "sub_wave_big_vars", "lcve_graph_calls", "sort_exact_lists",
              "ot_updates")


class Opts(ctypes.Structure):
    _fields_ = [("phases", ctypes.c_int32), ("mu_pos", ctypes.c_uint32), ("mu_neg", ctypes.c_uint32),
                ("lcve_min", ctypes.c_uint32), ("lits_min", ctypes.c_int64), ("shrink_rate", ctypes.c_int32),
                ("ve_clause_limit", ctypes.c_int32), ("ve_lbound_en", ctypes.c_int32),
                ("xor_max_arity", ctypes.c_int32), ("sub_limit", ctypes.c_int32), ("ere_limit", ctypes.c_int32),
                ("ere_max_resolvent", ctypes.c_int32), ("sub_en", ctypes.c_int32), ("ve_en", ctypes.c_int32),
                ("ere_en", ctypes.c_int32), ("lbd_tier1", ctypes.c_int32), ("stable_sort", ctypes.c_int32),
                ("bce_en", ctypes.c_int32), ("bce_limit", ctypes.c_int32), ("proof_en", ctypes.c_int32),
                ("ve_fun_en", ctypes.c_int32), ("ve_atomic_en", ctypes.c_int32),
                ("lcve_max", ctypes.c_uint32), ("lcve_clause_max", ctypes.c_uint32)]


class Times(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("upload", "hist", "ot", "sort_ot", "prop", "lcve", "sub", "ve", "ere", "gc", "count", "total", "bce")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Tuning(ctypes.Structure):
    """launch and algorithm choices (pfg_tuning); results never depend on them"""
    _fields_ = [("ot_build", ctypes.c_int32), ("clause_grid", ctypes.c_uint32), ("sub_group", ctypes.c_uint32),
                ("stage_words", ctypes.c_uint32), ("lcve_graph", ctypes.c_int32), ("ve_arena", ctypes.c_uint32),
                ("init_grid", ctypes.c_uint32)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class KStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("kernel", ctypes.c_char_p), ("ms", ctypes.c_double),
                ("launches", ctypes.c_uint64), ("bytes", ctypes.c_double)]


API = ["pfg_default_opts", "pfg_create", "pfg_destroy", "pfg_upload_cnf", "pfg_upload_cnf_device",
       "pfg_enqueue_units", "pfg_sigmify", "pfg_resize_cnf", "pfg_create_ot", "pfg_reduce_ot", "pfg_prop",
       "pfg_lcve", "pfg_sort_ot", "pfg_sub", "pfg_ve", "pfg_bce", "pfg_ere", "pfg_count", "pfg_filter_pvs", "pfg_gc",
       "pfg_stop", "pfg_next_phase", "pfg_hist", "pfg_download_cnf", "pfg_download_trail",
       "pfg_download_resolved", "pfg_download_vstate", "pfg_download_pvs", "pfg_download_ot_list",
       "pfg_times_get", "pfg_is_unsat", "pfg_last_error", "pfg_kernel_stats", "pfg_reset_stats", "pfg_reset",
       "pfg_set_accounting", "pfg_set_stage_words", "pfg_create_batch", "pfg_batch_ninst", "pfg_batch_relayout", "pfg_upload_batch",
       "pfg_sigmify_batch", "pfg_batch_download", "pfg_ifreeze", "pfg_batch_export", "pfg_download_mapped",
       "pfg_download_proof", "pfg_interrupt", "pfg_set_vorg", "pfg_path_counts", "pfg_set_tuning", "pfg_get_tuning",
       "pfg_batch_export_noalloc", "pfg_batch_view"]


def lib_path() -> str:
    return os.path.join(_HERE, "libpfgpu.so")


def lib():
    global _LIB
    if _LIB is None:
        # One HIP runtime per process: PyTorch ships its own libamdhip64 / libhsa-runtime64
        # with the same sonames as ROCm's. Loaded first, they also serve this library; if
        # this library loaded ROCm's copy first, PyTorch could no longer open the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        vp, i32, u32, u64, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64
        L.pfg_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.pfg_create.argtypes = [ctypes.POINTER(vp), i32, u64, u32, ctypes.POINTER(Opts)]
        L.pfg_upload_cnf.argtypes = [vp, vp, vp, u64, vp]
        L.pfg_upload_cnf_device.argtypes = [vp, vp, vp, u64, u64]
        L.pfg_enqueue_units.argtypes = [vp, vp, u32]
        L.pfg_ifreeze.argtypes = [vp, vp, u32, i32]
        L.pfg_download_proof.argtypes = [vp, vp, ctypes.POINTER(u64)]
        L.pfg_download_mapped.argtypes = [vp, vp, ctypes.POINTER(u64), ctypes.POINTER(u64), vp, ctypes.POINTER(u32), vp,
                                          ctypes.POINTER(u32), vp]
        for f in ("pfg_destroy", "pfg_sigmify", "pfg_resize_cnf", "pfg_create_ot", "pfg_reduce_ot", "pfg_prop",
                  "pfg_sort_ot", "pfg_sub", "pfg_ve", "pfg_bce", "pfg_ere", "pfg_filter_pvs", "pfg_gc", "pfg_next_phase",
                  "pfg_is_unsat"):
            getattr(L, f).argtypes = [vp]
        L.pfg_lcve.argtypes = [vp, ctypes.POINTER(u32)]
        L.pfg_count.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.pfg_stop.argtypes = [vp, i64, ctypes.POINTER(i32)]
        L.pfg_hist.argtypes = [vp, vp]
        L.pfg_download_cnf.argtypes = [vp, vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.pfg_download_trail.argtypes = [vp, vp, ctypes.POINTER(u32)]
        L.pfg_download_resolved.argtypes = [vp, vp, ctypes.POINTER(u64)]
        L.pfg_download_vstate.argtypes = [vp, vp]
        L.pfg_download_pvs.argtypes = [vp, vp, ctypes.POINTER(u32)]
        L.pfg_download_ot_list.argtypes = [vp, u32, vp, ctypes.POINTER(u32)]
        L.pfg_times_get.argtypes = [vp, ctypes.POINTER(Times)]
        L.pfg_kernel_stats.argtypes = [vp, ctypes.POINTER(KStat), i32, ctypes.POINTER(i32)]
        L.pfg_reset_stats.argtypes = [vp]
        L.pfg_set_accounting.argtypes = [vp, i32]
        L.pfg_set_stage_words.argtypes = [vp, u32]
        L.pfg_create_batch.argtypes = [ctypes.POINTER(vp), i32, u64, u32, vp, ctypes.POINTER(Opts)]
        L.pfg_batch_ninst.argtypes = [vp, ctypes.POINTER(u32)]
        L.pfg_batch_relayout.argtypes = [vp, u32, vp]
        L.pfg_upload_batch.argtypes = [vp, vp, vp, vp]
        L.pfg_sigmify_batch.argtypes = [vp, vp]
        L.pfg_batch_download.argtypes = [vp, u32, vp, ctypes.POINTER(u64), ctypes.POINTER(u64), vp,
                                         ctypes.POINTER(u64), vp, vp, ctypes.POINTER(u32)]
        L.pfg_batch_export.argtypes = [vp, u32, vp, vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.pfg_batch_export_noalloc.argtypes = [vp, u32, vp, vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        pp32 = ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))
        L.pfg_batch_view.argtypes = [vp, u32, pp32, ctypes.POINTER(u64), ctypes.POINTER(u64), pp32, ctypes.POINTER(u64)]
        L.pfg_reset.argtypes = [vp]
        L.pfg_interrupt.argtypes = [vp]
        L.pfg_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
        L.pfg_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
        L.pfg_set_vorg.argtypes = [vp, vp, u32]
        L.pfg_path_counts.argtypes = [vp, vp, i32, ctypes.POINTER(i32)]
        L.pfg_last_error.argtypes = [vp]
        L.pfg_last_error.restype = ctypes.c_char_p
        for f in API:
            if getattr(L, f).restype is None and f not in ("pfg_default_opts",):
                getattr(L, f).restype = ctypes.c_int
        L.pfg_default_opts.restype = None
        L.pfg_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().pfg_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class HostPool:
    """Grow-only host arrays reused from call to call, so that in a steady state downloads
    land in memory that is already mapped. Arrays taken since the last reset() do not overlap and stay
    valid until the next reset(); growing starts a new array (the old one lives on in
    the arrays taken from it)."""

    def __init__(self):
        self.bufs, self.cur = {}, {}

    def reset(self):
        self.cur = {k: 0 for k in self.cur}

    def take(self, name: str, n: int, dtype) -> np.ndarray:
        a, c = self.bufs.get(name), self.cur.get(name, 0)
        if a is None or a.dtype != np.dtype(dtype) or c + n > a.size:
            a = np.empty(max(n, 2 * (a.size if a is not None else 0), 1 << 16), dtype)
            self.bufs[name], c = a, 0
        self.cur[name] = c + n
        return a[c:c + n]


def _new(pool, name, n, dtype):
    if pool is None:
        return np.zeros(n, dtype) if np.dtype(dtype) == np.uint8 else np.empty(n, dtype)
    v = pool.take(name, n, dtype)
    if np.dtype(dtype) == np.uint8:
        v.fill(0)
    return v


class SimplifierError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class Unsat(SimplifierError):
    pass


class Interrupted(SimplifierError):
    """the call was stopped by interrupt(); the context holds a valid, partly simplified formula"""


class Simplifier:
    """One device-resident CNF instance (a pfg_ctx)."""

    def __init__(self, nvars: int, opts: Opts | None = None, device: int = 0, hbm_budget: int = 0):
        self.L = lib()
        self.nvars = nvars
        self.opts = opts or default_opts()
        h = ctypes.c_void_p()
        self._check(self.L.pfg_create(ctypes.byref(h), device, hbm_budget, nvars, ctypes.byref(self.opts)), None)
        self.h = h

    def _check(self, rc, h="self"):
        if rc == PFG_OK:
            return
        msg = self.L.pfg_last_error(self.h if h == "self" else None)
        msg = msg.decode() if msg else ""
        if rc == PFG_EUNSAT:
            raise Unsat(rc, "formula is unsatisfiable")
        if rc == PFG_EINTR:
            raise Interrupted(rc, "interrupted")
        raise SimplifierError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.L.pfg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- input
    def upload(self, cnf, hdr0=None):
        lits = np.ascontiguousarray(cnf.lits, np.uint32)
        offs = np.ascontiguousarray(cnf.offs, np.uint64)
        h = None if hdr0 is None else np.ascontiguousarray(hdr0, np.uint32)
        self._keep = (lits, offs, h)
        self._check(self.L.pfg_upload_cnf(self.h, lits.ctypes.data, offs.ctypes.data, cnf.ncls,
                                          None if h is None else h.ctypes.data))

    def upload_device(self, d_lits: int, d_offs: int, ncls: int, nlits: int):
        """Clauses already in HBM (device pointers, e.g. torch tensor data_ptr())."""
        self._check(self.L.pfg_upload_cnf_device(self.h, d_lits, d_offs, ncls, nlits))

    def enqueue_units(self, lits):
        a = np.ascontiguousarray(lits, np.uint32)
        self._check(self.L.pfg_enqueue_units(self.h, a.ctypes.data, len(a)))

    def mapped(self):
        """ParaFROST::map(true) write-back after sigmify, computed on the device: (clause
        images with renumbered literals, vorg[new] = variable, mapped trail, vstate)."""
        nw, nc = ctypes.c_uint64(), ctypes.c_uint64()
        nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self.L.pfg_download_mapped(self.h, None, ctypes.byref(nw), ctypes.byref(nc), None, ctypes.byref(nv),
                                               None, ctypes.byref(nt), None))
        w = np.zeros(max(1, nw.value), np.uint32)
        vo = np.zeros(nv.value + 1, np.uint32)
        tr = np.zeros(max(1, nt.value), np.uint32)
        vs = np.zeros(nv.value + 1, np.uint8)
        self._check(self.L.pfg_download_mapped(self.h, w.ctypes.data, ctypes.byref(nw), ctypes.byref(nc), vo.ctypes.data,
                                               ctypes.byref(nv), tr.ctypes.data, ctypes.byref(nt), vs.ctypes.data))
        return w[:nw.value], vo, tr[:nt.value], vs

    def set_vorg(self, vorg):
        """the solver's variable mapping (vorg[v] = original variable of v, nvars + 1
        entries; None = identity): model records and proof lines come out in original numbering"""
        if vorg is None:
            self._check(self.L.pfg_set_vorg(self.h, None, 0))
            return
        a = np.ascontiguousarray(vorg, np.uint32)
        self._check(self.L.pfg_set_vorg(self.h, a.ctypes.data, len(a)))

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self.L.pfg_get_tuning(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        """change some of the launch and algorithm choices (see Tuning)"""
        t = Tuning()
        self._check(self.L.pfg_get_tuning(self.h, ctypes.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        self._check(self.L.pfg_set_tuning(self.h, ctypes.byref(t)))

    def interrupt(self):
        """ask the running or next sigmify to stop at its next poll (thread-safe)"""
        self._check(self.L.pfg_interrupt(self.h))

    def path_counts(self) -> dict:
        """how often the less common code paths ran (names in PATH_NAMES)"""
        a = np.zeros(16, np.uint64)
        n = ctypes.c_int()
        self._check(self.L.pfg_path_counts(self.h, a.ctypes.data, 16, ctypes.byref(n)))
        return {PATH_NAMES[i] if i < len(PATH_NAMES) else f"path{i}": int(a[i]) for i in range(min(n.value, 16))}

    def ifreeze(self, vars, on: bool = True):
        """ifreeze / iunfreeze (src/cpu/assume.cpp:45-61): LCVE never elects these variables."""
        a = np.ascontiguousarray(vars, np.uint32)
        self._check(self.L.pfg_ifreeze(self.h, a.ctypes.data, len(a), 1 if on else 0))

    # ---- reference entry points
    def sigmify(self):
        self._check(self.L.pfg_sigmify(self.h))

    def resizeCNF(self): self._check(self.L.pfg_resize_cnf(self.h))
    def createOT(self): self._check(self.L.pfg_create_ot(self.h))
    def reduceOT(self): self._check(self.L.pfg_reduce_ot(self.h))
    def prop(self): self._check(self.L.pfg_prop(self.h))
    def sortOT(self): self._check(self.L.pfg_sort_ot(self.h))
    def SUB(self): self._check(self.L.pfg_sub(self.h))
    def VE(self): self._check(self.L.pfg_ve(self.h))
    def BCE(self): self._check(self.L.pfg_bce(self.h))
    def ERE(self): self._check(self.L.pfg_ere(self.h))
    def filterPVs(self): self._check(self.L.pfg_filter_pvs(self.h))
    def shrinkSimp(self): self._check(self.L.pfg_gc(self.h))
    def next_phase(self): self._check(self.L.pfg_next_phase(self.h))

    def LCVE(self) -> int:
        n = ctypes.c_uint32()
        self._check(self.L.pfg_lcve(self.h, ctypes.byref(n)))
        return n.value

    def countAll(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.L.pfg_count(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def stop(self, lr: int) -> bool:
        s = ctypes.c_int()
        self._check(self.L.pfg_stop(self.h, lr, ctypes.byref(s)))
        return bool(s.value)

    def hist(self) -> np.ndarray:
        h = np.empty(2 * (self.nvars + 1), np.uint32)
        self._check(self.L.pfg_hist(self.h, h.ctypes.data))
        return h

    # ---- output
    def cnf_words(self, pool=None) -> np.ndarray:
        """pool: a HostPool the array is taken from (reused host memory), else a new array"""
        nw, nc = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.pfg_download_cnf(self.h, None, ctypes.byref(nw), ctypes.byref(nc)))
        w = _new(pool, "words", nw.value, np.uint32)
        self._check(self.L.pfg_download_cnf(self.h, w.ctypes.data, ctypes.byref(nw), ctypes.byref(nc)))
        return w

    def trail(self, pool=None) -> np.ndarray:
        n = ctypes.c_uint32()
        self._check(self.L.pfg_download_trail(self.h, None, ctypes.byref(n)))
        a = _new(pool, "trail", n.value, np.uint32)
        self._check(self.L.pfg_download_trail(self.h, a.ctypes.data, ctypes.byref(n)))
        return a

    def resolved(self, pool=None) -> np.ndarray:
        n = ctypes.c_uint64()
        self._check(self.L.pfg_download_resolved(self.h, None, ctypes.byref(n)))
        a = _new(pool, "records", n.value, np.uint32)
        self._check(self.L.pfg_download_resolved(self.h, a.ctypes.data, ctypes.byref(n)))
        return a

    def proof(self) -> np.ndarray:
        """binary DRAT lines written so far (opts proof_en)"""
        n = ctypes.c_uint64()
        self._check(self.L.pfg_download_proof(self.h, None, ctypes.byref(n)))
        a = np.empty(n.value, np.uint8)
        if n.value:
            self._check(self.L.pfg_download_proof(self.h, a.ctypes.data, ctypes.byref(n)))
        return a

    def vstate(self) -> np.ndarray:
        a = np.empty(self.nvars + 1, np.uint8)
        self._check(self.L.pfg_download_vstate(self.h, a.ctypes.data))
        return a

    def pvs(self) -> np.ndarray:
        n = ctypes.c_uint32()
        self._check(self.L.pfg_download_pvs(self.h, None, ctypes.byref(n)))
        a = np.empty(n.value, np.uint32)
        self._check(self.L.pfg_download_pvs(self.h, a.ctypes.data, ctypes.byref(n)))
        return a

    def ot_list(self, lit: int) -> np.ndarray:
        n = ctypes.c_uint32()
        self._check(self.L.pfg_download_ot_list(self.h, lit, None, ctypes.byref(n)))
        a = np.empty(n.value, np.int64)
        self._check(self.L.pfg_download_ot_list(self.h, lit, a.ctypes.data, ctypes.byref(n)))
        return a

    def times(self) -> dict:
        t = Times()
        self._check(self.L.pfg_times_get(self.h, ctypes.byref(t)))
        return t.as_dict()

    def kernel_times(self) -> dict:
        arr = (KStat * 32)()
        n = ctypes.c_int()
        self._check(self.L.pfg_kernel_stats(self.h, arr, 32, ctypes.byref(n)))
        return {arr[i].name.decode(): {"kernel": arr[i].kernel.decode() if arr[i].kernel else None,
                                       "ms": arr[i].ms, "launches": int(arr[i].launches), "bytes": arr[i].bytes}
                for i in range(n.value)}

    def set_stage_words(self, words: int):
        """Cap the LDS words per block of the staged kernels (results do not depend on it)."""
        self._check(self.L.pfg_set_stage_words(self.h, int(words)))

    def set_accounting(self, on: bool):
        """Compute algorithmic bytes of the per-variable kernels (adds a reduction before each)."""
        self._check(self.L.pfg_set_accounting(self.h, 1 if on else 0))

    def reset_times(self):
        self._check(self.L.pfg_reset_stats(self.h))

    def reset_for_upload(self):
        self._check(self.L.pfg_reset(self.h))

    def stats(self) -> dict:
        ncls, nl = self.countAll()
        return {"clauses": ncls, "literals": nl, "eliminated": int((self.vstate() == 1).sum())}

    def is_unsat(self) -> bool:
        return bool(self.L.pfg_is_unsat(self.h))


class BatchSimplifier(Simplifier):
    """A batch of independent CNF instances in one context: every kernel launch of a
    round serves all of them. Instance i owns its own variables 1..nvars[i]; results
    equal simplifying each instance alone."""

    def __init__(self, nvars, opts: Opts | None = None, device: int = 0, hbm_budget: int = 0):
        self.L = lib()
        self.inst_nvars = np.ascontiguousarray(nvars, np.uint32)
        self.ninst = len(self.inst_nvars)
        self.opts = opts or default_opts()
        h = ctypes.c_void_p()
        self._check(self.L.pfg_create_batch(ctypes.byref(h), device, hbm_budget, self.ninst, self.inst_nvars.ctypes.data,
                                             ctypes.byref(self.opts)), None)
        self.h = h
        n = ctypes.c_uint32()
        # instance i's variable v is the context's variable base[i] + v (bases are multiples of 16)
        self.base = np.zeros(self.ninst + 1, np.uint64)
        at = 0
        for i, n in enumerate(self.inst_nvars):
            self.base[i] = at
            at += int(n)
            if i + 1 < self.ninst:
                at = (at + 15) & ~15
        self.base[self.ninst] = at
        self.nvars = at
        self.nvars_cap = at   # relayout() accepts instance sets up to this many (aligned) variables

    def relayout(self, nvars):
        """Re-lay the context out for another set of instances (allocations are kept)."""
        nv = np.ascontiguousarray(nvars, np.uint32)
        self._check(self.L.pfg_batch_relayout(self.h, len(nv), nv.ctypes.data))
        self.inst_nvars = nv
        self.ninst = len(nv)
        base = np.zeros(self.ninst + 1, np.uint64)
        at = 0
        for i, n in enumerate(nv):
            base[i] = at
            at += int(n)
            if i + 1 < self.ninst:
                at = (at + 15) & ~15
        base[self.ninst] = at
        self.base = base
        self.nvars = at

    def upload_instances(self, cnfs):
        """Instances in their own numbering, concatenated in order."""
        lits = np.concatenate([np.asarray(c.lits[:c.nlits], np.uint32) for c in cnfs]) if cnfs else np.zeros(0, np.uint32)
        offs = [np.zeros(1, np.uint64)]
        first = np.zeros(self.ninst + 1, np.uint64)
        at, lat = 0, 0
        for i, c in enumerate(cnfs):
            offs.append(np.asarray(c.offs[1:c.ncls + 1], np.uint64) + np.uint64(lat))
            at += c.ncls
            lat += int(c.nlits)
            first[i + 1] = at
        offs = np.concatenate(offs)
        self._keep = (lits, offs, first)
        self._check(self.L.pfg_upload_batch(self.h, lits.ctypes.data, offs.ctypes.data, first.ctypes.data))

    def sigmify_batch(self) -> np.ndarray:
        """Per-instance status: 0 (simplified) or 4 (proven unsatisfiable)."""
        st = np.zeros(self.ninst, np.int32)
        self._check(self.L.pfg_sigmify_batch(self.h, st.ctypes.data))
        return st

    def ifreeze_instance(self, i: int, vars, on: bool = True):
        """freeze instance i's variables (its own numbering)"""
        self.ifreeze(np.asarray(vars, np.uint64) + self.base[i], on)

    def enqueue_instance(self, i: int, lits):
        """root-level units of instance i (its own numbering)"""
        self.enqueue_units(np.asarray(lits, np.uint64) + 2 * self.base[i])

    def export(self, i: int, device=None):
        """Instance i's clauses after sigmify_batch as device CSR in its own numbering:
        (lits int32 tensor, offs int64 tensor of ncls + 1 entries), kept in HBM."""
        import torch
        nl, nc = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.pfg_batch_export(self.h, i, None, None, ctypes.byref(nl), ctypes.byref(nc)))
        dev = device or torch.device("cuda", torch.cuda.current_device())
        lits = torch.empty(max(1, nl.value), dtype=torch.int32, device=dev)
        offs = torch.empty(nc.value + 1, dtype=torch.int64, device=dev)
        torch.cuda.current_stream(dev).synchronize()   # the allocations are ready before the library writes
        self._check(self.L.pfg_batch_export(self.h, i, lits.data_ptr(), offs.data_ptr(), ctypes.byref(nl),
                                            ctypes.byref(nc)))
        return lits[:nl.value], offs

    def instance_view(self, i: int):
        """instance(i) without copying the clause words and model records: arrays that view
        the context's pinned host memory, valid until its next relayout, upload or sigmify
        (trail and variable states are copies)"""
        pw, pr = ctypes.POINTER(ctypes.c_uint32)(), ctypes.POINTER(ctypes.c_uint32)()
        nw, nc, nr = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.pfg_batch_view(self.h, i, ctypes.byref(pw), ctypes.byref(nw), ctypes.byref(nc),
                                          ctypes.byref(pr), ctypes.byref(nr)))
        w = np.ctypeslib.as_array(pw, (nw.value,)) if nw.value else np.zeros(0, np.uint32)
        r = np.ctypeslib.as_array(pr, (nr.value,)) if nr.value else np.zeros(0, np.uint32)
        nt = ctypes.c_uint32()
        self._check(self.L.pfg_batch_download(self.h, i, None, None, None, None, None, None, None, ctypes.byref(nt)))
        t = np.empty(nt.value, np.uint32)
        vs = np.zeros(int(self.inst_nvars[i]) + 1, np.uint8)
        self._check(self.L.pfg_batch_download(self.h, i, None, None, None, None, None, vs.ctypes.data, t.ctypes.data,
                                              ctypes.byref(nt)))
        return w, t, r, vs

    def instance(self, i: int, pool=None):
        """(clause words, trail, model records, variable states) of instance i in its own numbering;
        pool: a HostPool the arrays are taken from (reused host memory), else new arrays."""
        nw, nc, nr = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        nt = ctypes.c_uint32()
        self._check(self.L.pfg_batch_download(self.h, i, None, ctypes.byref(nw), ctypes.byref(nc), None, ctypes.byref(nr),
                                              None, None, ctypes.byref(nt)))
        w = _new(pool, "words", nw.value, np.uint32)
        r = _new(pool, "records", nr.value, np.uint32)
        t = _new(pool, "trail", nt.value, np.uint32)
        vs = _new(pool, "vstate", int(self.inst_nvars[i]) + 1, np.uint8)
        self._check(self.L.pfg_batch_download(self.h, i, w.ctypes.data, ctypes.byref(nw), ctypes.byref(nc), r.ctypes.data,
                                              ctypes.byref(nr), vs.ctypes.data, t.ctypes.data, ctypes.byref(nt)))
        return w, t, r, vs
