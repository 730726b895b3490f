"""ctypes view of the CPU restatement (liboracle.so) and of reference dumps.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg; never by the product.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import struct
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref", "pfref")


class Opts(ctypes.Structure):
    _fields_ = [("phases", ctypes.c_int32), ("mu_pos", ctypes.c_uint32), ("mu_neg", ctypes.c_uint32),
                ("lcve_min", ctypes.c_uint32), ("lits_min", ctypes.c_int64), ("shrink_rate", ctypes.c_int32),
                ("ve_clause_limit", ctypes.c_int32), ("ve_lbound_en", ctypes.c_int32),
                ("xor_max_arity", ctypes.c_int32), ("sub_limit", ctypes.c_int32), ("ere_limit", ctypes.c_int32),
                ("ere_max_resolvent", ctypes.c_int32), ("sub_en", ctypes.c_int32), ("ve_en", ctypes.c_int32),
                ("ere_en", ctypes.c_int32), ("lbd_tier1", ctypes.c_int32), ("stable_sort", ctypes.c_int32),
                ("bce_en", ctypes.c_int32), ("bce_limit", ctypes.c_int32), ("proof_en", ctypes.c_int32),
                ("ve_fun_en", ctypes.c_int32), ("lcve_max", ctypes.c_uint32), ("lcve_clause_max", ctypes.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", HERE, "all"], check=True, capture_output=True)
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
        L.pfo_default_opts.argtypes = [ctypes.POINTER(Opts)]
        L.pfo_create.restype = vp
        L.pfo_create.argtypes = [ctypes.c_uint32, ctypes.POINTER(Opts), u32p, u64p, ctypes.c_uint64, vp]
        L.pfo_destroy.argtypes = [vp]
        L.pfo_enqueue.argtypes = [vp, ctypes.c_uint32]
        L.pfo_enqueue.restype = ctypes.c_int
        L.pfo_ifreeze.argtypes = [vp, ctypes.c_uint32, ctypes.c_int]
        L.pfo_set_vorg.argtypes = [vp, u32p, ctypes.c_uint32]
        L.pfo_ifreeze.restype = ctypes.c_int
        L.pfo_vmap.restype = ctypes.c_uint64
        L.pfo_vmap.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ctypes.c_uint32), vp, ctypes.POINTER(ctypes.c_uint32), vp]
        for f in ("pfo_sigmify", "pfo_prop", "pfo_lcve", "pfo_unsat", "pfo_phase"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [vp]
        for f in ("pfo_resize_cnf", "pfo_create_ot", "pfo_sort_ot", "pfo_ere", "pfo_sub", "pfo_ve", "pfo_bce",
                  "pfo_filter_pvs", "pfo_shrink", "pfo_next_phase"):
            getattr(L, f).restype = None
            getattr(L, f).argtypes = [vp]
        L.pfo_stop.restype = ctypes.c_int
        L.pfo_stop.argtypes = [vp, ctypes.c_int64]
        L.pfo_count_all.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.pfo_hist.argtypes = [vp, u32p]
        L.pfo_mu_inc.restype = ctypes.c_uint32
        L.pfo_mu_inc.argtypes = [vp]
        L.pfo_num_refs.restype = ctypes.c_uint64
        L.pfo_num_refs.argtypes = [vp]
        L.pfo_dump_proof.restype = ctypes.c_uint64
        L.pfo_dump_proof.argtypes = [vp, vp]
        L.pfo_dump_cnf.restype = ctypes.c_uint64
        L.pfo_dump_cnf.argtypes = [vp, vp]
        L.pfo_trail.restype = ctypes.c_uint32
        L.pfo_trail.argtypes = [vp, vp]
        L.pfo_resolved.restype = ctypes.c_uint64
        L.pfo_resolved.argtypes = [vp, vp]
        L.pfo_vstate.argtypes = [vp, vp]
        L.pfo_pvs.restype = ctypes.c_uint32
        L.pfo_pvs.argtypes = [vp, vp]
        L.pfo_ot_list.restype = ctypes.c_int32
        L.pfo_ot_list.argtypes = [vp, ctypes.c_uint32, vp]
        L.pfo_ot_size.restype = ctypes.c_uint32
        L.pfo_ot_size.argtypes = [vp, ctypes.c_uint32]
        L.pfo_introsort_u32.argtypes = [u32p, ctypes.c_int64, u32p]
        L.pfo_extend_model.argtypes = [vp, vp]
        _lib = L
    return _lib


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().pfo_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@dataclass
class State:
    """Simplifier state in a comparable form."""
    nvars: int
    cnf: np.ndarray          # uint32 words, per clause [hdr0, sig, size, lits...] in scnf order
    trail: np.ndarray        # uint32
    resolved: np.ndarray     # uint32
    vstate: np.ndarray       # uint8, nvars + 1
    pvs: np.ndarray          # uint32
    lists: list = field(default_factory=list)   # per PV literal: np.int64 scnf indices
    unsat: int = 0
    proof: "np.ndarray | None" = None           # uint8 binary DRAT lines of the simplifier (proof_en)

    def clauses(self):
        w, i, out = self.cnf, 0, []
        while i < len(w):
            n = int(w[i + 2])
            out.append((int(w[i]), int(w[i + 1]), tuple(int(x) for x in w[i + 3:i + 3 + n])))
            i += 3 + n
        return out


@dataclasses.dataclass
class Mapped:
    """a mapped write-back: clauses (CSR) over variables 1..nvars, vorg[new] = old"""
    nvars: int
    lits: np.ndarray
    offs: np.ndarray
    vorg: np.ndarray
    trail: np.ndarray
    vstate: np.ndarray


def read_mapped(path: str) -> Mapped:
    """the harness's mapped dump (PFREF_MAP)"""
    b = open(path, "rb").read()
    off = 0

    def take(fmt, n=1):
        nonlocal off
        a = np.frombuffer(b, dtype=fmt, count=n, offset=off)
        off += a.nbytes
        return a
    assert int(take(np.uint32)[0]) == 0x4d524650
    nv = int(take(np.uint32)[0])
    nc = int(take(np.uint64)[0])
    lits, offs = [], [0]
    for _ in range(nc):
        k = int(take(np.uint32)[0])
        lits.append(take(np.uint32, k))
        offs.append(offs[-1] + k)
    nvo = int(take(np.uint32)[0])
    vorg = take(np.uint32, nvo).copy()
    nt = int(take(np.uint32)[0])
    trail = take(np.uint32, nt).copy()
    vs = take(np.uint8, nv + 1).copy()
    return Mapped(nv, np.concatenate(lits) if lits else np.zeros(0, np.uint32), np.array(offs, np.uint64), vorg,
                  trail, vs)


class Oracle:
    def __init__(self, cnf, opts: Opts | None = None, hdr0=None):
        self.L = lib()
        self.nvars = cnf.nvars
        self.opts = opts or default_opts()
        h = None if hdr0 is None else np.ascontiguousarray(hdr0, np.uint32).ctypes.data
        self.h = self.L.pfo_create(cnf.nvars, ctypes.byref(self.opts), np.ascontiguousarray(cnf.lits, np.uint32),
                                   np.ascontiguousarray(cnf.offs, np.uint64), cnf.ncls, h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.pfo_destroy(self.h)
            self.h = None

    def sigmify(self) -> int:
        return self.L.pfo_sigmify(self.h)

    def set_vorg(self, vorg):
        """the caller's variable mapping (vorg[v] = original variable of v): model
        records and proof lines are written in original numbering"""
        v = np.ascontiguousarray(vorg, np.uint32)
        if self.L.pfo_set_vorg(self.h, v, len(v)) != 0:
            raise ValueError("vorg needs nvars + 1 entries, all nonzero past entry 0")

    def ifreeze(self, vars, on: bool = True):
        """ifreeze / iunfreeze (src/cpu/assume.cpp:45-61): LCVE passes over these variables."""
        for v in np.asarray(vars, np.uint32).ravel():
            if self.L.pfo_ifreeze(self.h, int(v), 1 if on else 0) != 0:
                raise ValueError(f"variable {int(v)} out of range")

    def proof(self) -> np.ndarray:
        """the binary DRAT lines the simplifier has written so far (proof_en)"""
        n = self.L.pfo_dump_proof(self.h, None)
        out = np.zeros(n, np.uint8)
        if n:
            self.L.pfo_dump_proof(self.h, out.ctypes.data)
        return out

    def vmap(self) -> "Mapped":
        """variable mapping with write-back after sigmify (ParaFROST::map(true))"""
        st = self.state()
        nl = int(sum(len(c[2]) for c in st.clauses()))
        lits = np.zeros(max(1, nl), np.uint32)
        offs = np.zeros(len(st.clauses()) + 1, np.uint64)
        vorg = np.zeros(self.nvars + 1, np.uint32)
        trail = np.zeros(max(1, len(st.trail)), np.uint32)
        vs = np.zeros(self.nvars + 1, np.uint8)
        nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
        nc = self.L.pfo_vmap(self.h, lits.ctypes.data, offs.ctypes.data, vorg.ctypes.data, ctypes.byref(nv),
                             trail.ctypes.data, ctypes.byref(nt), vs.ctypes.data)
        n = int(nv.value)
        return Mapped(n, lits[:int(offs[nc])].copy(), offs[:nc + 1].copy(), vorg[:n + 1].copy(),
                      trail[:nt.value].copy(), vs[:n + 1].copy())

    def enqueue(self, lits) -> bool:
        """assign units before sigmify; False on a contradiction"""
        ok = True
        for l in np.asarray(lits, np.uint32).ravel():
            ok &= self.L.pfo_enqueue(self.h, int(l)) == 0
        return ok

    def state(self, lists=False) -> State:
        L, h = self.L, self.h
        nw = L.pfo_dump_cnf(h, None)
        cnf = np.empty(nw, np.uint32)
        L.pfo_dump_cnf(h, cnf.ctypes.data)
        nt = L.pfo_trail(h, None)
        trail = np.empty(nt, np.uint32)
        L.pfo_trail(h, trail.ctypes.data)
        nr = L.pfo_resolved(h, None)
        res = np.empty(nr, np.uint32)
        L.pfo_resolved(h, res.ctypes.data)
        vs = np.empty(self.nvars + 1, np.uint8)
        L.pfo_vstate(h, vs.ctypes.data)
        npv = L.pfo_pvs(h, None)
        pvs = np.empty(npv, np.uint32)
        L.pfo_pvs(h, pvs.ctypes.data)
        st = State(self.nvars, cnf, trail, res, vs, pvs, unsat=L.pfo_unsat(h))
        if lists:
            for v in pvs:
                for sg in (0, 1):
                    lit = 2 * int(v) + sg
                    n = L.pfo_ot_size(h, lit)
                    idx = np.empty(n, np.int64)
                    L.pfo_ot_list(h, lit, idx.ctypes.data)
                    st.lists.append(idx)
        return st

    def extend_model(self, value: np.ndarray) -> np.ndarray:
        v = np.ascontiguousarray(value, np.uint8).copy()
        self.L.pfo_extend_model(self.h, v.ctypes.data)
        return v


def read_dump(path: str) -> State:
    with open(path, "rb") as f:
        b = f.read()
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from(fmt, b, off)
        off += struct.calcsize(fmt)
        return v

    magic, nvars = take("<II")
    assert magic == 0x44524650
    _, words = take("<QQ")
    cnf = np.frombuffer(b, np.uint32, words, off).copy(); off += 4 * words
    (nt,) = take("<I")
    trail = np.frombuffer(b, np.uint32, nt, off).copy(); off += 4 * nt
    (nr,) = take("<Q")
    res = np.frombuffer(b, np.uint32, nr, off).copy(); off += 4 * nr
    vs = np.frombuffer(b, np.uint8, nvars + 1, off).copy(); off += nvars + 1
    (npv,) = take("<I")
    pvs = np.frombuffer(b, np.uint32, npv, off).copy(); off += 4 * npv
    (nl,) = take("<I")
    lists = []
    for _ in range(nl):
        (n,) = take("<I")
        lists.append(np.frombuffer(b, np.int64, n, off).copy()); off += 8 * n
    (status,) = take("<I")
    return State(nvars, cnf, trail, res, vs, pvs, lists, unsat=status)


def ref_opts_args(o: Opts) -> list[str]:
    """Reference command-line options equal to an Opts value (src/cpu/options.cpp names)."""
    a = [f"--phases={o.phases}", f"--mupos={o.mu_pos}", f"--muneg={o.mu_neg}", f"--lcvemin={o.lcve_min}",
         f"--vephasemin={o.lits_min}", f"--gcfreq={o.shrink_rate}", f"--veclausemax={o.ve_clause_limit}",
         f"--xormaxarity={o.xor_max_arity}", f"--submax={o.sub_limit}", f"--eremax={o.ere_limit}",
         f"--ereclausemax={o.ere_max_resolvent}", "-velbound" if o.ve_lbound_en else "-no-velbound",
         "-ere" if o.ere_en else "-no-ere", "-bce" if o.bce_en else "-no-bce", f"--bcemax={o.bce_limit}",
         # ve_en stands for the reference's ve_en || ve_plus_en (options.cpp:283)
         "-ve" if o.ve_en else "-no-ve", "-ve+" if o.ve_en else "-no-ve+", "-sub" if o.sub_en else "-no-sub",
         "--mapperc=1"]
    return a


def run_reference(cnf_path: str, out_path: str, o: Opts, trace_prefix: str | None = None,
                  timeout: float = 600, freeze=None, mapped: bool = False, second: bool = False):
    """freeze: variables frozen as by ifreeze() in incremental mode; mapped: run the
    reference's variable mapping with write-back after the simplifier and return its
    Mapped dump (None when the formula is refuted). With o.proof_en the reference
    writes its binary DRAT proof; the simplifier's part (from the start of the
    simplifier on) is returned as the state's `proof` bytes. second (with mapped):
    then run the simplifier a second time on the mapped formula (a second sigmify()
    call, phases lowered as sigmify() lowers them) and return (Mapped, State of the
    second call, model-stack size before it), or None when the first call refutes;
    the State's proof is the second call's part."""
    env = dict(os.environ)
    if second:
        env["PFREF_SECOND"] = "1"
        mapped = True
    proof_args = []
    if o.proof_en:
        proof_args = ["-proof", f"--proofout={out_path}.drat"]
        env["PFREF_PROOF_MARK"] = out_path + ".mark"
    if mapped:
        env["PFREF_MAP"] = "1"
    if trace_prefix:
        env["PFREF_TRACE"] = trace_prefix
    if freeze is not None:
        fz = out_path + ".freeze"
        np.asarray(freeze, np.uint32).tofile(fz)
        env["PFREF_FREEZE"] = fz
    if os.path.exists(out_path):
        os.remove(out_path)
    r = subprocess.run([REF_BIN, cnf_path, out_path] + ref_opts_args(o) + proof_args, env=env,
                       capture_output=True, timeout=timeout)

    def with_proof(st):
        if o.proof_en and st is not None:
            at = int(np.fromfile(out_path + ".mark", np.uint64)[0])
            st.proof = np.fromfile(out_path + ".drat", np.uint8)[at:]
        return st

    if not os.path.exists(out_path):
        # the reference ends the process on a proven contradiction (killSolver)
        if b"UNSATISFIABLE" in r.stdout:
            return with_proof(State(0, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                                    np.zeros(0, np.uint8), np.zeros(0, np.uint32), unsat=1))
        raise RuntimeError(f"reference run failed ({r.returncode}): {r.stdout[-500:]!r} {r.stderr[-500:]!r}")
    if mapped:
        with open(out_path, "rb") as f:
            magic = f.read(4)
        if magic != b"PFRM":
            return None
        m = read_mapped(out_path)
        if not second:
            return m
        if r.returncode != 0 or not os.path.exists(out_path + ".2"):
            raise RuntimeError(f"reference second call failed ({r.returncode}): {r.stderr[-500:]!r}")
        st = read_dump(out_path + ".2")
        with open(out_path + ".2", "rb") as f:
            nres_first = int(np.frombuffer(f.read()[-8:], np.uint64)[0])
        if o.proof_en:
            at = int(np.fromfile(out_path + ".mark.2", np.uint64)[0])
            st.proof = np.fromfile(out_path + ".drat", np.uint8)[at:]
        return m, st, nres_first
    return with_proof(read_dump(out_path))
