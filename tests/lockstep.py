"""Pass-by-pass comparison of the GPU simplifier with the oracle.

Both sides run the reference round loop (src/cpu/simplify.cpp:182-200); after
every pass the full state is compared: clause images in clause order, the unit
trail (as a set: units found by the batched device propagation enter the trail
in a different order, which no later pass reads), the model stack, the variable
states, the elected variables and the elected variables' occurrence lists.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
import parafrost_amd as pf  # noqa: E402
import drat  # noqa: E402


class Mismatch(AssertionError):
    pass


def _opts_pair(**kw):
    return po.default_opts(**kw), pf.default_opts(**kw)


def state_gpu(g: "pf.Simplifier", lists: bool = False) -> po.State:
    st = po.State(g.nvars, g.cnf_words(), g.trail(), g.resolved(), g.vstate(), g.pvs())
    if lists:
        for v in st.pvs:
            for sg in (0, 1):
                st.lists.append(g.ot_list(2 * int(v) + sg))
    return st


def compare(o: po.State, g: po.State, where: str, lists: bool = False):
    _compare(o, g, where, lists)


def _compare(o: po.State, g: po.State, where: str, lists: bool = False):
    def fail(what, a, b):
        msg = f"{where}: {what} differs"
        if len(a) != len(b):
            msg += f" (len oracle {len(a)} gpu {len(b)})"
        else:
            d = np.nonzero(np.asarray(a) != np.asarray(b))[0]
            msg += f" at {d[:8].tolist()}: oracle {np.asarray(a)[d[:8]].tolist()} gpu {np.asarray(b)[d[:8]].tolist()}"
        raise Mismatch(msg)

    if len(o.cnf) != len(g.cnf) or not np.array_equal(o.cnf, g.cnf):
        fail("clause database", o.cnf, g.cnf)
    to, tg = np.sort(o.trail), np.sort(g.trail)
    if not np.array_equal(to, tg):
        fail("unit trail", to, tg)
    if not np.array_equal(o.resolved, g.resolved):
        fail("model stack", o.resolved, g.resolved)
    if not np.array_equal(o.vstate, g.vstate):
        fail("variable states", o.vstate, g.vstate)
    if not np.array_equal(o.pvs, g.pvs):
        fail("elected variables", o.pvs, g.pvs)
    if lists:
        for k, (a, b) in enumerate(zip(o.lists, g.lists)):
            if not np.array_equal(a, b):
                fail(f"occurrence list of literal {2 * int(o.pvs[k // 2]) + (k & 1)}", a, b)


def run_lockstep(cnf, lists: bool = True, units=None, setup=None, freeze=None, vorg=None, inspect=None, keep=None,
                 **kw):
    """Run both engines pass by pass; raise Mismatch at the first difference.
    setup(g), if given, configures the GPU simplifier before the upload; freeze lists
    variables frozen by ifreeze (incremental mode); vorg is the caller's variable
    mapping (model records and proof lines in original numbering); inspect(g) is
    called on the GPU simplifier after the last comparison; keep (a dict), if given,
    receives the last GPU state compared ("gpu") and where ("where"). Returns the number of
    rounds and the final state. With proof_en the binary DRAT lines written so far are
    compared after every pass too (drat.normalized: runs of unit lines as sets)."""
    oo, go = _opts_pair(**kw)
    o = po.Oracle(cnf, oo)
    g = pf.Simplifier(cnf.nvars, go)

    def compare(a, gs, where, lists=False):
        _compare(a, gs, where, lists)
        if keep is not None:
            keep["gpu"], keep["where"] = gs, where

    if vorg is not None:
        o.set_vorg(vorg)
        g.set_vorg(vorg)

    def proof(where):
        if not oo.proof_en:
            return
        a, b = drat.normalized(drat.parse(o.proof())), drat.normalized(drat.parse(g.proof()))
        if a != b:
            k = next((i for i in range(min(len(a), len(b))) if a[i] != b[i]), min(len(a), len(b)))
            raise Mismatch(f"{where}: proof differs at line {k} of {len(a)}/{len(b)}: "
                           f"oracle {a[k] if k < len(a) else None} gpu {b[k] if k < len(b) else None}")

    if setup is not None:
        setup(g)
    g.upload(cnf)
    L = o.L
    if freeze is not None and len(freeze):
        o.ifreeze(freeze)
        g.ifreeze(freeze)
    if units is not None and len(units):
        for u in units:
            L.pfo_enqueue(o.h, int(u))
        g.enqueue_units(units)
    compare(o.state(), state_gpu(g), "upload")
    litsbefore = int(cnf.nlits)
    diff = (1 << 63) - 1
    rounds = 0
    try:
        while litsbefore:
            tag = f"round {rounds}"
            L.pfo_resize_cnf(o.h); g.resizeCNF()
            compare(o.state(), state_gpu(g), tag + " resizeCNF")
            L.pfo_create_ot(o.h); g.createOT()
            okp = L.pfo_prop(o.h)
            try:
                g.prop()
                okg = 1
            except pf.Unsat:
                okg = 0
            if okp != okg:
                raise Mismatch(f"{tag} prop: conflict oracle {not okp} gpu {not okg}")
            proof(tag + " prop")
            if not okp:
                return rounds, None
            compare(o.state(), state_gpu(g), tag + " prop")
            elect_o = L.pfo_lcve(o.h)
            npv = g.LCVE()
            elect_g = npv >= go.lcve_min
            compare(o.state(), state_gpu(g), tag + " LCVE")
            if bool(elect_o) != elect_g:
                raise Mismatch(f"{tag} LCVE decision differs")
            if not elect_o:
                break
            L.pfo_sort_ot(o.h); g.sortOT()
            compare(o.state(lists), state_gpu(g, lists), tag + " sortOT", lists)
            if L.pfo_stop(o.h, diff):
                if g.stop(diff) is not True:
                    raise Mismatch(f"{tag} stop differs")
                L.pfo_ere(o.h); g.ERE()
                compare(o.state(), state_gpu(g), tag + " ERE")
                proof(tag + " ERE")
                break
            L.pfo_sub(o.h)
            try:
                g.SUB()
            except pf.Unsat:
                pass
            if L.pfo_unsat(o.h) or g.is_unsat():
                if bool(L.pfo_unsat(o.h)) != g.is_unsat():
                    raise Mismatch(f"{tag} SUB: unsat oracle {L.pfo_unsat(o.h)} gpu {g.is_unsat()}")
                proof(tag + " SUB")
                return rounds, None
            compare(o.state(), state_gpu(g), tag + " SUB")
            proof(tag + " SUB")
            L.pfo_ve(o.h)
            try:
                g.VE()
            except pf.Unsat:
                pass
            if L.pfo_unsat(o.h) or g.is_unsat():
                if bool(L.pfo_unsat(o.h)) != g.is_unsat():
                    raise Mismatch(f"{tag} VE: unsat oracle {L.pfo_unsat(o.h)} gpu {g.is_unsat()}")
                proof(tag + " VE")
                return rounds, None
            compare(o.state(), state_gpu(g), tag + " VE")
            proof(tag + " VE")
            L.pfo_bce(o.h); g.BCE()
            compare(o.state(), state_gpu(g), tag + " BCE")
            import ctypes
            a, b = ctypes.c_int64(), ctypes.c_int64()
            L.pfo_count_all(o.h, ctypes.byref(a), ctypes.byref(b))
            nc, nl = g.countAll()
            if (a.value, b.value) != (nc, nl):
                raise Mismatch(f"{tag} countAll oracle {(a.value, b.value)} gpu {(nc, nl)}")
            L.pfo_filter_pvs(o.h); g.filterPVs()
            compare(o.state(), state_gpu(g), tag + " filterPVs")
            diff = litsbefore - nl
            litsbefore = nl
            L.pfo_next_phase(o.h); g.next_phase()
            rounds += 1
        okp = L.pfo_prop(o.h)
        try:
            g.prop()
        except pf.Unsat:
            if okp:
                raise Mismatch("final prop: conflict on the gpu only")
            proof("final prop")
            return rounds, None
        L.pfo_shrink(o.h); g.shrinkSimp()
        fo, fg = o.state(), state_gpu(g)
        compare(fo, fg, "final")
        proof("final")
        if inspect is not None:
            inspect(g)
        return rounds, fg
    finally:
        g.close()


def run_full(cnf, proof_cmp=None, **kw):
    """Whole sigmify on both sides, final states compared; with proof_en the proofs
    too, through proof_cmp(bytes) (default drat.normalized(drat.parse(bytes)))."""
    oo, go = _opts_pair(**kw)
    o = po.Oracle(cnf, oo)
    ro = o.sigmify()
    g = pf.Simplifier(cnf.nvars, go)
    g.upload(cnf)
    try:
        try:
            g.sigmify()
            rg = 0
        except pf.Unsat:
            rg = -1
        if (ro != 0) != (rg != 0):
            raise Mismatch(f"unsat oracle {ro} gpu {rg}")
        if ro == 0:
            compare(o.state(), state_gpu(g), "sigmify")
        if oo.proof_en:
            norm = proof_cmp or (lambda b: drat.normalized(drat.parse(b)))
            a, b = norm(o.proof()), norm(g.proof())
            if a != b:
                raise Mismatch(f"proof differs ({len(a)} / {len(b)} lines)")
        return o.state(), g.times()
    finally:
        g.close()
