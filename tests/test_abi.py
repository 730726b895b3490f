"""The C-ABI library loads without a GPU and exports every symbol the header declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pfgpu.h")).read()
    return sorted(set(re.findall(r"\b(pfg_[a-z_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("pfg_create", "pfg_upload_cnf", "pfg_sigmify", "pfg_lcve", "pfg_sub", "pfg_ve", "pfg_ere", "pfg_gc",
                 "pfg_download_cnf", "pfg_download_resolved", "pfg_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "parafrost_amd", "libpfgpu.so")
    if not os.path.exists(path):
        pytest.fail("libpfgpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    import parafrost_amd as pf
    assert set(pf.API) == set(_declared())


def test_default_opts_match_reference_defaults():
    import parafrost_amd as pf
    o = pf.default_opts()
    # src/cpu/options.cpp:24-46
    assert (o.phases, o.mu_pos, o.mu_neg, o.lcve_min, o.lits_min, o.shrink_rate) == (5, 32, 32, 1, 500, 2)
    assert (o.ve_clause_limit, o.xor_max_arity, o.sub_limit, o.ere_limit, o.ere_max_resolvent) == (100, 10, 3000, 3000, 1000)
    assert (o.sub_en, o.ve_en, o.ere_en, o.ve_lbound_en) == (1, 1, 1, 0)
    assert (o.bce_en, o.bce_limit, o.proof_en) == (0, 3000, 0)   # options.cpp:26, 35, 64


def test_batch_context_rejects_proof():
    """The proof is one instance's log: a batch context with proof_en is refused
    before any device work."""
    import ctypes
    import parafrost_amd as pf
    L = pf.lib()
    o = pf.default_opts(proof_en=1)
    h = ctypes.c_void_p()
    nv = (ctypes.c_uint32 * 2)(10, 20)
    assert L.pfg_create_batch(ctypes.byref(h), 0, 0, 2, nv, ctypes.byref(o)) == pf.PFG_EINVAL
    assert not h.value


def test_new_entry_points_reject_a_null_context():
    """pfg_interrupt, pfg_set_vorg and pfg_path_counts check their context before any
    device work."""
    import parafrost_amd as pf
    L = pf.lib()
    assert L.pfg_interrupt(None) == pf.PFG_EINVAL
    assert L.pfg_set_vorg(None, None, 0) == pf.PFG_EINVAL
    n = ctypes.c_int()
    assert L.pfg_path_counts(None, None, 0, ctypes.byref(n)) == pf.PFG_EINVAL
    assert pf.PFG_EINTR == 6 and pf.PATH_NAMES[0] == "ere_extra_rounds"


def test_tuning_entry_points_reject_a_null_context():
    import parafrost_amd as pf
    L = pf.lib()
    t = pf.Tuning()
    assert L.pfg_set_tuning(None, ctypes.byref(t)) == pf.PFG_EINVAL
    assert L.pfg_get_tuning(None, ctypes.byref(t)) == pf.PFG_EINVAL
