"""The no-allocation batch export (pfg_batch_export_noalloc) against pfg_batch_export,
through the C ABI: every instance's CSR (offsets and literals) equal on two calls in a
row on one device, the second on a smaller batch laid out into the same context, and
the size query equal too."""
import ctypes

import numpy as np
import pytest

from parafrost_amd import gen

pytestmark = pytest.mark.gpu


def _export(g, fn, i):
    """(literals, offsets) of instance i from export entry point fn (sizes queried first)"""
    import torch
    L = g.L
    nl, nc = ctypes.c_uint64(), ctypes.c_uint64()
    assert fn(g.h, i, None, None, ctypes.byref(nl), ctypes.byref(nc)) == 0, L.pfg_last_error(g.h)
    dev = torch.device("cuda", 0)
    lits = torch.full((max(1, nl.value),), -1, dtype=torch.int32, device=dev)
    offs = torch.full((nc.value + 1,), -1, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    nl2, nc2 = ctypes.c_uint64(), ctypes.c_uint64()
    assert fn(g.h, i, lits.data_ptr(), offs.data_ptr(), ctypes.byref(nl2), ctypes.byref(nc2)) == 0, L.pfg_last_error(g.h)
    assert (nl2.value, nc2.value) == (nl.value, nc.value)
    torch.cuda.synchronize()
    return lits.cpu().numpy()[:nl.value].view(np.uint32), offs.cpu().numpy().view(np.uint64)


def _batch(rng, k, lo, hi):
    return [gen.uniform_ksat(int(n), int(n * rng.uniform(3.8, 4.3)), 3, int(rng.integers(1, 1 << 30)))
            for n in rng.integers(lo, hi, k)]


def test_noalloc_export_equals_export_on_two_calls():
    import parafrost_amd as pf
    rng = np.random.default_rng(9700)
    big = _batch(rng, 14, 500, 6000)
    small = _batch(rng, 5, 100, 800)
    cap = int(gen.batch_bases([c.nvars for c in big])[-1])
    g = pf.BatchSimplifier(np.array([cap], np.uint32), pf.default_opts())
    try:
        for k, cnfs in enumerate((big, small)):
            print(f"no-alloc export: call {k + 1}, {len(cnfs)} instances", flush=True)
            g.relayout([c.nvars for c in cnfs])
            g.upload_instances(cnfs)
            g.sigmify_batch()
            for i in range(len(cnfs)):
                a_l, a_o = _export(g, g.L.pfg_batch_export, i)
                b_l, b_o = _export(g, g.L.pfg_batch_export_noalloc, i)
                assert np.array_equal(a_o, b_o), f"call {k + 1} instance {i}: offsets differ"
                assert np.array_equal(a_l, b_l), f"call {k + 1} instance {i}: literals differ"
        assert g.L.pfg_batch_export_noalloc(g.h, 99, None, None, None, None) == pf.PFG_EINVAL
    finally:
        g.close()
