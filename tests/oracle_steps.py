"""Round loop of the oracle with a state snapshot after every pass (the step
sequence of src/cpu/simplify.cpp:182-200, identical to oracle/ref_harness.cpp)."""
import ctypes

import pyoracle as po


def oracle_passes(o: "po.Oracle", lists_after=("sort",)):
    L, h = o.L, o.h
    litsbefore = int(sum(1 for _ in ()))  # placeholder, replaced below
    nw = L.pfo_dump_cnf(h, None)
    st = o.state()
    litsbefore = sum(len(c[2]) for c in st.clauses() if not (c[0] & 2))
    diff = (1 << 63) - 1
    rnd = 0
    while litsbefore:
        L.pfo_resize_cnf(h)
        L.pfo_create_ot(h)
        if not L.pfo_prop(h):
            return
        yield rnd, "prop", o.state()
        if not L.pfo_lcve(h):
            break
        yield rnd, "lcve", o.state()
        L.pfo_sort_ot(h)
        yield rnd, "sort", o.state(lists=True)
        if L.pfo_stop(h, diff):
            L.pfo_ere(h)
            yield rnd, "ere", o.state()
            break
        L.pfo_sub(h)
        if L.pfo_unsat(h):
            return
        yield rnd, "sub", o.state()
        L.pfo_ve(h)
        if L.pfo_unsat(h):
            return
        yield rnd, "ve", o.state()
        L.pfo_bce(h)
        a, b = ctypes.c_int64(), ctypes.c_int64()
        L.pfo_count_all(h, ctypes.byref(a), ctypes.byref(b))
        L.pfo_filter_pvs(h)
        diff = litsbefore - b.value
        litsbefore = b.value
        L.pfo_next_phase(h)
        rnd += 1
    L.pfo_prop(h)
    L.pfo_shrink(h)
    yield rnd, "final", o.state()
