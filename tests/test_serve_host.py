"""Host-side pieces of serving (no GPU): the request stream and the generator's
frame-by-frame growth."""
import numpy as np

from parafrost_amd import gen


def test_aig_frames_concatenate():
    for seed in (1, 2, 3):
        a, _ = gen.aig_bmc(seed, 700, 0, 3)
        b, nvb = gen.aig_bmc(seed, 700, 3, 7)
        w, nvw = gen.aig_bmc(seed, 700, 0, 7)
        assert nvb == nvw
        assert np.array_equal(np.concatenate([a.lits, b.lits]), w.lits)
        assert a.ncls + b.ncls == w.ncls


def test_aig_clauses_well_formed():
    c, nv = gen.aig_bmc(9, 1200, 0, 5)
    assert c.lits.min() >= 2 and (c.lits >> 1).max() <= nv
    for k in range(c.ncls):
        cl = c.lits[c.offs[k]:c.offs[k + 1]]
        assert 2 <= len(cl) <= 3
        assert len(set((cl >> 1).tolist())) == len(cl)   # distinct variables, no tautology


def test_request_stream_shape():
    from parafrost_amd import serve
    reqs = serve.make_requests(40, seed=3, gates=(200, 400), max_frames=16)
    frames = {}
    for r in reqs:
        assert r.t1 > r.t0
        assert frames.get(r.key, 0) == r.t0   # extensions continue where the instance stopped
        frames[r.key] = r.t1
    assert any(r.t0 > 0 for r in reqs) and any(r.t0 == 0 for r in reqs)
