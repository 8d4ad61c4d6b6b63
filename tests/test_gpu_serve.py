"""GPU: the continuous-batching serving loop.  Requests of mixed sizes,
encode and decode, 9/7 and 5/3, arrive while others run; each result must be
identical to the result of the same request served alone."""
import numpy as np
import pytest

from roger_amd import runtime as R
from roger_amd import serve

pytestmark = pytest.mark.gpu


def test_serving_results_match_alone(ctx):
    reqs = serve.make_requests(ctx, 14, seed=3, lo=64, hi=700)
    serve.poisson_arrivals(reqs, rate=2000.0, seed=3)
    # alone: one request at a time on a fresh slot
    alone = {}
    solo = serve.Slot()
    for r in reqs:
        solo.launch(r, 64)
        solo.ctx.sync()
        solo.collect()
        alone[r.rid] = r.result
    solo.close()
    srv = serve.Server(slots=3, hbm_budget=1 << 30)
    stats = srv.run(reqs, keep_results=True)
    srv.close()
    assert stats["requests"] == len(reqs)
    kinds = {r.kind for r in reqs}
    assert kinds == {"encode", "decode"}
    for r in reqs:
        if r.kind == "encode":
            assert r.result == alone[r.rid], "request %d" % r.rid
        else:
            np.testing.assert_array_equal(r.result, alone[r.rid])
    # requests overlapped: some request started before an earlier one finished
    by_start = sorted(reqs, key=lambda r: r.start)
    assert any(b.start < a.done for a, b in zip(by_start, by_start[1:]))


def test_budget_limits_admission(ctx):
    reqs = serve.make_requests(ctx, 6, seed=4, lo=64, hi=300, decode_share=0.0)
    for r in reqs:
        r.arrival = 0.0
    budget = max(r.bytes for r in reqs)  # the largest request fits only alone
    srv = serve.Server(slots=4, hbm_budget=budget)
    srv.run(reqs)
    srv.close()
    for r in reqs:  # working sets of requests in flight together never exceed the budget
        live = [o for o in reqs if o.start <= r.start < o.done]
        assert sum(o.bytes for o in live) <= budget
