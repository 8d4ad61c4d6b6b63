"""GPU: the continuous-batching serving loop against the CPU oracle.

Requests of mixed sizes -- encode and decode, 9/7 and 5/3, grayscale and
RGB -- arrive while others run, several slots run at once and waiting encodes
are merged into one batch plan.  Every served result must equal the oracle's:
an encode's codewords equal the oracle's encode of the request's image, a
decode's image equals the oracle's decode of the request's codewords (the
input image itself for 5/3, the oracle's FP32 inverse path for 9/7).  That
makes each result the same alone and in the mix; the loop is also checked to
have overlapped and merged requests, and to respect its HBM budget."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import serve
from roger_amd import synth

pytestmark = pytest.mark.gpu
NTHREADS = max(1, min(16, os.cpu_count() or 1))


def _host_image(ctx, r):
    """Request r's image, generated on the device (bit-identical to
    synth.image) and read back: much faster than numpy at 8192x8192."""
    buf = synth.device_image(ctx, r.w, r.h, r.ncomp, r.seed)
    img = buf.read(np.uint8, (r.h, r.w) if r.ncomp == 1 else (r.h, r.w, 3))
    buf.free()
    return img


def _expected(r, img, nthreads=1):
    """(codewords, image) the oracle gives for request r with image img."""
    rs, data, q = O.encode_image(img, levels=r.levels, reversible=r.reversible, nthreads=nthreads, style=r.style)
    if r.kind == "encode":
        return data, None
    if r.reversible:
        return data, img
    steps = np.array([s[2] for s in O.band_steps(r.w, r.h, r.levels, 1.0 / 64)], np.float32)
    planes = np.stack([O.dwt97_inv(O.dequantize(q[c], r.levels, steps), r.levels) for c in range(r.ncomp)])
    return data, (O.ict_inv_f32(planes) if r.ncomp == 3 else O.unshift(planes[0]))


def _check_one(r, data, img):
    if r.kind == "encode":
        assert r.result == data, "request %d (%dx%dx%d): codewords differ" % (r.rid, r.w, r.h, r.ncomp)
        return
    # the codewords the decode request carries are the oracle's too
    packed, _ = r.cbs[2]
    joined = b"".join(data)
    assert bytes(packed[:len(joined)]) == joined, "request %d: decode input differs" % r.rid
    np.testing.assert_array_equal(r.result, img, err_msg="request %d: decoded image" % r.rid)


def _check_against_oracle(ctx, reqs):
    # large images one at a time on every thread, the rest one per thread
    big = [r for r in reqs if r.w * r.h > 4 << 20]
    small = [r for r in reqs if r.w * r.h <= 4 << 20]
    for r in big:
        _check_one(r, *_expected(r, _host_image(ctx, r), NTHREADS))
    imgs = [_host_image(ctx, r) for r in small]
    with ThreadPoolExecutor(NTHREADS) as ex:
        for r, (data, img) in zip(small, ex.map(_expected, small, imgs)):
            _check_one(r, data, img)


@pytest.mark.parametrize("style", [O.STYLE_PARALLEL_BYPASS, 0])
def test_serving_results_match_oracle(ctx, style):
    reqs = serve.make_requests(ctx, 16, seed=3, lo=64, hi=700, style=style)
    serve.poisson_arrivals(reqs, rate=4000.0, seed=3)
    srv = serve.Server(slots=3, hbm_budget=1 << 30, merge=4)
    stats = srv.run(reqs, keep_results=True)
    srv.close()
    assert stats["requests"] == len(reqs)
    assert {r.kind for r in reqs} == {"encode", "decode"}
    _check_against_oracle(ctx, reqs)
    # requests overlapped: some request started before an earlier one finished
    by_start = sorted(reqs, key=lambda r: r.start)
    assert any(b.start < a.done for a, b in zip(by_start, by_start[1:]))


def test_serving_merges_waiting_encodes(ctx):
    """All requests waiting at t=0 with one slot: encodes behind the head are
    merged up to 5 per launch; results still equal the oracle's."""
    reqs = serve.make_requests(ctx, 12, seed=5, lo=64, hi=400, decode_share=0.25)
    for r in reqs:
        r.arrival = 0.0
    srv = serve.Server(slots=1, hbm_budget=1 << 30, merge=5)
    stats = srv.run(reqs, keep_results=True)
    srv.close()
    assert stats["launches"] < len(reqs)
    assert max(r.merged for r in reqs) > 1
    _check_against_oracle(ctx, reqs)


@pytest.mark.timeout(900)
def test_serving_bench_mix_vs_oracle(ctx):
    """The benchmark's serving workload (64 requests, sides log-uniform
    128..8192, seed 42) at saturation with 4 slots merging up to 8 encodes:
    every served result against the oracle."""
    reqs = serve.make_requests(ctx, 64, seed=42)
    print("serving mix: %d requests, %.1f Mpix" % (len(reqs), sum(r.mpix for r in reqs)), flush=True)
    for r in reqs:
        r.arrival = 0.0
    srv = serve.Server(slots=4, hbm_budget=96 << 30, merge=8)
    stats = srv.run(reqs, keep_results=True)
    srv.close()
    print("served: %s" % stats, flush=True)
    for r in reqs:
        if r.buf is not None:
            r.buf.free()
    _check_against_oracle(ctx, reqs)


def test_budget_limits_admission(ctx):
    reqs = serve.make_requests(ctx, 6, seed=4, lo=64, hi=300, decode_share=0.0)
    for r in reqs:
        r.arrival = 0.0
    budget = max(r.bytes for r in reqs)  # the largest request fits only alone
    for merge in (1, 4):
        srv = serve.Server(slots=4, hbm_budget=budget, merge=merge)
        srv.run(reqs)
        srv.close()
        for r in reqs:  # working sets of requests in flight together never exceed the budget
            live = [o for o in reqs if o.start <= r.start < o.done]
            assert sum(o.bytes for o in live) <= budget
