"""Serving loop: continuous batching of encode / decode requests on one GPU.

Requests arrive over time (Poisson).  The server owns `slots` execution slots,
each a context with its own HIP stream and its own resident encoder / decoder
buffers.  A request is admitted as soon as it has arrived, a slot is free and
its working set fits the remaining HBM budget; admitted requests from
different slots run concurrently on the device.  Completion is observed by
polling each busy slot's stream, so new arrivals are admitted while others are
still on the GPU.  With slots=1 the same loop is the plain per-request path.

The reference has no serving component (its asynchronous transfer manager,
src/OCLDataTransferManager.cpp:18-30, is a never-constructed stub); this is the
request loop around the hot path.
"""
import collections
import time

import numpy as np

from . import runtime as R
from . import synth


class Request:
    __slots__ = ("rid", "kind", "w", "h", "ncomp", "reversible", "levels", "arrival", "start", "done", "buf",
                 "cbs", "bytes", "result")

    def __init__(self, rid, kind, w, h, ncomp, reversible, levels=5):
        self.rid, self.kind, self.w, self.h, self.ncomp = rid, kind, w, h, ncomp
        self.reversible, self.levels = reversible, levels
        self.arrival = self.start = self.done = 0.0
        self.buf = None      # encode: input image in device memory
        self.cbs = None      # decode: (desc, res, blobs) of the image
        self.result = None
        # working set in HBM: planes + coefficients + scratch + code-block slots (~3 B/sample)
        n = w * h * ncomp
        self.bytes = int(n * (4 + 4 + 3) + n * 2 + w * h * ncomp)

    @property
    def mpix(self):
        return self.w * self.h / 1e6


def make_requests(ctx, n, seed=42, lo=128, hi=8192, lossy_share=0.7, decode_share=0.3, cb=64):
    """Request mix: sides log-uniform on [lo, hi], half grayscale / half RGB,
    70% 9/7 lossy / 30% 5/3 lossless, a share of decode requests (their
    code-blocks produced beforehand by the encoder).  Arrival times are set
    by poisson_arrivals()."""
    u = synth._u01(seed, 9, np.arange(5 * n))
    reqs = []
    enc = None
    for i in range(n):
        w = int(np.floor(lo * np.exp(u[5 * i] * np.log(hi / lo)) + 0.5))
        h = int(np.floor(lo * np.exp(u[5 * i + 1] * np.log(hi / lo)) + 0.5))
        nc = 1 if u[5 * i + 2] < 0.5 else 3
        rev = bool(u[5 * i + 3] >= lossy_share)
        kind = "decode" if u[5 * i + 4] < decode_share else "encode"
        r = Request(i, kind, w, h, nc, rev)
        buf = synth.device_image(ctx, w, h, nc, 5000 + i)
        if kind == "encode":
            r.buf = buf
        else:
            enc = enc or R.Encoder(ctx)
            enc.encode_u8(None, levels=r.levels, reversible=rev, cbw=cb, cbh=cb, d_src=buf, w=w, h=h, ncomp=nc)
            desc, res, blobs = enc.codeblocks()
            r.cbs = (desc, res, R.pack_codewords(blobs))
            r.result = blobs
            buf.free()
        reqs.append(r)
    if enc:
        enc.close()
    return reqs


def poisson_arrivals(reqs, rate, seed=42):
    """Exponential inter-arrival times at `rate` requests/s (deterministic)."""
    u = synth._u01(seed, 11, np.arange(len(reqs)))
    t = 0.0
    for r, x in zip(reqs, u):
        t += -np.log1p(-x) / rate
        r.arrival = t


class Slot:
    def __init__(self, device=0):
        self.ctx = R.Context(device)
        self.enc = R.Encoder(self.ctx)
        self.dec = R.Decoder(self.ctx)
        self.req = None

    def launch(self, r, cb):
        if r.kind == "encode":
            self.enc.encode_u8(None, levels=r.levels, reversible=r.reversible, cbw=cb, cbh=cb, d_src=r.buf, w=r.w,
                               h=r.h, ncomp=r.ncomp)
        else:
            desc, res, blobs = r.cbs
            self.dec.decode(desc, res, blobs, r.w, r.h, r.ncomp, levels=r.levels, reversible=r.reversible, cbw=cb,
                            cbh=cb)
        self.req = r

    def collect(self):
        r = self.req
        r.result = self.enc.codeblocks()[2] if r.kind == "encode" else self.dec.image()

    def close(self):
        self.enc.close()
        self.dec.close()
        self.ctx.close()


class Server:
    def __init__(self, slots=4, hbm_budget=64 << 30, device=0, cb=64):
        self.slots = [Slot(device) for _ in range(slots)]
        self.budget = hbm_budget
        self.cb = cb

    def warm(self, reqs):
        """Grows every slot's resident buffers to the largest request (untimed)."""
        big_e = max((r for r in reqs if r.kind == "encode"), key=lambda r: r.w * r.h * r.ncomp, default=None)
        big_d = max((r for r in reqs if r.kind == "decode"), key=lambda r: r.w * r.h * r.ncomp, default=None)
        for s in self.slots:
            for r in (big_e, big_d):
                if r is not None:
                    s.launch(r, self.cb)
                    s.ctx.sync()
            s.req = None

    def run(self, reqs, keep_results=False):
        """Serves `reqs` (arrival times set); returns latency / throughput stats."""
        pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
        free = list(self.slots)
        busy = []
        used = 0
        t0 = time.perf_counter()
        while pending or busy:
            now = time.perf_counter() - t0
            progressed = False
            while pending and pending[0].arrival <= now and free and used + pending[0].bytes <= self.budget:
                r = pending.popleft()
                s = free.pop()
                r.start = now
                s.launch(r, self.cb)
                used += r.bytes
                busy.append(s)
                progressed = True
            if pending and not busy and pending[0].bytes > self.budget:
                raise RuntimeError("request %d needs more HBM than the whole budget" % pending[0].rid)
            for s in list(busy):
                if s.ctx.query():
                    r = s.req
                    r.done = time.perf_counter() - t0
                    if keep_results:
                        s.collect()
                    used -= r.bytes
                    s.req = None
                    busy.remove(s)
                    free.append(s)
                    progressed = True
            if not progressed:
                time.sleep(20e-6)
        lat = np.array([r.done - r.arrival for r in reqs]) * 1e3
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        return {"requests": len(reqs), "slots": len(self.slots), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3),
                "mpix_s": round(sum(r.mpix for r in reqs) / span, 1), "span_s": round(span, 4)}

    def close(self):
        for s in self.slots:
            s.close()
