"""Serving loop: continuous batching of encode / decode requests on one GPU.

Requests arrive over time (Poisson).  The server owns `slots` execution
slots, each a context with its own HIP stream and its own resident encoder,
decoder and batch buffers (one stream per slot: the encoders run their
tier-1 in one phase, so slots x streams stay within the GPU's hardware
queues).  Whenever a slot is free, the request at the head of the queue is
admitted if it has arrived and its working set fits the remaining HBM budget;
with `merge` > 1, encode requests of the same code-block style that are
already waiting behind it join it, up to `merge` requests and the budget, and
the group runs as one batch plan (one launch per stage for all of them).
Admitted work from different slots runs concurrently on the device;
completion is observed by polling each busy slot's stream, so arrivals are
admitted while others are still on the GPU.  With slots=1, merge=1 the same
loop is the plain per-request path.

The reference has no serving component (its asynchronous transfer manager,
src/OCLDataTransferManager.cpp:18-30, is a never-constructed stub); this is the
request loop around the hot path.
"""
import collections
import time

import numpy as np

from . import runtime as R
from . import synth
from ._lib import RG_STYLE_PARALLEL_BYPASS


class Request:
    __slots__ = ("rid", "kind", "w", "h", "ncomp", "reversible", "levels", "style", "seed", "arrival", "start",
                 "done", "buf", "cbs", "bytes", "result", "merged")

    def __init__(self, rid, kind, w, h, ncomp, reversible, levels=5, style=RG_STYLE_PARALLEL_BYPASS, seed=0):
        self.rid, self.kind, self.w, self.h, self.ncomp = rid, kind, w, h, ncomp
        self.reversible, self.levels, self.style, self.seed = reversible, levels, int(style), seed
        self.arrival = self.start = self.done = 0.0
        self.buf = None      # encode: input image in device memory
        self.cbs = None      # decode: (desc, res, (data, offsets), seglen) of the image
        self.result = None
        self.merged = 1      # requests in the launch that served this one
        # working set in HBM: planes + coefficients + scratch + code-block
        # slots and tier-1 buffers (~3 B/sample), input
        n = w * h * ncomp
        self.bytes = int(n * (4 + 4 + 3) + n * 2 + w * h * ncomp)

    @property
    def mpix(self):
        return self.w * self.h / 1e6


def make_requests(ctx, n, seed=42, lo=128, hi=8192, lossy_share=0.7, decode_share=0.3, cb=64,
                  style=RG_STYLE_PARALLEL_BYPASS):
    """Request mix: sides log-uniform on [lo, hi], half grayscale / half RGB,
    70% 9/7 lossy / 30% 5/3 lossless, a share of decode requests (their
    code-blocks produced beforehand by the encoder), every request in code-
    block style `style`.  Request i's image is synth.image(w, h, ncomp, 5000 + i).
    Arrival times are set by poisson_arrivals()."""
    u = synth._u01(seed, 9, np.arange(5 * n))
    reqs = []
    enc = None
    for i in range(n):
        w = int(np.floor(lo * np.exp(u[5 * i] * np.log(hi / lo)) + 0.5))
        h = int(np.floor(lo * np.exp(u[5 * i + 1] * np.log(hi / lo)) + 0.5))
        nc = 1 if u[5 * i + 2] < 0.5 else 3
        rev = bool(u[5 * i + 3] >= lossy_share)
        kind = "decode" if u[5 * i + 4] < decode_share else "encode"
        r = Request(i, kind, w, h, nc, rev, style=style, seed=5000 + i)
        buf = synth.device_image(ctx, w, h, nc, r.seed)
        if kind == "encode":
            r.buf = buf
        else:
            enc = enc or R.Encoder(ctx)
            enc.encode_u8(None, levels=r.levels, reversible=rev, cbw=cb, cbh=cb, d_src=buf, w=w, h=h, ncomp=nc,
                          style=style)
            desc, res, blobs = enc.codeblocks()
            r.cbs = (desc, res, R.pack_codewords(blobs), enc.pass_lengths() if style else None)
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
    """One execution slot: a context (stream) with a resident encoder,
    decoder and batch plan."""

    def __init__(self, device=0, cb=64):
        self.ctx = R.Context(device)
        self.enc = R.Encoder(self.ctx)
        self.enc.set_t1_phases(1)
        self.dec = R.Decoder(self.ctx)
        self.batch = None
        self.cb = cb
        self.reqs = []

    def launch(self, reqs):
        """Queues `reqs` (one request, or several encodes of one style) on the slot's stream."""
        cb = self.cb
        if len(reqs) > 1:
            if self.batch is None:
                self.batch = R.Batch(self.ctx)
            images = [dict(buf=r.buf, w=r.w, h=r.h, ncomp=r.ncomp, reversible=r.reversible, levels=r.levels,
                           delta=1.0 / 64) for r in reqs]
            self.batch.plan(images, cbw=cb, cbh=cb, use_graph=False, style=reqs[0].style)
            self.batch.run()
        else:
            r = reqs[0]
            if r.kind == "encode":
                self.enc.encode_u8(None, levels=r.levels, reversible=r.reversible, cbw=cb, cbh=cb, d_src=r.buf,
                                   w=r.w, h=r.h, ncomp=r.ncomp, style=r.style)
            else:
                desc, res, blobs, seglen = r.cbs
                self.dec.decode(desc, res, blobs, r.w, r.h, r.ncomp, levels=r.levels, reversible=r.reversible,
                                cbw=cb, cbh=cb, style=r.style, seglen=seglen)
        for r in reqs:
            r.merged = len(reqs)
        self.reqs = list(reqs)

    def collect(self):
        """Results of the finished launch: codewords (encode) or the image (decode)."""
        if len(self.reqs) > 1:
            for i, r in enumerate(self.reqs):
                r.result = self.batch.codeblocks(i)[2]
        else:
            r = self.reqs[0]
            r.result = self.enc.codeblocks()[2] if r.kind == "encode" else self.dec.image()

    def close(self):
        if self.batch is not None:
            self.batch.close()
        self.enc.close()
        self.dec.close()
        self.ctx.close()


class Server:
    def __init__(self, slots=4, hbm_budget=64 << 30, device=0, cb=64, merge=1):
        self.slots = [Slot(device, cb) for _ in range(slots)]
        self.budget = hbm_budget
        self.cb = cb
        self.merge = max(1, int(merge))

    def warm(self, reqs):
        """Grows every slot's resident buffers to the largest request (untimed),
        and, with merging, its batch buffers to a group of the largest encodes."""
        big_e = max((r for r in reqs if r.kind == "encode"), key=lambda r: r.w * r.h * r.ncomp, default=None)
        big_d = max((r for r in reqs if r.kind == "decode"), key=lambda r: r.w * r.h * r.ncomp, default=None)
        enc = sorted((r for r in reqs if r.kind == "encode"), key=lambda r: -r.w * r.h * r.ncomp)
        group = []
        for r in enc:
            if len(group) < self.merge and sum(o.bytes for o in group) + r.bytes <= self.budget:
                group.append(r)
        for s in self.slots:
            for r in (big_e, big_d):
                if r is not None:
                    s.launch([r])
                    s.ctx.sync()
            if len(group) > 1:
                s.launch(group)
                s.ctx.sync()
            s.reqs = []
        for r in reqs:
            r.merged = 1

    def _take(self, pending, now, used):
        """The group admitted next: the head request, then (merging) waiting
        encodes of its style that fit the budget; None if the head cannot go."""
        head = pending[0]
        if head.arrival > now or used + head.bytes > self.budget:
            return None
        group = [pending.popleft()]
        if self.merge > 1 and head.kind == "encode":
            keep = collections.deque()
            load = used + head.bytes
            while pending and pending[0].arrival <= now and len(group) < self.merge:
                r = pending.popleft()
                if r.kind == "encode" and r.style == head.style and load + r.bytes <= self.budget:
                    group.append(r)
                    load += r.bytes
                else:
                    keep.append(r)
            pending.extendleft(reversed(keep))
        return group

    def run(self, reqs, keep_results=False):
        """Serves `reqs` (arrival times set); returns latency / throughput stats."""
        pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
        free = list(self.slots)
        busy = []
        used = 0
        launches = 0
        t0 = time.perf_counter()
        while pending or busy:
            now = time.perf_counter() - t0
            progressed = False
            while pending and free:
                group = self._take(pending, now, used)
                if group is None:
                    break
                s = free.pop()
                for r in group:
                    r.start = now
                s.launch(group)
                launches += 1
                used += sum(r.bytes for r in group)
                busy.append(s)
                progressed = True
            if pending and not busy and pending[0].bytes > self.budget:
                raise RuntimeError("request %d needs more HBM than the whole budget" % pending[0].rid)
            for s in list(busy):
                if s.ctx.query():
                    t = time.perf_counter() - t0
                    if keep_results:
                        s.collect()
                    for r in s.reqs:
                        r.done = t
                        used -= r.bytes
                    s.reqs = []
                    busy.remove(s)
                    free.append(s)
                    progressed = True
            if not progressed:
                time.sleep(20e-6)
        lat = np.array([r.done - r.arrival for r in reqs]) * 1e3
        span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        return {"requests": len(reqs), "slots": len(self.slots), "merge": self.merge, "launches": launches,
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3), "mpix_s": round(sum(r.mpix for r in reqs) / span, 1),
                "span_s": round(span, 4)}

    def close(self):
        for s in self.slots:
            s.close()
