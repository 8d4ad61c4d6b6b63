"""Python host side over the C ABI: device context, device buffers and the
stage entry points, taking numpy arrays on the host and DeviceBuffer objects
on the device.  No torch: device memory comes from rg_malloc."""
import ctypes
import weakref

import numpy as np

from ._lib import (CbDesc, CbResult, EncodeParams, RG_MAX_PASSES, RG_COLOUR_ICT, RG_COLOUR_NONE, RG_COLOUR_RCT, RG_FILTER_53,
                   RG_FILTER_97, check, lib)


def _vp(x):
    return ctypes.c_void_p(x)


class Context:
    """A device + HIP stream (the reference's ocl_args_d_t)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().rg_context_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        self._owned = weakref.WeakSet()   # encoders, decoders, batches, buffers made on this context

    def _own(self, obj):
        self._owned.add(obj)

    def sync(self):
        check(lib().rg_context_sync(self.handle))

    def query(self):
        """True when all work queued on this context's stream has finished."""
        return check(lib().rg_context_query(self.handle)) == 1

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        check(lib().rg_context_device_name(self.handle, buf, 256))
        return buf.value.decode()

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, arr.nbytes)
        b.write(arr)
        return b

    def host_alloc(self, nbytes):
        """Pinned host memory (rg_host_alloc) as a PinnedBuffer."""
        return PinnedBuffer(self, nbytes)

    def timer_start(self):
        check(lib().rg_timer_start(self.handle))

    def timer_stop(self):
        ms = ctypes.c_float()
        check(lib().rg_timer_stop(self.handle, ctypes.byref(ms)))
        return ms.value

    def close(self):
        """Releases what was made on this context (still alive), then the
        context: the library needs its objects destroyed before their
        context."""
        if self.handle:
            for o in list(self._owned):
                try:
                    o.close()
                except Exception:
                    pass
            lib().rg_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory owned by Python (rg_malloc / rg_free)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().rg_malloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        ctx._own(self)

    def write(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, "host array larger than the device buffer"
        check(lib().rg_memcpy_h2d(self.ctx.handle, _vp(self.ptr), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))

    def read(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes, "read larger than the device buffer"
        check(lib().rg_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p), _vp(self.ptr), out.nbytes))
        return out

    def zero(self):
        check(lib().rg_memset(self.ctx.handle, _vp(self.ptr), 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().rg_free(self.ctx.handle, _vp(self.ptr))
            self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory (rg_host_alloc / rg_host_free); .array(dtype,
    shape) is a numpy view of it."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().rg_host_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value
        ctx._own(self)

    def array(self, dtype, shape):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        assert n <= self.nbytes, "view larger than the pinned buffer"
        raw = (ctypes.c_uint8 * n).from_address(self.ptr)
        return np.frombuffer(raw, dtype=dt).reshape(shape)

    def free(self):
        if self.ptr:
            lib().rg_host_free(self.ctx.handle, _vp(self.ptr))
            self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- band steps
def band_steps(levels, delta, bitdepth=8):
    """Per-band quantisation steps (T.800 E.1 exponent/mantissa closest to delta):
    returns (step float32[3L+1], inv_step float32[3L+1], eps int[], mu int[])."""
    nb = 3 * levels + 1
    step = np.empty(nb, np.float32)
    inv = np.empty(nb, np.float32)
    eps = np.empty(nb, np.int32)
    mu = np.empty(nb, np.int32)
    for b in range(nb):
        orient = 0 if b == 0 else 1 + (b - 1) % 3
        Rb = bitdepth + (0 if orient == 0 else (2 if orient == 3 else 1))
        m, e = np.frexp(delta)
        m, e = m * 2.0, int(e) - 1
        u = int(np.floor((m - 1.0) * 2048.0 + 0.5))
        if u >= 2048:
            u, e = 0, e + 1
        ep = max(0, min(31, Rb - e))
        s = np.ldexp(1.0 + u / 2048.0, Rb - ep)
        step[b], inv[b], eps[b], mu[b] = np.float32(s), np.float32(1.0 / s), ep, u
    return step, inv, eps, mu


# ---------------------------------------------------------------- stages
def colour_fwd(ctx, d_src, w, h, ncomp, colour, out_float, d_planes):
    check(lib().rg_colour_fwd(ctx.handle, _vp(d_src.ptr), w, h, ncomp, colour, int(out_float), _vp(d_planes.ptr)))


def colour_inv(ctx, d_planes, w, h, ncomp, colour, in_float, d_dst):
    check(lib().rg_colour_inv(ctx.handle, _vp(d_planes.ptr), w, h, ncomp, colour, int(in_float), _vp(d_dst.ptr)))


def planar_split_i16(ctx, d_src, w, h, ncomp, d_planes):
    check(lib().rg_planar_split_i16(ctx.handle, _vp(d_src.ptr), w, h, ncomp, _vp(d_planes.ptr)))


def scratch_bytes(w, h, nplanes, filt):
    return lib().rg_dwt_scratch_bytes(w, h, nplanes, filt)


def h2d_async(ctx, d_dst, host):
    """Queued copy of a (pinned) host array into device memory on the context
    stream (for timing uploads)."""
    host = np.ascontiguousarray(host)
    check(lib().rg_memcpy_h2d_async(ctx.handle, _vp(d_dst.ptr), host.ctypes.data_as(ctypes.c_void_p), host.nbytes))


def dwt_fwd(ctx, filt, d_in, w, h, nplanes, levels, inv_step, ll_fp16, d_q, d_scratch):
    inv = None if inv_step is None else np.ascontiguousarray(inv_step, np.float32)
    check(lib().rg_dwt_fwd(ctx.handle, filt, _vp(d_in.ptr), w, h, nplanes, levels,
                           None if inv is None else inv.ctypes.data_as(ctypes.c_void_p), int(ll_fp16),
                           _vp(d_q.ptr), _vp(d_scratch.ptr) if d_scratch else None))


def dwt_fwd_u8_scratch_bytes(w, h, ncomp, colour, filt):
    return lib().rg_dwt_fwd_u8_scratch_bytes(w, h, ncomp, colour, filt)


def dwt_fwd_u8(ctx, d_src, w, h, ncomp, colour, filt, levels, inv_step, d_q, d_scratch):
    """Colour stage + forward transform from the interleaved uint8 image
    (first level fused where the shape allows)."""
    inv = None if inv_step is None else np.ascontiguousarray(inv_step, np.float32)
    check(lib().rg_dwt_fwd_u8(ctx.handle, _vp(d_src.ptr), w, h, ncomp, colour, filt, levels,
                              None if inv is None else inv.ctypes.data_as(ctypes.c_void_p), _vp(d_q.ptr),
                              _vp(d_scratch.ptr)))


def dwt_inv_u8_scratch_bytes(w, h, ncomp, colour, filt):
    return lib().rg_dwt_inv_u8_scratch_bytes(w, h, ncomp, colour, filt)


def dwt_inv_u8(ctx, d_q, w, h, ncomp, colour, filt, levels, step, d_dst, d_scratch):
    """Inverse transform + inverse colour stage into the interleaved uint8
    image (last level fused where the shape allows)."""
    st = None if step is None else np.ascontiguousarray(step, np.float32)
    check(lib().rg_dwt_inv_u8(ctx.handle, _vp(d_q.ptr), w, h, ncomp, colour, filt, levels,
                              None if st is None else st.ctypes.data_as(ctypes.c_void_p), _vp(d_dst.ptr),
                              _vp(d_scratch.ptr)))


def dwt_fwd_noalloc(ctx, filt, d_in, w, h, nplanes, levels, inv_step, ll_fp16, d_q):
    """rg_dwt_fwd with the library's static per-GPU scratch (no allocation)."""
    inv = None if inv_step is None else np.ascontiguousarray(inv_step, np.float32)
    check(lib().rg_dwt_fwd_noalloc(ctx.handle, filt, _vp(d_in.ptr), w, h, nplanes, levels,
                                   None if inv is None else inv.ctypes.data_as(ctypes.c_void_p), int(ll_fp16),
                                   _vp(d_q.ptr)))


def dwt97_fwd_float(ctx, d_in, w, h, nplanes, levels, d_out, d_scratch):
    check(lib().rg_dwt97_fwd_float(ctx.handle, _vp(d_in.ptr), w, h, nplanes, levels, _vp(d_out.ptr),
                                   _vp(d_scratch.ptr) if d_scratch else None))


def dwt_inv(ctx, filt, d_q, w, h, nplanes, levels, step, d_out, d_scratch):
    st = None if step is None else np.ascontiguousarray(step, np.float32)
    check(lib().rg_dwt_inv(ctx.handle, filt, _vp(d_q.ptr), w, h, nplanes, levels,
                           None if st is None else st.ctypes.data_as(ctypes.c_void_p), _vp(d_out.ptr),
                           _vp(d_scratch.ptr) if d_scratch else None))


def codeblocks(w, h, levels, cbw, cbh, ncomp=1):
    n = check(lib().rg_codeblocks(w, h, levels, cbw, cbh, None, 0))
    arr = (CbDesc * max(n, 1))()
    check(lib().rg_codeblocks(w, h, levels, cbw, cbh, ctypes.cast(arr, ctypes.c_void_p), n))
    one = np.ctypeslib.as_array(arr)[:n].copy()
    out = np.concatenate([one] * ncomp) if ncomp > 1 else one
    if ncomp > 1:
        out["comp"] = np.repeat(np.arange(ncomp, dtype=np.int32), n)
    return out


CB_DESC_DTYPE = np.dtype([("x0", "<i4"), ("y0", "<i4"), ("w", "<i4"), ("h", "<i4"), ("comp", "<i4"),
                          ("orient", "<i4"), ("band", "<i4"), ("res", "<i4")])
CB_RES_DTYPE = np.dtype([("len", "<i4"), ("npasses", "<i4"), ("nbp", "<i4"), ("flags", "<i4")])


def t1_encode(ctx, d_q, pitch, plane_stride, descs, slot_bytes):
    """Tier-1 encodes the blocks `descs` (structured array); returns (results, list of bytes)."""
    descs = np.ascontiguousarray(descs).view(CB_DESC_DTYPE)
    n = len(descs)
    d_desc = ctx.upload(descs)
    d_out = ctx.alloc(n * slot_bytes)
    d_res = ctx.alloc(n * 16)
    check(lib().rg_t1_encode(ctx.handle, _vp(d_q.ptr), pitch, plane_stride, _vp(d_desc.ptr), n, _vp(d_out.ptr),
                             slot_bytes, _vp(d_res.ptr)))
    res = d_res.read(CB_RES_DTYPE, (n,))
    raw = d_out.read(np.uint8, (n, slot_bytes))
    if (res["flags"] != 0).any():
        raise RuntimeError("tier-1 output overflow")
    data = [bytes(raw[i, :res["len"][i]]) for i in range(n)]
    return res, data


class Batch:
    """Batched encode of many images (rg_batch): one launch per stage for the
    whole batch, optionally replayed from a captured HIP graph."""

    def __init__(self, ctx):
        from ._lib import BatchImage
        self._BatchImage = BatchImage
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().rg_batch_create(ctx.handle, ctypes.byref(h)))
        self.handle = h
        ctx._own(self)

    def plan(self, images, cbw=64, cbh=64, use_graph=True, style=0, waves_per_cu=0):
        """images: list of dicts {buf: DeviceBuffer, w, h, ncomp, reversible, levels, delta}."""
        arr = (self._BatchImage * len(images))()
        for a, m in zip(arr, images):
            a.d_src = m["buf"].ptr
            a.w, a.h, a.ncomp = m["w"], m["h"], m["ncomp"]
            rev = m["reversible"]
            a.filter = RG_FILTER_53 if rev else RG_FILTER_97
            a.colour = RG_COLOUR_NONE if m["ncomp"] != 3 else (RG_COLOUR_RCT if rev else RG_COLOUR_ICT)
            a.levels, a.delta = m["levels"], float(m.get("delta", 1.0 / 64))
        self.images = images
        check(lib().rg_batch_plan(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(images), cbw, cbh,
                                  int(use_graph), int(style), int(waves_per_cu)))
        self.style = int(style)

    def run(self):
        check(lib().rg_batch_run(self.handle))

    def num_codeblocks(self, image=-1):
        return check(lib().rg_batch_num_codeblocks(self.handle, image))

    def coefficients(self, image):
        m = self.images[image]
        dq = ctypes.c_void_p()
        check(lib().rg_batch_coefficients(self.handle, image, ctypes.byref(dq)))
        out = np.empty((m["ncomp"], m["h"], m["w"]), np.int32)
        self.ctx.sync()
        check(lib().rg_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p), dq, out.nbytes))
        return out

    def codeblocks(self, image):
        n = self.num_codeblocks(image)
        desc = np.empty(n, CB_DESC_DTYPE)
        res = np.empty(n, CB_RES_DTYPE)
        offs = np.empty(n, np.int64)
        total = ctypes.c_long()
        self.ctx.sync()
        check(lib().rg_batch_get_codeblocks(self.handle, image, desc.ctypes.data_as(ctypes.c_void_p),
                                            res.ctypes.data_as(ctypes.c_void_p), None, None, 0, ctypes.byref(total)))
        data = np.empty(max(total.value, 1), np.uint8)
        check(lib().rg_batch_get_codeblocks(self.handle, image, None, None, offs.ctypes.data_as(ctypes.c_void_p),
                                            data.ctypes.data_as(ctypes.c_void_p), data.size, None))
        return desc, res, [bytes(data[offs[i]:offs[i] + res["len"][i]]) for i in range(n)]

    def pass_lengths(self, image):
        """TERMALL style: [ncb(image), RG_MAX_PASSES] segment lengths."""
        out = np.zeros((self.num_codeblocks(image), RG_MAX_PASSES), np.int32)
        check(lib().rg_batch_get_pass_lengths(self.handle, image, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def j2k(self, image):
        size = ctypes.c_long()
        self.ctx.sync()
        check(lib().rg_batch_write_j2k(self.handle, image, None, 0, ctypes.byref(size)))
        out = np.empty(size.value, np.uint8)
        check(lib().rg_batch_write_j2k(self.handle, image, out.ctypes.data_as(ctypes.c_void_p), out.size, None))
        return out.tobytes()

    def close(self):
        if self.handle:
            lib().rg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_codewords(blobs):
    """List of codewords -> (concatenated uint8 array, int64 offsets)."""
    offs = np.zeros(len(blobs), np.int64)
    pos = 0
    for i, b in enumerate(blobs):
        offs[i] = pos
        pos += len(b)
    return np.frombuffer(b"".join(blobs) + b"\0", np.uint8), offs


def t1_decode(ctx, descs, results, blobs, q_shape, style=0, seglen=None):
    """Tier-1 decodes code-blocks into zeroed int32 planes of shape (ncomp, h, w).
    style: code-block style (0, RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS);
    TERMALL styles take seglen, [ncb, RG_MAX_PASSES] pass segment lengths."""
    descs = np.ascontiguousarray(descs).view(CB_DESC_DTYPE)
    res = np.ascontiguousarray(results).view(CB_RES_DTYPE)
    n = len(descs)
    offs = np.zeros(n, np.int64)
    pos = 0
    for i, b in enumerate(blobs):
        offs[i] = pos
        pos += len(b)
    data = np.frombuffer(b"".join(blobs) + b"\0" * 16, np.uint8)
    d_data, d_offs, d_res, d_desc = ctx.upload(data), ctx.upload(offs), ctx.upload(res), ctx.upload(descs)
    nc, h, w = q_shape
    d_q = ctx.alloc(nc * h * w * 4)
    d_q.zero()
    if style:
        d_seg = ctx.upload(np.ascontiguousarray(seglen, np.int32))
        check(lib().rg_t1_decode_style(ctx.handle, _vp(d_data.ptr), _vp(d_offs.ptr), _vp(d_res.ptr),
                                       _vp(d_desc.ptr), n, _vp(d_q.ptr), w, h * w, int(style), _vp(d_seg.ptr)))
    else:
        check(lib().rg_t1_decode(ctx.handle, _vp(d_data.ptr), _vp(d_offs.ptr), _vp(d_res.ptr), _vp(d_desc.ptr), n,
                                 _vp(d_q.ptr), w, h * w))
    return d_q.read(np.int32, q_shape)


class Decoder:
    """The native decode path (rg_decoder): tier-1 decode -> dequantisation +
    inverse DWT -> inverse colour -> uint8."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().rg_decoder_create(ctx.handle, ctypes.byref(h)))
        self.handle = h
        ctx._own(self)

    def decode(self, descs, results, blobs, w, h, ncomp, levels=5, reversible=True, cbw=64, cbh=64,
               delta=1.0 / 64, style=0, seglen=None):
        """seglen: [ncb, RG_MAX_PASSES] pass segment lengths (TERMALL style)."""
        descs = np.ascontiguousarray(descs).view(CB_DESC_DTYPE)
        res = np.ascontiguousarray(results).view(CB_RES_DTYPE)
        data, offs = pack_codewords(blobs) if isinstance(blobs, list) else blobs
        pos = int(offs[-1] + res["len"][-1]) if len(offs) else 0
        p = EncodeParams()
        p.levels = levels
        p.filter = RG_FILTER_53 if reversible else RG_FILTER_97
        p.colour = RG_COLOUR_NONE if ncomp != 3 else (RG_COLOUR_RCT if reversible else RG_COLOUR_ICT)
        p.cbw, p.cbh, p.delta, p.ll_fp16, p.bitdepth, p.style = cbw, cbh, float(delta), 0, 8, int(style)
        sl = None
        if seglen is not None:
            sl = np.ascontiguousarray(seglen, np.int32)
            assert sl.shape == (len(res), RG_MAX_PASSES)
        check(lib().rg_decoder_decode_codeblocks(
            self.handle, descs.ctypes.data_as(ctypes.c_void_p), res.ctypes.data_as(ctypes.c_void_p),
            offs.ctypes.data_as(ctypes.c_void_p), data.ctypes.data_as(ctypes.c_void_p), pos, len(res), w, h,
            ncomp, ctypes.byref(p), None if sl is None else sl.ctypes.data_as(ctypes.c_void_p)))
        self.shape = (h, w) if ncomp == 1 else (h, w, ncomp)

    def decode_j2k(self, cs):
        """Decodes a .j2k codestream (bytes); returns its parameters."""
        from ._lib import J2KInfo
        buf = np.frombuffer(bytes(cs), np.uint8)
        info = J2KInfo()
        check(lib().rg_decoder_decode_j2k(self.handle, buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                          ctypes.byref(info)))
        self.shape = (info.h, info.w) if info.ncomp == 1 else (info.h, info.w, info.ncomp)
        return _info_dict(info)

    def image(self):
        dp = ctypes.c_void_p()
        check(lib().rg_decoder_image_u8(self.handle, ctypes.byref(dp)))
        out = np.empty(self.shape, np.uint8)
        self.ctx.sync()
        check(lib().rg_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p), dp, out.nbytes))
        return out

    def close(self):
        if self.handle:
            lib().rg_decoder_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Encoder:
    """The native encode path (rg_encoder): colour -> DWT -> quantisation -> tier-1."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().rg_encoder_create(ctx.handle, ctypes.byref(h)))
        self.handle = h
        ctx._own(self)

    def encode_u8(self, img, levels=5, reversible=True, cbw=64, cbh=64, delta=1.0 / 64, ll_fp16=False,
                  d_src=None, w=None, h=None, ncomp=None, style=0, colour=None):
        if d_src is None:
            img = np.ascontiguousarray(img, np.uint8)
            h, w = img.shape[:2]
            ncomp = 1 if img.ndim == 2 else img.shape[2]
            src, on_dev = img.ctypes.data_as(ctypes.c_void_p), 0
            # the upload is asynchronous: keep the host array alive until a
            # later call (every later call that reads results waits for it)
            self._host_src = img
        else:
            src, on_dev = _vp(d_src.ptr), 1
        p = EncodeParams()
        p.levels = levels
        p.filter = RG_FILTER_53 if reversible else RG_FILTER_97
        if colour is None:
            colour = RG_COLOUR_NONE if ncomp != 3 else (RG_COLOUR_RCT if reversible else RG_COLOUR_ICT)
        p.colour = colour
        p.cbw, p.cbh, p.delta, p.ll_fp16, p.bitdepth = cbw, cbh, float(delta), int(ll_fp16), 8
        p.style = int(style)
        self.shape = (w, h, ncomp)
        self.style = int(style)
        check(lib().rg_encoder_encode_u8(self.handle, src, on_dev, w, h, ncomp, ctypes.byref(p)))

    def coefficients(self):
        dq = ctypes.c_void_p()
        nc, w, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().rg_encoder_coefficients(self.handle, ctypes.byref(dq), ctypes.byref(nc), ctypes.byref(w),
                                            ctypes.byref(h)))
        out = np.empty((nc.value, h.value, w.value), np.int32)
        self.ctx.sync()
        check(lib().rg_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p), dq, out.nbytes))
        return out

    def dwt_in(self, level, dtype):
        """Input of DWT level `level` (getDwtIn): (ncomp, h, w) array of
        `dtype` (float32 for 9/7, int32 for 5/3)."""
        d = ctypes.c_void_p()
        w, h = ctypes.c_int(), ctypes.c_int()
        check(lib().rg_encoder_dwt_in(self.handle, level, ctypes.byref(d), ctypes.byref(w), ctypes.byref(h)))
        out = np.empty((self.shape[2], h.value, w.value), dtype)
        self.ctx.sync()
        check(lib().rg_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p), d, out.nbytes))
        return out

    def dwt(self, levels, delta=1.0 / 64):
        """Stage path: the forward transform of the loaded image again
        (OCLDWTForward::run over all levels), quantised."""
        check(lib().rg_encoder_dwt(self.handle, levels, 0, float(delta)))

    def codeblocks(self):
        n = check(lib().rg_encoder_num_codeblocks(self.handle))
        desc = np.empty(n, CB_DESC_DTYPE)
        res = np.empty(n, CB_RES_DTYPE)
        offs = np.empty(n, np.int64)
        total = ctypes.c_long()
        check(lib().rg_encoder_get_codeblocks(self.handle, desc.ctypes.data_as(ctypes.c_void_p),
                                              res.ctypes.data_as(ctypes.c_void_p), None, None, 0,
                                              ctypes.byref(total)))
        data = np.empty(max(total.value, 1), np.uint8)
        check(lib().rg_encoder_get_codeblocks(self.handle, None, None, offs.ctypes.data_as(ctypes.c_void_p),
                                              data.ctypes.data_as(ctypes.c_void_p), data.size, None))
        blobs = [bytes(data[offs[i]:offs[i] + res["len"][i]]) for i in range(n)]
        return desc, res, blobs

    def pass_lengths(self):
        """TERMALL styles: [ncb, RG_MAX_PASSES] codeword segment lengths per pass."""
        n = check(lib().rg_encoder_num_codeblocks(self.handle))
        out = np.zeros((n, RG_MAX_PASSES), np.int32)
        check(lib().rg_encoder_get_pass_lengths(self.handle, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def set_t1_phases(self, phases):
        """Modelling phases of the parallel-style tier-1 (1: one stream; 0: default)."""
        check(lib().rg_encoder_set_t1_phases(self.handle, int(phases)))

    def enable_rate_control(self, on=True):
        """Before encoding: record what rate control needs (parallel styles)."""
        check(lib().rg_encoder_enable_rate_control(self.handle, int(bool(on))))

    def _weights(self, weights):
        """(array, pointer) of per-block weights; the library reads exactly one
        double per code-block of the last encode, so the length is checked here."""
        if weights is None:
            return None, None
        weights = np.ascontiguousarray(weights, np.float64).ravel()
        n = check(lib().rg_encoder_num_codeblocks(self.handle))
        if weights.size != n:
            raise ValueError("rate control weights: %d given, the encode has %d code-blocks" % (weights.size, n))
        return weights, weights.ctypes.data_as(ctypes.c_void_p)

    def rate_control(self, target_bytes, weights=None):
        """Cuts the code-blocks at the PCRD-opt slope threshold for at most
        target_bytes of code-block data; returns the bytes kept."""
        kept = ctypes.c_long()
        weights, wp = self._weights(weights)
        check(lib().rg_encoder_rate_control(self.handle, int(target_bytes), wp, ctypes.byref(kept)))
        return kept.value

    def rate_control_layers(self, targets, weights=None):
        """Quality layers cut at the non-decreasing byte budgets `targets`;
        returns the bytes kept by the last layer."""
        kept = ctypes.c_long()
        t = np.ascontiguousarray(targets, np.int64)
        weights, wp = self._weights(weights)
        check(lib().rg_encoder_rate_control_layers(self.handle, len(t), t.ctypes.data_as(ctypes.c_void_p), wp,
                                                   ctypes.byref(kept)))
        return kept.value

    def j2k(self):
        """The JPEG 2000 codestream of the last encode (bytes)."""
        size = ctypes.c_long()
        self.ctx.sync()
        check(lib().rg_encoder_write_j2k(self.handle, None, 0, ctypes.byref(size)))
        out = np.empty(size.value, np.uint8)
        check(lib().rg_encoder_write_j2k(self.handle, out.ctypes.data_as(ctypes.c_void_p), out.size, None))
        return out.tobytes()

    def close(self):
        if self.handle:
            lib().rg_encoder_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _info_dict(info):
    d = {n: getattr(info, n) for n, _ in info._fields_ if n != "step"}
    d["step"] = list(info.step[:3 * info.levels + 1])
    return d


def j2k_parse(cs):
    """Tier-2 read of a .j2k codestream (host only): (info, descs, results,
    blobs, seglen or None), blocks in the encoder's order."""
    from ._lib import J2KInfo
    buf = np.frombuffer(bytes(cs), np.uint8)
    info = J2KInfo()
    vp = buf.ctypes.data_as(ctypes.c_void_p)
    check(lib().rg_j2k_parse(vp, buf.size, ctypes.byref(info), None, None, None, None, None))
    n = info.ncb
    desc = np.zeros(n, CB_DESC_DTYPE)
    res = np.zeros(n, CB_RES_DTYPE)
    offs = np.zeros(n, np.int64)
    data = np.zeros(max(info.data_bytes, 1), np.uint8)
    seg = np.zeros((n, RG_MAX_PASSES), np.int32) if info.style & 4 else None
    check(lib().rg_j2k_parse(vp, buf.size, ctypes.byref(info), desc.ctypes.data_as(ctypes.c_void_p),
                             res.ctypes.data_as(ctypes.c_void_p), offs.ctypes.data_as(ctypes.c_void_p),
                             data.ctypes.data_as(ctypes.c_void_p),
                             None if seg is None else seg.ctypes.data_as(ctypes.c_void_p)))
    blobs = [bytes(data[offs[i]:offs[i] + res["len"][i]]) for i in range(n)]
    return _info_dict(info), desc, res, blobs, seg
