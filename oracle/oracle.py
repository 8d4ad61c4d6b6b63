"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only the test-suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker: the product path in
``roger_amd`` never calls it.  The C sources next to this file restate
ITU-T T.800 for the stages of the reference's hot path (see j2k_oracle.h).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build():
    """Compiles the oracle with make (gcc); used by build() and tests."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class CbResult(ctypes.Structure):
    _fields_ = [("len", ctypes.c_int32), ("npasses", ctypes.c_int32),
                ("nbp", ctypes.c_int32), ("flags", ctypes.c_int32)]


class CbDesc(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_int), ("y0", ctypes.c_int), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("orient", ctypes.c_int), ("band", ctypes.c_int), ("res", ctypes.c_int),
                ("cbx", ctypes.c_int), ("cby", ctypes.c_int)]


class Band(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_int), ("y0", ctypes.c_int), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("orient", ctypes.c_int), ("level", ctypes.c_int), ("res", ctypes.c_int),
                ("band", ctypes.c_int)]


class J2KParams(ctypes.Structure):
    _fields_ = [("w", ctypes.c_int), ("h", ctypes.c_int), ("ncomp", ctypes.c_int), ("levels", ctypes.c_int),
                ("cbw_log2", ctypes.c_int), ("cbh_log2", ctypes.c_int), ("reversible", ctypes.c_int),
                ("mct", ctypes.c_int), ("guard_bits", ctypes.c_int), ("bitdepth", ctypes.c_int),
                ("eps", ctypes.c_int * 97), ("mu", ctypes.c_int * 97), ("cblk_style", ctypes.c_int)]


class EncodeCfg(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("cbw", ctypes.c_int), ("cbh", ctypes.c_int),
                ("reversible", ctypes.c_int), ("ncomp", ctypes.c_int), ("delta", ctypes.c_double),
                ("nthreads", ctypes.c_int), ("style", ctypes.c_int)]


P = ctypes.c_void_p
I = ctypes.c_int


def _declare(L):
    sig = {
        "orc_shift_i32": [P, I, I, I, P], "orc_shift_f32": [P, I, I, I, P],
        "orc_rct_fwd": [P, I, P, P, P], "orc_rct_inv": [P, P, P, I, P],
        "orc_ict_fwd_f32": [P, I, P, P, P], "orc_ict_inv_f32": [P, P, P, I, P],
        "orc_ict_fwd_f64": [P, I, P, P, P], "orc_ict_inv_f64": [P, P, P, I, P],
        "orc_unshift_i32": [P, I, I, I, P], "orc_unshift_f32": [P, I, I, I, P],
        "orc_dwt53_fwd": [P, I, I, I, I], "orc_dwt53_inv": [P, I, I, I, I],
        "orc_dwt97_fwd_f32": [P, I, I, I, I, I], "orc_dwt97_inv_f32": [P, I, I, I, I],
        "orc_dwt97_fwd_f64": [P, I, I, I, I], "orc_dwt97_inv_f64": [P, I, I, I, I],
        "orc_dwt53_fwd_1d": [P, I, P], "orc_dwt97_fwd_1d_f64": [P, I, P],
        "orc_bands": [I, I, I, P],
        "orc_quantize_f32": [P, I, I, I, I, P, P], "orc_dequantize_f32": [P, I, I, I, I, P, P],
        "orc_expmant_for_step": [I, ctypes.c_double, P, P],
        "orc_t1_encode": [P, I, I, I, I, P, I, P],
        "orc_t1_decode": [P, I, I, I, I, I, I, P, I],
        "orc_t1_encode_style": [P, I, I, I, I, I, P, I, P, P],
        "orc_t1_decode_style": [P, I, I, I, I, I, I, I, P, P, I],
        "orc_zc_context": [I, I, I, I], "orc_sc_context": [I, I, P],
        "orc_codeblocks": [I, I, I, I, I, P, I],
        "orc_band_range": [I, I],
        "orc_t2_write": [P, P, P, P, P, I, P, ctypes.c_long],
        "orc_t2_write_layers": [P, P, P, P, P, I, I, P, P, ctypes.c_long],
        "orc_encode_image": [P, I, I, P, P, P, I, I, P, P],
        "orc_round_f16": [ctypes.c_float],
        "orc_t1_encode_rd": [P, I, I, I, I, I, P, I, P, P, P],
        "orc_t1_dsig": [ctypes.c_uint32, I], "orc_t1_dref": [ctypes.c_uint32, I],
        "orc_rd_hull": [I, P, P, ctypes.c_double, P, P, P],
        "orc_rate_control": [I, P, P, P, I, P, ctypes.c_long, P, P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
    L.orc_step_from_expmant.argtypes = [I, I, I]
    L.orc_step_from_expmant.restype = ctypes.c_double
    L.orc_t2_write.restype = ctypes.c_long
    L.orc_t2_write_layers.restype = ctypes.c_long
    L.orc_t1_dsig.restype = ctypes.c_int64
    L.orc_t1_dref.restype = ctypes.c_int64
    L.orc_rate_control.restype = ctypes.c_long
    L.orc_round_f16.restype = ctypes.c_float


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------- colour ----------------
def rct_fwd(rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    n = rgb.size // 3
    out = np.empty((3, n), np.int32)
    lib().orc_rct_fwd(_p(rgb), n, _p(out[0]), _p(out[1]), _p(out[2]))
    return out.reshape((3,) + rgb.shape[:-1])


def rct_inv(planes):
    planes = np.ascontiguousarray(planes, dtype=np.int32)
    n = planes[0].size
    out = np.empty(planes.shape[1:] + (3,), np.uint8)
    lib().orc_rct_inv(_p(planes[0]), _p(planes[1]), _p(planes[2]), n, _p(out))
    return out


def ict_fwd(rgb, dtype=np.float32):
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    n = rgb.size // 3
    out = np.empty((3, n), dtype)
    f = lib().orc_ict_fwd_f32 if dtype == np.float32 else lib().orc_ict_fwd_f64
    f(_p(rgb), n, _p(out[0]), _p(out[1]), _p(out[2]))
    return out.reshape((3,) + rgb.shape[:-1])


def ict_inv_f32(planes):
    planes = np.ascontiguousarray(planes, dtype=np.float32)
    n = planes[0].size
    out = np.empty(planes.shape[1:] + (3,), np.uint8)
    lib().orc_ict_inv_f32(_p(planes[0]), _p(planes[1]), _p(planes[2]), n, _p(out))
    return out


def shift(img, dtype=np.int32):
    """Level shift of a single-channel uint8 image."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    out = np.empty(img.shape, dtype)
    f = lib().orc_shift_i32 if dtype == np.int32 else lib().orc_shift_f32
    f(_p(img), img.size, 1, 0, _p(out))
    return out


def unshift(plane):
    plane = np.ascontiguousarray(plane)
    out = np.empty(plane.shape, np.uint8)
    if plane.dtype == np.int32:
        lib().orc_unshift_i32(_p(plane), plane.size, 1, 0, _p(out))
    else:
        plane = plane.astype(np.float32)
        lib().orc_unshift_f32(_p(plane), plane.size, 1, 0, _p(out))
    return out


# ---------------- DWT ----------------
def dwt53_fwd(plane, levels):
    p = np.array(plane, dtype=np.int32, order="C", copy=True)
    h, w = p.shape
    lib().orc_dwt53_fwd(_p(p), w, h, w, levels)
    return p


def dwt53_inv(plane, levels):
    p = np.array(plane, dtype=np.int32, order="C", copy=True)
    h, w = p.shape
    lib().orc_dwt53_inv(_p(p), w, h, w, levels)
    return p


def dwt97_fwd(plane, levels, dtype=np.float32, ll_fp16=False):
    p = np.array(plane, dtype=dtype, order="C", copy=True)
    h, w = p.shape
    if dtype == np.float32:
        lib().orc_dwt97_fwd_f32(_p(p), w, h, w, levels, int(ll_fp16))
    else:
        lib().orc_dwt97_fwd_f64(_p(p), w, h, w, levels)
    return p


def dwt97_inv(plane, levels, dtype=np.float32):
    p = np.array(plane, dtype=dtype, order="C", copy=True)
    h, w = p.shape
    if dtype == np.float32:
        lib().orc_dwt97_inv_f32(_p(p), w, h, w, levels)
    else:
        lib().orc_dwt97_inv_f64(_p(p), w, h, w, levels)
    return p


def dwt53_fwd_1d(x):
    x = np.ascontiguousarray(x, dtype=np.int32)
    out = np.empty_like(x)
    lib().orc_dwt53_fwd_1d(_p(x), x.size, _p(out))
    return out


def dwt97_fwd_1d_f64(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_dwt97_fwd_1d_f64(_p(x), x.size, _p(out))
    return out


def round_f16(x):
    return lib().orc_round_f16(float(x))


def bands(w, h, levels):
    arr = (Band * (3 * levels + 1))()
    n = lib().orc_bands(w, h, levels, ctypes.cast(arr, P))
    return [dict(x0=b.x0, y0=b.y0, w=b.w, h=b.h, orient=b.orient, level=b.level, res=b.res, band=b.band)
            for b in arr[:n]]


# ---------------- quantisation ----------------
def band_range(bitdepth, orient):
    return lib().orc_band_range(bitdepth, orient)


def expmant_for_step(Rb, delta):
    e, m = ctypes.c_int(), ctypes.c_int()
    lib().orc_expmant_for_step(Rb, float(delta), ctypes.byref(e), ctypes.byref(m))
    return e.value, m.value


def step_from_expmant(Rb, eps, mu):
    return lib().orc_step_from_expmant(Rb, eps, mu)


def band_steps(w, h, levels, delta, bitdepth=8):
    """Per-band (eps, mu, step, inv_step as float32) for a base step delta."""
    out = []
    for b in bands(w, h, levels):
        Rb = band_range(bitdepth, b["orient"])
        e, m = expmant_for_step(Rb, delta)
        s = step_from_expmant(Rb, e, m)
        out.append((e, m, s, np.float32(1.0 / s)))
    return out


def quantize(plane, levels, inv_steps):
    p = np.ascontiguousarray(plane, dtype=np.float32)
    h, w = p.shape
    inv = np.ascontiguousarray(inv_steps, dtype=np.float32)
    q = np.empty((h, w), np.int32)
    lib().orc_quantize_f32(_p(p), w, h, w, levels, _p(inv), _p(q))
    return q


def dequantize(q, levels, steps):
    q = np.ascontiguousarray(q, dtype=np.int32)
    h, w = q.shape
    s = np.ascontiguousarray(steps, dtype=np.float32)
    y = np.empty((h, w), np.float32)
    lib().orc_dequantize_f32(_p(q), w, h, w, levels, _p(s), _p(y))
    return y


# ---------------- tier-1 ----------------
def zc_context(orient, h, v, d):
    return lib().orc_zc_context(orient, h, v, d)


def sc_context(H, V):
    x = ctypes.c_int()
    c = lib().orc_sc_context(H, V, ctypes.byref(x))
    return c, x.value


STYLE_BYPASS = 0x01    # T.800 Table A.19: raw SPP/MRP passes from the 5th bit-plane
STYLE_RESET = 0x02     # contexts reset after every pass
STYLE_TERMALL = 0x04   # codeword terminated after every pass
STYLE_PARALLEL = STYLE_RESET | STYLE_TERMALL
STYLE_PARALLEL_BYPASS = STYLE_PARALLEL | STYLE_BYPASS
MAX_PASSES = 96        # stride of the oracle's per-pass length arrays


def t1_encode(block, orient, cap=1 << 17, style=0):
    """(codeword bytes, npasses, nbp); with style, also the per-pass segment
    lengths (meaningful with STYLE_TERMALL) as a 4th element."""
    b = np.ascontiguousarray(block, dtype=np.int32)
    h, w = b.shape
    out = np.zeros(cap, np.uint8)
    r = CbResult()
    passlen = np.zeros(MAX_PASSES, np.int32)
    lib().orc_t1_encode_style(_p(b), w, h, w, orient, style, _p(out), cap, ctypes.byref(r), _p(passlen))
    if r.flags:
        raise RuntimeError("oracle t1 output overflow")
    if style:
        return bytes(out[:r.len]), r.npasses, r.nbp, passlen[:r.npasses].copy()
    return bytes(out[:r.len]), r.npasses, r.nbp


def t1_encode_rd(block, orient, style, cap=1 << 17):
    """(codeword bytes, npasses, nbp, per-pass segment lengths, per-pass
    distortion reductions as int64 in doubled index units)."""
    b = np.ascontiguousarray(block, dtype=np.int32)
    h, w = b.shape
    out = np.zeros(cap, np.uint8)
    r = CbResult()
    passlen = np.zeros(MAX_PASSES, np.int32)
    dist = np.zeros(MAX_PASSES, np.int64)
    lib().orc_t1_encode_rd(_p(b), w, h, w, orient, style, _p(out), cap, ctypes.byref(r), _p(passlen), _p(dist))
    if r.flags:
        raise RuntimeError("oracle t1 output overflow")
    return bytes(out[:r.len]), r.npasses, r.nbp, passlen[:r.npasses].copy(), dist[:r.npasses].copy()


def t1_dsig(m, p):
    return lib().orc_t1_dsig(int(m), int(p))


def t1_dref(m, p):
    return lib().orc_t1_dref(int(m), int(p))


def rd_hull(passlen, dist, weight):
    """Convex-hull truncation points of one block: (passes, bytes, slope) arrays."""
    pl = np.ascontiguousarray(passlen, np.int32)
    d = np.ascontiguousarray(dist, np.int64)
    n = len(pl)
    hn = np.zeros(n + 1, np.int32)
    hr = np.zeros(n + 1, np.int64)
    hs = np.zeros(n + 1, np.float64)
    m = lib().orc_rd_hull(n, _p(pl), _p(d), float(weight), _p(hn), _p(hr), _p(hs))
    return hn[:m].copy(), hr[:m].copy(), hs[:m].copy()


def rate_control(npasses, passlen, dist, weights, target):
    """PCRD-opt truncation of many blocks (passlen / dist: [nblocks,
    MAX_PASSES]) for a byte target: (passes kept, bytes kept, total)."""
    npasses = np.ascontiguousarray(npasses, np.int32)
    pl = np.ascontiguousarray(passlen, np.int32)
    d = np.ascontiguousarray(dist, np.int64)
    w = np.ascontiguousarray(weights, np.float64)
    nb = len(npasses)
    nt = np.zeros(nb, np.int32)
    rt = np.zeros(nb, np.int32)
    total = lib().orc_rate_control(nb, _p(npasses), _p(pl), _p(d), pl.shape[1], _p(w), int(target), _p(nt), _p(rt))
    return nt, rt, int(total)


def t1_decode(data, w, h, orient, nbp, npasses, style=0, passlen=None):
    buf = np.frombuffer(bytes(data) + b"\0", np.uint8).copy()
    q = np.empty((h, w), np.int32)
    if style:
        pl = np.zeros(MAX_PASSES, np.int32)
        if passlen is not None:
            pl[:len(passlen)] = passlen
        lib().orc_t1_decode_style(_p(buf), len(data), w, h, orient, nbp, npasses, style, _p(pl), _p(q), w)
    else:
        lib().orc_t1_decode(_p(buf), len(data), w, h, orient, nbp, npasses, _p(q), w)
    return q


def codeblocks(w, h, levels, cbw, cbh):
    n = lib().orc_codeblocks(w, h, levels, cbw, cbh, None, 0)
    arr = (CbDesc * max(n, 1))()
    lib().orc_codeblocks(w, h, levels, cbw, cbh, ctypes.cast(arr, P), n)
    return [dict(x0=d.x0, y0=d.y0, w=d.w, h=d.h, orient=d.orient, band=d.band, res=d.res,
                 cbx=d.cbx, cby=d.cby) for d in arr[:n]]


# ---------------- whole image + tier-2 ----------------
def encode_image(img, levels=5, reversible=True, cbw=64, cbh=64, delta=1.0 / 64, nthreads=0,
                 slot_bytes=1 << 15, style=0, return_passlen=False):
    """Returns (results[ncomp*ncb] as (len, npasses, nbp, flags), list of bytes,
    qplanes) and, with return_passlen, the per-pass segment lengths
    [ncomp*ncb, MAX_PASSES] as a 4th element."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[:2]
    nc = 1 if img.ndim == 2 else img.shape[2]
    ncb = lib().orc_codeblocks(w, h, levels, cbw, cbh, None, 0) * nc
    res = (CbResult * ncb)()
    slots = np.zeros((ncb, slot_bytes), np.uint8)
    q = np.empty((nc, h, w), np.int32)
    passlen = np.zeros((ncb, MAX_PASSES), np.int32)
    cfg = EncodeCfg(levels, cbw, cbh, int(reversible), nc, float(delta), int(nthreads), int(style))
    n = lib().orc_encode_image(_p(img), w, h, ctypes.byref(cfg), ctypes.cast(res, P), _p(slots),
                               slot_bytes, ncb, _p(q), _p(passlen))
    if n != ncb:
        raise RuntimeError("oracle encode failed: %d" % n)
    rs = np.array([(r.len, r.npasses, r.nbp, r.flags) for r in res], np.int32).reshape(-1, 4)
    if (rs[:, 3] != 0).any():
        raise RuntimeError("oracle code-block overflow")
    data = [bytes(slots[i, :rs[i, 0]]) for i in range(ncb)]
    if return_passlen:
        return rs, data, q, passlen
    return rs, data, q


def j2k_params(w, h, ncomp, levels, reversible, cbw=64, cbh=64, delta=1.0 / 64, guard_bits=2, bitdepth=8,
               style=0):
    p = J2KParams()
    p.cblk_style = int(style)
    p.w, p.h, p.ncomp, p.levels = w, h, ncomp, levels
    p.cbw_log2, p.cbh_log2 = int(cbw).bit_length() - 1, int(cbh).bit_length() - 1
    p.reversible, p.mct = int(reversible), int(ncomp == 3)
    p.guard_bits, p.bitdepth = guard_bits, bitdepth
    for b in bands(w, h, levels):
        Rb = band_range(bitdepth, b["orient"])
        if reversible:
            # no quantisation: exponent = nominal range (+1 bit of headroom for RCT chroma)
            p.eps[b["band"]], p.mu[b["band"]] = Rb + 1, 0
        else:
            e, m = expmant_for_step(Rb, delta)
            p.eps[b["band"]], p.mu[b["band"]] = e, m
    return p


def t2_write(params, results, datas, passlen=None, layers=None):
    """results: int array [ncomp*ncb, 4]; datas: list of bytes (same order);
    passlen: [ncomp*ncb, MAX_PASSES] segment lengths (TERMALL style);
    layers: [ncomp*ncb, nlayers] passes of each block after each quality
    layer (None: one layer)."""
    n = len(datas)
    ncb = n // params.ncomp
    res = (CbResult * n)()
    offs = np.zeros(n, np.int64)
    blob = bytearray()
    for i in range(n):
        res[i].len, res[i].npasses, res[i].nbp, res[i].flags = (int(v) for v in results[i])
        offs[i] = len(blob)
        blob += datas[i]
    blob_np = np.frombuffer(bytes(blob) + b"\0", np.uint8).copy()
    cap = len(blob) + 65536 + 64 * n
    out = np.zeros(cap, np.uint8)
    pl = None if passlen is None else np.ascontiguousarray(passlen, np.int32)
    if layers is None:
        sz = lib().orc_t2_write(ctypes.byref(params), ctypes.cast(res, P), _p(offs), _p(blob_np),
                                None if pl is None else _p(pl), ncb, _p(out), cap)
    else:
        ly = np.ascontiguousarray(layers, np.int32)
        cap += 8 * ly.size
        out = np.zeros(cap, np.uint8)
        sz = lib().orc_t2_write_layers(ctypes.byref(params), ctypes.cast(res, P), _p(offs), _p(blob_np),
                                       None if pl is None else _p(pl), ncb, ly.shape[1], _p(ly), _p(out), cap)
    if sz < 0:
        raise RuntimeError("t2 write failed")
    return bytes(out[:sz])


def encode_j2k(img, levels=5, reversible=True, cbw=64, cbh=64, delta=1.0 / 64, style=0):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[:2]
    nc = 1 if img.ndim == 2 else img.shape[2]
    rs, data, q, passlen = encode_image(img, levels, reversible, cbw, cbh, delta, style=style, return_passlen=True)
    params = j2k_params(w, h, nc, levels, reversible, cbw, cbh, delta, style=style)
    return t2_write(params, rs, data, passlen if style & STYLE_TERMALL else None), rs, data, q
