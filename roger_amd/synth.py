"""Seeded synthetic test images (the workloads are defined on these).

Base single-channel recipe (512x512 grayscale case and every other case):
four 2-D sinusoids with log-uniform periods in [8, 128] px, uniform
orientation in [0, pi) and phase in [0, 2 pi), amplitude 24 each, around 128,
plus integer noise uniform on -8..8, rounded half up and clamped to 0..255.
RGB images build each channel with its own seed (seed, seed+1, seed+2) and add
a shared luminance field W(x, y) = Rx(x) + Ry(y) made of two 1-D random walks
(steps -1/0/+1, seed + 3), each rescaled to +-16, before clamping.

Randomness is splitmix64 (Steele et al.) used as a counter-based generator:
draw k of stream s is mix(s*2^32 + k), so pixels are reproducible in any
order and vectorise in numpy.  The reference harness loads a JPEG instead
(src/OCLTest.cpp:30,54); no image ships with this package.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def _stream(seed, stream, idx):
    base = np.uint64((int(seed) * 0x100000001B3 + int(stream) * 0x9E37) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        return _mix(base * np.uint64(1 << 20) + np.asarray(idx, dtype=np.uint64))


def _u01(seed, stream, idx):
    return (_stream(seed, stream, idx) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def base_field(w, h, seed):
    """Float64 field 128 + sum of 4 sinusoids (no noise, no clamp)."""
    r = _u01(seed, 1, np.arange(12))
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    f = np.full((h, w), 128.0)
    for k in range(4):
        period = 8.0 * np.exp(r[3 * k] * np.log(128.0 / 8.0))
        theta = np.pi * r[3 * k + 1]
        phi = 2.0 * np.pi * r[3 * k + 2]
        f += 24.0 * np.sin(2.0 * np.pi * (xs * np.cos(theta) + ys * np.sin(theta)) / period + phi)
    return f


def noise(w, h, seed):
    u = _u01(seed, 2, np.arange(w * h, dtype=np.uint64)).reshape(h, w)
    return np.floor(u * 17.0) - 8.0


def _finish(f):
    return np.clip(np.floor(f + 0.5), 0, 255).astype(np.uint8)


def gray(w, h, seed=0):
    return _finish(base_field(w, h, seed) + noise(w, h, seed))


def _walk(n, seed, stream):
    steps = np.floor(_u01(seed, stream, np.arange(n)) * 3.0) - 1.0
    wk = np.cumsum(steps)
    m = np.abs(wk).max()
    return wk * (16.0 / m) if m > 0 else wk


def rgb(w, h, seed=1):
    lum = _walk(w, seed + 3, 3)[None, :] + _walk(h, seed + 3, 4)[:, None]
    chans = [_finish(base_field(w, h, seed + c) + noise(w, h, seed + c) + lum) for c in range(3)]
    return np.ascontiguousarray(np.stack(chans, axis=-1))


def image(w, h, ncomp, seed):
    return gray(w, h, seed) if ncomp == 1 else rgb(w, h, seed)


def batch_specs(n=256, seed=7, lo=256, hi=4096, levels=5):
    """Mixed-size batch: sides independently log-uniform on [lo, hi]; image i is
    grayscale when i is even, RGB when odd; 5/3 lossless when i//2 is even,
    9/7 otherwise; image seed 1000 + i."""
    u = _u01(seed, 5, np.arange(2 * n))
    sides = np.floor(lo * np.exp(u * np.log(hi / lo)) + 0.5).astype(int)
    return [dict(w=int(sides[2 * i]), h=int(sides[2 * i + 1]), ncomp=1 if i % 2 == 0 else 3,
                 reversible=(i // 2) % 2 == 0, levels=levels, delta=1.0 / 64, seed=1000 + i) for i in range(n)]


# ---- the same recipe generated on the device (rg_synth_u8) ----
def channel_params(seed):
    """16 doubles: 4 x (cos theta, sin theta, period, phase), as base_field uses them."""
    r = _u01(seed, 1, np.arange(12))
    out = []
    for k in range(4):
        period = 8.0 * np.exp(r[3 * k] * np.log(128.0 / 8.0))
        theta = np.pi * r[3 * k + 1]
        phi = 2.0 * np.pi * r[3 * k + 2]
        out += [np.cos(theta), np.sin(theta), period, phi]
    return np.array(out, np.float64)


def noise_base(seed):
    return np.uint64((int(seed) * 0x100000001B3 + 2 * 0x9E37) & 0xFFFFFFFFFFFFFFFF)


def device_image(ctx, w, h, ncomp, seed):
    """Generates image(w, h, ncomp, seed) directly in device memory."""
    import ctypes
    from ._lib import check, lib
    seeds = [seed] if ncomp == 1 else [seed + c for c in range(3)]
    par = np.concatenate([channel_params(s) for s in seeds])
    nb = np.array([noise_base(s) for s in seeds], np.uint64)
    buf = ctx.alloc(w * h * ncomp)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    if ncomp == 3:
        wx = np.ascontiguousarray(_walk(w, seed + 3, 3))
        wy = np.ascontiguousarray(_walk(h, seed + 3, 4))
        check(lib().rg_synth_u8(ctx.handle, ctypes.c_void_p(buf.ptr), w, h, 3, vp(par), vp(nb), vp(wx), vp(wy)))
    else:
        check(lib().rg_synth_u8(ctx.handle, ctypes.c_void_p(buf.ptr), w, h, 1, vp(par), vp(nb), None, None))
    return buf
