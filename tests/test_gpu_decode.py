"""GPU: EBCOT tier-1 decoder and the full native decode path against the CPU
oracle (bit-exact coefficients; lossless round trips; lossy reconstruction
identical to the oracle's FP32 inverse path)."""
import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth
from test_gpu_t1 import CASES, _block

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["wave", "lane"])
def t1dec(request, monkeypatch):
    """Runs a test under each tier-1 decoder (the library reads RG_T1_DEC on
    every launch)."""
    monkeypatch.setenv("RG_T1_DEC", request.param)
    return request.param


def test_t1_decode_blocks_vs_oracle(ctx, t1dec):
    rng = np.random.default_rng(77)
    pw = 64
    blocks, descs, blobs, res, y = [], [], [], [], 0
    for (w, h, kind) in CASES:
        for orient in range(4):
            b = _block(kind, w, h, rng)
            data, npass, nbp = O.t1_encode(b, orient)
            blocks.append(b)
            descs.append((0, y, w, h, 0, orient, 0, 0))
            res.append((len(data), npass, nbp, 0))
            blobs.append(data)
            y += h
    descs = np.array(descs, dtype=R.CB_DESC_DTYPE)
    res = np.array(res, dtype=R.CB_RES_DTYPE)
    q = R.t1_decode(ctx, descs, res, blobs, (1, y, pw))[0]
    for b, d in zip(blocks, descs):
        np.testing.assert_array_equal(q[d["y0"]:d["y0"] + d["h"], :d["w"]], b)


def test_t1_decode_truncated_passes(ctx, t1dec):
    """Decoding only the first passes gives what the oracle decoder gives."""
    rng = np.random.default_rng(5)
    b = rng.laplace(0, 40, size=(64, 64)).astype(np.int32)
    data, npass, nbp = O.t1_encode(b, 2)
    for keep in (1, 2, 3, 7, npass - 1):
        ref = O.t1_decode(data, 64, 64, 2, nbp, keep)
        got = R.t1_decode(ctx, np.array([(0, 0, 64, 64, 0, 2, 0, 0)], R.CB_DESC_DTYPE),
                          np.array([(len(data), keep, nbp, 0)], R.CB_RES_DTYPE), [data], (1, 64, 64))[0]
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("style", [O.STYLE_PARALLEL, O.STYLE_PARALLEL_BYPASS])
def test_t1_decode_style_blocks_vs_oracle(ctx, style, t1dec):
    """rg_t1_decode_style on device-resident codewords of the parallel styles
    (per-pass segments, raw passes under BYPASS): every block of every shape
    and orientation decodes to the oracle's input, and decoding only a
    block's first passes gives what the oracle decoder gives."""
    rng = np.random.default_rng(78)
    pw = 64
    blocks, descs, blobs, res, y = [], [], [], [], 0
    seg = []
    for (w, h, kind) in CASES:
        for orient in range(4):
            b = _block(kind, w, h, rng)
            data, npass, nbp, passlen = O.t1_encode(b, orient, style=style)
            blocks.append(b)
            descs.append((0, y, w, h, 0, orient, 0, 0))
            res.append((len(data), npass, nbp, 0))
            blobs.append(data)
            row = np.zeros(O.MAX_PASSES, np.int32)
            row[:npass] = passlen
            seg.append(row)
            y += h
    descs = np.array(descs, dtype=R.CB_DESC_DTYPE)
    res = np.array(res, dtype=R.CB_RES_DTYPE)
    q = R.t1_decode(ctx, descs, res, blobs, (1, y, pw), style=style, seglen=np.array(seg))[0]
    for b, d in zip(blocks, descs):
        np.testing.assert_array_equal(q[d["y0"]:d["y0"] + d["h"], :d["w"]], b)
    # truncation: the first `keep` passes (their segments) of a dense block
    b = rng.laplace(0, 400, size=(64, 64)).astype(np.int32)
    data, npass, nbp, passlen = O.t1_encode(b, 1, style=style)
    for keep in (1, 4, 10, 11, 12, npass - 1):
        n = int(passlen[:keep].sum())
        ref = O.t1_decode(data[:n], 64, 64, 1, nbp, keep, style=style, passlen=passlen[:keep])
        row = np.zeros((1, O.MAX_PASSES), np.int32)
        row[0, :keep] = passlen[:keep]
        got = R.t1_decode(ctx, np.array([(0, 0, 64, 64, 0, 1, 0, 0)], R.CB_DESC_DTYPE),
                          np.array([(n, keep, nbp, 0)], R.CB_RES_DTYPE), [data[:n]], (1, 64, 64), style=style,
                          seglen=row)[0]
        np.testing.assert_array_equal(got, ref, err_msg="keep %d" % keep)


@pytest.mark.parametrize("w,h,nc,rev,levels,cb", [
    (512, 512, 1, True, 5, 64), (300, 257, 3, True, 4, 64), (257, 190, 3, False, 6, 64), (128, 96, 1, False, 3, 32),
])
def test_full_decode(ctx, w, h, nc, rev, levels, cb, t1dec):
    img = synth.image(w, h, nc, 13)
    enc = R.Encoder(ctx)
    enc.encode_u8(img, levels=levels, reversible=rev, cbw=cb, cbh=cb)
    desc, res, blobs = enc.codeblocks()
    dec = R.Decoder(ctx)
    dec.decode(desc, res, blobs, w, h, nc, levels=levels, reversible=rev, cbw=cb, cbh=cb)
    out = dec.image()
    if rev:
        np.testing.assert_array_equal(out, img)
    else:
        # the oracle's own FP32 inverse path from the same indices
        q = enc.coefficients()
        steps = np.array([s[2] for s in O.band_steps(w, h, levels, 1.0 / 64)], np.float32)
        planes = np.stack([O.dwt97_inv(O.dequantize(q[c], levels, steps), levels) for c in range(nc)])
        ref = O.ict_inv_f32(planes) if nc == 3 else O.unshift(planes[0])
        np.testing.assert_array_equal(out, ref)
        assert np.abs(out.astype(int) - img.astype(int)).max() <= 1
    enc.close()
    dec.close()
