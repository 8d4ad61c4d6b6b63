"""GPU parity: EBCOT tier-1 encoder and the whole native encode path against
the CPU oracle (bit-exact codewords, pass counts and bit-plane counts), plus
an end-to-end decode of the GPU output by OpenJPEG (through Pillow)."""
import io

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth

pytestmark = pytest.mark.gpu


def _block(kind, w, h, rng):
    if kind == "zero":
        return np.zeros((h, w), np.int32)
    if kind == "single":
        b = np.zeros((h, w), np.int32)
        b[rng.integers(h), rng.integers(w)] = int(rng.integers(1, 5000)) * (1 if rng.random() < 0.5 else -1)
        return b
    if kind == "corners":
        b = np.zeros((h, w), np.int32)
        b[0, 0], b[-1, -1], b[0, -1], b[-1, 0] = 7, -9, 1, -1
        return b
    if kind == "sparse":
        b = rng.laplace(0, 4, size=(h, w)).astype(np.int32)
        b[rng.random((h, w)) < 0.7] = 0
        return b
    if kind == "laplace":
        return rng.laplace(0, 30, size=(h, w)).astype(np.int32)
    if kind == "dense":
        return rng.integers(-70000, 70000, size=(h, w), dtype=np.int32)
    if kind == "ones":
        return rng.choice(np.array([-1, 0, 1], np.int32), size=(h, w))
    raise ValueError(kind)


CASES = []
for (w, h) in [(64, 64), (32, 32), (1, 1), (4, 4), (37, 13), (64, 3), (3, 64), (5, 64), (64, 61), (16, 8)]:
    for kind in ["zero", "single", "corners", "sparse", "laplace", "dense", "ones"]:
        CASES.append((w, h, kind))


def test_t1_blocks_vs_oracle(ctx):
    rng = np.random.default_rng(2024)
    # pack every case (x 4 orientations) as a code-block of one tall plane
    pw = 64
    blocks, descs, y = [], [], 0
    for (w, h, kind) in CASES:
        for orient in range(4):
            b = _block(kind, w, h, rng)
            blocks.append((b, orient))
            descs.append((0, y, w, h, 0, orient, 0, 0))
            y += h
    plane = np.zeros((y, pw), np.int32)
    for (b, _), d in zip(blocks, descs):
        plane[d[1]:d[1] + d[3], :d[2]] = b
    d_q = ctx.upload(plane)
    desc = np.array(descs, dtype=R.CB_DESC_DTYPE)
    res, data = R.t1_encode(ctx, d_q, pw, plane.size, desc, slot_bytes=1 << 15)
    for i, ((b, orient), d) in enumerate(zip(blocks, descs)):
        ref, npass, nbp = O.t1_encode(b, orient)
        assert res["nbp"][i] == nbp, (i, d)
        assert res["npasses"][i] == npass, (i, d)
        assert data[i] == ref, "code-block %d (w=%d h=%d orient=%d) differs" % (i, d[2], d[3], orient)
        # the oracle decoder recovers the block from the GPU codeword
        if npass:
            np.testing.assert_array_equal(O.t1_decode(data[i], d[2], d[3], orient, nbp, npass), b)


@pytest.mark.parametrize("w,h,nc,rev,levels,cb", [
    (512, 512, 1, True, 5, 64),      # the 512x512 grayscale lossless case
    (300, 257, 3, True, 4, 64),
    (257, 190, 3, False, 6, 64),
    (128, 96, 1, False, 3, 32),
    (61, 33, 1, True, 2, 32),
])
def test_encoder_vs_oracle_and_openjpeg(ctx, w, h, nc, rev, levels, cb):
    img = synth.image(w, h, nc, 11)
    enc = R.Encoder(ctx)
    enc.encode_u8(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64)
    desc, res, blobs = enc.codeblocks()
    rs, data, q = O.encode_image(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64)
    np.testing.assert_array_equal(enc.coefficients(), q)
    np.testing.assert_array_equal(res["len"], rs[:, 0])
    np.testing.assert_array_equal(res["npasses"], rs[:, 1])
    np.testing.assert_array_equal(res["nbp"], rs[:, 2])
    assert blobs == data
    # wrap the GPU codewords in a codestream and decode it with OpenJPEG
    from PIL import Image
    params = O.j2k_params(w, h, nc, levels, rev, cb, cb, 1.0 / 64)
    gres = np.stack([res["len"], res["npasses"], res["nbp"], res["flags"]], axis=1)
    cs = O.t2_write(params, gres, blobs)
    # the library's own tier-2 writer produces the same codestream
    assert enc.j2k() == cs
    dec = np.array(Image.open(io.BytesIO(cs)))
    if rev:
        np.testing.assert_array_equal(dec, img)
    else:
        assert np.abs(dec.astype(int) - img.astype(int)).max() <= 1
    enc.close()
