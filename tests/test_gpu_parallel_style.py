"""GPU: the parallel code-block styles (RESET | TERMALL: one MQ chain per
coding pass; with BYPASS, raw low bit-plane passes) against the CPU oracle in
the same style -- bit-exact codewords and per-pass segment lengths, the same
codestream bytes from the library's tier-2 writer, OpenJPEG decoding it, the
native decoder inverting it, and the batched path."""
import io

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth
from roger_amd._lib import RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS

pytestmark = pytest.mark.gpu

assert RG_STYLE_PARALLEL == O.STYLE_PARALLEL and RG_STYLE_PARALLEL_BYPASS == O.STYLE_PARALLEL_BYPASS
STYLES = [RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS]


@pytest.mark.parametrize("PAR", STYLES)
@pytest.mark.parametrize("w,h,nc,rev,levels,cb", [
    (512, 512, 1, True, 5, 64),
    (300, 257, 3, True, 4, 64),
    (257, 190, 3, False, 6, 64),
    (128, 96, 1, False, 3, 32),
    (33, 1, 1, False, 1, 64),
])
def test_encoder_parallel_style_vs_oracle(ctx, w, h, nc, rev, levels, cb, PAR):
    img = synth.image(w, h, nc, 21)
    enc = R.Encoder(ctx)
    enc.encode_u8(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64, style=PAR)
    desc, res, blobs = enc.codeblocks()
    seg = enc.pass_lengths()
    rs, data, q, passlen = O.encode_image(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64,
                                          style=PAR, return_passlen=True)
    np.testing.assert_array_equal(enc.coefficients(), q)
    np.testing.assert_array_equal(res["npasses"], rs[:, 1])
    np.testing.assert_array_equal(res["nbp"], rs[:, 2])
    np.testing.assert_array_equal(res["flags"], 0)
    for i in range(len(res)):
        n = rs[i, 1]
        np.testing.assert_array_equal(seg[i, :n], passlen[i, :n])
    assert blobs == data
    # codestream: the library writer == the oracle writer; OpenJPEG decodes it
    params = O.j2k_params(w, h, nc, levels, rev, cb, cb, 1.0 / 64, style=PAR)
    cs = O.t2_write(params, rs, data, passlen)
    assert enc.j2k() == cs
    from PIL import Image
    dec = np.array(Image.open(io.BytesIO(cs)))
    if rev:
        np.testing.assert_array_equal(dec, img)
    else:
        assert np.abs(dec.astype(int) - img.astype(int)).max() <= 1
    # the native decoder reads the segments back
    d = R.Decoder(ctx)
    d.decode(desc, res, blobs, w, h, nc, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64,
             style=PAR, seglen=seg)
    if rev:
        np.testing.assert_array_equal(d.image(), img)
    else:
        # same coefficients as the standard style, so the same reconstruction
        enc0 = R.Encoder(ctx)
        enc0.encode_u8(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64)
        d0 = R.Decoder(ctx)
        d0.decode(*enc0.codeblocks(), w, h, nc, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64)
        np.testing.assert_array_equal(d.image(), d0.image())
        d0.close()
        enc0.close()
    d.close()
    enc.close()


@pytest.mark.parametrize("PAR", STYLES)
@pytest.mark.parametrize("wpc", [0, 4])
def test_batch_parallel_style_vs_oracle(ctx, wpc, PAR):
    specs = [(256, 256, 1, True, 5), (300, 190, 3, False, 5), (97, 61, 1, False, 3), (64, 64, 3, False, 2),
             (33, 1, 1, False, 1), (129, 257, 1, True, 4)]
    imgs = []
    for i, (w, h, nc, rev, lv) in enumerate(specs):
        buf = synth.device_image(ctx, w, h, nc, 70 + i)
        imgs.append(dict(buf=buf, w=w, h=h, ncomp=nc, reversible=rev, levels=lv, delta=1.0 / 64,
                         host=synth.image(w, h, nc, 70 + i)))
    b = R.Batch(ctx)
    b.plan(imgs, cbw=64, cbh=64, use_graph=True, style=PAR, waves_per_cu=wpc)
    for rep in range(2):
        b.run()
        ctx.sync()
        for i, m in enumerate(imgs):
            desc, res, blobs = b.codeblocks(i)
            seg = b.pass_lengths(i)
            rs, data, q, passlen = O.encode_image(m["host"], levels=m["levels"], reversible=m["reversible"],
                                                  cbw=64, cbh=64, delta=1.0 / 64, style=PAR, return_passlen=True)
            np.testing.assert_array_equal(res["npasses"], rs[:, 1])
            for k in range(len(res)):
                np.testing.assert_array_equal(seg[k, :rs[k, 1]], passlen[k, :rs[k, 1]])
            assert blobs == data, "image %d rep %d" % (i, rep)
            params = O.j2k_params(m["w"], m["h"], m["ncomp"], m["levels"], m["reversible"], 64, 64, 1.0 / 64,
                                  style=PAR)
            assert b.j2k(i) == O.t2_write(params, rs, data, passlen)
    b.close()


def test_bypass_dense_block_vs_oracle(ctx):
    """Uniform noise with no transform: one dense 64x64 block whose eight
    bit-planes are near-random, so the raw passes (from pass 10 on) carry
    most of the codeword."""
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (64, 64), dtype=np.uint8)
    enc = R.Encoder(ctx)
    enc.encode_u8(img, levels=0, reversible=True, cbw=64, cbh=64, style=RG_STYLE_PARALLEL_BYPASS)
    desc, res, blobs = enc.codeblocks()
    seg = enc.pass_lengths()
    rs, data, q, passlen = O.encode_image(img, levels=0, reversible=True, cbw=64, cbh=64,
                                          style=RG_STYLE_PARALLEL_BYPASS, return_passlen=True)
    assert blobs == data
    np.testing.assert_array_equal(seg[0, :rs[0, 1]], passlen[0, :rs[0, 1]])
    enc.close()
