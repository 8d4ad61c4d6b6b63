"""GPU: whole decodes of .j2k codestreams through the tier-2 reader
(rg_decoder_decode_j2k) -- the library's own codestreams in every style and
after rate control with quality layers (the same image as decoding its
code-blocks directly), and OpenJPEG-encoded ones (Pillow): lossless exactly,
lossy close to OpenJPEG's own decode."""
import io

import numpy as np
import pytest

from roger_amd import runtime as R
from roger_amd import synth
from roger_amd._lib import RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("style", [0, RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS])
@pytest.mark.parametrize("w,h,nc,rev,levels", [(300, 257, 3, True, 5), (257, 190, 3, False, 6), (96, 64, 1, False, 2)])
def test_decode_own_codestreams(ctx, w, h, nc, rev, levels, style):
    img = synth.image(w, h, nc, 14)
    enc = R.Encoder(ctx)
    enc.encode_u8(img, levels=levels, reversible=rev, cbw=64, cbh=64, delta=1.0 / 64, style=style)
    cs = enc.j2k()
    d = R.Decoder(ctx)
    info = d.decode_j2k(cs)
    assert (info["w"], info["h"], info["ncomp"], info["levels"]) == (w, h, nc, levels)
    got = d.image()
    d2 = R.Decoder(ctx)
    seg = enc.pass_lengths() if style else None
    d2.decode(*enc.codeblocks(), w, h, nc, levels=levels, reversible=rev, cbw=64, cbh=64, delta=1.0 / 64,
              style=style, seglen=seg)
    np.testing.assert_array_equal(got, d2.image())
    if rev:
        np.testing.assert_array_equal(got, img)
    d.close()
    d2.close()
    enc.close()


def test_decode_rate_controlled_layers(ctx):
    """Three quality layers from rate control decode (all layers) to the same
    image as the last layer's code-blocks decoded directly."""
    img = synth.image(320, 240, 3, 15)
    enc = R.Encoder(ctx)
    enc.enable_rate_control()
    enc.encode_u8(img, levels=5, reversible=False, style=RG_STYLE_PARALLEL_BYPASS)
    full = int(enc.codeblocks()[1]["len"].sum())
    enc.rate_control_layers([full // 40, full // 10, full // 4])
    d = R.Decoder(ctx)
    info = d.decode_j2k(enc.j2k())
    assert info["nlayers"] == 3
    d2 = R.Decoder(ctx)
    d2.decode(*enc.codeblocks(), 320, 240, 3, levels=5, reversible=False, style=RG_STYLE_PARALLEL_BYPASS,
              seglen=enc.pass_lengths())
    np.testing.assert_array_equal(d.image(), d2.image())
    d.close()
    d2.close()
    enc.close()


@pytest.mark.parametrize("irreversible", [False, True])
@pytest.mark.parametrize("nc", [1, 3])
def test_decode_openjpeg_codestreams(ctx, nc, irreversible):
    from PIL import Image
    img = synth.image(211, 163, nc, 16)
    buf = io.BytesIO()
    Image.fromarray(img).save(buf, "JPEG2000", irreversible=irreversible, num_resolutions=6, codeblock_size=(64, 64),
                              progression="LRCP", no_jp2=True)
    cs = buf.getvalue()
    d = R.Decoder(ctx)
    d.decode_j2k(cs)
    got = d.image().astype(int)
    ref = np.array(Image.open(io.BytesIO(cs))).astype(int)
    if not irreversible:
        np.testing.assert_array_equal(got, img)
    else:
        # the two decoders' arithmetic differs (FP32 lifting order,
        # rounding): measured at most 1 grey level, on 0.1% of the pixels
        diff = np.abs(got - ref)
        print("openjpeg 9/7: max %d mean %.4f" % (diff.max(), diff.mean()))
        assert diff.max() <= 1 and diff.mean() < 0.01
    d.close()
