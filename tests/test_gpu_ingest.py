"""GPU: double-buffered host ingest (rg_encoder_encode_u8 with a host source).

Back-to-back encodes of different host images, pageable and pinned, with no
wait between them: every upload runs on the context's upload stream into the
input buffer the encode before last used, and the result read after the last
call equals the oracle's encode of the last image."""
import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth

pytestmark = pytest.mark.gpu


def _check(enc, img, levels, rev, style=0):
    desc, res, blobs = enc.codeblocks()
    rs, data, q = O.encode_image(img, levels=levels, reversible=rev, style=style)
    np.testing.assert_array_equal(enc.coefficients(), q)
    np.testing.assert_array_equal(res["npasses"], rs[:, 1])
    assert blobs == data


@pytest.mark.parametrize("style", [0, O.STYLE_PARALLEL_BYPASS])
def test_back_to_back_host_encodes(ctx, style):
    imgs = [synth.image(300, 200, 3, 31), synth.image(512, 384, 3, 32), synth.image(257, 129, 1, 33),
            synth.image(300, 200, 3, 34)]
    enc = R.Encoder(ctx)
    for k in range(len(imgs)):
        for img in imgs[:k + 1]:
            enc.encode_u8(img, levels=4, reversible=False, style=style)
        _check(enc, imgs[k], 4, False, style)
    enc.close()


def test_pinned_source_reused_after_completion(ctx):
    """A pinned host buffer refilled with another image after each encode has
    completed; two encoders on one context share its upload stream."""
    w, h = 640, 480
    pin = ctx.host_alloc(w * h * 3)
    view = pin.array(np.uint8, (h, w, 3))
    e1, e2 = R.Encoder(ctx), R.Encoder(ctx)
    for seed in (41, 42, 43):
        img = synth.image(w, h, 3, seed)
        view[...] = img
        e1.encode_u8(view, levels=5, reversible=True)
        e2.encode_u8(view, levels=5, reversible=False)
        ctx.sync()
        _check(e1, img, 5, True)
        _check(e2, img, 5, False)
    e1.close()
    e2.close()
    pin.free()
