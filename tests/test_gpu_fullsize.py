"""GPU parity at the benchmark workloads' own sizes, against the CPU oracle.

* 4096x4096 RGB, ICT + 9/7 FP32 with 6 levels, step 1/64, 64x64 code-blocks
  (the headline encode) in the RESET|TERMALL|BYPASS and the standard
  code-block styles: quantisation indices, every block's codeword, pass and
  bit-plane counts, per-pass segment lengths and the whole tier-2 codestream
  bit-exact; the native decode of those codewords equal to the oracle's FP32
  inverse path.  At this size the tier-1 chain sort runs many workgroups
  (12,288 blocks, ~300K (block, pass) chains) and the phased tier-1 overlaps
  its streams, paths the small parity cases never reach.
* The 256-image mixed batch (sides log-uniform 256..4096, half grayscale /
  half RGB, half 5/3 / half 9/7, 5 levels) planned and run in one batch:
  every image's codewords equal the oracle's encode of that image.

The oracle runs on the host's cores (OpenMP over code-blocks, or one image per
thread for the batch)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth

pytestmark = pytest.mark.gpu

NTHREADS = max(1, min(16, os.cpu_count() or 1))
W = H = 4096
LEVELS, DELTA, CB = 6, 1.0 / 64, 64


@pytest.fixture(scope="module")
def c1(ctx):
    """The headline image (device-generated, read back) and its oracle
    encodes in both styles."""
    d_src = synth.device_image(ctx, W, H, 3, 1)
    img = d_src.read(np.uint8, (H, W, 3))
    print("4096x4096 RGB image generated", flush=True)
    yield d_src, img
    d_src.free()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("style", [O.STYLE_PARALLEL_BYPASS, 0])
def test_c1_encode_decode_4096_rgb(ctx, c1, style):
    d_src, img = c1
    print("C1 style %d: oracle encode on %d threads" % (style, NTHREADS), flush=True)
    rs, data, q, passlen = O.encode_image(img, levels=LEVELS, reversible=False, cbw=CB, cbh=CB, delta=DELTA,
                                          nthreads=NTHREADS, style=style, return_passlen=True)
    enc = R.Encoder(ctx)
    enc.encode_u8(None, levels=LEVELS, reversible=False, cbw=CB, cbh=CB, delta=DELTA, d_src=d_src, w=W, h=H,
                  ncomp=3, style=style)
    np.testing.assert_array_equal(enc.coefficients(), q)
    desc, res, blobs = enc.codeblocks()
    assert len(blobs) == len(data) == 3 * 4096
    np.testing.assert_array_equal(res["npasses"], rs[:, 1])
    np.testing.assert_array_equal(res["nbp"], rs[:, 2])
    bad = [i for i in range(len(data)) if blobs[i] != data[i]]
    assert not bad, "%d code-blocks differ, first %d" % (len(bad), bad[0])
    if style:
        np.testing.assert_array_equal(enc.pass_lengths(), passlen)
    cs = enc.j2k()
    params = O.j2k_params(W, H, 3, LEVELS, False, CB, CB, DELTA, style=style)
    assert cs == O.t2_write(params, rs, data, passlen if style else None)
    print("C1 style %d: encode bit-exact, %d codestream bytes" % (style, len(cs)), flush=True)
    # native decode of the same codewords vs the oracle's FP32 inverse path
    dec = R.Decoder(ctx)
    dec.decode(desc, res, blobs, W, H, 3, levels=LEVELS, reversible=False, cbw=CB, cbh=CB, delta=DELTA,
               style=style, seglen=enc.pass_lengths() if style else None)
    out = dec.image()
    steps = np.array([s[2] for s in O.band_steps(W, H, LEVELS, DELTA)], np.float32)
    planes = np.stack([O.dwt97_inv(O.dequantize(q[c], LEVELS, steps), LEVELS) for c in range(3)])
    np.testing.assert_array_equal(out, O.ict_inv_f32(planes))
    assert np.abs(out.astype(int) - img.astype(int)).max() <= 1
    enc.close()
    dec.close()


@pytest.mark.timeout(900)
def test_c2_full_batch_vs_oracle(ctx):
    """The benchmark's whole 256-image batch in the headline code-block style
    (64x64 blocks, per-block tier-1 grid, graph replay)."""
    specs = synth.batch_specs(256)
    images = [dict(s, buf=synth.device_image(ctx, s["w"], s["h"], s["ncomp"], s["seed"])) for s in specs]
    print("C2: %d images, %.1f Mpix" % (len(images), sum(m["w"] * m["h"] for m in images) / 1e6), flush=True)
    b = R.Batch(ctx)
    b.plan(images, cbw=CB, cbh=CB, use_graph=True, style=O.STYLE_PARALLEL_BYPASS)
    b.run()
    b.run()
    ctx.sync()

    def oracle(i, img):
        m = images[i]
        rs, data, _ = O.encode_image(img, levels=m["levels"], reversible=m["reversible"], cbw=CB, cbh=CB,
                                     delta=m["delta"], nthreads=1, style=O.STYLE_PARALLEL_BYPASS)
        return i, rs, data

    # the device results are read here, the oracle encodes run in a pool
    # (ctypes releases the GIL), a bounded number at a time
    got = {}
    with ThreadPoolExecutor(NTHREADS) as ex:
        futs = []
        for i, m in enumerate(images):
            shape = (m["h"], m["w"]) if m["ncomp"] == 1 else (m["h"], m["w"], 3)
            img = m["buf"].read(np.uint8, shape)
            desc, res, blobs = b.codeblocks(i)
            got[i] = (res["npasses"].copy(), res["nbp"].copy(), blobs)
            futs.append(ex.submit(oracle, i, img))
            if len(futs) >= 2 * NTHREADS or i == len(images) - 1:
                for f in futs:
                    k, rs, data = f.result()
                    npasses, nbp, blobs = got.pop(k)
                    np.testing.assert_array_equal(npasses, rs[:, 1], err_msg="image %d" % k)
                    np.testing.assert_array_equal(nbp, rs[:, 2], err_msg="image %d" % k)
                    assert blobs == data, "image %d: codewords differ" % k
                futs = []
            m["buf"].free()
    assert not got
    print("C2: all 256 images bit-exact", flush=True)
    b.close()
