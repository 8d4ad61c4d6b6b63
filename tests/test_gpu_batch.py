"""GPU: the batched encode path (one launch per stage over many images of
uneven sizes, HIP-graph replay, tier-1 job list sorted largest first, per-block
grid or persistent work-queue grid) produces, for
every image, exactly the code-blocks the CPU oracle produces; and the device
image generator reproduces the numpy generator bit-exactly."""
import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,nc,seed", [(64, 48, 1, 0), (333, 97, 3, 1), (1000, 17, 1, 7), (257, 256, 3, 1000)])
def test_device_synth_matches_numpy(ctx, w, h, nc, seed):
    buf = synth.device_image(ctx, w, h, nc, seed)
    shape = (h, w) if nc == 1 else (h, w, 3)
    np.testing.assert_array_equal(buf.read(np.uint8, shape), synth.image(w, h, nc, seed))


SPECS = [  # w, h, ncomp, reversible, levels
    (256, 256, 1, True, 5), (300, 190, 3, False, 5), (97, 61, 1, False, 3), (512, 130, 3, True, 5),
    (64, 64, 3, False, 2), (129, 257, 1, True, 4), (33, 1, 1, False, 1), (400, 300, 3, False, 5),
]


@pytest.mark.parametrize("use_graph,wpc", [(True, 0), (False, 0), (True, 4), (False, 2)])
@pytest.mark.parametrize("cb", [64, 32])
def test_batch_matches_oracle(ctx, use_graph, wpc, cb):
    """wpc > 0: tier-1 as a persistent grid pulling from the device queue."""
    imgs = []
    for i, (w, h, nc, rev, lv) in enumerate(SPECS):
        buf = synth.device_image(ctx, w, h, nc, 50 + i)
        imgs.append(dict(buf=buf, w=w, h=h, ncomp=nc, reversible=rev, levels=lv, delta=1.0 / 64,
                         host=synth.image(w, h, nc, 50 + i)))
    b = R.Batch(ctx)
    b.plan(imgs, cbw=cb, cbh=cb, use_graph=use_graph, waves_per_cu=wpc)
    for rep in range(2):  # the second run replays the plan (counter reset inside)
        b.run()
        ctx.sync()
        for i, m in enumerate(imgs):
            desc, res, blobs = b.codeblocks(i)
            rs, data, q = O.encode_image(m["host"], levels=m["levels"], reversible=m["reversible"], cbw=cb, cbh=cb,
                                         delta=1.0 / 64)
            np.testing.assert_array_equal(b.coefficients(i), q, err_msg="image %d rep %d" % (i, rep))
            np.testing.assert_array_equal(res["nbp"], rs[:, 2])
            np.testing.assert_array_equal(res["npasses"], rs[:, 1])
            assert blobs == data, "image %d rep %d" % (i, rep)
    b.close()


def test_batch_refuses_zero_levels(ctx):
    """The batch has no level-0 (copy / quantise) job, so an image with 0 DWT
    levels beside others with levels is refused on the host at plan time
    rather than tier-1 coding an unwritten coefficient plane."""
    from roger_amd._lib import RogerError
    imgs = [dict(buf=synth.device_image(ctx, w, h, 1, 60 + i), w=w, h=h, ncomp=1, reversible=True, levels=lv)
            for i, (w, h, lv) in enumerate([(96, 80, 3), (64, 64, 0)])]
    b = R.Batch(ctx)
    with pytest.raises(RogerError):
        b.plan(imgs)
    b.close()
