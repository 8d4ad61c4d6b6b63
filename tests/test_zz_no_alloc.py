"""GPU: rg_dwt_fwd_noalloc (static per-GPU scratch, no allocation) gives the
same coefficients as rg_dwt_fwd with a caller-allocated workspace, on two
calls in a row (the second smaller, over the first call's scratch contents),
for both filters and the FP16 inter-level storage; sizes beyond the static
buffer are refused on the host before any work is queued."""
import numpy as np
import pytest

from roger_amd import runtime as R
from roger_amd import synth
from roger_amd._lib import RG_COLOUR_ICT, RG_COLOUR_NONE, RG_COLOUR_RCT, RG_FILTER_53, RG_FILTER_97, RogerError

pytestmark = pytest.mark.gpu


def _planes(ctx, w, h, nc, filt, seed):
    img = synth.image(w, h, nc, seed)
    d_src = ctx.upload(img)
    d_pl = ctx.alloc(nc * w * h * 4)
    colour = RG_COLOUR_NONE if nc == 1 else (RG_COLOUR_RCT if filt == RG_FILTER_53 else RG_COLOUR_ICT)
    R.colour_fwd(ctx, d_src, w, h, nc, colour, filt == RG_FILTER_97, d_pl)
    return d_pl


@pytest.mark.parametrize("filt,ll_fp16", [(RG_FILTER_97, 0), (RG_FILTER_53, 0), (RG_FILTER_97, 1)])
def test_noalloc_matches_workspace_entry(ctx, filt, ll_fp16):
    for (w, h, nc, levels, seed) in [(1031, 777, 3, 6, 1), (300, 129, 1, 4, 2)]:
        d_pl = _planes(ctx, w, h, nc, filt, seed)
        inv = R.band_steps(levels, 1.0 / 64)[1] if filt == RG_FILTER_97 else None
        d_q1, d_q2 = ctx.alloc(nc * w * h * 4), ctx.alloc(nc * w * h * 4)
        d_s = ctx.alloc(R.scratch_bytes(w, h, nc, filt))
        R.dwt_fwd(ctx, filt, d_pl, w, h, nc, levels, inv, ll_fp16, d_q1, d_s)
        R.dwt_fwd_noalloc(ctx, filt, d_pl, w, h, nc, levels, inv, ll_fp16, d_q2)
        ref = d_q1.read(np.int32, (nc, h, w))
        np.testing.assert_array_equal(d_q2.read(np.int32, (nc, h, w)), ref, err_msg="%dx%dx%d" % (w, h, nc))
        for b in (d_pl, d_q1, d_q2, d_s):
            b.free()


def test_noalloc_refuses_oversize(ctx):
    # 8192x8192 x 3 planes needs ~256 MiB of LL scratch: refused before any
    # launch (the buffers are full size all the same)
    n = 8192 * 8192 * 3 * 4
    d_in, d_q = ctx.alloc(n), ctx.alloc(n)
    with pytest.raises(RogerError):
        R.dwt_fwd_noalloc(ctx, RG_FILTER_53, d_in, 8192, 8192, 3, 5, None, 0, d_q)
    d_in.free()
    d_q.free()
