"""GPU parity: colour and DWT kernels against the CPU oracle (bit-exact).

Sizes cover the odd / tiny shapes T.800 allows (1-sample signals, odd
lengths, multi-tile images) plus the benchmark's 4096x4096 shape via
size-independent properties (5/3 round trip)."""
import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth
from roger_amd._lib import RG_COLOUR_ICT, RG_COLOUR_NONE, RG_COLOUR_RCT, RG_FILTER_53, RG_FILTER_97

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (2, 3), (5, 1), (1, 7), (17, 5), (33, 31), (64, 64), (65, 129), (200, 77), (257, 300)]


def rand_img(w, h, nc, seed):
    rng = np.random.default_rng(seed)
    shape = (h, w) if nc == 1 else (h, w, nc)
    return rng.integers(0, 256, size=shape, dtype=np.uint8)


@pytest.mark.parametrize("w,h", SIZES + [(1000, 3)])
def test_colour_rct_ict(ctx, w, h):
    img = rand_img(w, h, 3, w * 7 + h)
    d_src = ctx.upload(img)
    d_pl = ctx.alloc(3 * w * h * 4)
    # RCT, int32 planes
    R.colour_fwd(ctx, d_src, w, h, 3, RG_COLOUR_RCT, 0, d_pl)
    got = d_pl.read(np.int32, (3, h, w))
    np.testing.assert_array_equal(got, O.rct_fwd(img))
    d_back = ctx.alloc(w * h * 3)
    R.colour_inv(ctx, d_pl, w, h, 3, RG_COLOUR_RCT, 0, d_back)
    np.testing.assert_array_equal(d_back.read(np.uint8, img.shape), img)
    # ICT, float planes, bit-exact with the oracle's FP32 op order
    R.colour_fwd(ctx, d_src, w, h, 3, RG_COLOUR_ICT, 1, d_pl)
    gotf = d_pl.read(np.float32, (3, h, w))
    np.testing.assert_array_equal(gotf.view(np.uint32), O.ict_fwd(img).view(np.uint32))
    R.colour_inv(ctx, d_pl, w, h, 3, RG_COLOUR_ICT, 1, d_back)
    np.testing.assert_array_equal(d_back.read(np.uint8, img.shape), O.ict_inv_f32(gotf))


@pytest.mark.parametrize("w,h", [(1, 1), (15, 3), (64, 64), (513, 65)])
def test_colour_gray_shift(ctx, w, h):
    img = rand_img(w, h, 1, w + h)
    d_src = ctx.upload(img)
    d_pl = ctx.alloc(w * h * 4)
    R.colour_fwd(ctx, d_src, w, h, 1, RG_COLOUR_NONE, 0, d_pl)
    np.testing.assert_array_equal(d_pl.read(np.int32, (h, w)), O.shift(img, np.int32))
    R.colour_fwd(ctx, d_src, w, h, 1, RG_COLOUR_NONE, 1, d_pl)
    np.testing.assert_array_equal(d_pl.read(np.float32, (h, w)), O.shift(img, np.float32))
    d_back = ctx.alloc(w * h)
    R.colour_inv(ctx, d_pl, w, h, 1, RG_COLOUR_NONE, 1, d_back)
    np.testing.assert_array_equal(d_back.read(np.uint8, (h, w)), img)


@pytest.mark.parametrize("w,h", [(7, 3), (257, 300)])
def test_colour_rgb_shift_only(ctx, w, h):
    """Three components without a colour transform: plain level shift."""
    img = rand_img(w, h, 3, w * 3 + h)
    d_src = ctx.upload(img)
    d_pl = ctx.alloc(3 * w * h * 4)
    want = np.moveaxis(img.astype(np.int32) - 128, 2, 0)
    R.colour_fwd(ctx, d_src, w, h, 3, RG_COLOUR_NONE, 0, d_pl)
    np.testing.assert_array_equal(d_pl.read(np.int32, (3, h, w)), want)
    R.colour_fwd(ctx, d_src, w, h, 3, RG_COLOUR_NONE, 1, d_pl)
    np.testing.assert_array_equal(d_pl.read(np.float32, (3, h, w)), want.astype(np.float32))


@pytest.mark.parametrize("w,h", [(1, 1), (5, 3), (17, 5), (64, 64), (257, 300)])
@pytest.mark.parametrize("nc", [1, 2, 3, 4])
def test_planar_split_i16(ctx, w, h, nc):
    """rg_planar_split_i16 (the reference's OCLRGBtoPlanar de-interleave,
    src/CL/oclplanar.cl:23-39): interleaved int16 channels -> one int16 plane
    each, over the whole int16 range."""
    rng = np.random.default_rng(w * 31 + h * 7 + nc)
    src = rng.integers(-32768, 32768, size=(h, w, nc), dtype=np.int16)
    d_src = ctx.upload(src)
    d_pl = ctx.alloc(nc * w * h * 2)
    R.planar_split_i16(ctx, d_src, w, h, nc, d_pl)
    np.testing.assert_array_equal(d_pl.read(np.int16, (nc, h, w)), np.moveaxis(src, 2, 0))


@pytest.mark.parametrize("w,h", [(1, 1), (3, 1), (5, 7), (33, 31), (1031, 65)])
def test_colour_inv_out_of_range(ctx, w, h):
    """Inverse colour on planes a decoder can produce (values beyond the
    8-bit range, fractional float samples): rounding and clamping equal the
    oracle's for ICT, RCT and the plain level shift."""
    rng = np.random.default_rng(w + 5 * h)
    f = (rng.uniform(-400.0, 400.0, size=(3, h, w)).astype(np.float32))
    f[:, ::3] = np.round(f[:, ::3]) + 0.5          # exact halves
    i = rng.integers(-700, 700, size=(3, h, w)).astype(np.int32)
    d_f, d_i = ctx.upload(f), ctx.upload(i)
    d_out = ctx.alloc(3 * w * h)
    R.colour_inv(ctx, d_f, w, h, 3, RG_COLOUR_ICT, 1, d_out)
    np.testing.assert_array_equal(d_out.read(np.uint8, (h, w, 3)), O.ict_inv_f32(f))
    R.colour_inv(ctx, d_i, w, h, 3, RG_COLOUR_RCT, 0, d_out)
    np.testing.assert_array_equal(d_out.read(np.uint8, (h, w, 3)), O.rct_inv(i))
    R.colour_inv(ctx, d_f, w, h, 3, RG_COLOUR_NONE, 1, d_out)
    np.testing.assert_array_equal(d_out.read(np.uint8, (h, w, 3)), np.stack([O.unshift(f[c]) for c in range(3)], 2))
    R.colour_inv(ctx, d_i, w, h, 1, RG_COLOUR_NONE, 0, d_out)
    np.testing.assert_array_equal(d_out.read(np.uint8, (h, w)), O.unshift(i[0]))


def _levels_for(w, h):
    return [0, 1, 3, 6]


@pytest.mark.parametrize("w,h", SIZES)
def test_dwt53_forward_inverse(ctx, w, h):
    rng = np.random.default_rng(w * 1000 + h)
    planes = rng.integers(-256, 256, size=(2, h, w), dtype=np.int32)
    for levels in _levels_for(w, h):
        d_in = ctx.upload(planes)
        d_q = ctx.alloc(planes.nbytes)
        d_s = ctx.alloc(R.scratch_bytes(w, h, 2, RG_FILTER_53))
        R.dwt_fwd(ctx, RG_FILTER_53, d_in, w, h, 2, levels, None, 0, d_q, d_s)
        got = d_q.read(np.int32, planes.shape)
        for c in range(2):
            np.testing.assert_array_equal(got[c], O.dwt53_fwd(planes[c], levels), err_msg="level %d" % levels)
        d_out = ctx.alloc(planes.nbytes)
        R.dwt_inv(ctx, RG_FILTER_53, d_q, w, h, 2, levels, None, d_out, d_s)
        np.testing.assert_array_equal(d_out.read(np.int32, planes.shape), planes)


@pytest.mark.parametrize("w,h", SIZES)
def test_dwt97_forward_quant_inverse(ctx, w, h):
    img = rand_img(w, h, 1, 3 * w + h)
    plane = O.shift(img, np.float32)
    for levels in _levels_for(w, h):
        for delta in (1.0 / 64, 0.7):
            step, inv, _, _ = R.band_steps(levels, delta)
            d_in = ctx.upload(plane)
            d_q = ctx.alloc(plane.nbytes)
            d_s = ctx.alloc(R.scratch_bytes(w, h, 1, RG_FILTER_97))
            R.dwt_fwd(ctx, RG_FILTER_97, d_in, w, h, 1, levels, inv, 0, d_q, d_s)
            got = d_q.read(np.int32, (h, w))
            ref = O.quantize(O.dwt97_fwd(plane, levels), levels, inv)
            np.testing.assert_array_equal(got, ref, err_msg="levels %d delta %g" % (levels, delta))
            # inverse with dequantisation, bit-exact with the oracle FP32 inverse
            d_out = ctx.alloc(plane.nbytes)
            R.dwt_inv(ctx, RG_FILTER_97, d_q, w, h, 1, levels, step, d_out, d_s)
            rec = d_out.read(np.float32, (h, w))
            ref_rec = O.dwt97_inv(O.dequantize(ref, levels, step), levels)
            np.testing.assert_array_equal(rec.view(np.uint32), ref_rec.view(np.uint32))


@pytest.mark.parametrize("w,h", [(1, 1), (33, 31), (64, 64), (200, 77), (257, 300)])
def test_dwt97_float_and_fp16(ctx, w, h):
    img = rand_img(w, h, 1, w + 5 * h)
    plane = O.shift(img, np.float32)
    levels = 4
    d_in = ctx.upload(plane)
    d_out = ctx.alloc(plane.nbytes)
    d_s = ctx.alloc(R.scratch_bytes(w, h, 1, RG_FILTER_97))
    R.dwt97_fwd_float(ctx, d_in, w, h, 1, levels, d_out, d_s)
    got = d_out.read(np.float32, (h, w))
    np.testing.assert_array_equal(got.view(np.uint32), O.dwt97_fwd(plane, levels).view(np.uint32))
    # float inverse (no dequantisation) reproduces the oracle FP32 inverse
    d_rec = ctx.alloc(plane.nbytes)
    R.dwt_inv(ctx, RG_FILTER_97, d_out, w, h, 1, levels, None, d_rec, d_s)
    np.testing.assert_array_equal(d_rec.read(np.float32, (h, w)).view(np.uint32),
                                  O.dwt97_inv(got, levels).view(np.uint32))
    # FP16 inter-level storage
    step, inv, _, _ = R.band_steps(levels, 1.0 / 64)
    d_q = ctx.alloc(plane.nbytes)
    R.dwt_fwd(ctx, RG_FILTER_97, d_in, w, h, 1, levels, inv, 1, d_q, d_s)
    ref = O.quantize(O.dwt97_fwd(plane, levels, ll_fp16=True), levels, inv)
    np.testing.assert_array_equal(d_q.read(np.int32, (h, w)), ref)


def test_dwt53_roundtrip_4096(ctx):
    """Full-size property: 5/3 forward then inverse is the identity at 4096x4096x3."""
    w = h = 4096
    img = synth.rgb(w, h, 1)
    d_src = ctx.upload(img)
    d_pl = ctx.alloc(3 * w * h * 4)
    R.colour_fwd(ctx, d_src, w, h, 3, RG_COLOUR_RCT, 0, d_pl)
    d_q = ctx.alloc(3 * w * h * 4)
    d_s = ctx.alloc(R.scratch_bytes(w, h, 3, RG_FILTER_53))
    R.dwt_fwd(ctx, RG_FILTER_53, d_pl, w, h, 3, 6, None, 0, d_q, d_s)
    R.dwt_inv(ctx, RG_FILTER_53, d_q, w, h, 3, 6, None, d_pl, d_s)
    d_back = ctx.alloc(w * h * 3)
    R.colour_inv(ctx, d_pl, w, h, 3, RG_COLOUR_RCT, 0, d_back)
    np.testing.assert_array_equal(d_back.read(np.uint8, img.shape), img)
