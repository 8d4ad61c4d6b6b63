"""GPU: rate control (PCRD-opt) against the oracle -- the same truncation
point in every code-block for the same weights and targets (per-pass
distortions, convex hulls and the slope threshold all agree), the same cut
codewords and codestream bytes, the kept bytes within the target, OpenJPEG
decoding every cut with quality rising with the target, and recutting from
the full codewords on repeated calls."""
import io

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth
from roger_amd._lib import RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS, RogerError

pytestmark = pytest.mark.gpu

STYLES = [RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS]


def _psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 99.0 if mse == 0 else 10 * np.log10(255.0 ** 2 / mse)


def _oracle_dists(desc, q, style, passlen, rs):
    d_all = np.zeros((len(desc), O.MAX_PASSES), np.int64)
    for i, d in enumerate(desc):
        blk = q[d["comp"], d["y0"]:d["y0"] + d["h"], d["x0"]:d["x0"] + d["w"]]
        _, npass, nbp, pl, dist = O.t1_encode_rd(blk, int(d["orient"]), style)
        assert npass == rs[i, 1]
        np.testing.assert_array_equal(pl, passlen[i, :npass])
        d_all[i, :npass] = dist
    return d_all


@pytest.mark.parametrize("PAR", STYLES)
@pytest.mark.parametrize("w,h,nc,rev,levels", [
    (512, 512, 1, True, 5),
    (300, 257, 3, False, 4),
    (128, 96, 1, False, 3),
])
def test_rate_control_vs_oracle(ctx, w, h, nc, rev, levels, PAR):
    from PIL import Image
    img = synth.image(w, h, nc, 33)
    enc = R.Encoder(ctx)
    enc.enable_rate_control()
    enc.encode_u8(img, levels=levels, reversible=rev, cbw=64, cbh=64, delta=1.0 / 64, style=PAR)
    desc, res, blobs = enc.codeblocks()
    rs, data, q, passlen = O.encode_image(img, levels=levels, reversible=rev, cbw=64, cbh=64, delta=1.0 / 64,
                                          style=PAR, return_passlen=True)
    assert blobs == data
    dists = _oracle_dists(desc, q, PAR, passlen, rs)
    weights = np.random.default_rng(8).uniform(0.2, 3.0, len(desc))
    full = int(rs[:, 0].sum())
    params = O.j2k_params(w, h, nc, levels, rev, 64, 64, 1.0 / 64, style=PAR)
    last = None
    for target in [0, full // 40, full // 10, full // 4, full // 2, full - 1, full, full + 999, full // 10]:
        kept = enc.rate_control(target, weights)
        nt, rt, total = O.rate_control(rs[:, 1], passlen, dists, weights, target)
        assert kept == total <= target
        _, res2, blobs2 = enc.codeblocks()
        np.testing.assert_array_equal(res2["npasses"], nt, err_msg="target %d" % target)
        np.testing.assert_array_equal(res2["len"], rt)
        cut = [data[i][:rt[i]] for i in range(len(data))]
        assert blobs2 == cut
        rs2 = rs.copy()
        rs2[:, 0], rs2[:, 1] = rt, nt
        cs = O.t2_write(params, rs2, cut, passlen)
        assert enc.j2k() == cs
        if target >= full:
            # every block at its last hull point (trailing passes that add no
            # distortion reduction are never worth their bytes)
            last_pt = [int(O.rd_hull(passlen[i, :rs[i, 1]], dists[i, :rs[i, 1]], weights[i])[0][-1])
                       if rs[i, 1] else 0 for i in range(len(rs))]
            np.testing.assert_array_equal(nt, last_pt)
        if target > full // 40 and target <= full:
            dec = np.array(Image.open(io.BytesIO(cs)))
            p = _psnr(dec, img)
            if last is not None and target > last[0]:
                assert p >= last[1] - 0.05, (target, p, last)
            last = (target, p)
    enc.close()


def test_rate_control_default_weights(ctx):
    """Band weights from the quantisation steps and synthesis energies:
    targets met, the codestream decodes, and at a given size the weighted
    cut is at least as good as weighting every block alike."""
    from PIL import Image
    img = synth.image(384, 320, 3, 44)
    enc = R.Encoder(ctx)
    enc.enable_rate_control()
    enc.encode_u8(img, levels=5, reversible=False, cbw=64, cbh=64, delta=1.0 / 64, style=RG_STYLE_PARALLEL_BYPASS)
    desc, res, _ = enc.codeblocks()
    full = int(res["len"].sum())
    prev = -1.0
    for frac in (0.03, 0.1, 0.3):
        target = int(full * frac)
        kept = enc.rate_control(target)
        assert 0.9 * target <= kept <= target
        dec = np.array(Image.open(io.BytesIO(enc.j2k())))
        p = _psnr(dec, img)
        assert p > prev
        prev = p
        enc.rate_control(target, np.ones(len(desc)))
        p1 = _psnr(np.array(Image.open(io.BytesIO(enc.j2k()))), img)
        assert p >= p1 - 0.1, (frac, p, p1)
    enc.close()


def test_rate_control_errors(ctx):
    img = synth.image(64, 64, 1, 3)
    enc = R.Encoder(ctx)
    enc.encode_u8(img, levels=2, reversible=True, style=RG_STYLE_PARALLEL)
    with pytest.raises(RogerError):
        enc.rate_control(1000)          # not enabled before tier-1
    enc.enable_rate_control()
    enc.encode_u8(img, levels=2, reversible=True, style=0)
    with pytest.raises(RogerError):
        enc.rate_control(1000)          # standard style: passes are not terminated
    enc.encode_u8(img, levels=2, reversible=True, style=RG_STYLE_PARALLEL)
    with pytest.raises(RogerError):
        enc.rate_control(-1)
    with pytest.raises(RogerError):
        enc.rate_control(1000, np.zeros(enc.codeblocks()[0].shape[0]))
    with pytest.raises(ValueError):
        enc.rate_control(1000, np.ones(3))   # one weight per code-block
    kept = enc.rate_control(1 << 30)
    assert 0 < kept == int(enc.codeblocks()[1]["len"].sum())
    enc.close()
    # enabled only after the encode: no masks were recorded for these codewords
    # (the check is on the host, before any launch)
    enc = R.Encoder(ctx)
    enc.encode_u8(img, levels=2, reversible=True, style=RG_STYLE_PARALLEL)
    enc.enable_rate_control()
    with pytest.raises(RogerError):
        enc.rate_control(1000)
    # masks recorded for a small image, then off for a larger encode, then on
    # again: the stale masks must not be used
    enc.encode_u8(img, levels=2, reversible=True, style=RG_STYLE_PARALLEL)
    enc.rate_control(1000)
    enc.enable_rate_control(False)
    big = synth.image(256, 192, 1, 4)
    enc.encode_u8(big, levels=3, reversible=True, style=RG_STYLE_PARALLEL)
    enc.enable_rate_control(True)
    with pytest.raises(RogerError):
        enc.rate_control(1000)
    enc.close()


@pytest.mark.parametrize("PAR", STYLES)
def test_quality_layers_vs_oracle(ctx, PAR):
    """Three quality layers: every block's passes per layer equal the oracle's
    cut at each layer's target, the codestream equals the oracle's
    multi-layer writer, and OpenJPEG stopping after layer l decodes exactly
    the single-layer cut at layer l's target."""
    from PIL import Image
    from test_rate import decode_layers
    w, h, nc, levels = 300, 257, 3, 4
    img = synth.image(w, h, nc, 61)
    enc = R.Encoder(ctx)
    enc.enable_rate_control()
    enc.encode_u8(img, levels=levels, reversible=False, cbw=64, cbh=64, delta=1.0 / 64, style=PAR)
    desc, res, blobs = enc.codeblocks()
    rs, data, q, passlen = O.encode_image(img, levels=levels, reversible=False, cbw=64, cbh=64, delta=1.0 / 64,
                                          style=PAR, return_passlen=True)
    dists = _oracle_dists(desc, q, PAR, passlen, rs)
    weights = np.random.default_rng(9).uniform(0.2, 3.0, len(desc))
    full = int(rs[:, 0].sum())
    targets = [full // 30, full // 8, full // 3]
    kept = enc.rate_control_layers(targets, weights)
    cuts = [O.rate_control(rs[:, 1], passlen, dists, weights, t) for t in targets]
    layers = np.stack([c[0] for c in cuts], 1)
    nt, rt, total = cuts[-1]
    assert kept == total
    _, res2, blobs2 = enc.codeblocks()
    np.testing.assert_array_equal(res2["npasses"], nt)
    np.testing.assert_array_equal(res2["len"], rt)
    params = O.j2k_params(w, h, nc, levels, False, 64, 64, 1.0 / 64, style=PAR)
    r3 = rs.copy()
    r3[:, 0], r3[:, 1] = rt, nt
    cs = O.t2_write(params, r3, [data[i][:rt[i]] for i in range(len(data))], passlen, layers=layers)
    assert enc.j2k() == cs
    for l, (nl, rl, _) in enumerate(cuts):
        r1 = rs.copy()
        r1[:, 0], r1[:, 1] = rl, nl
        single = O.t2_write(params, r1, [data[i][:rl[i]] for i in range(len(data))], passlen)
        np.testing.assert_array_equal(decode_layers(cs, l + 1), np.array(Image.open(io.BytesIO(single))),
                                      err_msg="layer %d" % l)
    # a single cut afterwards, and a new encode, go back to one layer
    enc.rate_control(targets[1], weights)
    r1 = rs.copy()
    r1[:, 0], r1[:, 1] = cuts[1][1], cuts[1][0]
    assert enc.j2k() == O.t2_write(params, r1, [data[i][:cuts[1][1][i]] for i in range(len(data))], passlen)
    with pytest.raises(RogerError):
        enc.rate_control_layers([targets[1], targets[0]], weights)
    enc.close()


@pytest.mark.parametrize("shape", [(64, 64, 1), (33, 1, 1), (130, 70, 3)])
def test_rate_control_flat_and_tiny_images(ctx, shape):
    """Blocks with no bit-planes (a flat image has only its LL block coded),
    single-row images: the cut matches the oracle for every target."""
    w, h, nc = shape
    img = np.full((h, w, nc) if nc > 1 else (h, w), 77, np.uint8)
    img.reshape(-1)[::97] = 200
    enc = R.Encoder(ctx)
    enc.enable_rate_control()
    enc.encode_u8(img, levels=3 if h > 1 else 1, reversible=True, cbw=64, cbh=64, style=RG_STYLE_PARALLEL)
    desc, res, blobs = enc.codeblocks()
    rs, data, q, passlen = O.encode_image(img, levels=3 if h > 1 else 1, reversible=True, cbw=64, cbh=64,
                                          style=RG_STYLE_PARALLEL, return_passlen=True)
    dists = _oracle_dists(desc, q, RG_STYLE_PARALLEL, passlen, rs)
    weights = np.ones(len(desc))
    full = int(rs[:, 0].sum())
    for target in (0, 1, full // 3, full):
        kept = enc.rate_control(target, weights)
        nt, rt, total = O.rate_control(rs[:, 1], passlen, dists, weights, target)
        assert kept == total <= target
        np.testing.assert_array_equal(enc.codeblocks()[1]["npasses"], nt)
    enc.close()
