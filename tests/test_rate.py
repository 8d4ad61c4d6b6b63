"""CPU: the oracle's rate control (PCRD-opt) -- per-pass distortion
reductions, convex hulls, threshold selection -- pinned by what the
mathematics fixes: the reductions of all passes add up to the block's
energy, each bit-plane's to its closed form, the significance-propagation
share to what the oracle decoder shows became significant in that pass; the
hull against brute force; the threshold as the smallest one that fits,
against brute force over every candidate."""
import numpy as np
import pytest

from oracle import oracle as O

STYLES = [0, O.STYLE_PARALLEL, O.STYLE_PARALLEL_BYPASS]


def _blocks(rng):
    out = []
    for k, scale in enumerate([0.7, 3, 40, 900, 25000]):
        b = np.round(rng.laplace(0, scale, (64, 64))).astype(np.int32)
        out.append((b, k % 4))
    sp = np.round(rng.laplace(0, 60, (37, 13))).astype(np.int32)
    sp[rng.random(sp.shape) < 0.9] = 0
    out.append((sp, 1))
    one = np.zeros((5, 7), np.int32)
    one[2, 3] = -77
    out.append((one, 3))
    return out


def _plane_sums(mag, nbp):
    """Closed form per bit-plane: significance events (msb == p) and
    refinements (msb > p), with python integers."""
    m = [int(x) for x in mag.ravel()]
    out = {}
    for p in range(nbp):
        sig = sum(O.t1_dsig(x, p) for x in m if x >> p == 1)
        ref = sum(O.t1_dref(x, p) for x in m if x >> (p + 1))
        out[p] = (sig, ref)
    return out


def test_distortion_units_match_the_reconstruction_error():
    """dsig / dref: squared-error drops, in doubled index units, for a
    decoder that reconstructs the known bits plus half of what is left --
    checked against that error computed from the interval of magnitudes the
    known bits allow (its midpoint), for every small magnitude and plane."""
    for m in range(1, 300):
        msb = m.bit_length() - 1
        for p in range(msb + 1):
            def err2(known_from):
                # known bits: those >= known_from; none known -> reconstruct 0
                if known_from is None:
                    return (2 * m) ** 2
                lo = (m >> known_from) << known_from
                mid2 = 2 * lo + (1 << known_from) if known_from > 0 else 2 * m
                return (2 * m - mid2) ** 2
            if p == msb:
                assert O.t1_dsig(m, p) == err2(None) - err2(p)
            else:
                assert O.t1_dref(m, p) == err2(p + 1) - err2(p)


@pytest.mark.parametrize("style", STYLES)
def test_pass_distortion_adds_up(style):
    rng = np.random.default_rng(17 + style)
    for b, orient in _blocks(rng):
        data, npass, nbp, passlen, dist = O.t1_encode_rd(b, orient, style)
        mag = np.abs(b).astype(np.int64)
        # all passes: the whole energy (every sample ends exactly reconstructed)
        assert int(dist.sum()) == int((4 * mag * mag).sum())
        if nbp == 0:
            continue
        sums = _plane_sums(mag, nbp)
        # top plane: one cleanup pass carrying every significance event
        assert int(dist[0]) == sums[nbp - 1][0]
        for p in range(nbp - 2, -1, -1):
            k = 3 * (nbp - 1 - p) - 2          # SPP, MRP, CUP of plane p
            sig, ref = sums[p]
            assert int(dist[k + 1]) == ref
            assert int(dist[k] + dist[k + 2]) == sig


@pytest.mark.parametrize("style", STYLES)
def test_spp_share_matches_the_decoder(style):
    """The significance-propagation pass's reduction is the sum over exactly
    the samples that the oracle decoder shows becoming significant in it."""
    rng = np.random.default_rng(5)
    b = np.round(rng.laplace(0, 200, (64, 64))).astype(np.int32)
    data, npass, nbp, passlen, dist = O.t1_encode_rd(b, 2, style)
    mag = np.abs(b)

    def decoded(n):
        if style:
            nb = int(passlen[:n].sum())
            return O.t1_decode(data[:nb], 64, 64, 2, nbp, n, style=style, passlen=passlen[:n])
        return O.t1_decode(data, 64, 64, 2, nbp, n)

    for p in range(nbp - 2, max(nbp - 7, -1), -1):
        k = 3 * (nbp - 1 - p) - 2
        before, after = decoded(k) != 0, decoded(k + 1) != 0
        new = after & ~before
        assert int(dist[k]) == sum(O.t1_dsig(int(x), p) for x in mag[new])


def _brute_hull(passlen, dist, w):
    """Upper concave envelope of the points (R_n, w * D_n), n = 0..N, by
    brute force: a point is on the hull if no chord between two other points
    lies strictly above it and it adds distortion."""
    R = np.concatenate([[0], np.cumsum(passlen)]).astype(np.int64)
    D = np.concatenate([[0], np.cumsum(dist)]).astype(np.int64)
    pts = []
    best = 0
    for n in range(1, len(R)):
        if D[n] <= best:
            continue
        best = D[n]
        pts.append(n)
    # remove points under a chord of two others (exact integer cross products)
    keep = []
    cand = [0] + pts
    for idx, n in enumerate(cand[1:], 1):
        under = False
        for i in cand[:idx]:
            for j in cand[idx + 1:]:
                # (R_n, D_n) on or under the chord i -> j: the chord's D at R_n >= D_n
                if (D[j] - D[i]) * (R[n] - R[i]) >= (D[n] - D[i]) * (R[j] - R[i]) and R[i] <= R[n] <= R[j]:
                    under = True
        if not under:
            keep.append(n)
    return keep


@pytest.mark.parametrize("style", [O.STYLE_PARALLEL, O.STYLE_PARALLEL_BYPASS])
def test_hull_against_brute_force(style):
    rng = np.random.default_rng(29)
    for b, orient in _blocks(rng):
        data, npass, nbp, passlen, dist = O.t1_encode_rd(b, orient, style)
        if npass == 0:
            continue
        hn, hr, hs = O.rd_hull(passlen, dist, 1.0)
        assert list(hn) == _brute_hull(passlen, dist, 1.0)
        assert (np.diff(hs) < 0).all()                       # slopes strictly decrease
        assert hr[-1] == np.cumsum(passlen)[hn[-1] - 1]


def test_threshold_is_the_smallest_that_fits():
    """Every candidate threshold (each hull slope, and just above it) by
    brute force: the chosen truncation is that of the smallest threshold
    whose total stays within the target, and totals never exceed it."""
    rng = np.random.default_rng(41)
    blocks = _blocks(rng) + _blocks(np.random.default_rng(42))
    npasses, pls, dists, weights, hulls = [], [], [], [], []
    for i, (b, orient) in enumerate(blocks):
        data, npass, nbp, passlen, dist = O.t1_encode_rd(b, orient, O.STYLE_PARALLEL_BYPASS)
        row_p = np.zeros(O.MAX_PASSES, np.int32)
        row_d = np.zeros(O.MAX_PASSES, np.int64)
        row_p[:npass] = passlen
        row_d[:npass] = dist
        npasses.append(npass)
        pls.append(row_p)
        dists.append(row_d)
        weights.append(0.5 + (i % 5) * 0.37)
        hulls.append(O.rd_hull(passlen, dist, weights[-1]))
    pls, dists = np.array(pls), np.array(dists)
    full = int(sum(pl.sum() for pl in pls))
    keys = sorted({float(s) for h in hulls for s in h[2]})

    def at(lam):
        nt, rt = [], []
        for hn, hr, hs in hulls:
            k = int((hs >= lam).sum())
            nt.append(int(hn[k - 1]) if k else 0)
            rt.append(int(hr[k - 1]) if k else 0)
        return nt, rt

    prev = None
    for target in [-1, 0, 50, 300, 1000, 3000, full // 2, full - 1, full, full + 100]:
        nt, rt, total = O.rate_control(npasses, pls, dists, weights, target)
        assert total == int(rt.sum()) and total <= max(target, 0)
        # brute force: the lowest candidate slope whose truncation fits
        best = None
        for lam in [np.inf] + keys[::-1]:
            bn, br = at(lam)
            if sum(br) <= target:
                best = (bn, br)
        if target < 0:
            best = ([0] * len(hulls), [0] * len(hulls))
        assert list(nt) == best[0] and list(rt) == best[1], target
        if prev is not None:
            assert (nt >= prev).all()                        # monotone in the target
        prev = nt


def _oracle_image_rd(img, levels, rev, style):
    """Oracle encode of an image plus every block's per-pass distortions."""
    h, w = img.shape[:2]
    nc = 1 if img.ndim == 2 else img.shape[2]
    rs, data, q, passlen = O.encode_image(img, levels=levels, reversible=rev, cbw=64, cbh=64, style=style,
                                          return_passlen=True)
    cbs = O.codeblocks(w, h, levels, 64, 64)
    dists = np.zeros((len(rs), O.MAX_PASSES), np.int64)
    for c in range(nc):
        for i, d in enumerate(cbs):
            k = c * len(cbs) + i
            blk = q[c, d["y0"]:d["y0"] + d["h"], d["x0"]:d["x0"] + d["w"]]
            _, npass, _, pl, dist = O.t1_encode_rd(blk, d["orient"], style)
            assert npass == rs[k, 1] and (pl == passlen[k, :npass]).all()
            dists[k, :npass] = dist
    return rs, data, passlen, dists


def test_quality_layers_decode_like_single_cuts():
    """Multi-layer codestreams from the oracle writer: one layer is the
    single-layer writer's bytes; OpenJPEG decoding the first l layers of a
    three-layer stream gives exactly the image of the single-layer stream cut
    at layer l's target."""
    import io
    from PIL import Image
    from roger_amd import synth
    img = synth.image(192, 160, 3, 12)
    style = O.STYLE_PARALLEL_BYPASS
    rs, data, passlen, dists = _oracle_image_rd(img, 4, False, style)
    params = O.j2k_params(192, 160, 3, 4, False, 64, 64, 1.0 / 64, style=style)
    weights = np.ones(len(rs))
    full = int(rs[:, 0].sum())
    targets = [full // 30, full // 8, full // 3]
    cuts, layers = [], []
    for t in targets:
        nt, rt, _ = O.rate_control(rs[:, 1], passlen, dists, weights, t)
        r2 = rs.copy()
        r2[:, 0], r2[:, 1] = rt, nt
        cuts.append(O.t2_write(params, r2, [data[i][:rt[i]] for i in range(len(data))], passlen))
        layers.append(nt)
    layers = np.stack(layers, 1)
    assert (np.diff(layers, axis=1) >= 0).all()
    # one layer == the single-layer writer
    assert O.t2_write(params, rs, data, passlen, layers=rs[:, 1:2]) == O.t2_write(params, rs, data, passlen)
    r3 = rs.copy()
    r3[:, 1] = layers[:, -1]
    r3[:, 0] = [int(passlen[i, :layers[i, -1]].sum()) for i in range(len(rs))]
    cs = O.t2_write(params, r3, data, passlen, layers=layers)
    for l in range(3):
        got = decode_layers(cs, l + 1)
        ref = np.array(Image.open(io.BytesIO(cuts[l])))
        np.testing.assert_array_equal(got, ref, err_msg="layer %d" % l)


def decode_layers(cs, n):
    """OpenJPEG (Pillow) decode of the first n quality layers.  Pillow reads
    its `layers` setting only together with a resolution reduction, so the
    decoder tile's arguments are set directly."""
    import io
    from PIL import Image, ImageFile
    im = Image.open(io.BytesIO(cs))
    t = im.tile[0]
    im.tile = [ImageFile._Tile(t[0], t[1], t[2], (t[3][0], 0, n, t[3][3], t[3][4]))]
    return np.array(im)
