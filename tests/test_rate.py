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
