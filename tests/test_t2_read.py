"""CPU: the library's tier-2 reader (host code, no device work) -- round
trips of the oracle writer's codestreams in every style and with quality
layers, and OpenJPEG-encoded lossless codestreams (Pillow) decoded on the
CPU through the reader and the oracle's tier-1 decoder, inverse 5/3 and
inverse RCT back to the original image."""
import io

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth
from roger_amd._lib import RogerError

STYLES = [0, O.STYLE_PARALLEL, O.STYLE_PARALLEL_BYPASS]


@pytest.mark.parametrize("style", STYLES)
@pytest.mark.parametrize("w,h,nc,rev,levels,cb", [
    (200, 150, 3, True, 4, 64),
    (129, 67, 1, False, 3, 32),
    (64, 64, 3, False, 0, 64),
    (33, 1, 1, True, 1, 64),
])
def test_reader_round_trip(w, h, nc, rev, levels, cb, style):
    img = synth.image(w, h, nc, 5)
    rs, data, q, passlen = O.encode_image(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, style=style,
                                          return_passlen=True)
    params = O.j2k_params(w, h, nc, levels, rev, cb, cb, 1.0 / 64, style=style)
    cs = O.t2_write(params, rs, data, passlen if style else None)
    info, desc, res, blobs, seg = R.j2k_parse(cs)
    assert (info["w"], info["h"], info["ncomp"], info["levels"], info["cbw"], info["style"], info["nlayers"]) == \
        (w, h, nc, levels, cb, style, 1)
    np.testing.assert_array_equal(res["npasses"], rs[:, 1])
    np.testing.assert_array_equal(res["nbp"], np.where(rs[:, 1] > 0, rs[:, 2], 0))
    assert blobs == data
    if style:
        for i in range(len(rs)):
            np.testing.assert_array_equal(seg[i, :rs[i, 1]], passlen[i, :rs[i, 1]])
            assert (seg[i, rs[i, 1]:] == 0).all()
    cbs = O.codeblocks(w, h, levels, cb, cb)
    for c in range(nc):
        for i, d in enumerate(cbs):
            e = desc[c * len(cbs) + i]
            assert (e["x0"], e["y0"], e["w"], e["h"], e["comp"], e["orient"]) == (d["x0"], d["y0"], d["w"], d["h"],
                                                                                 c, d["orient"])
    if not rev:
        steps = O.band_steps(w, h, levels, 1.0 / 64)
        np.testing.assert_allclose(info["step"], np.asarray(steps, np.float64)[:3 * levels + 1, 2], rtol=1e-6)


def test_reader_quality_layers():
    """Three layers: the blocks read back with every layer's passes joined."""
    from test_rate import _oracle_image_rd
    img = synth.image(192, 160, 3, 12)
    style = O.STYLE_PARALLEL_BYPASS
    rs, data, passlen, dists = _oracle_image_rd(img, 4, False, style)
    params = O.j2k_params(192, 160, 3, 4, False, 64, 64, 1.0 / 64, style=style)
    full = int(rs[:, 0].sum())
    cuts = [O.rate_control(rs[:, 1], passlen, dists, np.ones(len(rs)), t) for t in (full // 30, full // 8, full // 3)]
    layers = np.stack([c[0] for c in cuts], 1)
    nt, rt, _ = cuts[-1]
    r3 = rs.copy()
    r3[:, 0], r3[:, 1] = rt, nt
    cs = O.t2_write(params, r3, [data[i][:rt[i]] for i in range(len(data))], passlen, layers=layers)
    info, desc, res, blobs, seg = R.j2k_parse(cs)
    assert info["nlayers"] == 3
    np.testing.assert_array_equal(res["npasses"], nt)
    assert blobs == [data[i][:rt[i]] for i in range(len(data))]


def _cpu_decode(cs):
    """Reader + oracle tier-1 decode + inverse 5/3 + inverse RCT / shift."""
    info, desc, res, blobs, seg = R.j2k_parse(cs)
    w, h, nc, lv = info["w"], info["h"], info["ncomp"], info["levels"]
    planes = np.zeros((nc, h, w), np.int32)
    for i, d in enumerate(desc):
        n = res["npasses"][i]
        if n == 0:
            continue
        blk = O.t1_decode(blobs[i], d["w"], d["h"], d["orient"], res["nbp"][i], n, style=info["style"],
                          passlen=None if seg is None else seg[i, :n])
        planes[d["comp"], d["y0"]:d["y0"] + d["h"], d["x0"]:d["x0"] + d["w"]] = blk
    planes = np.stack([O.dwt53_inv(p, lv) for p in planes])
    if nc == 3 and info["colour"] == 1:
        return O.rct_inv(planes)
    return np.stack([O.unshift(p) for p in planes], -1).squeeze()


@pytest.mark.parametrize("nc,res_levels,cb,layers,sop", [
    (3, 6, (64, 64), None, False),
    (1, 4, (32, 32), None, True),
    (3, 3, (64, 64), [40, 20, 1], False),
])
def test_reader_openjpeg_lossless(nc, res_levels, cb, layers, sop):
    """Codestreams that OpenJPEG (through Pillow) wrote: lossless 5/3, RCT for
    RGB, code-block style 0, optional SOP/EPH markers and rate-layered
    packets (the last layer lossless); decoded on the CPU through the reader
    to the exact image."""
    from PIL import Image
    img = synth.image(173, 141, nc, 8)
    im = Image.fromarray(img)
    buf = io.BytesIO()
    kw = dict(irreversible=False, num_resolutions=res_levels, codeblock_size=cb, progression="LRCP", no_jp2=True)
    if layers:
        kw.update(quality_mode="rates", quality_layers=layers)
    if sop:
        kw.update(sop=True)
    im.save(buf, "JPEG2000", **kw)
    cs = buf.getvalue()
    info = R.j2k_parse(cs)[0]
    assert info["levels"] == res_levels - 1 and info["nlayers"] == (len(layers) if layers else 1)
    np.testing.assert_array_equal(_cpu_decode(cs), img)


def test_reader_refuses_unsupported():
    from PIL import Image
    img = synth.image(64, 64, 1, 2)
    buf = io.BytesIO()
    Image.fromarray(img).save(buf, "JPEG2000", irreversible=False, progression="RLCP", no_jp2=True)
    with pytest.raises(RogerError):
        R.j2k_parse(buf.getvalue())
    with pytest.raises(RogerError):
        R.j2k_parse(b"\xff\x4f\xff\x51\x00")


def test_reader_survives_damaged_codestreams():
    """Truncations and byte flips of valid codestreams: the reader either
    reads them or refuses them with an error, never reads out of bounds."""
    img = synth.image(150, 120, 3, 9)
    streams = []
    for style in STYLES:
        rs, data, q, passlen = O.encode_image(img, levels=3, reversible=False, style=style, return_passlen=True)
        streams.append(O.t2_write(O.j2k_params(150, 120, 3, 3, False, style=style), rs, data,
                                  passlen if style else None))
    rng = np.random.default_rng(3)
    outcomes = {"ok": 0, "refused": 0}
    for k in range(300):
        cs = bytearray(streams[k % len(streams)])
        if k % 3 == 0:
            cs = cs[:int(rng.integers(2, len(cs)))]
        else:
            for _ in range(int(rng.integers(1, 6))):
                cs[int(rng.integers(0, len(cs)))] = int(rng.integers(0, 256))
        try:
            R.j2k_parse(bytes(cs))
            outcomes["ok"] += 1
        except RogerError:
            outcomes["refused"] += 1
    assert outcomes["ok"] + outcomes["refused"] == 300 and outcomes["refused"] > 0


def _header_only_stream(w, h, ncomp=1, levels=0, xcb=0, ycb=0):
    """SOC, SIZ, COD (LRCP, one layer, 5/3), QCD, one empty tile-part, EOC:
    a codestream whose headers alone describe a w x h image."""
    import struct
    siz = struct.pack(">HHHIIIIIIIIH", 0xFF51, 38 + 3 * ncomp, 0, w, h, 0, 0, w, h, 0, 0, ncomp)
    siz += bytes([7, 1, 1]) * ncomp
    cod = struct.pack(">HHBBHBBBBBB", 0xFF52, 12, 0, 0, 1, 0, levels, xcb, ycb, 0, 1)
    nb = 3 * levels + 1
    qcd = struct.pack(">HHB", 0xFF5C, 3 + nb, 2 << 5) + bytes([9 << 3] * nb)
    body = b"\x00" * (levels + 1)        # one empty packet per resolution
    sot = struct.pack(">HHHIBB", 0xFF90, 10, 0, 12 + 2 + len(body), 0, 1)
    return b"\xff\x4f" + siz + cod + qcd + sot + b"\xff\x93" + body + b"\xff\xd9"


def test_reader_refuses_oversized_headers():
    """An 82-byte codestream whose SIZ/COD ask for 65536x65536 samples in 4x4
    code-blocks (268M blocks) is refused before any allocation, with an error
    code rather than an exception or an abort; a small stream with the same
    headers reads."""
    huge = _header_only_stream(65536, 65536)
    assert len(huge) < 100
    with pytest.raises(RogerError):
        R.j2k_parse(huge)
    with pytest.raises(RogerError):
        R.j2k_parse(_header_only_stream(40000, 40000, ncomp=1, levels=5, xcb=4, ycb=4))
    info = R.j2k_parse(_header_only_stream(64, 48, levels=2, xcb=2, ycb=2))[0]
    assert (info["w"], info["h"], info["levels"], info["cbw"]) == (64, 48, 2, 16)
