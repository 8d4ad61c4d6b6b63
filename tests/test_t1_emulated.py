"""CPU: the tier-1 kernels' own code (roger_amd/csrc/t1_core.h, built for the
host against the lock-step wavefront emulator in tests/emu/) produces exactly
the oracle's code-block bytes, pass counts and bit-plane counts, in all three
ways the kernels drive it (inline MQ, split with flat arrays, split over a job
list).  The emulator also fails a run whose cross-lane calls are not
wave-uniform or whose declared-uniform values differ between lanes."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import synth
from roger_amd.runtime import CB_DESC_DTYPE, CB_RES_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_SRC = os.path.join(HERE, "emu", "t1_emu.cpp")
MODES = ["inline", "split", "queue_split"]


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emu") / "t1_emu")
    # -fno-tree-ch: loop-header copying can give one cross-lane call of the
    # source two code addresses, which the emulator's call-site check (by
    # return address) would report as non-uniform control flow
    subprocess.run([gxx, "-O2", "-fno-tree-ch", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", EMU_SRC,
                    "-o", exe], check=True)
    return exe


def _run(emu, tmp_path, mode, q, desc, pitch, plane_stride, slot, style=None):
    d = str(tmp_path)
    np.ascontiguousarray(q, np.int32).tofile(os.path.join(d, "q.bin"))
    np.ascontiguousarray(desc, CB_DESC_DTYPE).tofile(os.path.join(d, "desc.bin"))
    args = [emu, mode, d, str(pitch), str(plane_stride), str(slot)] + ([str(style)] if style is not None else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = np.fromfile(os.path.join(d, "res.bin"), CB_RES_DTYPE)
    out = np.fromfile(os.path.join(d, "out.bin"), np.uint8).reshape(len(desc), slot)
    return res, [bytes(out[i, :res["len"][i]]) for i in range(len(desc))]


def _image_case(w, h, nc, rev, levels, seed, cb):
    img = synth.image(w, h, nc, seed)
    rs, data, q = O.encode_image(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64)
    one = O.codeblocks(w, h, levels, cb, cb)
    desc = np.zeros(len(one) * nc, CB_DESC_DTYPE)
    k = 0
    for c in range(nc):
        for d in one:
            desc[k] = (d["x0"], d["y0"], d["w"], d["h"], c, d["orient"], d["band"], d["res"])
            k += 1
    return q, desc, rs, data


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("w,h,nc,rev,levels,seed,cb", [
    (97, 61, 1, False, 3, 52, 64),    # 9/7, ragged bands
    (33, 1, 1, False, 1, 56, 64),     # one-row image: empty high bands, 1-row blocks
    (64, 64, 3, False, 2, 54, 32),    # RGB, 32x32 blocks
    (129, 57, 1, True, 4, 55, 64),    # 5/3, odd sizes
])
def test_image_blocks_match_oracle(emu, tmp_path, mode, w, h, nc, rev, levels, seed, cb):
    q, desc, rs, data = _image_case(w, h, nc, rev, levels, seed, cb)
    res, blobs = _run(emu, tmp_path, mode, q, desc, w, w * h, max(256, cb * cb * 3))
    assert (res["flags"] == 0).all()
    np.testing.assert_array_equal(res["nbp"], rs[:, 2])
    np.testing.assert_array_equal(res["npasses"], rs[:, 1])
    assert blobs == data


def _skewed_lps_blocks():
    """Mostly-zero blocks whose rare 1s are LPS decisions in very skewed
    states (hundreds of run-length / ZC zeros first), so that an LPS
    renormalisation shifts 9-15 bits and, for these patterns and
    orientations, needs two BYTEOUTs in one coding step.  Returns (block,
    orientation) pairs."""
    cases = [(3, [(41, 31, 1), (41, 46, 3), (51, 40, 1)]), (0, [(44, 53, 1), (53, 0, 3)]),
             (3, [(38, 44, 1), (47, 24, 3)]), (1, [(33, 5, -1), (33, 34, 3), (38, 45, 1)]),
             (3, [(47, 55, -1), (54, 16, 3), (58, 61, 1)]), (2, [(36, 16, -1), (52, 15, 3), (53, 47, 1)]),
             (0, [(63, 63, 1)])]
    out = []
    for orient, pts in cases:
        b = np.zeros((64, 64), np.int32)
        for y, x, v in pts:
            b[y, x] = v
        out.append((b, orient))
    return out


@pytest.mark.parametrize("mode", MODES)
def test_dense_blocks_match_oracle(emu, tmp_path, mode):
    """Full 64x64 blocks of wide-range Laplacian-like values (many bit-planes,
    long codewords, frequent carries into the byte register) plus sparse and
    single-sample blocks, all four orientations."""
    rng = np.random.default_rng(11)
    blocks = []
    for k in range(6):
        scale = [3, 40, 900, 20000, 5, 250][k]
        b = np.round(rng.laplace(0, scale, (64, 64))).astype(np.int32)
        if k == 4:
            b[rng.random((64, 64)) < 0.97] = 0          # sparse: run-length paths
        blocks.append(b)
    one = np.zeros((64, 64), np.int32)
    one[37, 21] = -5
    blocks.append(one)
    orients = [i % 4 for i in range(len(blocks))]
    for b, o in _skewed_lps_blocks():
        blocks.append(b)
        orients.append(o)
    plane = np.concatenate(blocks, axis=1)               # 64 x (64 * nblocks)
    n = len(blocks)
    desc = np.zeros(n, CB_DESC_DTYPE)
    for i in range(n):
        desc[i] = (64 * i, 0, 64, 64, 0, orients[i], 0, 0)
    res, blobs = _run(emu, tmp_path, mode, plane, desc, plane.shape[1], plane.size, 64 * 64 * 4)
    for i, b in enumerate(blocks):
        data, npasses, nbp = O.t1_encode(b, orients[i])
        assert res["flags"][i] == 0
        assert (res["nbp"][i], res["npasses"][i]) == (nbp, npasses), i
        assert blobs[i] == data, i


@pytest.mark.parametrize("mode", MODES)
def test_small_slots_flag_overflow(emu, tmp_path, mode):
    """Slots smaller than the codeword: the block is flagged, nothing is
    written past its slot, and blocks that fit are still exact."""
    rng = np.random.default_rng(5)
    blocks = [np.round(rng.laplace(0, s, (64, 64))).astype(np.int32) for s in (0.4, 1.0, 30, 2000)]
    plane = np.concatenate(blocks, axis=1)
    desc = np.zeros(len(blocks), CB_DESC_DTYPE)
    for i in range(len(blocks)):
        desc[i] = (64 * i, 0, 64, 64, 0, 0, 0, 0)
    slot = 160
    res, blobs = _run(emu, tmp_path, mode, plane, desc, plane.shape[1], plane.size, slot)
    raw = np.fromfile(os.path.join(str(tmp_path), "out.bin"), np.uint8).reshape(len(blocks), slot)
    for i, b in enumerate(blocks):
        data, npasses, nbp = O.t1_encode(b, 0)
        assert res["flags"][i] == (1 if len(data) > slot else 0), i
        if len(data) <= slot:
            assert blobs[i] == data
        else:
            assert bytes(raw[i]) == data[:slot]   # the slot holds the codeword's first bytes


def _run_parallel(emu, tmp_path, mode, q, desc, pitch, plane_stride, slot, style=O.STYLE_PARALLEL):
    res, blobs = _run(emu, tmp_path, mode, q, desc, pitch, plane_stride, slot, style)
    seglen = np.fromfile(os.path.join(str(tmp_path), "seglen.bin"), np.int32).reshape(len(desc), O.MAX_PASSES)
    return res, blobs, seglen


@pytest.mark.parametrize("style", [O.STYLE_PARALLEL, O.STYLE_PARALLEL_BYPASS])
@pytest.mark.parametrize("mode", ["parallel", "queue_parallel", "sorted", "two_phase", "three_phase"])
@pytest.mark.parametrize("w,h,nc,rev,levels,seed,cb", [
    (97, 61, 1, False, 3, 52, 64),
    (33, 1, 1, False, 1, 56, 64),
    (64, 64, 3, False, 2, 54, 32),
    (129, 57, 1, True, 4, 55, 64),
])
def test_parallel_style_matches_oracle(emu, tmp_path, mode, w, h, nc, rev, levels, seed, cb, style):
    """RESET|TERMALL (and with BYPASS): pass-parallel MQ / raw lanes + segment
    packing give the oracle's codeword and per-pass segment lengths."""
    img = synth.image(w, h, nc, seed)
    rs, data, q, passlen = O.encode_image(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64,
                                          style=style, return_passlen=True)
    one = O.codeblocks(w, h, levels, cb, cb)
    desc = np.zeros(len(one) * nc, CB_DESC_DTYPE)
    k = 0
    for c in range(nc):
        for d in one:
            desc[k] = (d["x0"], d["y0"], d["w"], d["h"], c, d["orient"], d["band"], d["res"])
            k += 1
    res, blobs, seglen = _run_parallel(emu, tmp_path, mode, q, desc, w, w * h, max(256, cb * cb * 3), style)
    assert (res["flags"] == 0).all()
    np.testing.assert_array_equal(res["nbp"], rs[:, 2])
    np.testing.assert_array_equal(res["npasses"], rs[:, 1])
    for i in range(len(desc)):
        n = rs[i, 1]
        np.testing.assert_array_equal(seglen[i, :n], passlen[i, :n])
    assert blobs == data


@pytest.mark.parametrize("mode", ["parallel", "sorted", "two_phase", "three_phase"])
@pytest.mark.parametrize("style", [O.STYLE_PARALLEL, O.STYLE_PARALLEL_BYPASS])
def test_parallel_style_dense_blocks(emu, tmp_path, style, mode):
    rng = np.random.default_rng(23)
    blocks = [np.round(rng.laplace(0, s, (64, 64))).astype(np.int32) for s in (2, 60, 3000, 40000)]
    blocks.append(np.zeros((64, 64), np.int32))
    # raw segments with many 0xFF bytes (bit stuffing): all-ones refinement bits
    blocks.append(np.full((64, 64), (1 << 14) - 1, np.int32) * np.where(rng.random((64, 64)) < 0.5, 1, -1))
    orients = [(i + 1) % 4 for i in range(len(blocks))]
    for b, o in _skewed_lps_blocks():
        blocks.append(b)
        orients.append(o)
    plane = np.concatenate(blocks, axis=1)
    desc = np.zeros(len(blocks), CB_DESC_DTYPE)
    for i in range(len(blocks)):
        desc[i] = (64 * i, 0, 64, 64, 0, orients[i], 0, 0)
    res, blobs, seglen = _run_parallel(emu, tmp_path, mode, plane, desc, plane.shape[1], plane.size,
                                       64 * 64 * 4, style)
    for i, b in enumerate(blocks):
        if not b.any():
            assert (res["len"][i], res["npasses"][i], res["nbp"][i]) == (0, 0, 0)
            continue
        data, npasses, nbp, passlen = O.t1_encode(b, orients[i], style=style)
        assert (res["nbp"][i], res["npasses"][i], res["flags"][i]) == (nbp, npasses, 0)
        np.testing.assert_array_equal(seglen[i, :npasses], passlen)
        assert blobs[i] == data


# ---------------- decoder core (t1_dec_core.h) ----------------
def _run_decode(emu, tmp_path, desc, res_rows, blobs, pitch, plane_stride, nplanes, style, passlen=None,
                mode="decode"):
    d = str(tmp_path)
    np.ascontiguousarray(desc, CB_DESC_DTYPE).tofile(os.path.join(d, "desc.bin"))
    res = np.zeros(len(desc), CB_RES_DTYPE)
    res["len"], res["npasses"], res["nbp"] = res_rows[:, 0], res_rows[:, 1], res_rows[:, 2]
    res.tofile(os.path.join(d, "res.bin"))
    lens = np.array([len(b) for b in blobs], np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    offs.tofile(os.path.join(d, "offs.bin"))
    with open(os.path.join(d, "cw.bin"), "wb") as f:
        f.write(b"".join(blobs))
    if passlen is not None:
        np.ascontiguousarray(passlen, np.int32).tofile(os.path.join(d, "seglen.bin"))
    r = subprocess.run([emu, mode, d, str(pitch), str(plane_stride), str(nplanes), str(style)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return np.fromfile(os.path.join(d, "dq.bin"), np.int32).reshape(nplanes, -1)


DEC_STYLES = [0, O.STYLE_PARALLEL, O.STYLE_PARALLEL_BYPASS]
# decode: the wavefront decoder core (t1_dec_core.h) under the emulator;
# decode_lane: the lane decoder (t1_dec_lane.h) with plain host state
DEC_MODES = ["decode", "decode_lane"]


@pytest.mark.parametrize("mode", DEC_MODES)
@pytest.mark.parametrize("style", DEC_STYLES)
@pytest.mark.parametrize("w,h,nc,rev,levels,seed,cb", [
    (97, 61, 1, False, 3, 52, 64),
    (33, 1, 1, False, 1, 56, 64),
    (64, 64, 3, False, 2, 54, 32),
    (129, 57, 1, True, 4, 55, 64),
])
def test_decoder_core_inverts_oracle(emu, tmp_path, mode, style, w, h, nc, rev, levels, seed, cb):
    """The decoder kernel's core on the oracle's codewords of whole images, in
    every code-block style: exactly the quantisation indices."""
    img = synth.image(w, h, nc, seed)
    rs, data, q, passlen = O.encode_image(img, levels=levels, reversible=rev, cbw=cb, cbh=cb, delta=1.0 / 64,
                                          style=style, return_passlen=True)
    one = O.codeblocks(w, h, levels, cb, cb)
    desc = np.zeros(len(one) * nc, CB_DESC_DTYPE)
    k = 0
    for c in range(nc):
        for d in one:
            desc[k] = (d["x0"], d["y0"], d["w"], d["h"], c, d["orient"], d["band"], d["res"])
            k += 1
    dq = _run_decode(emu, tmp_path, desc, rs, data, w, w * h, nc, style, passlen if style else None, mode)
    np.testing.assert_array_equal(dq.reshape(q.shape), q)


@pytest.mark.parametrize("mode", DEC_MODES)
@pytest.mark.parametrize("style", DEC_STYLES)
def test_decoder_core_dense_blocks(emu, tmp_path, mode, style):
    """Dense and sparse 64x64 blocks (many bit-planes, run-length paths,
    skewed LPS states, raw segments full of 0xFF bytes) and a ragged 64x7
    block, every orientation: decoded back exactly."""
    rng = np.random.default_rng(29)
    blocks = [np.round(rng.laplace(0, s, (64, 64))).astype(np.int32) for s in (2, 60, 3000, 40000)]
    sp = np.round(rng.laplace(0, 5, (64, 64))).astype(np.int32)
    sp[rng.random((64, 64)) < 0.97] = 0
    blocks.append(sp)
    blocks.append(np.full((64, 64), (1 << 14) - 1, np.int32) * np.where(rng.random((64, 64)) < 0.5, 1, -1))
    orients = [(i + 2) % 4 for i in range(len(blocks))]
    for b, o in _skewed_lps_blocks():
        blocks.append(b)
        orients.append(o)
    rag = np.zeros((64, 64), np.int32)
    rag[:7, :] = np.round(rng.laplace(0, 300, (7, 64))).astype(np.int32)
    blocks.append(rag)
    orients.append(1)
    hs = [64] * (len(blocks) - 1) + [7]
    plane = np.concatenate(blocks, axis=1)
    n = len(blocks)
    desc = np.zeros(n, CB_DESC_DTYPE)
    rows, blobs, pls = [], [], np.zeros((n, O.MAX_PASSES), np.int32)
    for i, b in enumerate(blocks):
        desc[i] = (64 * i, 0, 64, hs[i], 0, orients[i], 0, 0)
        r = O.t1_encode(b[:hs[i]], orients[i], style=style)
        rows.append((len(r[0]), r[1], r[2]))
        blobs.append(r[0])
        if style:
            pls[i, :r[1]] = r[3]
    dq = _run_decode(emu, tmp_path, desc, np.array(rows, np.int32), blobs, plane.shape[1], plane.size, 1, style,
                     pls if style else None, mode)
    dq = dq.reshape(plane.shape)
    for i, b in enumerate(blocks):
        np.testing.assert_array_equal(dq[:hs[i], 64 * i:64 * (i + 1)], b[:hs[i]], err_msg="block %d" % i)
