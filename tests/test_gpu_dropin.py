"""GPU: the C++ drop-in classes (include/roger/OCLEncoder.h etc.) driven as the
reference harness drives its own (4 level-shifted float components with A = B,
5 levels, precision 8, 32x32 code-blocks), compared with the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import runtime as R
from roger_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dropin_bin(tmp_path_factory):
    out = tmp_path_factory.mktemp("dropin") / "dropin_main"
    lib = os.path.join(ROOT, "roger_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp"), "-o", str(out), "-L" + lib,
                    "-lroger_amd", "-Wl,-rpath," + lib], check=True)
    return str(out)


@pytest.mark.parametrize("w,h", [(256, 192), (131, 77)])
def test_dropin_encoder_decoder(dropin_bin, tmp_path, w, h):
    img = synth.rgb(w, h, 21)
    src = tmp_path / "in.rgb"
    img.tofile(src)
    r = subprocess.run([dropin_bin, str(src), str(w), str(h), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    n = w * h
    chans = [img[..., c if c < 3 else 2].astype(np.float32) - 128 for c in range(4)]
    levels = 5
    step, inv, _, _ = R.band_steps(levels, 1.0 / 64)
    # quantised indices, int16-saturated and interleaved like mapDWTOut
    q = np.fromfile(tmp_path / "lossy_q.i16", np.int16).reshape(h, w, 4)
    qref = [O.quantize(O.dwt97_fwd(c, levels), levels, inv) for c in chans]
    for c in range(4):
        np.testing.assert_array_equal(q[..., c], np.clip(qref[c], -32768, 32767))
    # OCLRGBtoPlanar: one int16 plane per channel, read back with mapBuffer
    for c in range(4):
        plane = np.fromfile(tmp_path / ("planar_c%d.i16" % c), np.int16).reshape(h, w)
        np.testing.assert_array_equal(plane, np.clip(qref[c], -32768, 32767), err_msg="channel %d" % c)
    # getDwtIn(1): the LL band after one level, component 0
    ll1 = np.fromfile(tmp_path / "dwtin1_c0.f32", np.float32).reshape((h + 1) // 2, (w + 1) // 2)
    np.testing.assert_array_equal(ll1.view(np.uint32),
                                  O.dwt97_fwd(chans[0], 1)[:(h + 1) // 2, :(w + 1) // 2].view(np.uint32))
    # tier-1 on 32x32 code-blocks, all four components
    res = np.fromfile(tmp_path / "lossy_cb.res", np.int32).reshape(-1, 4)
    blob = np.fromfile(tmp_path / "lossy_cb.bin", np.uint8).tobytes()
    cbs = O.codeblocks(w, h, levels, 32, 32)
    assert len(res) == 4 * len(cbs)
    off = 0
    for c in range(4):
        for i, d in enumerate(cbs):
            k = c * len(cbs) + i
            blk = qref[c][d["y0"]:d["y0"] + d["h"], d["x0"]:d["x0"] + d["w"]]
            data, npass, nbp = O.t1_encode(blk, d["orient"])
            assert (res[k][1], res[k][2]) == (npass, nbp)
            assert blob[off:off + res[k][0]] == data
            off += res[k][0]
    # float DWT output (outputDwt) and the decoder's reconstruction, bit-exact
    coef = np.fromfile(tmp_path / "lossy_coef.f32", np.float32).reshape(h, w)
    cref = O.dwt97_fwd(chans[0], levels)
    np.testing.assert_array_equal(coef.view(np.uint32), cref.view(np.uint32))
    rec = np.fromfile(tmp_path / "lossy_rec.f32", np.float32).reshape(h, w, 4)
    np.testing.assert_array_equal(rec[..., 0].view(np.uint32), O.dwt97_inv(cref, levels).view(np.uint32))
    assert np.abs(rec[..., 1] - chans[1]).max() < 2e-3
    # 5/3 coefficients
    c53 = np.fromfile(tmp_path / "lossless_coef.i16", np.int16).reshape(h, w)
    np.testing.assert_array_equal(c53, O.dwt53_fwd(chans[0].astype(np.int32), levels))
    assert n > 0
