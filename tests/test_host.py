"""CPU tests of the host side: the C-ABI library loads and exports every
entry point include/roger_amd.h declares, host-only logic (code-block
enumeration, quantisation step signalling, argument validation) agrees with
the oracle, and the synthetic inputs are deterministic.  No GPU call."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from roger_amd import _lib
from roger_amd import runtime as R
from roger_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    src = open(os.path.join(ROOT, "include", "roger_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rg_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), "missing export " + n
        assert n in _lib.SIGNATURES, "no ctypes signature for " + n
    assert set(_lib.SIGNATURES) == set(names)


def test_error_strings():
    L = _lib.lib()
    assert L.rg_error_string(0) == b"success"
    assert L.rg_error_string(_lib.RG_ERR_OVERFLOW) == b"code-block output overflow"


@pytest.mark.parametrize("w,h,levels,cb", [(512, 512, 5, 64), (4096, 4096, 6, 64), (97, 61, 3, 32), (1, 1, 0, 64),
                                           (300, 2, 7, 64), (8192, 8192, 6, 32)])
def test_codeblocks_match_oracle(w, h, levels, cb):
    mine = R.codeblocks(w, h, levels, cb, cb)
    ref = O.codeblocks(w, h, levels, cb, cb)
    assert len(mine) == len(ref)
    for a, b in zip(mine, ref):
        assert (a["x0"], a["y0"], a["w"], a["h"], a["orient"], a["band"], a["res"]) == \
               (b["x0"], b["y0"], b["w"], b["h"], b["orient"], b["band"], b["res"])
    assert int((mine["w"] * mine["h"]).sum()) == w * h


def test_codeblocks_rejects_bad_sizes():
    L = _lib.lib()
    assert L.rg_codeblocks(64, 64, 2, 128, 32, None, 0) == _lib.RG_ERR_INVALID
    assert L.rg_codeblocks(64, 64, 2, 48, 32, None, 0) == _lib.RG_ERR_INVALID
    assert b"not supported" in L.rg_last_error()
    assert L.rg_codeblocks(0, 64, 2, 64, 64, None, 0) == _lib.RG_ERR_INVALID


@pytest.mark.parametrize("levels,delta", [(6, 1.0 / 64), (5, 0.7), (3, 2.0), (1, 0.001)])
def test_band_steps_match_oracle(levels, delta):
    step, inv, eps, mu = R.band_steps(levels, delta)
    ref = O.band_steps(64, 64, levels, delta)
    for b, (e, m, s, i) in enumerate(ref):
        assert (eps[b], mu[b]) == (e, m)
        assert step[b] == np.float32(s) and inv[b] == i


def test_synthetic_images_deterministic():
    a = synth.gray(64, 48, 0)
    assert a.shape == (48, 64) and a.dtype == np.uint8
    assert np.array_equal(a, synth.gray(64, 48, 0))
    assert not np.array_equal(a, synth.gray(64, 48, 1))
    rgb = synth.rgb(32, 16, 1)
    assert rgb.shape == (16, 32, 3)
    assert 20 < a.std() < 60  # 4 sinusoids of amplitude 24 plus noise
    # the noise term is uniform on -8..8
    n = synth.noise(256, 256, 3)
    assert n.min() == -8 and n.max() == 8


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_no_oracle_in_product():
    """The product package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "roger_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                for pat in ("import oracle", "from oracle", "liboracle", "oracle/"):
                    assert pat not in text, (os.path.join(dirpath, f), pat)


def test_dwt_fwd_u8_scratch_covers_the_path_taken():
    """rg_dwt_fwd_u8's workspace: the LL pyramid when the first level fuses
    with the colour stage (rows of a multiple of 4 bytes, sides beyond the
    filter halo), the pyramid plus 256-byte aligned colour planes otherwise;
    the choice depends on the shape alone (no environment setting)."""
    from roger_amd._lib import RG_COLOUR_ICT, RG_COLOUR_NONE, RG_COLOUR_RCT, RG_FILTER_53, RG_FILTER_97
    cases = [(4096, 4096, 3, RG_COLOUR_ICT, RG_FILTER_97, True), (132, 40, 3, RG_COLOUR_RCT, RG_FILTER_53, True),
             (8, 5, 1, RG_COLOUR_NONE, RG_FILTER_97, True), (4, 100, 1, RG_COLOUR_NONE, RG_FILTER_97, False),
             (257, 300, 3, RG_COLOUR_ICT, RG_FILTER_97, False), (1000, 3, 3, RG_COLOUR_RCT, RG_FILTER_53, True),
             (1000, 3, 3, RG_COLOUR_ICT, RG_FILTER_97, False),
             (5, 5, 1, RG_COLOUR_NONE, RG_FILTER_53, False)]
    for w, h, nc, col, filt, fused in cases:
        base = R.scratch_bytes(w, h, nc, filt)
        for env in ("1", "0"):
            os.environ["RG_DWT_FUSED"] = env
            try:
                got = R.dwt_fwd_u8_scratch_bytes(w, h, nc, col, filt)
            finally:
                del os.environ["RG_DWT_FUSED"]
            if fused:
                assert got == base, (w, h, nc)
            else:
                assert got >= ((base + 255) & ~255) + w * h * nc * 4, (w, h, nc)


def test_dwt_fwd_u8_refuses_zero_levels():
    ctx_free = _lib.lib().rg_dwt_fwd_u8(None, None, 8, 8, 1, 0, 1, 0, None, None, None)
    assert ctx_free == _lib.RG_ERR_INVALID
