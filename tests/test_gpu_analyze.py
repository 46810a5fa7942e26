"""The raw-data analyzer on the GPU (imop_hip_analyze_data and
imop_hip_guess_pattern) against the oracle (oracle/isp_oracle.c
isp_analyze_data / isp_guess_pattern, pinned against the reference's compiled
src/analyzer.cpp): the same DataInfo and the same BayerPattern for every
branch of the bit-depth detection, packed and unpacked layouts, odd, tiny and
full-sensor buffers, and frames of every pattern."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import analyzer_cases as AC
import imop_amd
import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
AN_BUFS = AC.analyze_buffers(big=True)
PAT_FRAMES = AC.pattern_frames(big=True)


def _dev(buf, offset=0):
    """buf on the device; offset > 0 places it that many bytes into a larger
    allocation (a view whose address is not 16-byte aligned)."""
    b = np.ascontiguousarray(buf).view(np.uint8).ravel()
    t = torch.zeros(b.size + offset + 1, dtype=torch.uint8, device=DEV)
    t[offset:offset + b.size] = torch.from_numpy(b.copy()).to(DEV)
    return t[offset:offset + b.size]


@pytest.mark.parametrize("k", range(len(AN_BUFS)), ids=[b[0] for b in AN_BUFS])
def test_analyze_data_matches_oracle(k):
    name, make = AN_BUFS[k]
    buf = make()
    print(f"analyze {name} ({buf.size} bytes)", flush=True)
    e, info = imop_amd.analyze_data(_dev(buf))
    assert e == 0
    assert info == O.analyze_data(buf)


@pytest.mark.parametrize("offset", [1, 2, 5, 8])
def test_analyze_data_unaligned_buffers(offset):
    for name, make in AN_BUFS[:12] + [b for b in AN_BUFS if b[0].startswith(("random", "lsb", "generic"))]:
        buf = make()
        e, info = imop_amd.analyze_data(_dev(buf, offset))
        assert e == 0 and info == O.analyze_data(buf), name


def test_analyze_data_empty_and_null():
    e, info = imop_amd.analyze_data(torch.zeros(0, dtype=torch.uint8, device=DEV))
    assert e == 0 and info["detected_bit_depth"] == 0 and info["possible_dimensions"] == []
    assert imop_amd.lib().imop_hip_analyze_data(None, 16, None, None) == 1


def test_analyze_data_repeated_calls_do_not_leak_state():
    """The per-device statistics are cleared for every call."""
    big, small = AN_BUFS[2][1](), dict(AN_BUFS)["random-12345"]()
    for buf in (big, small, big, small[:7]):
        e, info = imop_amd.analyze_data(_dev(buf))
        assert e == 0 and info == O.analyze_data(buf)


@pytest.mark.parametrize("k", range(len(PAT_FRAMES)), ids=[f[0] for f in PAT_FRAMES])
def test_guess_pattern_matches_oracle(k):
    name, make, w, h, bd, packed = PAT_FRAMES[k]
    buf = make()
    print(f"guess_pattern {name}", flush=True)
    e, pat = imop_amd.guess_pattern(_dev(buf), w, h, bd, packed)
    assert e == 0
    assert pat == O.guess_pattern(buf, w, h, bd, packed)


def test_guess_pattern_recovers_ramp_patterns():
    for pat in range(4):
        f = AC.ramp_frame(pat)
        assert imop_amd.guess_pattern(_dev(f), 256, 256, 16) == (0, pat)


def test_guess_pattern_errors():
    L = imop_amd.lib()
    d = _dev(np.zeros(64 * 64 * 2, np.uint8))
    p = ctypes.c_int(-1)
    ptr = ctypes.c_void_p(d.data_ptr())
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 8, 8, 8, 0, None, None) == 1
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 8, 8, 0, 0, ctypes.byref(p), None) == 3
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 8, 8, 17, 0, ctypes.byref(p), None) == 3
    assert L.imop_hip_guess_pattern(ptr, 100, 64, 64, 12, 0, ctypes.byref(p), None) == 6
    # too small or no data: RGGB, as the reference (analyzer.cpp:218)
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 3, 64, 8, 0, ctypes.byref(p), None) == 0 and p.value == 0
    assert L.imop_hip_guess_pattern(None, 0, 64, 64, 8, 0, ctypes.byref(p), None) == 0 and p.value == 0


def test_guess_pattern_packing_argument():
    """is_packed as every entry point takes it: 2 (GigE Vision 10-bit) with
    another bit depth is InvalidBitDepth, values outside 0..2 InternalError;
    the valid forms still answer."""
    L = imop_amd.lib()
    d = _dev(np.zeros(64 * 64 * 2, np.uint8))
    p = ctypes.c_int(-1)
    ptr = ctypes.c_void_p(d.data_ptr())
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 64, 64, 12, 2, ctypes.byref(p), None) == 3
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 64, 64, 8, 2, ctypes.byref(p), None) == 3
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 64, 64, 10, 3, ctypes.byref(p), None) == 6
    assert L.imop_hip_guess_pattern(ptr, d.numel(), 64, 64, 10, -1, ctypes.byref(p), None) == 6
    for packed in (0, 1, 2):
        p.value = -1
        assert L.imop_hip_guess_pattern(ptr, d.numel(), 64, 64, 10, packed, ctypes.byref(p), None) == 0
        assert 0 <= p.value <= 3
