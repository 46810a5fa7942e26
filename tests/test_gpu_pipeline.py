"""The whole single-frame path on the GPU (imop_hip_demosaic_pipeline):
unpack -> black level -> polynomial lens shading -> demosaic -> white
balance / colour matrix / gamma, against the oracle's composition of the
same stages (oracle.pipeline; pinned on the CPU against the reference's own
functions run in that order).

HQLI / MG frames on the register-resident path apply the lens gain inside
the demosaic's decode; every other frame (other algorithms, widths the
register path does not take, gamma tables above 4096 entries, fractional
black levels) goes through the separate prologue pass into a stream-ordered
unpacked frame.  Both are covered, with every packing, the four patterns,
u16 samples above the bit depth's maximum, and graph capture."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
A = imop_amd.DemosaicAlgorithm
M1 = [1.6, -0.4, -0.2, -0.3, 1.5, -0.2, 0.05, -0.55, 1.5]
GAINS = [1.9, 0.8, 2.3]
LENS = [([0.8, 0.1, -0.05] * 4, (0.5, 0.5)),
        ([0.3, 0.0, 0.0, 0.5, 0.1, 0.0, 0.45, 0.1, 0.0, 0.9, -0.2, 0.3], (0.4, 0.6)),
        ([-3.0, 1.0, 0.5] * 4, (1.5, -0.25)),
        ([0.0] * 12, (0.0, 0.0)),            # max_r = 0: the reference divides by zero
        ([1e9, 0.0, 0.0] * 4, (0.5, 0.5)),   # gains past 2^31
        ([0.5, 0.1, 0.02] * 4, (1e-9, 0.5)),  # centre x below 2^-20 px: the separate pass
        ([0.5, 0.1, 0.02] * 4, (1e-5, 0.5)),  # just inside the fused range
        ([0.5, 0.1, 0.02] * 4, (1e6, -1e6))]  # far outside the frame


def _epi(flags, gamma=2.2):
    return imop_amd.rgb_epilogue(gains=GAINS if flags & 1 else None, ccm=M1 if flags & 2 else None,
                                 gamma=gamma if flags & 4 else None)


def _frame(w, h, bd, pattern, packed, seed, above=False):
    if above:  # raw u16 words anywhere in [0, 65535]
        return np.random.default_rng(seed).integers(0, 65536, size=(h, w)).astype("<u2").view(np.uint8).ravel()
    return scene.raw_bytes(scene.make_scene(h, w, bd, pattern, seed=seed), bd, packed)


def _want(raw, w, h, pattern, algo, bd, packed, bl, per_channel, li, flags, gamma=2.2):
    lens = LENS[li] if li is not None else None
    return O.pipeline(raw, w, h, pattern, int(algo), bd, packed, bl, per_channel, lens,
                      (flags, GAINS, M1, gamma) if flags else None)


def _got(raw, w, h, pattern, algo, bd, packed, bl, per_channel, li, flags, gamma=2.2):
    d_in = torch.from_numpy(raw.copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
    pro = imop_amd.bayer_prologue(*LENS[li]) if li is not None else None
    e = imop_amd.demosaic_pipeline(d_in, d_out, w, h, pattern, algo, bd, packed, bl, per_channel, prologue=pro,
                                   epilogue=_epi(flags, gamma) if flags else None)
    torch.cuda.synchronize()
    assert e == 0
    out = d_out.cpu().numpy()
    return (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)


def _check(w, h, bd, packed, pattern, algo, bl, per_channel, li, flags, seed, above=False, gamma=2.2):
    raw = _frame(w, h, bd, pattern, packed, seed, above)
    print(f"pipeline {w}x{h} bd {bd} packed {packed} pat {pattern} algo {int(algo)} bl {bl} lens {li} "
          f"flags {flags}", flush=True)
    want = _want(raw, w, h, pattern, algo, bd, packed, bl, per_channel, li, flags, gamma)
    got = _got(raw, w, h, pattern, algo, bd, packed, bl, per_channel, li, flags, gamma)
    np.testing.assert_array_equal(got, want)


# ------------------------------------------------ fused into the demosaic's decode

@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("flags", [0, 7])
@pytest.mark.parametrize("li", [0, 1, 2])
def test_fused_mg12(pattern, flags, li):
    _check(264, 70, 12, True, pattern, A.MG, [64.0] * 4, False, li, flags, seed=pattern * 8 + li)


@pytest.mark.parametrize("li", [3, 4, 5, 6, 7])
def test_lens_centre_and_gain_ranges(li):
    # 3, 4, 5 take the separate prologue pass, 6 and 7 the fused decode
    _check(200, 40, 12, True, 1, A.MG, [64.0] * 4, False, li, 7, seed=li)


# (w, h, bd, packed, algo, black level, per channel)
FUSED = [(256, 64, 10, True, A.HQLI, [16.0] * 4, False),
         (258, 66, 10, 2, A.MG, [16.0, 17.0, 18.0, 19.0], True),     # GigE 10-bit, w % 4 == 2
         (264, 70, 8, False, A.MG, [8.0] * 4, False),
         (264, 70, 8, False, A.HQLI, None, False),
         (262, 68, 12, False, A.MG, [200.0, 210.0, 220.0, 230.0], True),
         (256, 64, 14, True, A.MG, None, False),
         (256, 64, 16, False, A.HQLI, [1000.0] * 4, False),
         (12, 6, 12, True, A.MG, [64.0] * 4, False),                # rim groups only
         (1030, 514, 12, True, A.MG, [64.0] * 4, False)]


@pytest.mark.parametrize("k", range(len(FUSED)))
@pytest.mark.parametrize("flags", [0, 3])
def test_fused_formats(k, flags):
    w, h, bd, packed, algo, bl, pc = FUSED[k]
    _check(w, h, bd, packed, k % 4, algo, bl, pc, k % 3, flags, seed=100 + k)


@pytest.mark.parametrize("bd", [10, 12, 14])
@pytest.mark.parametrize("bl", [None, [32.0] * 4])
def test_u16_values_above_the_maximum(bd, bl):
    # the lens stage reads the stored word (no MSB alignment) and clamps
    _check(256, 64, bd, False, 0, A.MG, bl, False, 0, 0, seed=bd, above=True)


def test_fused_full_frame():
    _check(4024, 3032, 12, True, 0, A.MG, [64.0] * 4, False, 0, 7, seed=7)


# ------------------------------------------------ the separate prologue pass

SEPARATE = [(200, 120, 12, True, A.PRISM, [64.0] * 4, False, 7),
            (130, 98, 10, True, A.RCD, [16.0, 17.0, 18.0, 19.0], True, 4),
            (96, 64, 8, False, A.AHD, None, False, 3),
            (101, 67, 12, True, A.MG, [64.0] * 4, False, 7),             # odd width
            (98, 66, 16, False, A.HQLI, [100.0] * 4, False, 4),          # 65536-entry gamma table
            (96, 64, 12, True, A.MG, [64.5] * 4, False, 1),              # fractional black level
            (64, 48, 12, True, A.AMAZE, None, False, 0),
            (64, 48, 12, True, A.L7, [64.0] * 4, False, 2),
            (64, 48, 10, True, A.DFPD, None, False, 0),
            (64, 48, 12, True, A.SUPER_FAST, [64.0] * 4, False, 7)]


@pytest.mark.parametrize("k", range(len(SEPARATE)))
def test_separate_pass(k):
    w, h, bd, packed, algo, bl, pc, flags = SEPARATE[k]
    _check(w, h, bd, packed, k % 4, algo, bl, pc, k % 3, flags, seed=200 + k)


def test_separate_pass_values_above_the_maximum():
    _check(98, 66, 12, False, 2, A.PRISM, None, False, 1, 7, seed=9, above=True)


# ------------------------------------------------ equivalences and arguments

def test_no_prologue_equals_demosaic_isp():
    w, h, bd = 264, 70, 12
    raw = _frame(w, h, bd, 0, True, 3)
    d_in = torch.from_numpy(raw.copy()).to(DEV)
    a = torch.zeros(imop_amd.rgb_byte_size(w, h, bd), dtype=torch.uint8, device=DEV)
    b = torch.ones_like(a)
    c = torch.full_like(a, 7)
    assert imop_amd.demosaic_isp(d_in, a, w, h, 0, A.MG, bd, True, [64.0] * 4, epilogue=_epi(7)) == 0
    assert imop_amd.demosaic_pipeline(d_in, b, w, h, 0, A.MG, bd, True, [64.0] * 4, epilogue=_epi(7)) == 0
    assert imop_amd.demosaic_pipeline(d_in, c, w, h, 0, A.MG, bd, True, [64.0] * 4,
                                      prologue=imop_amd.bayer_prologue(), epilogue=_epi(7)) == 0
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(a, c)


def test_graph_replays_match_eager():
    cases = [(264, 70, 12, True, A.MG, 0), (200, 120, 12, True, A.PRISM, 1)]
    for w, h, bd, packed, algo, li in cases:
        raw = _frame(w, h, bd, 0, packed, 21)
        want = _want(raw, w, h, 0, algo, bd, packed, [64.0] * 4, False, li, 7)
        d_in = torch.from_numpy(raw.copy()).to(DEV)
        d_out = torch.zeros(imop_amd.rgb_byte_size(w, h, bd), dtype=torch.uint8, device=DEV)
        pro, ep = imop_amd.bayer_prologue(*LENS[li]), _epi(7)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            assert imop_amd.demosaic_pipeline(d_in, d_out, w, h, 0, algo, bd, packed, [64.0] * 4,
                                              prologue=pro, epilogue=ep) == 0
        for rep in range(3):
            d_out.fill_(0x5A)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            print(f"replay {rep} algo {int(algo)}", flush=True)
            np.testing.assert_array_equal(d_out.cpu().numpy().view(np.uint16).reshape(h, w, 3), want)
        del g


def test_invalid_arguments():
    w, h, bd = 64, 64, 12
    d_in = torch.zeros(w * h * 2, dtype=torch.uint8, device=DEV)
    d_out = torch.zeros(w * h * 6, dtype=torch.uint8, device=DEV)
    pro = imop_amd.bayer_prologue(*LENS[0])
    L = imop_amd.lib()
    assert L.imop_hip_demosaic_pipeline(None, 0, None, 0, w, h, 0, 2, bd, 0, None, 0, ctypes.byref(pro), None,
                                        None) == 1
    assert imop_amd.demosaic_pipeline(d_in, d_out, 0, h, 0, A.MG, bd, prologue=pro) == 2
    assert imop_amd.demosaic_pipeline(d_in, d_out, w, h, 0, A.MG, 17, prologue=pro) == 3
    assert imop_amd.demosaic_pipeline(d_in, d_out, w, h, 9, A.MG, bd, prologue=pro) == 4
    bad = imop_amd.bayer_prologue(*LENS[0])
    bad.flags = 3  # reserved bit
    assert imop_amd.demosaic_pipeline(d_in, d_out, w, h, 0, A.MG, bd, prologue=bad) == 6
    assert imop_amd.demosaic_pipeline(d_in[: w * h], d_out, w, h, 0, A.MG, bd, prologue=pro) == 6
    assert imop_amd.demosaic_pipeline(d_in, d_out[: w * h * 3], w, h, 0, A.MG, bd, prologue=pro) == 6
    torch.cuda.synchronize()
    assert not d_out.any()
