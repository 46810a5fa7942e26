"""DFPD (src/algorithms/dfpd.cpp:7-263) on the GPU against the oracle, byte for
byte: the integer 8-bit core and the exact float path, the directional green
with its tie rule, the chroma means with C division, and the width-5 border
fill.  The oracle is pinned against the compiled reference
(tests/test_oracle.py)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DFPD = imop_amd.DemosaicAlgorithm.DFPD


def _check(raw, w, h, pattern, bd, packed, bl=None, per_channel=False):
    d_in = torch.from_numpy(np.ascontiguousarray(raw).view(np.uint8).ravel().copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
    e = imop_amd.demosaic(d_in, d_out, w, h, pattern, DFPD, bd, packed, bl, per_channel)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    got = (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)
    eo, want = O.demosaic(raw, w, h, pattern, O.DFPD, bd, packed, bl, per_channel)
    assert e == eo == 0
    d = np.argwhere(got != want)
    assert d.size == 0, f"{len(d)} samples differ, first {d[:4].tolist()}"


@pytest.mark.parametrize("bd,packed", [(8, False), (8, True), (10, True), (12, True), (12, False), (16, False),
                                       (6, False)])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("w,h", [(12, 12), (13, 15), (65, 17), (130, 33)])
def test_matches_oracle(bd, packed, pattern, w, h):
    print(f"dfpd {w}x{h} bd {bd} packed {packed} pattern {pattern}", flush=True)
    rng = np.random.default_rng(w * 3 + h + bd + pattern)
    vals = rng.integers(0, 1 << bd, size=(h, w)) if (w + pattern) % 2 else scene.make_scene(h, w, bd, pattern, seed=w)
    _check(scene.raw_bytes(vals, bd, packed), w, h, pattern, bd, packed)


def test_direction_ties():
    """Flat and striped frames: equal gradients take the mean of both averages."""
    _check(np.full((20, 24), 77, np.uint8).ravel(), 24, 20, 0, 8, False)
    stripes = np.repeat(np.random.default_rng(2).integers(0, 4, size=(1, 40)) * 900, 20, axis=0)
    _check(scene.raw_bytes(stripes, 12, False), 40, 20, 2, 12, False)


@pytest.mark.parametrize("bd", [8, 12])
def test_black_level(bd):
    w, h = 96, 40
    vals = np.random.default_rng(bd).integers(0, 1 << bd, size=(h, w))
    _check(scene.raw_bytes(vals, bd, bd > 8), w, h, 3, bd, bd > 8, [10.0, 20.0, 30.0, 40.0], True)


def test_values_above_the_bit_depth_maximum():
    w, h = 40, 16
    vals = np.random.default_rng(4).integers(0, 256, size=(h, w)).astype(np.uint8)
    _check(vals.ravel(), w, h, 1, 5, False)


def test_full_frame_12bit_packed():
    w, h, bd = 4024, 3032, 12
    print("dfpd full frame", flush=True)
    vals = scene.make_scene(h, w, bd, 0, seed=2)
    _check(scene.raw_bytes(vals, bd, True), w, h, 0, bd, True, [64.0] * 4)


def test_minimum_size():
    d_in = torch.zeros(512, dtype=torch.uint8, device=DEV)
    d_out = torch.zeros(2048, dtype=torch.uint8, device=DEV)
    assert imop_amd.demosaic(d_in, d_out, 11, 12, 0, DFPD, 8) == 5     # dfpd.cpp:149
    assert imop_amd.demosaic(d_in, d_out, 12, 11, 0, DFPD, 8) == 5
