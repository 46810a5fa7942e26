"""SUPER_FAST (src/algorithms/super_fast.cpp:9-43) on the GPU against the
oracle, byte for byte, including the reference's 8-bit vector routine
(avx2.cpp:99-207: a row's first 32 pixels take themselves as their left
neighbour, a last vector starting at w - 32 takes itself as its right one, no
clamp) and the clamped generic path of the other layouts.  The oracle is
pinned against the compiled reference (tests/test_oracle.py)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
SF = imop_amd.DemosaicAlgorithm.SUPER_FAST


def _check(raw, w, h, pattern, bd, packed, bl=None, per_channel=False):
    d_in = torch.from_numpy(np.ascontiguousarray(raw).view(np.uint8).ravel().copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
    e = imop_amd.demosaic(d_in, d_out, w, h, pattern, SF, bd, packed, bl, per_channel)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    got = (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)
    eo, want = O.demosaic(raw, w, h, pattern, O.SUPER_FAST, bd, packed, bl, per_channel)
    assert e == eo == 0
    d = np.argwhere(got != want)
    assert d.size == 0, f"{len(d)} samples differ, first {d[:4].tolist()}"


@pytest.mark.parametrize("bd,packed", [(8, False), (8, True), (10, True), (12, True), (12, False), (16, False),
                                       (6, False)])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("w,h", [(2, 2), (3, 5), (31, 7), (32, 9), (33, 6), (64, 5), (70, 12), (300, 9)])
def test_matches_oracle(bd, packed, pattern, w, h):
    print(f"super_fast {w}x{h} bd {bd} packed {packed} pattern {pattern}", flush=True)
    vals = np.random.default_rng(w * 7 + h + bd + pattern).integers(0, 1 << bd, size=(h, w))
    _check(scene.raw_bytes(vals, bd, packed), w, h, pattern, bd, packed)


@pytest.mark.parametrize("bd", [8, 12])
def test_black_level(bd):
    w, h = 96, 40
    vals = np.random.default_rng(bd).integers(0, 1 << bd, size=(h, w))
    _check(scene.raw_bytes(vals, bd, bd > 8), w, h, 3, bd, bd > 8, [10.0, 20.0, 30.0, 40.0], True)
    _check(scene.raw_bytes(vals, bd, bd > 8), w, h, 3, bd, bd > 8, [33.0] * 4)


@pytest.mark.parametrize("bd", [1, 4, 7])
def test_values_above_the_bit_depth_maximum(bd):
    """Bytes above 2^bd - 1 in a shallow 8-bit frame pass through unclamped, as
    in the reference's 8-bit routine."""
    w, h = 64, 8
    vals = np.random.default_rng(bd).integers(0, 256, size=(h, w)).astype(np.uint8)
    _check(scene.raw_bytes(vals, 8, False), w, h, 0, bd, False)


def test_full_frame_12bit_packed():
    w, h, bd = 4024, 3032, 12
    print("super_fast full frame", flush=True)
    vals = scene.make_scene(h, w, bd, 0, seed=2)
    _check(scene.raw_bytes(vals, bd, True), w, h, 0, bd, True, [64.0] * 4)


def test_minimum_size():
    d_in = torch.zeros(16, dtype=torch.uint8, device=DEV)
    d_out = torch.zeros(64, dtype=torch.uint8, device=DEV)
    assert imop_amd.demosaic(d_in, d_out, 1, 4, 0, SF, 8) == 5     # super_fast.cpp:13
    assert imop_amd.demosaic(d_in, d_out, 4, 1, 0, SF, 8) == 5
