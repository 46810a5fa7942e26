"""L7 (src/algorithms/l7.cpp:65-128) on the GPU against the oracle, byte for
byte: the 7 x 7 weighted means inside [3, w-3) x [3, h-3), the clamps of the
8-bit and generic paths, and the width-3 border fill with zero corners.  The
oracle is pinned against the compiled reference (tests/test_oracle.py)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
L7 = imop_amd.DemosaicAlgorithm.L7


def _check(raw, w, h, pattern, bd, packed, bl=None, per_channel=False, offset=0):
    d_in = torch.from_numpy(np.ascontiguousarray(raw).view(np.uint8).ravel().copy()).to(DEV)
    nbytes = imop_amd.rgb_byte_size(w, h, bd)
    buf = torch.full((nbytes + 32,), 0xA5, dtype=torch.uint8, device=DEV)
    d_out = buf[offset:offset + nbytes]
    e = imop_amd.demosaic(d_in, d_out, w, h, pattern, L7, bd, packed, bl, per_channel)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    got = (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)
    eo, want = O.demosaic(raw, w, h, pattern, O.L7, bd, packed, bl, per_channel)
    assert e == eo == 0
    d = np.argwhere(got != want)
    assert d.size == 0, f"{len(d)} samples differ, first {d[:4].tolist()}"
    rest = buf.cpu().numpy()
    assert (rest[:offset] == 0xA5).all() and (rest[offset + nbytes:] == 0xA5).all()


@pytest.mark.parametrize("bd,packed", [(8, False), (8, True), (10, True), (12, True), (12, False), (16, False),
                                       (6, False)])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("w,h", [(8, 8), (9, 13), (31, 17), (40, 20), (259, 11), (300, 19)])
def test_matches_oracle(bd, packed, pattern, w, h):
    print(f"l7 {w}x{h} bd {bd} packed {packed} pattern {pattern}", flush=True)
    vals = np.random.default_rng(w * 5 + h + bd + pattern).integers(0, 1 << bd, size=(h, w))
    _check(scene.raw_bytes(vals, bd, packed), w, h, pattern, bd, packed)


@pytest.mark.parametrize("offset", [2, 6, 10])
def test_unaligned_output(offset):
    """Row segments that do not start on 16 bytes take the head / tail stores."""
    w, h = 300, 20
    vals = np.random.default_rng(offset).integers(0, 4096, size=(h, w))
    _check(scene.raw_bytes(vals, 12, True), w, h, 1, 12, True, offset=offset)


@pytest.mark.parametrize("bd", [8, 12])
def test_black_level(bd):
    w, h = 96, 40
    vals = np.random.default_rng(bd).integers(0, 1 << bd, size=(h, w))
    _check(scene.raw_bytes(vals, bd, bd > 8), w, h, 2, bd, bd > 8, [10.0, 20.0, 30.0, 40.0], True)


def test_values_above_the_bit_depth_maximum():
    w, h = 64, 12
    vals = np.random.default_rng(3).integers(0, 256, size=(h, w)).astype(np.uint8)
    _check(scene.raw_bytes(vals, 8, False), w, h, 0, 5, False)


def test_full_frame_12bit_packed():
    w, h, bd = 4024, 3032, 12
    print("l7 full frame", flush=True)
    vals = scene.make_scene(h, w, bd, 0, seed=2)
    _check(scene.raw_bytes(vals, bd, True), w, h, 0, bd, True, [64.0] * 4)


def test_minimum_size():
    d_in = torch.zeros(128, dtype=torch.uint8, device=DEV)
    d_out = torch.zeros(512, dtype=torch.uint8, device=DEV)
    assert imop_amd.demosaic(d_in, d_out, 7, 9, 0, L7, 8) == 5     # l7.cpp:69
    assert imop_amd.demosaic(d_in, d_out, 9, 7, 0, L7, 8) == 5
