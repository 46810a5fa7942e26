"""AMAZE (src/algorithms/amaze.cpp:7-200) on the GPU against the oracle, byte
for byte: the gradient-weighted green with its -1 sentinel outside the
interior, the chroma differences that read it, the intermediate border fill,
and the 8-bit core's extra output border fill.  The oracle is pinned against
the compiled reference (tests/test_oracle.py)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
AMAZE = imop_amd.DemosaicAlgorithm.AMAZE


def _check(raw, w, h, pattern, bd, packed, bl=None, per_channel=False):
    d_in = torch.from_numpy(np.ascontiguousarray(raw).view(np.uint8).ravel().copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
    e = imop_amd.demosaic(d_in, d_out, w, h, pattern, AMAZE, bd, packed, bl, per_channel)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    got = (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)
    eo, want = O.demosaic(raw, w, h, pattern, O.AMAZE, bd, packed, bl, per_channel)
    assert e == eo == 0
    d = np.argwhere(got != want)
    assert d.size == 0, f"{len(d)} samples differ, first {d[:4].tolist()}"


@pytest.mark.parametrize("bd,packed", [(8, False), (8, True), (10, True), (12, True), (12, False), (16, False),
                                       (6, False)])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("w,h", [(6, 6), (7, 9), (33, 17), (64, 32), (70, 35)])
def test_matches_oracle(bd, packed, pattern, w, h):
    print(f"amaze {w}x{h} bd {bd} packed {packed} pattern {pattern}", flush=True)
    rng = np.random.default_rng(w * 3 + h + bd + pattern)
    vals = rng.integers(0, 1 << bd, size=(h, w)) if (w + pattern) % 2 else scene.make_scene(h, w, bd, pattern, seed=w)
    _check(scene.raw_bytes(vals, bd, packed), w, h, pattern, bd, packed)


@pytest.mark.parametrize("bd", [8, 12])
def test_black_level(bd):
    w, h = 96, 40
    vals = np.random.default_rng(bd).integers(0, 1 << bd, size=(h, w))
    _check(scene.raw_bytes(vals, bd, bd > 8), w, h, 3, bd, bd > 8, [10.0, 20.0, 30.0, 40.0], True)


def test_values_above_the_bit_depth_maximum():
    w, h = 40, 12
    vals = np.random.default_rng(4).integers(0, 256, size=(h, w)).astype(np.uint8)
    _check(vals.ravel(), w, h, 1, 5, False)


def test_full_frame_12bit_packed():
    w, h, bd = 4024, 3032, 12
    print("amaze full frame", flush=True)
    vals = scene.make_scene(h, w, bd, 0, seed=2)
    _check(scene.raw_bytes(vals, bd, True), w, h, 0, bd, True, [64.0] * 4)


def test_batch_with_amaze_frames():
    """AMAZE frames in a ragged batch share the plan's workspace with PRISM."""
    A = imop_amd.DemosaicAlgorithm
    frames = [(96, 64, 12, True, AMAZE), (130, 66, 10, True, A.PRISM), (67, 45, 8, False, AMAZE)]
    keep, descs, wants = [], [], []
    for k, (w, h, bd, packed, algo) in enumerate(frames):
        raw = scene.raw_bytes(scene.make_scene(h, w, bd, k, seed=k), bd, packed)
        d_in = torch.from_numpy(raw.copy()).to(DEV)
        d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
        keep += [d_in, d_out]
        descs.append(imop_amd.frame_desc(d_in, d_out, w, h, k, algo, bd, packed))
        wants.append(O.demosaic(raw, w, h, k, int(algo), bd, packed)[1])
    assert imop_amd.demosaic_batch(descs) == (0, -1)
    torch.cuda.synchronize()
    for k, (w, h, bd, packed, algo) in enumerate(frames):
        out = keep[2 * k + 1].cpu().numpy()
        np.testing.assert_array_equal((out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3), wants[k])


def test_minimum_size():
    d_in = torch.zeros(128, dtype=torch.uint8, device=DEV)
    d_out = torch.zeros(512, dtype=torch.uint8, device=DEV)
    assert imop_amd.demosaic(d_in, d_out, 5, 8, 0, AMAZE, 8) == 5     # amaze.cpp:162
    assert imop_amd.demosaic(d_in, d_out, 8, 5, 0, AMAZE, 8) == 5
