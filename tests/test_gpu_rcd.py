"""RCD (src/algorithms/rcd.cpp:7-248) on the GPU against the oracle, byte for
byte.  The oracle's RCD is pinned against the compiled reference
(tests/test_oracle.py); RCD shares PRISM's two LDS-tiled passes without the
Laplacian term and the colour-difference median, so the float operation order
and fused multiply-adds are the ones PRISM's bit-exactness already holds to."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
RCD = imop_amd.DemosaicAlgorithm.RCD


def _gpu(raw, w, h, pattern, bd, packed, bl=None, per_channel=False):
    d_in = torch.from_numpy(np.ascontiguousarray(raw).view(np.uint8).ravel().copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
    e = imop_amd.demosaic(d_in, d_out, w, h, pattern, RCD, bd, packed, bl, per_channel)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    return e, (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)


def _check(raw, w, h, pattern, bd, packed, bl=None, per_channel=False):
    e, got = _gpu(raw, w, h, pattern, bd, packed, bl, per_channel)
    eo, want = O.demosaic(raw, w, h, pattern, O.RCD, bd, packed, bl, per_channel)
    assert e == eo == 0
    d = np.argwhere(got != want)
    assert d.size == 0, f"{len(d)} samples differ, first {d[:4].tolist()}"


@pytest.mark.parametrize("bd,packed", [(8, False), (10, True), (12, True), (12, False), (14, False), (16, False)])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("w,h", [(10, 10), (13, 11), (67, 45), (130, 66)])
def test_matches_oracle(bd, packed, pattern, w, h):
    print(f"rcd {w}x{h} bd {bd} packed {packed} pattern {pattern}", flush=True)
    vals = scene.make_scene(h, w, bd, pattern, seed=w + h + bd + pattern)
    _check(scene.raw_bytes(vals, bd, packed), w, h, pattern, bd, packed)


@pytest.mark.parametrize("bd", [8, 12, 16])
@pytest.mark.parametrize("per_channel", [False, True])
def test_noise_and_black_level(bd, per_channel):
    w, h = 200, 120
    vals = np.random.default_rng(bd).integers(0, 1 << bd, size=(h, w))
    bl = [10.0, 20.0, 30.0, 40.0] if per_channel else [48.0] * 4
    _check(scene.raw_bytes(vals, bd, False), w, h, 2, bd, False, bl, per_channel)


def test_full_frame_10bit_packed():
    """4024 x 3032 10-bit packed (the PRISM workload's frame), black level 64."""
    w, h, bd = 4024, 3032, 10
    print("rcd full frame", flush=True)
    vals = scene.make_scene(h, w, bd, 0, seed=31, n_lines=64)
    _check(scene.raw_bytes(vals, bd, True), w, h, 0, bd, True, [64.0] * 4)


def test_size_limits_and_host_entry():
    d_in = torch.zeros(64 * 64 * 2, dtype=torch.uint8, device=DEV)
    d_out = torch.zeros(64 * 64 * 6, dtype=torch.uint8, device=DEV)
    assert imop_amd.demosaic(d_in, d_out, 9, 12, 0, RCD, 8) == 5     # rcd.cpp:11
    assert imop_amd.demosaic(d_in, d_out, 12, 9, 0, RCD, 8) == 5
    w, h = 96, 64
    raw = scene.raw_bytes(scene.make_scene(h, w, 12, 1, seed=4), 12, True)
    e1, r1 = _gpu(raw, w, h, 1, 12, True, [32.0] * 4)
    e2, r2 = imop_amd.demosaic_host(raw, w, h, 1, RCD, 12, True, [32.0] * 4)
    assert e1 == e2 == 0
    np.testing.assert_array_equal(r1, r2)


def test_batch_with_rcd_frames():
    """RCD frames in a ragged batch (run as their own launches) next to MG."""
    frames = [(96, 64, 12, True, RCD), (130, 66, 10, True, imop_amd.DemosaicAlgorithm.MG), (67, 45, 8, False, RCD)]
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


def test_epilogue_after_rcd():
    """White balance, colour matrix and gamma after RCD run as the separate pass."""
    w, h, bd = 130, 66, 12
    raw = scene.raw_bytes(scene.make_scene(h, w, bd, 0, seed=8), bd, True)
    gains, ccm = [1.9, 0.8, 2.3], [1.6, -0.4, -0.2, -0.3, 1.5, -0.2, 0.05, -0.55, 1.5]
    ep = imop_amd.rgb_epilogue(gains=gains, ccm=ccm, gamma=2.2)
    d_in = torch.from_numpy(raw.copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
    assert imop_amd.demosaic_isp(d_in, d_out, w, h, 0, RCD, bd, True, [64.0] * 4, epilogue=ep) == 0
    torch.cuda.synchronize()
    got = d_out.cpu().numpy().view(np.uint16).reshape(h, w, 3)
    e, rgb = O.demosaic(raw, w, h, 0, O.RCD, bd, True, [64.0] * 4)
    assert e == 0
    np.testing.assert_array_equal(got, O.isp_epilogue(rgb, bd, 7, gains, ccm, 2.2))
