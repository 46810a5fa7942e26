"""The GigE Vision 10-bit packing (is_packed = 2; 2 px in 3 bytes) in the
oracle, pinned against the reference's own GigE-packed frame
(tests/golden/rg10_gige.npz, row crops of test_data/*_pBayerRG10_Packed.raw):
its byte layout, the placement of the high and low bits (a planted swap of
the low-bit fields is detected), the 10-bit scale against the unpacked RG10
frame of the same camera, round trips, and demosaic of a GigE frame being
the reference's demosaic of the decoded samples."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as O
from imop_amd import scene

W = 1006
GIGE = 2
FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "rg10_gige.npz"))


def _rough(v):
    """Mean |v - mean of the same-colour horizontal neighbours|."""
    v = v.astype(np.float64)
    return float(np.abs(v[:, 2:-2] - (v[:, :-4] + v[:, 4:]) / 2).mean())


def _swapped_low_bits(rows):
    """A planted mistake: the two 2-bit fields of the middle byte exchanged."""
    g = rows.reshape(-1, 3).astype(np.int32)
    v = np.empty(g.shape[0] * 2, np.int32)
    v[0::2] = (g[:, 0] << 2) | ((g[:, 1] >> 4) & 3)
    v[1::2] = (g[:, 2] << 2) | (g[:, 1] & 3)
    return v.reshape(rows.shape[0], W)


@pytest.mark.parametrize("key", ["top", "bottom"])
def test_fixture_has_the_gige_layout(key):
    rows = FIX[key]
    assert rows.shape[1] == W // 2 * 3
    # only bits 0-1 and 4-5 of every middle byte carry data
    assert not (rows.reshape(-1, 3)[:, 1] & 0xCC).any()


@pytest.mark.parametrize("key", ["top", "bottom"])
def test_high_bits_are_the_outer_bytes(key):
    rows = FIX[key]
    h = rows.shape[0]
    v = O.unpack(rows, W, h, 10, GIGE).astype(np.int32)
    g = rows.reshape(-1, 3)
    assert np.array_equal(v.ravel()[0::2] >> 2, g[:, 0]) and np.array_equal(v.ravel()[1::2] >> 2, g[:, 2])
    assert v.max() <= 1023


@pytest.mark.parametrize("key", ["top", "bottom"])
def test_low_bit_fields_follow_the_frame(key):
    """The decoded frame is smoother than with the low-bit fields swapped, and
    far smoother than the reference's own 10-bit rule (4 px / 5 B) reads it."""
    rows = FIX[key]
    h = rows.shape[0]
    v = O.unpack(rows, W, h, 10, GIGE)
    ref_rule = O.unpack(rows, W, h, 10, True)
    assert _rough(v) < _rough(_swapped_low_bits(rows)) - 0.05
    assert _rough(ref_rule) > 5 * _rough(v)


def test_scale_matches_the_unpacked_rg10_frame():
    """Same camera, same rows, a minute apart: the GigE frame's decoded means
    per CFA parity are within 3 % of the unpacked RG10 frame's (dropping the
    low bits' shift would put them 4x off)."""
    v = O.unpack(FIX["top"], W, 64, 10, GIGE).astype(np.float64)
    u = FIX["u16_top"].astype(np.float64)
    for py in range(2):
        for px in range(2):
            a, b = v[py::2, px::2].mean(), u[py::2, px::2].mean()
            assert abs(a - b) < 0.03 * b, (py, px, a, b)


@pytest.mark.parametrize("shape", [(W, 64), (37, 23), (6, 6), (11, 9)])
def test_round_trips(shape):
    w, h = shape
    rng = np.random.default_rng(w * h)
    vals = rng.integers(0, 1024, size=(h, w)).astype(np.uint16)
    packed = scene.pack_gige10(vals)
    assert packed.size == (w * h + 1) // 2 * 3
    assert np.array_equal(O.unpack(packed, w, h, 10, GIGE), vals)
    rows = FIX["top"]
    assert np.array_equal(scene.pack_gige10(O.unpack(rows, W, 64, 10, GIGE)), rows.ravel())


@pytest.mark.parametrize("algo", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("bl", [None, [64.0] * 4, [64.0, 3.0, 9.0, 40.0]])
def test_demosaic_of_gige_frame_is_demosaic_of_decoded_samples(algo, bl):
    rows = FIX["top"]
    u16 = O.unpack(rows, W, 64, 10, GIGE)
    pc = bl is not None and len(set(bl)) > 1
    e1, a = O.demosaic(rows, W, 64, 0, algo, 10, GIGE, bl, pc)
    e2, b = O.demosaic(scene.raw_bytes(u16, 10, False), W, 64, 0, algo, 10, False, bl, pc)
    assert e1 == e2 == 0
    assert np.array_equal(a, b)
    if O.reference_available() and bl is None:
        e3, c = O.ref_demosaic(scene.raw_bytes(u16, 10, False), W, 64, 0, algo, 10, False)
        assert e3 == 0 and np.array_equal(a, c)


def test_gige_packing_is_10_bit_only():
    vals = np.zeros((8, 8), np.uint16)
    for bd in (8, 12, 16):
        e, _ = O.demosaic(scene.pack_gige10(vals), 8, 8, 0, 2, bd, GIGE)
        assert e == 3
