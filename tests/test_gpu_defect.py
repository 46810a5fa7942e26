"""Adaptive defect correction on the GPU (imop_hip_defect_adaptive) against the
oracle (oracle/isp_oracle.c isp_defect_adaptive, pinned against the reference's
compiled process_adaptive): byte-identical frames for every bit depth, frame
shape and defect density, including frames whose corrections chain across
many tiles (pure noise at the lowest threshold), and the reference's error
codes."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _img(kind, bd, h, w, seed):
    rng = np.random.default_rng(seed)
    mx = (1 << bd) - 1
    if kind == "noise":
        img = rng.integers(0, mx + 1, size=(h, w))
    elif kind == "sparse":
        yy, xx = np.mgrid[0:h, 0:w]
        img = (mx // 4 + (xx + 2 * yy) % max(1, mx // 8)) + rng.integers(0, max(1, mx // 64) + 1, size=(h, w))
        img[rng.random((h, w)) < 0.002] = mx
        img[rng.random((h, w)) < 0.001] = 0
    elif kind == "clusters":
        img = rng.integers(0, mx // 2 + 1, size=(h, w)) + (rng.random((h, w)) < 0.3) * (mx // 2)
    else:
        img = np.full((h, w), mx // 3)
    return np.clip(img, 0, mx).astype(np.uint8 if bd <= 8 else np.uint16)


def _run(img, bd, thr, pattern=0, stream=None):
    h, w = img.shape
    d = torch.from_numpy(img.view(np.uint8).ravel().copy()).to(DEV)
    e = imop_amd.defect_correct(d, w, h, pattern, bd, thr, stream=stream)
    assert e == 0
    out = d.cpu().numpy()
    return (out.view(np.uint16) if bd > 8 else out).reshape(h, w)


def _check(img, bd, thr, **kw):
    e, want = O.defect_adaptive(img, bd, thr)
    assert e == 0
    np.testing.assert_array_equal(_run(img, bd, thr, **kw), want)


@pytest.mark.parametrize("bd", [1, 8, 10, 12, 16])
@pytest.mark.parametrize("kind", ["noise", "sparse", "clusters", "flat"])
@pytest.mark.parametrize("thr", [0.3, 0.05, 1.0, float("nan")])
def test_matches_oracle(bd, kind, thr):
    h, w = 70, 301   # 3 x 3 tiles, odd width
    print(f"defect bd {bd} {kind} thr {thr}", flush=True)
    _check(_img(kind, bd, h, w, bd * 11 + len(kind)), bd, thr)


@pytest.mark.parametrize("h,w", [(5, 5), (5, 6), (6, 5), (7, 9), (5, 300), (300, 5), (8, 8), (33, 129),
                                 (32, 128), (34, 130), (1000, 7), (64, 1030)])
def test_shapes(h, w):
    print(f"defect shape {w}x{h}", flush=True)
    for bd, kind, thr in ((8, "noise", 0.05), (12, "clusters", 0.1), (16, "sparse", 0.3)):
        _check(_img(kind, bd, h, w, h * 1000 + w), bd, thr)


@pytest.mark.parametrize("pattern", [0, 1, 2, 3, 7])
def test_pattern_is_unused(pattern):
    img = _img("clusters", 12, 40, 200, 3)
    _check(img, 12, 0.1, pattern=pattern)


def test_reference_unit_tests():
    """tests/unit/test_defect_correct.cpp:23-72 of the reference, on the GPU."""
    img = np.full((8, 8), 50, np.uint8)
    img[4, 4] = 255
    assert _run(img, 8, 0.3)[4, 4] < 200
    assert np.abs(_run(np.full((8, 8), 128, np.uint8), 8, 0.3).astype(int) - 128).max() <= 5
    ramp = (np.arange(256) % 256).astype(np.uint8).reshape(16, 16)
    _check(ramp, 8, 0.3)


@pytest.mark.parametrize("kind,thr", [("sparse", 0.3), ("noise", 0.05)])
def test_full_frame(kind, thr):
    """A 12 MP 12-bit frame: sparse hot / cold pixels, and pure noise at the
    lowest threshold (corrections chained across the frame)."""
    w, h = 4024, 3032
    print(f"defect full frame {kind}", flush=True)
    _check(_img(kind, 12, h, w, 17), 12, thr)


def test_repeated_calls_and_stream():
    """Calls of different sizes in a row reuse and grow the scratch; a caller's
    stream works."""
    s = torch.cuda.Stream()
    for (h, w), seed in (((300, 500), 1), ((40, 60), 2), ((600, 900), 3), ((300, 500), 4)):
        img = _img("clusters", 10, h, w, seed)
        with torch.cuda.stream(s):
            _check(img, 10, 0.1, stream=s)


@pytest.mark.parametrize("bd,offset", [(8, 1), (12, 2)])
def test_unaligned_buffer(bd, offset):
    """A frame that starts off a pixel-pair boundary takes the per-pixel loads."""
    h, w = 70, 300
    img = _img("clusters", bd, h, w, 5)
    raw = img.view(np.uint8).ravel()
    buf = torch.zeros(raw.size + 8, dtype=torch.uint8, device=DEV)
    buf[offset:offset + raw.size] = torch.from_numpy(raw.copy()).to(DEV)
    d = buf[offset:offset + raw.size]
    assert imop_amd.defect_correct(d, w, h, 0, bd, 0.1) == 0
    out = d.cpu().numpy()
    got = (out.view(np.uint16) if bd > 8 else out).reshape(h, w)
    np.testing.assert_array_equal(got, O.defect_adaptive(img, bd, 0.1)[1])
    assert not buf[:offset].any() and not buf[offset + raw.size:].any()


def test_invalid_arguments():
    d = torch.zeros(64 * 64 * 2, dtype=torch.uint8, device=DEV)
    L = imop_amd.lib()
    assert L.imop_hip_defect_adaptive(None, 0, 64, 64, 0, 12, ctypes.c_float(0.3), None) == 1   # NullInput
    assert imop_amd.defect_correct(d, 0, 64, 0, 12) == 2                                         # InvalidDimensions
    assert imop_amd.defect_correct(d, 64, -1, 0, 12) == 2
    assert imop_amd.defect_correct(d, 64, 64, 0, 0) == 3                                         # InvalidBitDepth
    assert imop_amd.defect_correct(d, 64, 64, 0, 17) == 3
    assert imop_amd.defect_correct(d, 4, 64, 0, 12) == 5                                         # ImageTooSmall
    assert imop_amd.defect_correct(d, 64, 4, 0, 12) == 5
    assert imop_amd.defect_correct(d, 64, 65, 0, 12) == 6                                        # buffer too small
    assert not d.any()


# ---------------------------------------------------------------- stream-ordered form

def _run_async(img, bd, thr, max_passes, stream=None):
    h, w = img.shape
    d = torch.from_numpy(img.view(np.uint8).ravel().copy()).to(DEV)
    status = torch.full((1,), -1, dtype=torch.int32, device=DEV)
    assert imop_amd.defect_correct_async(d, w, h, 0, bd, status, thr, max_passes, stream=stream) == 0
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    return (out.view(np.uint16) if bd > 8 else out).reshape(h, w), int(status.item())


@pytest.mark.parametrize("bd", [8, 12, 16])
@pytest.mark.parametrize("kind,thr", [("noise", 0.05), ("sparse", 0.3), ("clusters", 0.1), ("flat", 0.3)])
def test_async_matches_oracle(bd, kind, thr):
    """Enough passes: status 0 and the oracle's frame, byte for byte."""
    h, w = 70, 301
    img = _img(kind, bd, h, w, bd * 7 + len(kind))
    print(f"defect async bd {bd} {kind} thr {thr}", flush=True)
    got, status = _run_async(img, bd, thr, 256)
    assert status == 0
    np.testing.assert_array_equal(got, O.defect_adaptive(img, bd, thr)[1])


def test_async_reports_unconverged_frames():
    """Pure noise at the lowest threshold needs several passes: one pass
    reports 1; the passes the synchronous form takes report 0."""
    img = _img("noise", 12, 96, 256, 3)
    want = O.defect_adaptive(img, 12, 0.05)[1]
    got, status = _run_async(img, 12, 0.05, 1)
    assert status == 1
    assert not np.array_equal(got, want)
    got, status = _run_async(img, 12, 0.05, 512)
    assert status == 0
    np.testing.assert_array_equal(got, want)


def test_async_full_frame_noise():
    """12 MP 12-bit pure noise at threshold 0.05 within 64 passes."""
    img = _img("noise", 12, 3032, 4024, 17)
    print("defect async full frame noise", flush=True)
    got, status = _run_async(img, 12, 0.05, 64)
    assert status == 0
    np.testing.assert_array_equal(got, O.defect_adaptive(img, 12, 0.05)[1])


def test_async_two_streams_no_sync():
    """Two frames corrected on two streams with no synchronisation between the
    calls: each call's scratch is its own."""
    imgs = [_img("clusters", 12, 300, 500, 1), _img("noise", 12, 320, 480, 2)]
    thrs = [0.1, 0.05]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    ds, sts = [], []
    for img, thr, s in zip(imgs, thrs, streams):
        h, w = img.shape
        d = torch.from_numpy(img.view(np.uint8).ravel().copy()).to(DEV)
        st = torch.full((1,), -1, dtype=torch.int32, device=DEV)
        torch.cuda.synchronize()
        assert imop_amd.defect_correct_async(d, w, h, 0, 12, st, thr, 128, stream=s) == 0
        ds.append(d)
        sts.append(st)
    torch.cuda.synchronize()
    for img, thr, d, st in zip(imgs, thrs, ds, sts):
        assert int(st.item()) == 0
        got = d.cpu().numpy().view(np.uint16).reshape(img.shape)
        np.testing.assert_array_equal(got, O.defect_adaptive(img, 12, thr)[1])


def test_async_graph_capture_replays():
    """Input copy + stream-ordered correction captured into a graph, replayed
    three times with eager work between the replays; every replay equals the
    oracle and reports 0."""
    img = _img("clusters", 12, 200, 333, 9)
    h, w = img.shape
    want = O.defect_adaptive(img, 12, 0.1)[1]
    src = torch.from_numpy(img.view(np.uint8).ravel().copy()).to(DEV)
    d = torch.empty_like(src)
    st = torch.zeros((1,), dtype=torch.int32, device=DEV)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        d.copy_(src)
        assert imop_amd.defect_correct_async(d, w, h, 0, 12, st, 0.1, 64) == 0
    for rep in range(3):
        print(f"defect graph replay {rep}", flush=True)
        d.fill_(0x33)
        st.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(st.item()) == 0
        np.testing.assert_array_equal(d.cpu().numpy().view(np.uint16).reshape(h, w), want)


def test_async_invalid_arguments():
    d = torch.zeros(64 * 64 * 2, dtype=torch.uint8, device=DEV)
    st = torch.zeros((1,), dtype=torch.int32, device=DEV)
    L = imop_amd.lib()
    f = ctypes.c_float(0.3)
    assert L.imop_hip_defect_adaptive_async(None, 0, 64, 64, 0, 12, f, 8, ctypes.c_void_p(st.data_ptr()), None) == 1
    assert imop_amd.defect_correct_async(d, 0, 64, 0, 12, st) == 2
    assert imop_amd.defect_correct_async(d, 64, 64, 0, 17, st) == 3
    assert imop_amd.defect_correct_async(d, 4, 64, 0, 12, st) == 5
    assert imop_amd.defect_correct_async(d, 64, 65, 0, 12, st) == 6
    assert L.imop_hip_defect_adaptive_async(ctypes.c_void_p(d.data_ptr()), d.numel(), 64, 64, 0, 12, f, 8, None,
                                            None) == 1
    assert imop_amd.defect_correct_async(d, 64, 64, 0, 12, st, max_passes=0) == 6
    assert imop_amd.defect_correct_async(d, 64, 64, 0, 12, st, max_passes=65537) == 6
    torch.cuda.synchronize()
    assert not d.any() and not st.any()
