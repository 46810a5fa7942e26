"""Host-buffer batches through pinned, double-buffered staging
(imop_hip_demosaic_batch_host): ragged frames against the oracle, long runs
of chunks against the single-frame host entry point, unaligned host buffers,
invalid frames (nothing written), repeated calls that grow the staging, and
bytes past each frame's output left alone."""
from __future__ import annotations

import numpy as np
import pytest

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

HQLI, MG, PRISM = imop_amd.DemosaicAlgorithm.HQLI, imop_amd.DemosaicAlgorithm.MG, imop_amd.DemosaicAlgorithm.PRISM


def _spec_frames(seed: int, n: int):
    """Random specs: every layout, odd and even widths, PRISM, fractional and
    per-channel black level, u16 output of 8-bit frames."""
    rng = np.random.default_rng(seed)
    layouts = [(8, False), (10, True), (12, True), (12, False), (16, False), (14, True), (10, False)]
    widths = [10, 14, 16, 38, 64, 126, 250, 256, 258, 502, 1006, 13, 31]
    specs = []
    for k in range(n):
        bd, packed = layouts[int(rng.integers(0, len(layouts)))]
        specs.append(dict(w=int(widths[int(rng.integers(0, len(widths)))]), h=int(rng.integers(10, 70)), bd=bd,
                          packed=packed, algo=[HQLI, MG, MG, PRISM][int(rng.integers(0, 4))],
                          pattern=int(rng.integers(0, 4)),
                          bl=[None, [16.0] * 4, [64.0, 3.0, 9.0, 40.0], [12.5] * 4][int(rng.integers(0, 4))],
                          out_elem=2 if (bd <= 8 and rng.integers(0, 2)) else 0, seed=seed * 100 + k))
    return specs


def _host_frames(specs, in_offset=0, out_offset=0, slack=0):
    """(spec, raw, in view, out view, out backing) with the views placed
    in_offset / out_offset bytes into larger arrays; out is 0xA5-filled and
    `slack` bytes longer than the frame's RGB."""
    frames = []
    for s in specs:
        rng = np.random.default_rng(s["seed"])
        vals = rng.integers(0, 1 << s["bd"], size=(s["h"], s["w"])).astype(np.uint8 if s["bd"] <= 8 else np.uint16)
        raw = scene.raw_bytes(vals, s["bd"], s["packed"])
        ib = np.zeros(raw.size + in_offset, np.uint8)
        ib[in_offset:] = raw
        elem = s["out_elem"] or (2 if s["bd"] > 8 else 1)
        n_out = s["w"] * s["h"] * 3 * elem
        ob = np.full(n_out + slack + out_offset, 0xA5, np.uint8)
        frames.append((s, raw, ib[in_offset:], ob[out_offset:], ob))
    return frames


def _descs(frames):
    return [imop_amd.host_frame_desc(iv, ov, s["w"], s["h"], s["pattern"], s["algo"], s["bd"], s["packed"], s["bl"],
                                     s["bl"] is not None and len(set(s["bl"])) > 1, s["out_elem"])
            for s, _, iv, ov, _ in frames]


def _check_oracle(frames, slack=0):
    for s, raw, _, ov, _ in frames:
        pc = s["bl"] is not None and len(set(s["bl"])) > 1
        e, want = O.demosaic(raw, s["w"], s["h"], s["pattern"], int(s["algo"]), s["bd"], s["packed"], s["bl"], pc)
        assert e == 0
        elem = s["out_elem"] or (2 if s["bd"] > 8 else 1)
        n_out = s["w"] * s["h"] * 3 * elem
        got = ov[:n_out].view(np.uint16 if elem == 2 else np.uint8).reshape(s["h"], s["w"], 3)
        bad = np.argwhere(got.astype(np.int64) != want.astype(np.int64))
        assert bad.size == 0, f"frame {s}: {len(bad)} samples differ, first {bad[:4].tolist()}"
        assert (ov[n_out:n_out + slack] == 0xA5).all(), f"frame {s}: bytes past the output written"


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ragged_host_batch_matches_oracle(seed):
    frames = _host_frames(_spec_frames(seed, 24), slack=37)
    print(f"stream ragged seed {seed}", flush=True)
    assert imop_amd.demosaic_batch_host(_descs(frames)) == (0, -1)
    _check_oracle(frames, slack=37)


@pytest.mark.parametrize("in_off,out_off", [(1, 0), (0, 1), (3, 5)])
def test_unaligned_host_buffers(in_off, out_off):
    frames = _host_frames(_spec_frames(7, 12), in_offset=in_off, out_offset=out_off)
    print(f"stream unaligned {in_off} {out_off}", flush=True)
    assert imop_amd.demosaic_batch_host(_descs(frames)) == (0, -1)
    _check_oracle(frames)


def test_many_chunks_match_single_frame_entry():
    """Alternating 12 MP and 1006x758 frames: the 12 MP frames each fill a
    chunk, so chunks of every size pass through both staging slots."""
    rng = np.random.default_rng(5)
    cases = []
    for k in range(9):
        w, h = (4024, 3032) if k % 3 == 1 else (1006, 758)
        bd, packed = [(12, True), (10, True), (8, False)][k % 3]
        vals = rng.integers(0, 1 << bd, size=(h, w)).astype(np.uint8 if bd <= 8 else np.uint16)
        raw = scene.raw_bytes(vals, bd, packed)
        out = np.full(imop_amd.rgb_byte_size(w, h, bd), 0xA5, np.uint8)
        cases.append((w, h, bd, packed, raw, out, [64.0] * 4 if k % 2 else None))
    descs = [imop_amd.host_frame_desc(raw, out, w, h, 0, MG, bd, packed, bl)
             for w, h, bd, packed, raw, out, bl in cases]
    print("stream 9 frames, 3 at 12 MP", flush=True)
    assert imop_amd.demosaic_batch_host(descs) == (0, -1)
    for w, h, bd, packed, raw, out, bl in cases:
        e, want = imop_amd.demosaic_host(raw, w, h, 0, MG, bd, packed, bl)
        assert e == 0
        np.testing.assert_array_equal(out.view(np.uint16 if bd > 8 else np.uint8).reshape(h, w, 3), want)


def test_invalid_frame_writes_nothing():
    frames = _host_frames(_spec_frames(21, 6))
    descs = _descs(frames)
    descs[3].height = 4                     # below every algorithm's minimum
    assert imop_amd.demosaic_batch_host(descs) == (int(imop_amd.DemosaicError.ImageTooSmall), 3)
    descs = _descs(frames)
    descs[4].in_bytes -= 1                  # input shorter than the frame
    assert imop_amd.demosaic_batch_host(descs) == (int(imop_amd.DemosaicError.InternalError), 4)
    descs = _descs(frames)
    descs[1].in_ = None
    assert imop_amd.demosaic_batch_host(descs) == (int(imop_amd.DemosaicError.NullInput), 1)
    for _, _, _, ov, _ in frames:
        assert (ov == 0xA5).all()
    assert imop_amd.demosaic_batch_host([]) == (0, -1)


def test_repeated_calls_grow_and_reuse_staging():
    """A small batch, a large one (the staging grows), then small again: every
    call's outputs are its own."""
    for seed, n in ((31, 4), (32, 40), (33, 3)):
        frames = _host_frames(_spec_frames(seed, n))
        print(f"stream call seed {seed} n {n}", flush=True)
        assert imop_amd.demosaic_batch_host(_descs(frames)) == (0, -1)
        _check_oracle(frames)
