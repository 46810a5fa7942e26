"""PRISM / RCD planes under concurrency and graph capture.

Each PRISM / RCD call takes its four FP32 planes as its own stream-ordered
allocation (imop_hip.cpp: planes_alloc), so:
* calls on two streams with different frames, enqueued with no
  synchronisation between them, each match the oracle;
* synchronous host-buffer calls from two host threads at once each match;
* a graph captured over a small frame still matches on replays after an
  eager call on a larger frame (which, with one shared per-device buffer,
  grew and freed the buffer the graph pointed at).
The direct batch call refuses a capturing stream (its work list is copied
from the call's host memory); graphs use the batch plan.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
PRISM = imop_amd.DemosaicAlgorithm.PRISM
RCD = imop_amd.DemosaicAlgorithm.RCD

# (w, h, bd, packed, algorithm, pattern, seed)
FRAMES = [(200, 120, 12, True, PRISM, 0, 1), (257, 131, 10, True, RCD, 1, 2), (130, 98, 8, False, PRISM, 2, 3),
          (322, 180, 16, False, RCD, 3, 4)]


def _raw(f):
    w, h, bd, packed, algo, pattern, seed = f
    return scene.raw_bytes(scene.make_scene(h, w, bd, pattern, seed=seed, n_lines=8), bd, packed)


def _want(f, raw):
    w, h, bd, packed, algo, pattern, seed = f
    e, rgb = O.demosaic(raw, w, h, pattern, int(algo), bd, packed, [16.0] * 4)
    assert e == 0
    return rgb


def _as_rgb(t, f):
    w, h, bd = f[0], f[1], f[2]
    out = t.cpu().numpy()
    return (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)


def test_two_streams_no_sync_between():
    raws = [_raw(f) for f in FRAMES]
    wants = [_want(f, r) for f, r in zip(FRAMES, raws)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    d_in = [torch.from_numpy(r.copy()).to(DEV) for r in raws]
    outs = [[torch.full((imop_amd.rgb_byte_size(f[0], f[1], f[2]),), 0xA5, dtype=torch.uint8, device=DEV)
             for f in FRAMES] for _ in range(3)]
    torch.cuda.synchronize()
    # frame k on stream k % 2, three rounds, all enqueued before any wait
    for r in range(3):
        for k, f in enumerate(FRAMES):
            w, h, bd, packed, algo, pattern, _ = f
            e = imop_amd.demosaic(d_in[k], outs[r][k], w, h, pattern, algo, bd, packed, [16.0] * 4,
                                  stream=streams[(k + r) % 2])
            assert e == 0
    torch.cuda.synchronize()
    for r in range(3):
        for k, f in enumerate(FRAMES):
            print(f"round {r} frame {k} {f[:5]}", flush=True)
            np.testing.assert_array_equal(_as_rgb(outs[r][k], f), wants[k])


def test_two_host_threads():
    raws = [_raw(f) for f in FRAMES]
    wants = [_want(f, r) for f, r in zip(FRAMES, raws)]
    errors = []

    def worker(idx):
        try:
            for it in range(4):
                k = (idx + 2 * it) % len(FRAMES)
                w, h, bd, packed, algo, pattern, _ = FRAMES[k]
                e, rgb = imop_amd.demosaic_host(raws[k], w, h, pattern, algo, bd, packed, [16.0] * 4)
                if e != 0 or not np.array_equal(rgb, wants[k]):
                    errors.append((idx, it, k, int(e)))
        except Exception as exc:  # noqa: BLE001
            errors.append((idx, repr(exc)))

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors


def test_two_threads_own_streams_device_buffers():
    raws = [_raw(f) for f in FRAMES]
    wants = [_want(f, r) for f, r in zip(FRAMES, raws)]
    d_in = [torch.from_numpy(r.copy()).to(DEV) for r in raws]
    torch.cuda.synchronize()
    errors = []

    def worker(idx):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for it in range(3):
                    k = (idx + it) % len(FRAMES)
                    f = FRAMES[k]
                    w, h, bd, packed, algo, pattern, _ = f
                    out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
                    e = imop_amd.demosaic(d_in[k], out, w, h, pattern, algo, bd, packed, [16.0] * 4, stream=s)
                    s.synchronize()
                    if e != 0 or not np.array_equal(_as_rgb(out, f), wants[k]):
                        errors.append((idx, it, k, int(e)))
        except Exception as exc:  # noqa: BLE001
            errors.append((idx, repr(exc)))

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors


def test_graph_survives_larger_eager_call():
    small = (96, 64, 12, True, PRISM, 0, 11)
    large = (1024, 768, 12, True, RCD, 1, 12)
    r_small, r_large = _raw(small), _raw(large)
    w_small, w_large = _want(small, r_small), _want(large, r_large)
    d_small = torch.from_numpy(r_small.copy()).to(DEV)
    d_large = torch.from_numpy(r_large.copy()).to(DEV)
    o_small = torch.full((imop_amd.rgb_byte_size(96, 64, 12),), 0xA5, dtype=torch.uint8, device=DEV)
    o_large = torch.full((imop_amd.rgb_byte_size(1024, 768, 12),), 0xA5, dtype=torch.uint8, device=DEV)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        assert imop_amd.demosaic(d_small, o_small, 96, 64, 0, PRISM, 12, True, [16.0] * 4) == 0
    for rep in range(3):
        o_small.fill_(0xA5)
        o_large.fill_(0xA5)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"replay {rep}", flush=True)
        np.testing.assert_array_equal(_as_rgb(o_small, small), w_small)
        # eager work on a larger frame between replays
        assert imop_amd.demosaic(d_large, o_large, 1024, 768, 1, RCD, 12, True, [16.0] * 4) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_as_rgb(o_large, large), w_large)
    # replay right behind an eager call on the same stream, no wait between
    o_small.fill_(0xA5)
    assert imop_amd.demosaic(d_large, o_large, 1024, 768, 1, RCD, 12, True, [16.0] * 4) == 0
    g.replay()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_as_rgb(o_small, small), w_small)
    np.testing.assert_array_equal(_as_rgb(o_large, large), w_large)
    del g


def test_direct_batch_refuses_capture():
    w, h, bd = 96, 64, 12
    raw = _raw((w, h, bd, True, PRISM, 0, 5))
    d_in = torch.from_numpy(raw.copy()).to(DEV)
    d_out = torch.zeros(imop_amd.rgb_byte_size(w, h, bd), dtype=torch.uint8, device=DEV)
    d_out2 = torch.zeros_like(d_out)
    desc = [imop_amd.frame_desc(d_in, d_out, w, h, 0, imop_amd.DemosaicAlgorithm.MG, bd, True)]
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        e, failed = imop_amd.demosaic_batch(desc)
        assert imop_amd.demosaic(d_in, d_out2, w, h, 0, imop_amd.DemosaicAlgorithm.MG, bd, True) == 0
    assert e == 6
    g.replay()
    torch.cuda.synchronize()
    assert imop_amd.demosaic_batch(desc) == (0, -1)
    torch.cuda.synchronize()
    assert torch.equal(d_out, d_out2)
