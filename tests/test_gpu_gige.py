"""GigE Vision 10-bit packed input (is_packed = 2) on the GPU through the C
ABI: the reference's own GigE-packed frame (row crops) and random frames,
every algorithm, with and without black level, odd pixel counts, the host,
batch and host-batch entry points, a full 12 MP frame against the same frame
given unpacked, and the bit-depth check."""
from __future__ import annotations

import os

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
GIGE = imop_amd.PACKED_GIGE10
W = 1006
FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "rg10_gige.npz"))
ALGOS = list(imop_amd.DemosaicAlgorithm)


def _gpu(raw, w, h, pattern, algo, bl=None, pc=False):
    d_in = torch.from_numpy(np.ascontiguousarray(raw).view(np.uint8).ravel().copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, 10),), 0xA5, dtype=torch.uint8, device=DEV)
    e = imop_amd.demosaic(d_in, d_out, w, h, pattern, algo, 10, GIGE, bl, pc)
    torch.cuda.synchronize()
    return e, d_out.cpu().numpy().view(np.uint16).reshape(h, w, 3)


@pytest.mark.parametrize("algo", ALGOS, ids=[a.name for a in ALGOS])
@pytest.mark.parametrize("pattern", [0, 3])
@pytest.mark.parametrize("key", ["top", "bottom"])
def test_reference_gige_frame_matches_oracle(algo, pattern, key):
    rows = FIX[key]
    h = rows.shape[0]
    print(f"gige {key} {algo.name} pattern {pattern}", flush=True)
    e, got = _gpu(rows, W, h, pattern, algo)
    eo, want = O.demosaic(rows, W, h, pattern, int(algo), 10, GIGE)
    assert e == eo == 0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("algo", ALGOS, ids=[a.name for a in ALGOS])
@pytest.mark.parametrize("bl", [[64.0] * 4, [64.0, 3.0, 9.0, 40.0]])
def test_black_level_matches_oracle(algo, bl):
    pc = len(set(bl)) > 1
    print(f"gige bl {algo.name} {bl}", flush=True)
    e, got = _gpu(FIX["top"], W, 64, 1, algo, bl, pc)
    eo, want = O.demosaic(FIX["top"], W, 64, 1, int(algo), 10, GIGE, bl, pc)
    assert e == eo == 0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [(37, 23), (13, 13), (12, 12), (258, 31), (1001, 17)])
@pytest.mark.parametrize("algo", [imop_amd.DemosaicAlgorithm.HQLI, imop_amd.DemosaicAlgorithm.MG,
                                  imop_amd.DemosaicAlgorithm.PRISM, imop_amd.DemosaicAlgorithm.DFPD])
def test_random_frames_with_odd_pixel_counts(shape, algo):
    w, h = shape
    vals = np.random.default_rng(w * 31 + h).integers(0, 1024, size=(h, w)).astype(np.uint16)
    raw = scene.pack_gige10(vals)
    print(f"gige random {w}x{h} {algo.name}", flush=True)
    e, got = _gpu(raw, w, h, 2, algo)
    eo, want = O.demosaic(raw, w, h, 2, int(algo), 10, GIGE)
    assert e == eo == 0
    assert np.array_equal(got, want)


def test_host_batch_and_host_batch_entries():
    MG = imop_amd.DemosaicAlgorithm.MG
    rows = FIX["top"]
    eo, want = O.demosaic(rows, W, 64, 0, int(MG), 10, GIGE, [16.0] * 4)
    e, got = imop_amd.demosaic_host(rows, W, 64, 0, MG, 10, GIGE, [16.0] * 4)
    assert e == eo == 0 and np.array_equal(got, want)
    # device batch: a GigE frame next to a reference-packed one
    other = scene.raw_bytes(np.random.default_rng(3).integers(0, 1024, size=(64, W)), 10, True)
    d_in = [torch.from_numpy(r.copy()).to(DEV) for r in (rows.ravel(), other)]
    d_out = [torch.full((W * 64 * 6,), 0xA5, dtype=torch.uint8, device=DEV) for _ in range(2)]
    descs = [imop_amd.frame_desc(d_in[0], d_out[0], W, 64, 0, MG, 10, GIGE, [16.0] * 4),
             imop_amd.frame_desc(d_in[1], d_out[1], W, 64, 0, MG, 10, True, [16.0] * 4)]
    assert imop_amd.demosaic_batch(descs) == (0, -1)
    torch.cuda.synchronize()
    assert np.array_equal(d_out[0].cpu().numpy().view(np.uint16).reshape(64, W, 3), want)
    _, want2 = O.demosaic(other, W, 64, 0, int(MG), 10, True, [16.0] * 4)
    assert np.array_equal(d_out[1].cpu().numpy().view(np.uint16).reshape(64, W, 3), want2)
    # host batch
    outs = [np.full(W * 64 * 6, 0xA5, np.uint8) for _ in range(2)]
    hd = [imop_amd.host_frame_desc(np.ascontiguousarray(rows.ravel()), outs[0], W, 64, 0, MG, 10, GIGE, [16.0] * 4),
          imop_amd.host_frame_desc(other, outs[1], W, 64, 0, MG, 10, True, [16.0] * 4)]
    assert imop_amd.demosaic_batch_host(hd) == (0, -1)
    assert np.array_equal(outs[0].view(np.uint16).reshape(64, W, 3), want)
    assert np.array_equal(outs[1].view(np.uint16).reshape(64, W, 3), want2)


@pytest.mark.parametrize("algo", [imop_amd.DemosaicAlgorithm.MG, imop_amd.DemosaicAlgorithm.PRISM])
def test_full_frame_equals_the_same_frame_unpacked(algo):
    w, h = 4024, 3032
    vals = scene.make_scene(h, w, 10, 0, seed=9, n_lines=200)
    print(f"gige full frame {algo.name}", flush=True)
    e1, a = _gpu(scene.pack_gige10(vals), w, h, 0, algo, [64.0] * 4)
    d_in = torch.from_numpy(scene.raw_bytes(vals, 10, False).copy()).to(DEV)
    d_out = torch.empty(imop_amd.rgb_byte_size(w, h, 10), dtype=torch.uint8, device=DEV)
    e2 = imop_amd.demosaic(d_in, d_out, w, h, 0, algo, 10, False, [64.0] * 4)
    torch.cuda.synchronize()
    assert e1 == e2 == 0
    assert np.array_equal(a, d_out.cpu().numpy().view(np.uint16).reshape(h, w, 3))


def test_gige_packing_is_10_bit_only():
    d_in = torch.zeros(64 * 64 * 2, dtype=torch.uint8, device=DEV)
    d_out = torch.zeros(64 * 64 * 6, dtype=torch.uint8, device=DEV)
    for bd in (8, 12, 16):
        assert imop_amd.demosaic(d_in, d_out, 64, 64, 0, imop_amd.DemosaicAlgorithm.MG, bd, GIGE) == \
            int(imop_amd.DemosaicError.InvalidBitDepth)
    # is_packed outside 0..2 through the C ABI directly
    import ctypes
    assert imop_amd.lib().imop_hip_demosaic(ctypes.c_void_p(d_in.data_ptr()), d_in.numel(),
                                            ctypes.c_void_p(d_out.data_ptr()), d_out.numel(), 64, 64, 0,
                                            int(imop_amd.DemosaicAlgorithm.MG), 10, 3, None, 0,
                                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)) == \
        int(imop_amd.DemosaicError.InternalError)


@pytest.mark.parametrize("algo", [imop_amd.DemosaicAlgorithm.MG, imop_amd.DemosaicAlgorithm.HQLI,
                                  imop_amd.DemosaicAlgorithm.PRISM])
@pytest.mark.parametrize("flags", [7, 4])
def test_with_rgb_epilogue(algo, flags):
    """GigE input through imop_hip_demosaic_isp: the epilogue fused into the
    register-resident kernel (MG, HQLI) or run as a separate pass (PRISM),
    against the oracle's demosaic followed by its restated RGB stages."""
    gains, ccm, gamma = [1.9, 0.8, 2.3], [1.6, -0.4, -0.2, -0.3, 1.5, -0.2, 0.05, -0.55, 1.5], 2.2
    ep = imop_amd.rgb_epilogue(gains=gains if flags & 1 else None, ccm=ccm if flags & 2 else None,
                               gamma=gamma if flags & 4 else None)
    rows = FIX["top"]
    print(f"gige epilogue {algo.name} flags {flags}", flush=True)
    d_in = torch.from_numpy(np.ascontiguousarray(rows.ravel()).copy()).to(DEV)
    d_out = torch.full((W * 64 * 6,), 0xA5, dtype=torch.uint8, device=DEV)
    e = imop_amd.demosaic_isp(d_in, d_out, W, 64, 0, algo, 10, GIGE, [16.0] * 4, epilogue=ep)
    torch.cuda.synchronize()
    eo, rgb = O.demosaic(rows, W, 64, 0, int(algo), 10, GIGE, [16.0] * 4)
    assert e == eo == 0
    want = O.isp_epilogue(rgb, 10, flags, gains, ccm, gamma)
    assert np.array_equal(d_out.cpu().numpy().view(np.uint16).reshape(64, W, 3), want)


@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_guess_pattern_reads_the_gige_layout(pattern):
    """imop_hip_guess_pattern with is_packed = 2 sees the same samples as with
    the frame given unpacked, and answers as the oracle does on that frame."""
    import analyzer_cases as AC
    vals = AC.ramp_frame(pattern).astype(np.uint32) >> 6   # 256 x 256, values < 1024
    h, w = vals.shape
    g = torch.from_numpy(scene.pack_gige10(vals)).to(DEV)
    u16 = scene.raw_bytes(vals.astype(np.uint16), 10, False)
    u = torch.from_numpy(u16.copy()).to(DEV)
    e1, p1 = imop_amd.guess_pattern(g, w, h, 10, GIGE)
    e2, p2 = imop_amd.guess_pattern(u, w, h, 10, False)
    assert e1 == e2 == 0
    assert p1 == p2 == O.guess_pattern(u16, w, h, 10, False)
