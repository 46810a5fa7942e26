"""Batched form: frames of mixed size, layout, bit depth, black level,
pattern and algorithm in one call (imop_hip_demosaic_batch) and as a
graph-captured plan (imop_hip_batch_plan_*), compared frame by frame with
the oracle and with the single-frame entry point, through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
HQLI, MG, PRISM = imop_amd.DemosaicAlgorithm.HQLI, imop_amd.DemosaicAlgorithm.MG, imop_amd.DemosaicAlgorithm.PRISM


def _ragged_frames(seed: int, n: int):
    """Random frame specs covering every layout the batch sorts: batched
    (even widths, u8/u16/12p/10p) and single-frame (odd widths, 14-bit packed,
    PRISM, fractional black level)."""
    rng = np.random.default_rng(seed)
    layouts = [(8, False), (10, True), (12, True), (12, False), (16, False), (14, True), (10, False)]
    widths = [10, 14, 16, 38, 64, 126, 250, 252, 256, 258, 500, 502, 1006, 13, 31]
    specs = []
    for k in range(n):
        w = int(widths[int(rng.integers(0, len(widths)))])
        h = int(rng.integers(10, 70))
        bd, packed = layouts[int(rng.integers(0, len(layouts)))]
        algo = [HQLI, MG, MG, PRISM][int(rng.integers(0, 4))]
        bl = [None, [16.0] * 4, [64.0, 3.0, 9.0, 40.0], [12.5] * 4][int(rng.integers(0, 4))]
        out16 = 2 if (bd <= 8 and rng.integers(0, 2)) else 0
        specs.append(dict(w=w, h=h, bd=bd, packed=packed, algo=algo, pattern=int(rng.integers(0, 4)), bl=bl,
                          out_elem=out16, seed=seed * 100 + k))
    return specs


def _materialise(specs):
    """Host raw frames, device inputs and 0xA5-filled device outputs."""
    frames = []
    for s in specs:
        rng = np.random.default_rng(s["seed"])
        vals = rng.integers(0, 1 << s["bd"], size=(s["h"], s["w"])).astype(np.uint8 if s["bd"] <= 8 else np.uint16)
        raw = scene.raw_bytes(vals, s["bd"], s["packed"])
        elem = s["out_elem"] or (2 if s["bd"] > 8 else 1)
        d_in = torch.from_numpy(raw.copy()).to(DEV)
        d_out = torch.full((s["w"] * s["h"] * 3 * elem,), 0xA5, dtype=torch.uint8, device=DEV)
        frames.append((s, raw, d_in, d_out))
    return frames


def _descs(frames):
    out = []
    for s, raw, d_in, d_out in frames:
        pc = s["bl"] is not None and len(set(s["bl"])) > 1
        out.append(imop_amd.frame_desc(d_in, d_out, s["w"], s["h"], s["pattern"], s["algo"], s["bd"], s["packed"],
                                       s["bl"], pc, s["out_elem"]))
    return out


def _check_against_oracle(frames):
    for s, raw, d_in, d_out in frames:
        pc = s["bl"] is not None and len(set(s["bl"])) > 1
        e, ro = O.demosaic(raw, s["w"], s["h"], s["pattern"], int(s["algo"]), s["bd"], s["packed"], s["bl"], pc)
        assert e == 0
        got = d_out.cpu().numpy()
        got = (got.view(np.uint16) if (s["out_elem"] == 2 or s["bd"] > 8) else got).reshape(s["h"], s["w"], 3)
        bad = np.argwhere(got.astype(np.int64) != ro.astype(np.int64))
        assert bad.size == 0, f"frame {s}: {len(bad)} samples differ, first {bad[:4].tolist()}"


@pytest.fixture
def restore_batch_shape():
    saved = imop_amd.get_batch_shape()
    yield
    imop_amd.set_batch_shape(*saved)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ragged_batch_matches_oracle(seed):
    frames = _materialise(_ragged_frames(seed, 40))
    print(f"ragged batch seed {seed}", flush=True)
    e, failed = imop_amd.demosaic_batch(_descs(frames))
    torch.cuda.synchronize()
    assert (e, failed) == (0, -1)
    _check_against_oracle(frames)


def test_ragged_batch_plan_replays_match_oracle():
    """The graph-captured plan, replayed several times with eager work on the
    stream between replays; every replay is checked."""
    frames = _materialise(_ragged_frames(7, 30))
    with imop_amd.BatchPlan(_descs(frames)) as plan:
        assert (plan.error, plan.failed_frame) == (0, -1)
        for rep in range(3):
            print(f"plan replay {rep}", flush=True)
            for _, _, _, d_out in frames:
                d_out.fill_(0x5A)          # eager work between replays
            assert plan.launch() == 0
            torch.cuda.synchronize()
            _check_against_oracle(frames)


@pytest.mark.parametrize("rows,wpb", [(2, 4), (4, 4), (4, 1), (8, 1)])
@pytest.mark.parametrize("bpc", [0, 1, 3, 16])
@pytest.mark.parametrize("order", [0, 2, 4, 6])
def test_every_batch_shape(rows, wpb, bpc, order, restore_batch_shape):
    imop_amd.set_batch_shape(rows, wpb, bpc, order)
    assert imop_amd.get_batch_shape() == (rows, wpb, bpc, order)
    frames = _materialise(_ragged_frames(11, 24))
    print(f"batch shape {rows} {wpb} {bpc} {order}", flush=True)
    e, failed = imop_amd.demosaic_batch(_descs(frames))
    torch.cuda.synchronize()
    assert (e, failed) == (0, -1)
    _check_against_oracle(frames)


@pytest.mark.parametrize("wpb", [4, 1])
def test_frame_major_small_frames_cover_every_item(wpb, restore_batch_shape):
    """Frame-major placement over frames of 1 to 23 items (fewer than, equal
    to and not multiples of the 8 XCDs), so every frame leaves some XCDs idle
    slots, between larger ones; eager and as a replayed plan."""
    imop_amd.set_batch_shape(4, wpb, 0, 6)
    sizes = [(256, 16), (224, 8), (1006, 16), (256, 64), (512, 40), (2000, 24), (1006, 758), (480, 52), (256, 16)]
    specs = [dict(w=w, h=h, bd=[8, 10, 12][k % 3], packed=k % 3 != 0, algo=MG if k % 2 else HQLI, pattern=k % 4,
                  bl=None if k % 2 else [16.0] * 4, out_elem=2, seed=40 + k) for k, (w, h) in enumerate(sizes)]
    frames = _materialise(specs)
    print(f"frame-major small frames wpb {wpb}", flush=True)
    e, failed = imop_amd.demosaic_batch(_descs(frames))
    torch.cuda.synchronize()
    assert (e, failed) == (0, -1)
    _check_against_oracle(frames)
    with imop_amd.BatchPlan(_descs(frames)) as plan:
        assert (plan.error, plan.failed_frame) == (0, -1)
        for rep in range(2):
            for _, _, _, d_out in frames:
                d_out.fill_(0x3C)
            assert plan.launch() == 0
            torch.cuda.synchronize()
            _check_against_oracle(frames)


def test_tiny_grid_covers_every_item(restore_batch_shape):
    """Fewer workgroups than XCDs: every work item still runs (a 256 x 16
    frame is 2 strips x one 16-row group row: 2 workgroups)."""
    imop_amd.set_batch_shape(4, 4, 1, 2)
    frames = _materialise([dict(w=256, h=16, bd=12, packed=True, algo=MG, pattern=0, bl=None, out_elem=0, seed=5)])
    e, failed = imop_amd.demosaic_batch(_descs(frames))
    torch.cuda.synchronize()
    assert (e, failed) == (0, -1)
    _check_against_oracle(frames)


def test_mixed_size_workload_matches_single_frame_path():
    """48 x 1006x758 + 16 x 4024x3032, depths {8u, 10p, 12p}, black level
    {0, 16, 64}, u16 output: every frame of the batch (direct and plan) equals
    the single-frame entry point's output, value for value."""
    spec = scene.mixed_batch_spec(seed=4)
    rng = np.random.default_rng(44)
    frames, singles = [], []
    for k, f in enumerate(spec):
        nb = imop_amd.bayer_byte_size(f["w"], f["h"], f["bd"], f["packed"])
        d_in = torch.from_numpy(rng.integers(0, 256, size=nb, dtype=np.uint8)).to(DEV)
        d_out = torch.full((f["w"] * f["h"] * 6,), 0xA5, dtype=torch.uint8, device=DEV)
        ref = torch.full((imop_amd.rgb_byte_size(f["w"], f["h"], f["bd"]),), 0x5A, dtype=torch.uint8, device=DEV)
        assert imop_amd.demosaic(d_in, ref, f["w"], f["h"], 0, MG, f["bd"], f["packed"], [f["bl"]] * 4) == 0
        frames.append((f, d_in, d_out))
        singles.append(ref)
    descs = [imop_amd.frame_desc(d_in, d_out, f["w"], f["h"], 0, MG, f["bd"], f["packed"], [f["bl"]] * 4,
                                 False, 2) for f, d_in, d_out in frames]

    def check():
        torch.cuda.synchronize()
        for (f, _, d_out), ref in zip(frames, singles):
            got = d_out.view(torch.int16).to(torch.int32) & 0xFFFF
            want = ref.to(torch.int32) if f["bd"] <= 8 else (ref.view(torch.int16).to(torch.int32) & 0xFFFF)
            assert torch.equal(got, want), f"frame {f}"

    print("mixed workload: direct", flush=True)
    assert imop_amd.demosaic_batch(descs) == (0, -1)
    check()
    for _, _, d_out in frames:
        d_out.fill_(0xA5)
    print("mixed workload: plan", flush=True)
    with imop_amd.BatchPlan(descs) as plan:
        assert plan.error == 0
        for _ in range(2):
            assert plan.launch() == 0
            check()


def test_invalid_frame_enqueues_nothing():
    frames = _materialise(_ragged_frames(21, 6))
    descs = _descs(frames)
    descs[3].height = 4                      # below the MG / HQLI / PRISM minimum
    e, failed = imop_amd.demosaic_batch(descs)
    torch.cuda.synchronize()
    assert (e, failed) == (int(imop_amd.DemosaicError.ImageTooSmall), 3)
    for _, _, _, d_out in frames:
        assert bool((d_out == 0xA5).all())
    descs = _descs(frames)
    descs[2].out_elem_size = 1
    descs[2].bit_depth = 12
    descs[2].is_packed = 1
    plan = imop_amd.BatchPlan(descs)
    assert (plan.error, plan.failed_frame) == (int(imop_amd.DemosaicError.InternalError), 2)
    assert imop_amd.demosaic_batch([]) == (0, -1)


@pytest.mark.parametrize("w,h,bd,algo", [(1006, 758, 8, MG), (640, 480, 12, HQLI), (131, 67, 10, PRISM),
                                         (4024, 3032, 12, MG)])
def test_cpp_demosaic_cuda_batch_matches_per_frame(w, h, bd, algo):
    """imop::demosaic_cuda_batch (the reference's C++ entry point, host
    buffers) equals imop::demosaic_cuda frame by frame."""
    import ctypes
    fn = getattr(imop_amd.lib(), "_ZN4imop19demosaic_cuda_batchEPKPKhPKPhiiiNS_12BayerPatternENS_17DemosaicAlgorithmEi")
    fn.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)] + [ctypes.c_int] * 6
    fn.restype = ctypes.c_int
    n = 3 if w * h > 1_000_000 else 5
    rng = np.random.default_rng(w + bd)
    raws = [rng.integers(0, 1 << bd, size=(h, w)).astype(np.uint8 if bd <= 8 else np.uint16) for _ in range(n)]
    outs = [np.zeros(imop_amd.rgb_byte_size(w, h, bd), dtype=np.uint8) for _ in range(n)]
    ins = (ctypes.c_void_p * n)(*[r.ctypes.data for r in raws])
    ops = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    print(f"demosaic_cuda_batch {w}x{h} bd{bd} algo {int(algo)}", flush=True)
    assert fn(ins, ops, n, w, h, 1, int(algo), bd) == 0
    for r, o in zip(raws, outs):
        e, want = imop_amd.demosaic_host(r, w, h, 1, algo, bd)
        assert e == 0
        np.testing.assert_array_equal(o.view(np.uint16 if bd > 8 else np.uint8).reshape(h, w, 3), want)
