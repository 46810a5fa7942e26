"""The RGB epilogue on the GPU (imop_hip_demosaic_isp, imop_hip_rgb_epilogue_apply)
against the oracle: the demosaic oracle followed by the restated white balance,
colour matrix and gamma (oracle/isp_oracle.c, pinned against the reference's
compiled CPU stages).  Fused into the register-resident kernel (HQLI / MG of
even widths) and as a separate pass (PRISM, odd widths, other layouts).  Also
the in-place polynomial lens-shading pass (imop_hip_lens_shading_poly)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
A = imop_amd.DemosaicAlgorithm
M1 = [1.6, -0.4, -0.2, -0.3, 1.5, -0.2, 0.05, -0.55, 1.5]
FULL = dict(gains=[1.9, 0.8, 2.3], ccm=M1, gamma=2.2)


def _epi(flags, gains=FULL["gains"], ccm=FULL["ccm"], gamma=FULL["gamma"]):
    return imop_amd.rgb_epilogue(gains=gains if flags & 1 else None, ccm=ccm if flags & 2 else None,
                                 gamma=gamma if flags & 4 else None)


def _oracle(raw, w, h, pattern, algo, bd, packed, bl, flags, gains=FULL["gains"], ccm=FULL["ccm"],
            gamma=FULL["gamma"]):
    e, rgb = O.demosaic(raw, w, h, pattern, int(algo), bd, packed, bl)
    assert e == 0
    return O.isp_epilogue(rgb, bd, flags, gains, ccm, gamma)


def _run(raw, w, h, pattern, algo, bd, packed, bl, ep):
    d_in = torch.from_numpy(raw.copy()).to(DEV)
    d_out = torch.full((imop_amd.rgb_byte_size(w, h, bd),), 0xA5, dtype=torch.uint8, device=DEV)
    e = imop_amd.demosaic_isp(d_in, d_out, w, h, pattern, algo, bd, packed, bl, epilogue=ep)
    torch.cuda.synchronize()
    assert e == 0
    out = d_out.cpu().numpy()
    return (out.view(np.uint16) if bd > 8 else out).reshape(h, w, 3)


def _frame(w, h, bd, pattern, packed, seed):
    vals = scene.make_scene(h, w, bd, pattern, seed=seed)
    return scene.raw_bytes(vals, bd, packed)


@pytest.mark.parametrize("flags", [1, 2, 4, 7])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_fused_mg12_small(flags, pattern):
    w, h, bd = 256 + 8, 66, 12
    raw = _frame(w, h, bd, pattern, True, seed=flags * 4 + pattern)
    print(f"isp mg12 {w}x{h} pat {pattern} flags {flags}", flush=True)
    got = _run(raw, w, h, pattern, A.MG, bd, True, [64.0] * 4, _epi(flags))
    np.testing.assert_array_equal(got, _oracle(raw, w, h, pattern, A.MG, bd, True, [64.0] * 4, flags))


def test_fused_mg12_full_frame():
    """The headline frame: 4024x3032 RAW12 packed, black level 64, MG, all three stages."""
    w, h, bd = 4024, 3032, 12
    raw = _frame(w, h, bd, 0, True, seed=3)
    print("isp mg12 full frame", flush=True)
    got = _run(raw, w, h, 0, A.MG, bd, True, [64.0] * 4, _epi(7))
    np.testing.assert_array_equal(got, _oracle(raw, w, h, 0, A.MG, bd, True, [64.0] * 4, 7))


@pytest.mark.parametrize("w", [1006, 1008])  # w % 4 == 2 and 0
@pytest.mark.parametrize("algo", [A.HQLI, A.MG])
def test_fused_u8(w, algo):
    h, bd = 70, 8
    raw = _frame(w, h, bd, 1, False, seed=w + int(algo))
    print(f"isp u8 {w} algo {int(algo)}", flush=True)
    got = _run(raw, w, h, 1, algo, bd, False, [16.0] * 4, _epi(7))
    np.testing.assert_array_equal(got, _oracle(raw, w, h, 1, algo, bd, False, [16.0] * 4, 7))


@pytest.mark.parametrize("case", [
    dict(w=131, h=67, bd=10, packed=True, algo=A.PRISM),   # PRISM: separate pass
    dict(w=640, h=482, bd=10, packed=True, algo=A.PRISM),
    dict(w=131, h=67, bd=12, packed=False, algo=A.MG),      # odd width: LDS-tiled path
    dict(w=250, h=40, bd=14, packed=True, algo=A.HQLI),     # generic packed depth
    dict(w=97, h=31, bd=16, packed=False, algo=A.MG),
])
def test_separate_pass_paths(case):
    w, h, bd, packed, algo = case["w"], case["h"], case["bd"], case["packed"], case["algo"]
    raw = _frame(w, h, bd, 2, packed, seed=w * h)
    print(f"isp separate {case}", flush=True)
    got = _run(raw, w, h, 2, algo, bd, packed, None, _epi(7))
    np.testing.assert_array_equal(got, _oracle(raw, w, h, 2, algo, bd, packed, None, 7))


@pytest.mark.parametrize("algo", [A.HQLI, A.MG])
@pytest.mark.parametrize("bd,gamma", [(6, 2.2), (4, 0.1), (7, 0.3)])
def test_values_above_the_bit_depth_maximum(algo, bd, gamma):
    """Frames below 8 bits stored as bytes: raw values up to 255 demosaic past
    the maximum, and gamma maps them as the reference does -- to the maximum,
    or to 0 where powf(v / max, 1 / gamma) * max reaches 2^31 (gamma 0.1 at
    4 bits)."""
    w, h = 264, 20
    raw = np.random.default_rng(5 + bd).integers(0, 256, size=w * h, dtype=np.uint8)
    for flags in (4, 2, 7):
        got = _run(raw, w, h, 0, algo, bd, False, None, _epi(flags, gamma=gamma))
        np.testing.assert_array_equal(got, _oracle(raw, w, h, 0, algo, bd, False, None, flags, gamma=gamma))


@pytest.mark.parametrize("bd", [10, 12])
@pytest.mark.parametrize("gamma", [2.2, 0.1])
def test_standalone_values_above_the_maximum(bd, gamma):
    """The standalone pass on u16 data over the whole 0..65535 range: values
    up to the maximum come from the LDS copy of the table, larger ones from
    the whole table in global memory."""
    w, h = 256, 64
    img = np.random.default_rng(bd).integers(0, 65536, size=(h, w, 3)).astype(np.uint16)
    img[0, :3] = [[(1 << bd) - 1] * 3, [1 << bd] * 3, [65535] * 3]
    d = torch.from_numpy(img.view(np.uint8).ravel().copy()).to(DEV)
    ep = _epi(4, gamma=gamma)
    assert imop_amd.rgb_epilogue_apply(d, w, h, bd, ep) == 0
    torch.cuda.synchronize()
    got = d.cpu().numpy().view(np.uint16).reshape(h, w, 3)
    np.testing.assert_array_equal(got, O.isp_epilogue(img, bd, 4, gamma=gamma))


@pytest.mark.parametrize("params", [
    dict(gains=[-1.0, 3e9, float("nan")], ccm=M1, gamma=2.2),
    dict(gains=[1.0, 1.0, 1.0], ccm=[2.5, 1e-3, -7.25, 0.0, 1.0, 0.0, 1e6, -1e6, 0.5], gamma=0.45),
    dict(gains=[0.5, 0.5, 0.5], ccm=M1, gamma=float("nan")),
    dict(gains=[1.2, 1.0, 1.4], ccm=M1, gamma=12.0),
    dict(gains=[1.2, 1.0, 1.4], ccm=M1, gamma=0.05),
])
def test_extreme_parameters(params):
    w, h, bd = 264, 34, 12
    raw = _frame(w, h, bd, 3, True, seed=11)
    ep = imop_amd.rgb_epilogue(**params)
    got = _run(raw, w, h, 3, A.MG, bd, True, None, ep)
    np.testing.assert_array_equal(got, _oracle(raw, w, h, 3, A.MG, bd, True, None, 7, **params))


@pytest.mark.parametrize("bd", [8, 10, 16])
@pytest.mark.parametrize("offset", [0, 2])   # 2: not 16-byte aligned, per-pixel path
@pytest.mark.parametrize("w,h", [(64, 8), (67, 5), (1, 1)])
def test_standalone_epilogue(bd, offset, w, h):
    rng = np.random.default_rng(bd + offset + w)
    mx = (1 << bd) - 1
    img = rng.integers(0, mx + 1, size=(h, w, 3)).astype(np.uint8 if bd <= 8 else np.uint16)
    raw = img.view(np.uint8).ravel()
    buf = torch.zeros(raw.size + 64, dtype=torch.uint8, device=DEV)
    view = buf[offset:offset + raw.size]
    view.copy_(torch.from_numpy(raw.copy()).to(DEV))
    for flags in (1, 2, 4, 7):
        view.copy_(torch.from_numpy(raw.copy()).to(DEV))
        assert imop_amd.rgb_epilogue_apply(view, w, h, bd, _epi(flags)) == 0
        torch.cuda.synchronize()
        got = view.cpu().numpy()
        got = (got.view(np.uint16) if bd > 8 else got).reshape(h, w, 3)
        np.testing.assert_array_equal(got, O.isp_epilogue(img, bd, flags, FULL["gains"], FULL["ccm"], FULL["gamma"]))
    assert not buf[:offset].any() and not buf[offset + raw.size:].any()  # nothing outside the image written


def test_invalid_arguments():
    d = torch.zeros(64 * 64 * 6, dtype=torch.uint8, device=DEV)
    bad = imop_amd.RgbEpilogue()
    bad.flags = 8
    assert imop_amd.rgb_epilogue_apply(d, 64, 64, 12, bad) == 6
    assert imop_amd.lib().imop_hip_rgb_epilogue_apply(None, 0, 64, 64, 12, None, None) == 1
    assert imop_amd.rgb_epilogue_apply(d, 64, 64, 17, _epi(1)) == 3
    assert imop_amd.rgb_epilogue_apply(d, 0, 64, 12, _epi(1)) == 2
    # demosaic checks come first, in the reference's order
    assert imop_amd.demosaic_isp(d, d, 64, 64, 9, A.MG, 12, False, epilogue=_epi(7)) == 4


def test_graph_replays_match_eager():
    """demosaic_isp captured into a HIP graph (its gamma table built by an
    eager call first) and replayed with eager work between the replays."""
    w, h, bd = 1008, 66, 12
    raw = _frame(w, h, bd, 0, True, seed=21)
    want = _oracle(raw, w, h, 0, A.MG, bd, True, [64.0] * 4, 7)
    d_in = torch.from_numpy(raw.copy()).to(DEV)
    d_out = torch.zeros(imop_amd.rgb_byte_size(w, h, bd), dtype=torch.uint8, device=DEV)
    ep = _epi(7)
    assert imop_amd.demosaic_isp(d_in, d_out, w, h, 0, A.MG, bd, True, [64.0] * 4, epilogue=ep) == 0
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            assert imop_amd.demosaic_isp(d_in, d_out, w, h, 0, A.MG, bd, True, [64.0] * 4, epilogue=ep,
                                         stream=s) == 0
    for rep in range(3):
        d_out.fill_(0x5A)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_out.cpu().numpy().view(np.uint16).reshape(h, w, 3), want)


# ------------------------------------------------------------ lens shading

LS = [([0.8, 0.1, -0.05] * 4, (0.5, 0.5)),
      ([0.3, 0.0, 0.0, 0.5, 0.1, 0.0, 0.45, 0.1, 0.0, 0.9, -0.2, 0.3], (0.4, 0.6)),
      ([0.0] * 12, (0.0, 0.0)),            # max_r = 0: the reference divides by zero
      ([1e9, 0.0, 0.0] * 4, (0.5, 0.5)),   # gains past 2^31
      ([-3.0, 1.0, 0.5] * 4, (1.5, -0.25))]


def _ls_run(img, pattern, bd, coef, center):
    h, w = img.shape
    d = torch.from_numpy(img.view(np.uint8).ravel().copy()).to(DEV)
    assert imop_amd.lens_shading(d, w, h, pattern, bd, imop_amd.lens_shading_params(coef, center)) == 0
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    return (out.view(np.uint16) if bd > 8 else out).reshape(h, w)


@pytest.mark.parametrize("bd", [1, 8, 10, 12, 16])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("ci", range(len(LS)))
def test_lens_shading_matches_oracle(bd, pattern, ci):
    coef, ctr = LS[ci]
    img = np.random.default_rng(bd * 7 + pattern).integers(0, 1 << bd, size=(37, 131)).astype(
        np.uint8 if bd <= 8 else np.uint16)
    np.testing.assert_array_equal(_ls_run(img, pattern, bd, coef, ctr), O.lens_shading(img, pattern, bd, coef, ctr))


@pytest.mark.parametrize("w,h", [(1, 1), (3, 2), (4096 + 3, 5)])
def test_lens_shading_shapes(w, h):
    img = np.random.default_rng(w).integers(0, 4096, size=(h, w)).astype(np.uint16)
    coef, ctr = LS[1]
    np.testing.assert_array_equal(_ls_run(img, 2, 12, coef, ctr), O.lens_shading(img, 2, 12, coef, ctr))


def test_lens_shading_full_frame():
    w, h = 4024, 3032
    img = np.random.default_rng(9).integers(0, 4096, size=(h, w)).astype(np.uint16)
    coef, ctr = LS[0]
    print("lens shading full frame", flush=True)
    np.testing.assert_array_equal(_ls_run(img, 0, 12, coef, ctr), O.lens_shading(img, 0, 12, coef, ctr))


def test_lens_shading_invalid_arguments():
    d = torch.zeros(64 * 64 * 2, dtype=torch.uint8, device=DEV)
    p = imop_amd.lens_shading_params(LS[0][0])
    L = imop_amd.lib()
    assert L.imop_hip_lens_shading_poly(None, 0, 64, 64, 0, 12, ctypes.byref(p), None) == 1
    assert imop_amd.lens_shading(d, 0, 64, 0, 12, p) == 2
    assert imop_amd.lens_shading(d, 64, 64, 0, 17, p) == 3
    assert imop_amd.lens_shading(d, 64, 65, 0, 12, p) == 6   # buffer too small
    assert not d.any()
