"""The reference's own GPU operator entry points, called by their mangled
C++ names (the symbols its dispatchers link against) on the GPU, against the
oracle: imop::demosaic_cuda (src/algorithms/cuda_dispatch.cpp:68-141),
black_level::process_black_level_cuda (src/gpu_dispatch.cpp:421-448),
white_balance::process_white_balance_cuda (:343-360), ccm::process_ccm_cuda
(:319-336), tone::process_tone_cuda (:292-312),
lens_shading::process_lens_shading_cuda (:457-486) and
defect_correct::process_defect_correct_cuda (:529-571).  Host buffers,
synchronous, results equal to the reference's CPU functions; algorithms the
library does not take return InternalError (6), on which the reference's
dispatchers run their CPU path."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

U8P = ctypes.POINTER(ctypes.c_uint8)
I, F = ctypes.c_int, ctypes.c_float


class ToneParams(ctypes.Structure):  # include/tone/types.hpp:50-58
    _fields_ = [(n, F) for n in ("gamma", "contrast", "shadows", "highlights", "black_point", "white_point",
                                 "mid_point")]


class LensShadingParams(ctypes.Structure):  # include/lens_shading/types.hpp:43-60
    _fields_ = [("coef", F * 12), ("center_x", F), ("center_y", F), ("flat_field", ctypes.c_void_p),
                ("flat_field_width", I), ("flat_field_height", I)]


class DefectCorrectParams(ctypes.Structure):  # include/defect_correct/types.hpp:45-49
    _fields_ = [("threshold", F), ("map", ctypes.c_void_p), ("map_count", I)]


class BlackLevelParams(ctypes.Structure):  # include/black_level/types.hpp
    _fields_ = [(n, F) for n in ("r_offset", "gr_offset", "gb_offset", "b_offset")]


def fn(name, argtypes, restype=I):
    f = getattr(imop_amd.lib(), name)
    f.argtypes, f.restype = argtypes, restype
    return f


def p(a):
    return a.view(np.uint8).ctypes.data_as(U8P)


HAS_CUDA = ["_ZN4imop8has_cudaEv", "_ZN11black_level8has_cudaEv", "_ZN13white_balance8has_cudaEv", "_ZN3ccm8has_cudaEv",
            "_ZN4tone8has_cudaEv", "_ZN12lens_shading8has_cudaEv", "_ZN14defect_correct8has_cudaEv"]


def test_has_cuda_true_on_the_gpu():
    for n in HAS_CUDA:
        assert fn(n, [], ctypes.c_bool)(), n


@pytest.mark.parametrize("algo", [1, 2, 8])
@pytest.mark.parametrize("bd,packed", [(8, False), (12, False), (12, True), (10, True)])
def test_demosaic_cuda(algo, bd, packed):
    f = fn("_ZN4imop13demosaic_cudaEPKhPhiiNS_12BayerPatternENS_17DemosaicAlgorithmEib",
           [U8P, U8P, I, I, I, I, I, ctypes.c_bool])
    w, h = 130, 66
    raw = scene.raw_bytes(scene.make_scene(h, w, bd, 2, seed=bd + algo), bd, packed)
    out = np.zeros(imop_amd.rgb_byte_size(w, h, bd), np.uint8)
    assert f(p(raw), p(out), w, h, 2, algo, bd, packed) == 0
    e, want = O.demosaic(raw, w, h, 2, algo, bd, packed)
    assert e == 0
    np.testing.assert_array_equal(out.view(np.uint16 if bd > 8 else np.uint8).reshape(h, w, 3), want)


@pytest.mark.parametrize("algorithm", [0, 1])  # PER_CHANNEL, GLOBAL
def test_process_black_level_cuda(algorithm):
    f = fn("_ZN11black_level24process_black_level_cudaEPhiiN4imop12BayerPatternENS_19BlackLevelAlgorithmEiRKNS_16BlackLevelParamsE",
           [U8P, I, I, I, I, I, ctypes.POINTER(BlackLevelParams)])
    w, h, bd = 96, 40, 12
    raw = np.random.default_rng(3).integers(0, 4096, size=(h, w)).astype(np.uint16)
    got = raw.copy()
    bp = BlackLevelParams(60.0, 30.0, 40.0, 70.5)
    assert f(p(got), w, h, 1, algorithm, bd, ctypes.byref(bp)) == 0
    want = raw.copy()
    O.oracle().orc_black_level(p(want), w, h, 1, bd, np.array([60.0, 30.0, 40.0, 70.5], np.float32).ctypes.data_as(O._fp),
                               1 if algorithm == 0 else 0)
    np.testing.assert_array_equal(got, want)


def _rgb(w, h, bd, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << bd, size=(h, w, 3)).astype(np.uint16 if bd > 8 else np.uint8)


@pytest.mark.parametrize("bd", [8, 10, 12, 16])
def test_process_white_balance_cuda(bd):
    f = fn("_ZN13white_balance26process_white_balance_cudaEPKhPhiiiifff", [U8P, U8P, I, I, I, I, F, F, F])
    w, h = 77, 31
    rgb = _rgb(w, h, bd, bd)
    out = np.zeros_like(rgb)
    assert f(p(rgb), p(out), w, h, 3, bd, 1.9, 0.8, 2.3) == 0
    np.testing.assert_array_equal(out, O.isp_wb(rgb, bd, [1.9, 0.8, 2.3]))
    # in place, as the CPU function allows
    inplace = rgb.copy()
    assert f(p(inplace), p(inplace), w, h, 3, bd, 1.9, 0.8, 2.3) == 0
    np.testing.assert_array_equal(inplace, out)
    assert f(p(rgb), p(out), w, h, 4, bd, 1.0, 1.0, 1.0) == 4  # InvalidChannels


@pytest.mark.parametrize("bd", [8, 12, 16])
def test_process_ccm_cuda(bd):
    f = fn("_ZN3ccm16process_ccm_cudaEPKhPhiiiiPKf", [U8P, U8P, I, I, I, I, ctypes.POINTER(F)])
    w, h = 64, 45
    rgb = _rgb(w, h, bd, 10 + bd)
    m = np.array([1.6, -0.4, -0.2, -0.3, 1.5, -0.2, 0.05, -0.55, 1.5], np.float32)
    out = np.zeros_like(rgb)
    assert f(p(rgb), p(out), w, h, 3, bd, m.ctypes.data_as(ctypes.POINTER(F))) == 0
    np.testing.assert_array_equal(out, O.isp_ccm(rgb, bd, m))
    assert f(p(rgb), p(out), w, h, 3, bd, None) == 6


@pytest.mark.parametrize("bd", [8, 12, 16])
@pytest.mark.parametrize("gamma", [2.2, 0.45, 20.0])
def test_process_tone_cuda_gamma(bd, gamma):
    f = fn("_ZN4tone17process_tone_cudaEPKhPhiiiNS_13ToneAlgorithmEiRKNS_10ToneParamsE",
           [U8P, U8P, I, I, I, I, I, ctypes.POINTER(ToneParams)])
    w, h = 50, 33
    rgb = _rgb(w, h, bd, 20 + bd)
    out = np.zeros_like(rgb)
    tp = ToneParams(gamma, 0.0, 0.0, 0.0, 0.0, 1.0, 0.5)
    assert f(p(rgb), p(out), w, h, 3, 0, bd, ctypes.byref(tp)) == 0
    np.testing.assert_array_equal(out, O.isp_gamma(rgb, bd, gamma))
    for algo in (1, 2, 3, 4):  # S_CURVE, LEVELS, CURVES_3POINT, SHADOWS_HIGHLIGHTS: the CPU path
        assert f(p(rgb), p(out), w, h, 3, algo, bd, ctypes.byref(tp)) == 6


@pytest.mark.parametrize("bd", [8, 12, 16])
@pytest.mark.parametrize("pattern", [0, 3])
def test_process_lens_shading_cuda(bd, pattern):
    f = fn("_ZN12lens_shading25process_lens_shading_cudaEPhiiN4imop12BayerPatternENS_20LensShadingAlgorithmEiRKNS_17LensShadingParamsE",
           [U8P, I, I, I, I, I, ctypes.POINTER(LensShadingParams)])
    w, h = 123, 70
    raw = np.random.default_rng(bd).integers(0, 1 << bd, size=(h, w)).astype(np.uint16 if bd > 8 else np.uint8)
    coef = [0.31, 0.05, -0.01, 0.22, 0.0, 0.02, 0.2, 0.01, 0.0, 0.4, -0.1, 0.03]
    lp = LensShadingParams((F * 12)(*coef), 0.45, 0.55, None, 0, 0)
    got = raw.copy()
    assert f(p(got), w, h, pattern, 0, bd, ctypes.byref(lp)) == 0
    np.testing.assert_array_equal(got, O.lens_shading(raw, pattern, bd, coef, (0.45, 0.55)))
    assert f(p(got), w, h, pattern, 1, bd, ctypes.byref(lp)) == 6  # FLAT_FIELD: the CPU path


@pytest.mark.parametrize("bd", [8, 12])
def test_process_defect_correct_cuda(bd):
    f = fn("_ZN14defect_correct27process_defect_correct_cudaEPhiiN4imop12BayerPatternENS_22DefectCorrectAlgorithmEiRKNS_19DefectCorrectParamsE",
           [U8P, I, I, I, I, I, ctypes.POINTER(DefectCorrectParams)])
    w, h = 140, 90
    rng = np.random.default_rng(bd + 1)
    raw = scene.make_scene(h, w, bd, 0, seed=bd).astype(np.uint16 if bd > 8 else np.uint8)
    hot = rng.integers(0, w * h, size=60)
    raw.ravel()[hot] = (1 << bd) - 1
    dp = DefectCorrectParams(0.25, None, 0)
    got = raw.copy()
    assert f(p(got), w, h, 0, 0, bd, ctypes.byref(dp)) == 0
    e, want = O.defect_adaptive(raw, bd, 0.25)
    assert e == 0
    np.testing.assert_array_equal(got, want)
    assert f(p(got), w, h, 0, 1, bd, ctypes.byref(dp)) == 6  # MAP_BASED: the CPU path
