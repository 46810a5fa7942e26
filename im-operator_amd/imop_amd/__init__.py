"""imop_amd -- Python binding of the gfx950 Bayer demosaic library.

The product is libimop_hip.so (C ABI in include/imop_hip.h).  This package
loads it with ctypes and exposes the reference operator interface for the
demosaic path:

  * demosaic(...)          imop::demosaic_cuda with device tensors (stream-ordered)
  * demosaic_host(...)     imop::demosaic_cuda with host arrays (synchronous)
  * black_level(...)       black_level::process_black_level_cuda
  * demosaic_batch(...)    frames of mixed size / depth in one persistent launch
  * BatchPlan(...)         the same batch captured into a HIP graph

Names, enumerations and error codes follow the reference
(include/imop/types.hpp:12-39, include/imop/algorithms.hpp:19-84).  If the
library is missing every call raises: there is no CPU fallback in this
package.
"""
from __future__ import annotations

import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libimop_hip.so")


class BayerPattern(enum.IntEnum):
    """include/imop/types.hpp:12-17"""
    RGGB = 0
    BGGR = 1
    GRBG = 2
    GBRG = 3


class DemosaicAlgorithm(enum.IntEnum):
    """include/imop/types.hpp:19-29"""
    SUPER_FAST = 0
    HQLI = 1
    MG = 2
    L7 = 3
    DFPD = 4
    AHD = 5
    AMAZE = 6
    RCD = 7
    PRISM = 8


class DemosaicError(enum.IntEnum):
    """include/imop/types.hpp:31-39"""
    Ok = 0
    NullInput = 1
    InvalidDimensions = 2
    InvalidBitDepth = 3
    InvalidPattern = 4
    ImageTooSmall = 5
    InternalError = 6


class BlackLevelAlgorithm(enum.IntEnum):
    """include/black_level/types.hpp"""
    PER_CHANNEL = 0
    GLOBAL = 1


class LibraryMissing(RuntimeError):
    pass


class FrameDesc(ctypes.Structure):
    """imop_hip_frame_desc (include/imop_hip.h)."""
    _fields_ = [
        ("in_", ctypes.c_void_p),
        ("in_bytes", ctypes.c_uint64),
        ("out", ctypes.c_void_p),
        ("out_bytes", ctypes.c_uint64),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("pattern", ctypes.c_int32),
        ("algorithm", ctypes.c_int32),
        ("bit_depth", ctypes.c_int32),
        ("is_packed", ctypes.c_int32),
        ("out_elem_size", ctypes.c_int32),
        ("black_level_mode", ctypes.c_int32),
        ("black_level", ctypes.c_float * 4),
    ]


class RgbEpilogue(ctypes.Structure):
    """imop_hip_rgb_epilogue (include/imop_hip.h): flags 1 white balance,
    2 colour matrix, 4 gamma, applied in that order."""
    _fields_ = [
        ("flags", ctypes.c_int32),
        ("wb_gains", ctypes.c_float * 3),
        ("ccm", ctypes.c_float * 9),
        ("gamma", ctypes.c_float),
    ]


def rgb_epilogue(gains=None, ccm=None, gamma=None) -> RgbEpilogue:
    """Epilogue parameters; a stage is on when its argument is given."""
    e = RgbEpilogue()
    if gains is not None:
        e.flags |= 1
        e.wb_gains[:] = [float(v) for v in gains]
    if ccm is not None:
        e.flags |= 2
        e.ccm[:] = [float(v) for v in np.asarray(ccm, dtype=np.float64).ravel()]
    if gamma is not None:
        e.flags |= 4
        e.gamma = float(gamma)
    return e


class LensShadingParams(ctypes.Structure):
    """imop_hip_lens_shading (include/imop_hip.h): {a2, a4, a6} for R, Gr, Gb, B
    and the normalised optical centre."""
    _fields_ = [
        ("coef", (ctypes.c_float * 3) * 4),
        ("center_x", ctypes.c_float),
        ("center_y", ctypes.c_float),
    ]


def lens_shading_params(coef, center=(0.5, 0.5)) -> LensShadingParams:
    p = LensShadingParams()
    c = np.asarray(coef, dtype=np.float64).reshape(4, 3)
    for i in range(4):
        for k in range(3):
            p.coef[i][k] = float(c[i, k])
    p.center_x, p.center_y = float(center[0]), float(center[1])
    return p


class BayerPrologue(ctypes.Structure):
    """imop_hip_bayer_prologue (include/imop_hip.h): flag 1 polynomial lens
    shading with `lens`."""
    _fields_ = [
        ("flags", ctypes.c_int32),
        ("lens", LensShadingParams),
    ]


def bayer_prologue(lens_coef=None, lens_center=(0.5, 0.5)) -> BayerPrologue:
    """Prologue parameters; lens shading is on when its coefficients are given."""
    p = BayerPrologue()
    if lens_coef is not None:
        p.flags |= 1
        p.lens = lens_shading_params(lens_coef, lens_center)
    return p


class DataInfo(ctypes.Structure):
    """imop_hip_data_info (include/imop_hip.h): imop::DataInfo
    (include/imop/types.hpp:76-86) with up to 8 possible dimensions."""
    _fields_ = [
        ("detected_bit_depth", ctypes.c_int32),
        ("suggested_width", ctypes.c_int32),
        ("suggested_height", ctypes.c_int32),
        ("pixel_count", ctypes.c_int32),
        ("max_value", ctypes.c_int32),
        ("min_value", ctypes.c_int32),
        ("is_likely_16bit", ctypes.c_int32),
        ("is_packed", ctypes.c_int32),
        ("n_dimensions", ctypes.c_int32),
        ("dimensions", (ctypes.c_int32 * 2) * 8),
    ]

    def as_dict(self) -> dict:
        return dict(detected_bit_depth=self.detected_bit_depth, suggested_width=self.suggested_width,
                    suggested_height=self.suggested_height, pixel_count=self.pixel_count,
                    max_value=self.max_value, min_value=self.min_value,
                    is_likely_16bit=bool(self.is_likely_16bit), is_packed=bool(self.is_packed),
                    possible_dimensions=[(self.dimensions[k][0], self.dimensions[k][1])
                                         for k in range(self.n_dimensions)])


_lib = None


def lib() -> ctypes.CDLL:
    """Load libimop_hip.so; raise LibraryMissing when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} not built (run __graft_entry__.build() or make -C im-operator_amd)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u8p, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_float)
        i, u64 = ctypes.c_int, ctypes.c_uint64
        L.imop_hip_device_count.restype = i
        L.imop_hip_device_name.restype = ctypes.c_char_p
        L.imop_hip_synchronize.restype = i
        L.imop_hip_demosaic.argtypes = [vp, u64, vp, u64, i, i, i, i, i, i, fp, i, vp]
        L.imop_hip_demosaic.restype = i
        L.imop_hip_demosaic_host.argtypes = [u8p, u8p, i, i, i, i, i, i, fp, i]
        L.imop_hip_demosaic_host.restype = i
        L.imop_hip_black_level.argtypes = [vp, i, i, i, i, i, fp, vp]
        L.imop_hip_black_level.restype = i
        L.imop_hip_black_level_host.argtypes = [u8p, i, i, i, i, i, fp]
        L.imop_hip_black_level_host.restype = i
        L.imop_hip_set_launch_shape.argtypes = [i, i, i]
        L.imop_hip_set_launch_shape.restype = i
        L.imop_hip_get_launch_shape.argtypes = [ctypes.POINTER(i)] * 3
        L.imop_hip_get_launch_shape.restype = i
        L.imop_hip_demosaic_batch.argtypes = [ctypes.POINTER(FrameDesc), i, vp, ctypes.POINTER(i)]
        L.imop_hip_demosaic_batch.restype = i
        L.imop_hip_demosaic_batch_host.argtypes = [ctypes.POINTER(FrameDesc), i, ctypes.POINTER(i)]
        L.imop_hip_demosaic_batch_host.restype = i
        L.imop_hip_batch_plan_create.argtypes = [ctypes.POINTER(FrameDesc), i, ctypes.POINTER(vp), ctypes.POINTER(i)]
        L.imop_hip_batch_plan_create.restype = i
        L.imop_hip_batch_plan_launch.argtypes = [vp, vp]
        L.imop_hip_batch_plan_launch.restype = i
        L.imop_hip_batch_plan_destroy.argtypes = [vp]
        L.imop_hip_batch_plan_destroy.restype = i
        L.imop_hip_demosaic_isp.argtypes = [vp, u64, vp, u64, i, i, i, i, i, i, fp, i, ctypes.POINTER(RgbEpilogue), vp]
        L.imop_hip_demosaic_isp.restype = i
        L.imop_hip_demosaic_pipeline.argtypes = [vp, u64, vp, u64, i, i, i, i, i, i, fp, i,
                                                 ctypes.POINTER(BayerPrologue), ctypes.POINTER(RgbEpilogue), vp]
        L.imop_hip_demosaic_pipeline.restype = i
        L.imop_hip_rgb_epilogue_apply.argtypes = [vp, u64, i, i, i, ctypes.POINTER(RgbEpilogue), vp]
        L.imop_hip_rgb_epilogue_apply.restype = i
        L.imop_hip_lens_shading_poly.argtypes = [vp, u64, i, i, i, i, ctypes.POINTER(LensShadingParams), vp]
        L.imop_hip_lens_shading_poly.restype = i
        L.imop_hip_defect_adaptive.argtypes = [vp, u64, i, i, i, i, ctypes.c_float, vp]
        L.imop_hip_defect_adaptive.restype = i
        L.imop_hip_defect_adaptive_async.argtypes = [vp, u64, i, i, i, i, ctypes.c_float, i, vp, vp]
        L.imop_hip_defect_adaptive_async.restype = i
        L.imop_hip_analyze_data.argtypes = [vp, u64, ctypes.POINTER(DataInfo), vp]
        L.imop_hip_analyze_data.restype = i
        L.imop_hip_guess_pattern.argtypes = [vp, u64, i, i, i, i, ctypes.POINTER(i), vp]
        L.imop_hip_guess_pattern.restype = i
        L.imop_hip_set_batch_shape.argtypes = [i, i, i, i]
        L.imop_hip_set_batch_shape.restype = i
        L.imop_hip_get_batch_shape.argtypes = [ctypes.POINTER(i)] * 4
        L.imop_hip_get_batch_shape.restype = i
        L.imop_hip_valu_issue_rate.argtypes = [ctypes.POINTER(ctypes.c_double), vp]
        L.imop_hip_valu_issue_rate.restype = i
        _lib = L
    return _lib


def has_cuda() -> bool:
    """imop::has_cuda (src/algorithms/cuda_dispatch.cpp:26-33)."""
    return lib().imop_hip_device_count() > 0


def cuda_device_name() -> str:
    return lib().imop_hip_device_name().decode()


def cuda_synchronize() -> None:
    lib().imop_hip_synchronize()


PACKED_GIGE10 = 2
"""is_packed value of the GigE Vision 10-bit packing (BayerXX10Packed: 2 px in
3 bytes, [P0 9:2][P1 1:0 << 4 | P0 1:0][P1 9:2]); True / 1 is the reference's
packed layouts."""


def _packing(is_packed) -> int:
    """The C ABI's is_packed: 0 unpacked, 1 the reference's packed layouts,
    2 GigE Vision 10-bit packing."""
    return PACKED_GIGE10 if (not isinstance(is_packed, bool) and is_packed == PACKED_GIGE10) else int(bool(is_packed))


def bayer_byte_size(w: int, h: int, bd: int, packed: bool) -> int:
    """pixel::compute_bayer_byte_size (include/imop/pixel_utils.hpp:25-28);
    packed = PACKED_GIGE10: 3 bytes per pixel pair."""
    if _packing(packed) == PACKED_GIGE10:
        return (w * h + 1) // 2 * 3
    if packed:
        return (w * h * bd + 7) // 8
    return w * h * (1 if bd <= 8 else 2)


def rgb_byte_size(w: int, h: int, bd: int) -> int:
    """pixel::compute_rgb_byte_size (include/imop/pixel_utils.hpp:30-34)."""
    return w * h * 3 * (2 if bd > 8 else 1)


def _bl_args(black_level, per_channel: bool):
    if black_level is None:
        return None, 0
    arr = np.asarray(black_level, dtype=np.float32).ravel()
    if arr.size == 1:
        arr = np.repeat(arr, 4)
    arr = np.ascontiguousarray(arr[:4], dtype=np.float32)
    return arr, (2 if per_channel else 1)


def demosaic(d_in, d_out, width: int, height: int, pattern: int, algorithm: int, bit_depth: int = 8,
             is_packed: bool = False, black_level=None, per_channel: bool = False, stream=None) -> int:
    """Stream-ordered demosaic of device tensors (torch, any dtype, contiguous).

    Returns the DemosaicError value; kernels are enqueued on `stream`
    (a torch.cuda.Stream, a raw hipStream_t int, or None for the current
    torch stream)."""
    bl, mode = _bl_args(black_level, per_channel)
    blp = bl.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if bl is not None else None
    if stream is None:
        import torch
        stream = torch.cuda.current_stream().cuda_stream
    elif hasattr(stream, "cuda_stream"):
        stream = stream.cuda_stream
    in_bytes = d_in.numel() * d_in.element_size()
    out_bytes = d_out.numel() * d_out.element_size()
    return lib().imop_hip_demosaic(ctypes.c_void_p(d_in.data_ptr()), in_bytes, ctypes.c_void_p(d_out.data_ptr()),
                                   out_bytes, width, height, int(pattern), int(algorithm), bit_depth,
                                   _packing(is_packed), blp, mode, ctypes.c_void_p(stream))


def _stream_arg(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else stream


def demosaic_isp(d_in, d_out, width: int, height: int, pattern: int, algorithm: int, bit_depth: int = 8,
                 is_packed: bool = False, black_level=None, per_channel: bool = False,
                 epilogue: RgbEpilogue | None = None, stream=None) -> int:
    """imop_hip_demosaic_isp: the demosaic followed by the RGB epilogue
    (white balance, colour matrix, gamma) on device tensors."""
    bl, mode = _bl_args(black_level, per_channel)
    blp = bl.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if bl is not None else None
    ep = ctypes.byref(epilogue) if epilogue is not None else None
    return lib().imop_hip_demosaic_isp(ctypes.c_void_p(d_in.data_ptr()), d_in.numel() * d_in.element_size(),
                                       ctypes.c_void_p(d_out.data_ptr()), d_out.numel() * d_out.element_size(),
                                       width, height, int(pattern), int(algorithm), bit_depth,
                                       _packing(is_packed), blp, mode, ep, ctypes.c_void_p(_stream_arg(stream)))


def demosaic_pipeline(d_in, d_out, width: int, height: int, pattern: int, algorithm: int, bit_depth: int = 8,
                      is_packed: bool = False, black_level=None, per_channel: bool = False,
                      prologue: BayerPrologue | None = None, epilogue: RgbEpilogue | None = None,
                      stream=None) -> int:
    """imop_hip_demosaic_pipeline: unpack, black level, the Bayer prologue
    (lens shading), the demosaic and the RGB epilogue in one call on device
    tensors."""
    bl, mode = _bl_args(black_level, per_channel)
    blp = bl.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if bl is not None else None
    pp = ctypes.byref(prologue) if prologue is not None else None
    ep = ctypes.byref(epilogue) if epilogue is not None else None
    return lib().imop_hip_demosaic_pipeline(ctypes.c_void_p(d_in.data_ptr()), d_in.numel() * d_in.element_size(),
                                            ctypes.c_void_p(d_out.data_ptr()), d_out.numel() * d_out.element_size(),
                                            width, height, int(pattern), int(algorithm), bit_depth,
                                            _packing(is_packed), blp, mode, pp, ep,
                                            ctypes.c_void_p(_stream_arg(stream)))


def rgb_epilogue_apply(d_rgb, width: int, height: int, bit_depth: int, epilogue: RgbEpilogue, stream=None) -> int:
    """imop_hip_rgb_epilogue_apply: the epilogue alone, in place, on a device RGB tensor."""
    return lib().imop_hip_rgb_epilogue_apply(ctypes.c_void_p(d_rgb.data_ptr()), d_rgb.numel() * d_rgb.element_size(),
                                             width, height, bit_depth, ctypes.byref(epilogue),
                                             ctypes.c_void_p(_stream_arg(stream)))


def lens_shading(d_bayer, width: int, height: int, pattern: int, bit_depth: int, params: LensShadingParams,
                 stream=None) -> int:
    """imop_hip_lens_shading_poly: polynomial lens shading in place on an unpacked
    Bayer frame (device tensor, u8 or u16 samples)."""
    return lib().imop_hip_lens_shading_poly(ctypes.c_void_p(d_bayer.data_ptr()),
                                            d_bayer.numel() * d_bayer.element_size(), width, height, int(pattern),
                                            bit_depth, ctypes.byref(params), ctypes.c_void_p(_stream_arg(stream)))


def defect_correct(d_bayer, width: int, height: int, pattern: int, bit_depth: int, threshold: float = 0.3,
                   stream=None) -> int:
    """imop_hip_defect_adaptive: adaptive defect correction in place on an
    unpacked Bayer frame (device tensor, u8 or u16 samples); threshold as
    DefectCorrectParams::threshold (default 0.3).  Synchronises the stream."""
    return lib().imop_hip_defect_adaptive(ctypes.c_void_p(d_bayer.data_ptr()),
                                          d_bayer.numel() * d_bayer.element_size(), width, height, int(pattern),
                                          bit_depth, ctypes.c_float(threshold), ctypes.c_void_p(_stream_arg(stream)))


def defect_correct_async(d_bayer, width: int, height: int, pattern: int, bit_depth: int, d_status,
                         threshold: float = 0.3, max_passes: int = 32, stream=None) -> int:
    """imop_hip_defect_adaptive_async: the same correction stream-ordered
    (capturable into a graph); writes 0 / 1 / 2 (converged / not within
    max_passes / error) to d_status, a one-element int32 device tensor, when
    the passes end."""
    return lib().imop_hip_defect_adaptive_async(ctypes.c_void_p(d_bayer.data_ptr()),
                                                d_bayer.numel() * d_bayer.element_size(), width, height,
                                                int(pattern), bit_depth, ctypes.c_float(threshold), int(max_passes),
                                                ctypes.c_void_p(d_status.data_ptr()),
                                                ctypes.c_void_p(_stream_arg(stream)))


def valu_issue_rate(stream=None) -> float:
    """imop_hip_valu_issue_rate: wave64 VALU instructions per second this GPU
    sustains (a probe kernel); RuntimeError on a HIP failure.  Synchronises
    the stream."""
    r = ctypes.c_double()
    # the null stream unless one is given (the probe synchronises it anyway)
    e = lib().imop_hip_valu_issue_rate(ctypes.byref(r), ctypes.c_void_p(0 if stream is None else _stream_arg(stream)))
    if e != 0:
        raise RuntimeError(f"imop_hip_valu_issue_rate failed: {e}")
    return r.value


def analyze_data(d_data, stream=None) -> tuple[int, dict]:
    """imop_hip_analyze_data: imop::analyze_data of a raw buffer in device
    memory (a device tensor, its bytes); (error code, DataInfo as a dict).
    Synchronises the stream."""
    info = DataInfo()
    e = lib().imop_hip_analyze_data(ctypes.c_void_p(d_data.data_ptr()), d_data.numel() * d_data.element_size(),
                                    ctypes.byref(info), ctypes.c_void_p(_stream_arg(stream)))
    return e, info.as_dict()


def guess_pattern(d_data, width: int, height: int, bit_depth: int, is_packed: bool = False,
                  stream=None) -> tuple[int, int]:
    """imop_hip_guess_pattern: imop::guess_pattern of a frame in device memory;
    (error code, BayerPattern value).  Synchronises the stream."""
    pat = ctypes.c_int(-1)
    e = lib().imop_hip_guess_pattern(ctypes.c_void_p(d_data.data_ptr()), d_data.numel() * d_data.element_size(),
                                     width, height, bit_depth, _packing(is_packed), ctypes.byref(pat),
                                     ctypes.c_void_p(_stream_arg(stream)))
    return e, pat.value


def demosaic_host(bayer: np.ndarray, width: int, height: int, pattern: int, algorithm: int, bit_depth: int = 8,
                  is_packed: bool = False, black_level=None, per_channel: bool = False,
                  out: np.ndarray | None = None):
    """imop::demosaic_cuda on host buffers: returns (DemosaicError, rgb)."""
    raw = np.ascontiguousarray(bayer).view(np.uint8).ravel()
    if out is None:
        out = np.zeros(rgb_byte_size(width, height, bit_depth), dtype=np.uint8)
    bl, mode = _bl_args(black_level, per_channel)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    blp = bl.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if bl is not None else None
    e = lib().imop_hip_demosaic_host(raw.ctypes.data_as(u8p), out.ctypes.data_as(u8p), width, height, int(pattern),
                                     int(algorithm), bit_depth, _packing(is_packed), blp, mode)
    rgb = out.view(np.uint16) if bit_depth > 8 else out
    return DemosaicError(e), rgb.reshape(height, width, 3)


def black_level(d_data, width: int, height: int, pattern: int, algorithm: int, bit_depth: int, offsets,
                stream=None) -> int:
    """In-place black level of a device tensor (black_level::process_black_level_cuda)."""
    o = np.ascontiguousarray(np.asarray(offsets, dtype=np.float32).ravel()[:4])
    if stream is None:
        import torch
        stream = torch.cuda.current_stream().cuda_stream
    elif hasattr(stream, "cuda_stream"):
        stream = stream.cuda_stream
    return lib().imop_hip_black_level(ctypes.c_void_p(d_data.data_ptr()), width, height, int(pattern),
                                      int(algorithm), bit_depth, o.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                      ctypes.c_void_p(stream))


def _stream_handle(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return stream


def frame_desc(d_in, d_out, width: int, height: int, pattern: int, algorithm: int, bit_depth: int,
               is_packed: bool = False, black_level=None, per_channel: bool = False, out_elem_size: int = 0):
    """One imop_hip_frame_desc over device tensors (the tensors must outlive its use)."""
    d = FrameDesc()
    d.in_ = d_in.data_ptr()
    d.in_bytes = d_in.numel() * d_in.element_size()
    d.out = d_out.data_ptr()
    d.out_bytes = d_out.numel() * d_out.element_size()
    d.width, d.height = width, height
    d.pattern, d.algorithm, d.bit_depth = int(pattern), int(algorithm), bit_depth
    d.is_packed = _packing(is_packed)
    d.out_elem_size = out_elem_size
    bl, mode = _bl_args(black_level, per_channel)
    d.black_level_mode = mode
    if bl is not None:
        for k in range(4):
            d.black_level[k] = float(bl[k])
    return d


def _desc_array(descs):
    arr = (FrameDesc * max(1, len(descs)))()
    for k, d in enumerate(descs):
        arr[k] = d
    return arr


def demosaic_batch(descs, stream=None) -> tuple[int, int]:
    """imop_hip_demosaic_batch over a list of FrameDesc: (DemosaicError, failed frame or -1)."""
    arr = _desc_array(descs)
    failed = ctypes.c_int(-1)
    e = lib().imop_hip_demosaic_batch(arr, len(descs), ctypes.c_void_p(_stream_handle(stream)),
                                      ctypes.byref(failed))
    return e, failed.value


def host_frame_desc(bayer: np.ndarray, out: np.ndarray, width: int, height: int, pattern: int, algorithm: int,
                    bit_depth: int, is_packed: bool = False, black_level=None, per_channel: bool = False,
                    out_elem_size: int = 0):
    """One imop_hip_frame_desc over contiguous host arrays (they must outlive its use)."""
    d = FrameDesc()
    if not (bayer.flags.c_contiguous and out.flags.c_contiguous):
        raise ValueError("host frames must be C-contiguous")
    d.in_ = bayer.ctypes.data
    d.in_bytes = bayer.nbytes
    d.out = out.ctypes.data
    d.out_bytes = out.nbytes
    d.width, d.height = width, height
    d.pattern, d.algorithm, d.bit_depth = int(pattern), int(algorithm), bit_depth
    d.is_packed = _packing(is_packed)
    d.out_elem_size = out_elem_size
    bl, mode = _bl_args(black_level, per_channel)
    d.black_level_mode = mode
    if bl is not None:
        for k in range(4):
            d.black_level[k] = float(bl[k])
    return d


def demosaic_batch_host(descs) -> tuple[int, int]:
    """imop_hip_demosaic_batch_host over FrameDescs of host buffers (pinned,
    triple-buffered staging; synchronous): (DemosaicError, failed frame or -1)."""
    arr = _desc_array(descs)
    failed = ctypes.c_int(-1)
    e = lib().imop_hip_demosaic_batch_host(arr, len(descs), ctypes.byref(failed))
    return e, failed.value


class BatchPlan:
    """A batch captured into a HIP graph (imop_hip_batch_plan_*)."""

    def __init__(self, descs):
        self._h = ctypes.c_void_p()
        failed = ctypes.c_int(-1)
        self.error = lib().imop_hip_batch_plan_create(_desc_array(descs), len(descs), ctypes.byref(self._h),
                                                      ctypes.byref(failed))
        self.failed_frame = failed.value
        if self.error != 0:
            self._h = None

    def launch(self, stream=None) -> int:
        if self._h is None:
            raise RuntimeError(f"batch plan was not created (error {self.error})")
        return lib().imop_hip_batch_plan_launch(self._h, ctypes.c_void_p(_stream_handle(stream)))

    def close(self) -> None:
        if self._h is not None:
            lib().imop_hip_batch_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_batch_shape(rows_per_wave: int = 0, waves_per_block: int = 0, blocks_per_cu: int = -1,
                    order: int = -1) -> None:
    """Launch shape of the batched kernel (see imop_hip.h)."""
    lib().imop_hip_set_batch_shape(rows_per_wave, waves_per_block, blocks_per_cu, order)


def get_batch_shape() -> tuple[int, int, int, int]:
    v = [ctypes.c_int() for _ in range(4)]
    lib().imop_hip_get_batch_shape(*[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def set_launch_shape(rows_per_wave: int = -1, waves_per_block: int = 0, xcd_remap: int = -1) -> None:
    """Launch shape of the register-resident kernel (imop_hip_set_launch_shape):
    rows_per_wave 0 = automatic band height, -1 / 0 / -1 = keep a setting."""
    lib().imop_hip_set_launch_shape(rows_per_wave, waves_per_block, xcd_remap)


def get_launch_shape() -> tuple[int, int, int]:
    r, w, x = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().imop_hip_get_launch_shape(ctypes.byref(r), ctypes.byref(w), ctypes.byref(x))
    return r.value, w.value, x.value
