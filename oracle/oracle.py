"""ctypes binding of the CPU oracle (liboracle.so) and, when it has been
built, of the compiled reference (oracle/_ref/libimop_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.

The C restatement lives in imop_oracle.c; every function there cites the
reference file:line it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libimop_ref.so")
REF_ISP_PATH = os.path.join(HERE, "_ref", "libimop_ref_isp.so")

SUPER_FAST, HQLI, MG, L7, DFPD, AHD, AMAZE, RCD, PRISM = 0, 1, 2, 3, 4, 5, 6, 7, 8
RGGB, BGGR, GRBG, GBRG = 0, 1, 2, 3

_u8p = ctypes.POINTER(ctypes.c_uint8)
_fp = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and the reference library when its sources are
    present) with oracle/Makefile / oracle/build_ref.sh."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    subprocess.run([os.path.join(HERE, "build_ref.sh")], check=True,
                   stdout=subprocess.DEVNULL)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u8p)


_oracle = None
_ref = None
_ref_isp = None


def oracle() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_demosaic.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 6
        lib.orc_pipeline.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 6 + [_fp, ctypes.c_int]
        lib.orc_black_level.argtypes = [_u8p] + [ctypes.c_int] * 4 + [_fp, ctypes.c_int]
        lib.orc_packed_pixel.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t]
        lib.orc_unpack.argtypes = [_u8p] + [ctypes.c_int] * 4 + [_u8p]
        lib.orc_packed_bytes.restype = ctypes.c_size_t
        lib.orc_packed_bytes.argtypes = [ctypes.c_int] * 3
        lib.orc_set_mg8_scalar.argtypes = [ctypes.c_int]
        lib.orc_set_num_threads.argtypes = [ctypes.c_int]
        _dp = ctypes.POINTER(ctypes.c_double)
        _lp = ctypes.POINTER(ctypes.c_long)
        _ip = ctypes.POINTER(ctypes.c_int)
        lib.orc_bilinear_csr.restype = ctypes.c_long
        lib.orc_bilinear_csr.argtypes = [ctypes.c_int] * 3 + [_lp, _ip, _dp]
        lib.orc_chain_fwd64.argtypes = [_fp] + [ctypes.c_int] * 4 + [ctypes.c_double, _dp, _dp, ctypes.c_double,
                                                                      _dp, _fp, _dp]
        lib.orc_chain_vjp64.argtypes = [_fp] + [ctypes.c_int] * 4 + [ctypes.c_double, _dp, _dp, ctypes.c_double,
                                                                      _fp, _fp, _dp, _dp, _dp]
        lib.orc_chain_set_plant.argtypes = [ctypes.c_int]
        lib.orc_chain_set_plant.restype = None
        lib.orc_bilinear_apply64.argtypes = [_dp] + [ctypes.c_int] * 3 + [_dp]
        lib.orc_bilinear_apply_t64.argtypes = [_dp] + [ctypes.c_int] * 3 + [_dp]
        lib.isp_wb_manual.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 3 + [_fp]
        lib.isp_ccm_linear.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 3 + [_fp]
        lib.isp_gamma.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 3 + [ctypes.c_float]
        lib.isp_epilogue.argtypes = [_u8p] + [ctypes.c_int] * 4 + [_fp, _fp, ctypes.c_float]
        lib.isp_lens_shading_poly.argtypes = [_u8p] + [ctypes.c_int] * 4 + [_fp, ctypes.c_float, ctypes.c_float]
        lib.isp_defect_adaptive.argtypes = [_u8p] + [ctypes.c_int] * 3 + [ctypes.c_float]
        for fn in ("isp_wb_manual", "isp_ccm_linear", "isp_gamma", "isp_epilogue", "isp_lens_shading_poly"):
            getattr(lib, fn).restype = None
        lib.isp_gamma_clamp.argtypes = [ctypes.c_float]
        lib.isp_gamma_clamp.restype = ctypes.c_float
        _oracle = lib
    return _oracle


def reference_available() -> bool:
    return os.path.exists(REF_PATH)


def reference() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_PATH)
        lib.ref_demosaic_cpu.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 6
        lib.ref_black_level.argtypes = [_u8p] + [ctypes.c_int] * 4 + [_fp, ctypes.c_int]
        lib.ref_get_packed_raw.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_size_t, ctypes.c_int]
        lib.ref_get_raw.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_size_t, ctypes.c_int]
        lib.ref_set_num_threads.argtypes = [ctypes.c_int]
        lib.ref_unpack.argtypes = [_u8p] + [ctypes.c_int] * 4 + [_u8p]
        _ref = lib
    return _ref


def reference_isp_available() -> bool:
    return os.path.exists(REF_ISP_PATH)


def reference_isp() -> ctypes.CDLL:
    """The reference's compiled white balance / colour matrix / gamma stages."""
    global _ref_isp
    if _ref_isp is None:
        lib = ctypes.CDLL(REF_ISP_PATH)
        lib.ref_wb_manual.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 3 + [_fp]
        lib.ref_ccm_linear.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 3 + [_fp]
        lib.ref_gamma.argtypes = [_u8p, _u8p] + [ctypes.c_int] * 3 + [ctypes.c_float]
        lib.ref_lens_shading_poly.argtypes = [_u8p] + [ctypes.c_int] * 4 + [_fp, ctypes.c_float, ctypes.c_float]
        lib.ref_defect_adaptive.argtypes = [_u8p] + [ctypes.c_int] * 4 + [ctypes.c_float]
        lib.ref_analyze_data.argtypes = [_u8p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
        lib.ref_analyze_data.restype = None
        lib.ref_guess_pattern.argtypes = [_u8p] + [ctypes.c_int] * 4
        _ref_isp = lib
    return _ref_isp


def _rgb_stage(lib, fn: str, rgb: np.ndarray, bd: int, arg):
    """Run one RGB stage of `lib` on an [h, w, 3] u8 / u16 image; returns a new image."""
    h, w, _ = rgb.shape
    src = np.ascontiguousarray(rgb)
    dst = np.zeros_like(src)
    a = np.ascontiguousarray(arg, dtype=np.float32).ctypes.data_as(_fp) if not isinstance(arg, float) else ctypes.c_float(arg)
    e = getattr(lib, fn)(src.view(np.uint8).ctypes.data_as(_u8p), dst.view(np.uint8).ctypes.data_as(_u8p), w, h, bd, a)
    if e not in (None, 0):
        raise RuntimeError(f"{fn} returned {e}")
    return dst


def isp_wb(rgb, bd, gains):
    """white_balance::process_manual_wb (src/white_balance/manual.cpp:8-41), restated."""
    return _rgb_stage(oracle(), "isp_wb_manual", rgb, bd, gains)


def isp_ccm(rgb, bd, m):
    """ccm::process_linear_3x3 (src/ccm/linear_3x3.cpp:9-46), restated."""
    return _rgb_stage(oracle(), "isp_ccm_linear", rgb, bd, m)


def isp_gamma(rgb, bd, gamma):
    """tone::process_gamma (src/tone/gamma.cpp:10-34), restated."""
    return _rgb_stage(oracle(), "isp_gamma", rgb, bd, float(gamma))


def isp_epilogue(rgb, bd, flags, gains=(1.0, 1.0, 1.0), m=(1, 0, 0, 0, 1, 0, 0, 0, 1), gamma=1.0):
    """White balance, colour matrix, gamma in that order (flags 1 / 2 / 4), restated."""
    out = np.ascontiguousarray(rgb).copy()
    h, w, _ = out.shape
    oracle().isp_epilogue(out.view(np.uint8).ctypes.data_as(_u8p), w, h, bd, flags,
                          np.asarray(gains, dtype=np.float32).ctypes.data_as(_fp),
                          np.asarray(m, dtype=np.float32).ctypes.data_as(_fp),
                          ctypes.c_float(gamma))
    return out


def _lens_shading(lib, fn, bayer, pattern, bd, coef, center):
    out = np.ascontiguousarray(bayer).copy()
    h, w = out.shape
    e = getattr(lib, fn)(out.view(np.uint8).ctypes.data_as(_u8p), w, h, pattern, bd,
                         np.asarray(coef, dtype=np.float32).ravel().ctypes.data_as(_fp),
                         ctypes.c_float(center[0]), ctypes.c_float(center[1]))
    if e not in (None, 0):
        raise RuntimeError(f"{fn} returned {e}")
    return out


def lens_shading(bayer, pattern, bd, coef, center=(0.5, 0.5)):
    """lens_shading::process_polynomial (src/lens_shading/polynomial.cpp:11-52),
    restated; bayer [h, w] u8 / u16, coef 4 x (a2, a4, a6) for R, Gr, Gb, B."""
    return _lens_shading(oracle(), "isp_lens_shading_poly", bayer, pattern, bd, coef, center)


def ref_lens_shading(bayer, pattern, bd, coef, center=(0.5, 0.5)):
    return _lens_shading(reference_isp(), "ref_lens_shading_poly", bayer, pattern, bd, coef, center)


def defect_adaptive(bayer, bd, threshold=0.3):
    """defect_correct::process_adaptive (src/defect_correct/adaptive.cpp:8-61),
    restated: returns (error, corrected copy)."""
    out = np.ascontiguousarray(bayer).copy()
    h, w = out.shape
    e = oracle().isp_defect_adaptive(out.view(np.uint8).ctypes.data_as(_u8p), w, h, bd, ctypes.c_float(threshold))
    return e, out


def ref_defect_adaptive(bayer, bd, threshold=0.3, pattern=0):
    out = np.ascontiguousarray(bayer).copy()
    h, w = out.shape
    e = reference_isp().ref_defect_adaptive(out.view(np.uint8).ctypes.data_as(_u8p), w, h, pattern, bd,
                                            ctypes.c_float(threshold))
    return e, out


def analyze_data(data):
    """analyze_data (src/analyzer.cpp:151-214), restated: the same dict as ref_analyze_data."""
    buf = np.ascontiguousarray(data).view(np.uint8).ravel()
    out = (ctypes.c_int * 25)()
    lib = oracle()
    lib.isp_analyze_data.argtypes = [_u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
    lib.isp_analyze_data.restype = None
    lib.isp_analyze_data(buf.ctypes.data_as(_u8p), buf.size, out)
    v = list(out)
    return dict(detected_bit_depth=v[0], suggested_width=v[1], suggested_height=v[2], pixel_count=v[3],
                max_value=v[4], min_value=v[5], is_likely_16bit=bool(v[6]), is_packed=bool(v[7]),
                possible_dimensions=[(v[9 + 2 * k], v[10 + 2 * k]) for k in range(v[8])])


def guess_pattern(data, w, h, bd, packed=False):
    """guess_pattern (src/analyzer.cpp:216-321), restated."""
    buf = np.ascontiguousarray(data).view(np.uint8).ravel()
    lib = oracle()
    lib.isp_guess_pattern.argtypes = [_u8p] + [ctypes.c_int] * 4
    return lib.isp_guess_pattern(buf.ctypes.data_as(_u8p), w, h, bd, int(packed))


def ref_analyze_data(data):
    """imop::analyze_data of the compiled reference (src/analyzer.cpp:151-214):
    a dict of DataInfo's fields, possible_dimensions as a list of (w, h)."""
    buf = np.ascontiguousarray(data).view(np.uint8).ravel()
    out = (ctypes.c_int * 25)()
    reference_isp().ref_analyze_data(buf.ctypes.data_as(_u8p), buf.size, out)
    v = list(out)
    return dict(detected_bit_depth=v[0], suggested_width=v[1], suggested_height=v[2], pixel_count=v[3],
                max_value=v[4], min_value=v[5], is_likely_16bit=bool(v[6]), is_packed=bool(v[7]),
                possible_dimensions=[(v[9 + 2 * k], v[10 + 2 * k]) for k in range(v[8])])


def ref_guess_pattern(data, w, h, bd, packed=False):
    """imop::guess_pattern of the compiled reference (src/analyzer.cpp:216-321)."""
    buf = np.ascontiguousarray(data).view(np.uint8).ravel()
    return reference_isp().ref_guess_pattern(buf.ctypes.data_as(_u8p), w, h, bd, int(packed))


def ref_wb(rgb, bd, gains):
    return _rgb_stage(reference_isp(), "ref_wb_manual", rgb, bd, gains)


def ref_ccm(rgb, bd, m):
    return _rgb_stage(reference_isp(), "ref_ccm_linear", rgb, bd, m)


def ref_gamma(rgb, bd, gamma):
    return _rgb_stage(reference_isp(), "ref_gamma", rgb, bd, float(gamma))


def rgb_bytes(w: int, h: int, bd: int) -> int:
    return w * h * 3 * (2 if bd > 8 else 1)


def _as_rgb(buf: np.ndarray, w: int, h: int, bd: int) -> np.ndarray:
    if bd > 8:
        return buf.view(np.uint16).reshape(h, w, 3)
    return buf.reshape(h, w, 3)


def demosaic(raw: np.ndarray, w: int, h: int, pattern: int, algo: int, bd: int,
             packed: bool = False, black_level=None, per_channel: bool = False,
             out: np.ndarray | None = None) -> tuple[int, np.ndarray]:
    """Oracle demosaic of a byte buffer; returns (error_code, rgb array).

    black_level None -> demosaic only; else unpack -> black level -> demosaic.
    The output buffer starts zeroed (the reference's corner rule copies the
    caller's prior bytes, and zero is the convention every reference test
    uses)."""
    lib = oracle()
    raw = np.ascontiguousarray(raw).view(np.uint8).ravel()
    if out is None:
        out = np.zeros(rgb_bytes(w, h, bd), dtype=np.uint8)
    if black_level is None:
        e = lib.orc_demosaic(_ptr(raw), _ptr(out), w, h, pattern, algo, bd, int(packed))
    else:
        bl = np.asarray(black_level, dtype=np.float32)
        if bl.size == 1:
            bl = np.repeat(bl.reshape(1), 4)
        bl = np.ascontiguousarray(bl, dtype=np.float32)
        e = lib.orc_pipeline(_ptr(raw), _ptr(out), w, h, pattern, algo, bd, int(packed),
                             bl.ctypes.data_as(_fp), int(per_channel))
    return e, _as_rgb(out, w, h, bd)


def ref_demosaic(raw: np.ndarray, w: int, h: int, pattern: int, algo: int, bd: int,
                 packed: bool = False, black_level=None, per_channel: bool = False
                 ) -> tuple[int, np.ndarray]:
    """The compiled reference, composed the same way as `demosaic`."""
    lib = reference()
    raw = np.ascontiguousarray(raw).view(np.uint8).ravel().copy()
    out = np.zeros(rgb_bytes(w, h, bd), dtype=np.uint8)
    if black_level is not None:
        bl = np.asarray(black_level, dtype=np.float32)
        if bl.size == 1:
            bl = np.repeat(bl.reshape(1), 4)
        bl = np.ascontiguousarray(bl, dtype=np.float32)
        if packed:
            un = np.zeros(w * h * (2 if bd > 8 else 1), dtype=np.uint8)
            lib.ref_unpack(_ptr(raw), w, h, bd, 1, _ptr(un))
            raw = un
        e = lib.ref_black_level(_ptr(raw), w, h, pattern, bd, bl.ctypes.data_as(_fp),
                                int(per_channel))
        if e != 0:
            return e, _as_rgb(out, w, h, bd)
        packed = False
    e = lib.ref_demosaic_cpu(_ptr(raw), _ptr(out), w, h, pattern, algo, bd, int(packed))
    return e, _as_rgb(out, w, h, bd)


def unpack(raw: np.ndarray, w: int, h: int, bd: int, packed: bool) -> np.ndarray:
    """Oracle unpack to u8 (bd <= 8) or u16 values, shape (h, w)."""
    lib = oracle()
    raw = np.ascontiguousarray(raw).view(np.uint8).ravel()
    out = np.zeros(w * h * (2 if bd > 8 else 1), dtype=np.uint8)
    lib.orc_unpack(_ptr(raw), w, h, bd, int(packed), _ptr(out))
    return (out.view(np.uint16) if bd > 8 else out).reshape(h, w)


def black_level(bayer: np.ndarray, pattern: int, bd: int, offsets, per_channel: bool = False) -> np.ndarray:
    """Oracle black level (process_global / process_per_channel) on an
    unpacked (h, w) u8 / u16 frame; returns a copy."""
    out = np.ascontiguousarray(bayer).copy()
    h, w = out.shape
    bl = np.asarray(offsets, dtype=np.float32).ravel()
    if bl.size == 1:
        bl = np.repeat(bl, 4)
    bl = np.ascontiguousarray(bl, dtype=np.float32)
    oracle().orc_black_level(out.view(np.uint8).ctypes.data_as(_u8p), w, h, pattern, bd, bl.ctypes.data_as(_fp),
                             int(per_channel))
    return out


def pipeline(raw: np.ndarray, w: int, h: int, pattern: int, algo: int, bd: int, packed: bool = False,
             black_level_=None, per_channel: bool = False, lens=None, epilogue=None) -> np.ndarray:
    """The single-frame path as the reference's CPU functions in order:
    unpack (packed input only) -> black level -> polynomial lens shading (lens = (coef, center))
    -> demosaic -> white balance / colour matrix / gamma (epilogue = (flags,
    gains, ccm, gamma)).  Returns the RGB array."""
    if packed:
        u = unpack(raw, w, h, bd, packed)
    else:  # unpacked words go to the Bayer stages as stored (no MSB alignment)
        u = np.ascontiguousarray(raw).view(np.uint8).ravel()[: w * h * (2 if bd > 8 else 1)].copy()
        u = (u.view(np.uint16) if bd > 8 else u).reshape(h, w)
    if black_level_ is not None:
        u = black_level(u, pattern, bd, black_level_, per_channel)
    if lens is not None:
        u = lens_shading(u, pattern, bd, lens[0], lens[1])
    e, rgb = demosaic(u.view(np.uint8).ravel(), w, h, pattern, algo, bd, False)
    if e != 0:
        raise RuntimeError(f"oracle demosaic returned {e}")
    if epilogue is not None and epilogue[0]:
        rgb = isp_epilogue(rgb, bd, *epilogue)
    return rgb


def ref_pipeline(raw: np.ndarray, w: int, h: int, pattern: int, algo: int, bd: int, packed: bool = False,
                 black_level_=None, per_channel: bool = False, lens=None, epilogue=None) -> np.ndarray:
    """`pipeline` composed from the compiled reference's own functions."""
    lib = reference()
    raw = np.ascontiguousarray(raw).view(np.uint8).ravel()
    if packed:
        u = np.zeros(w * h * (2 if bd > 8 else 1), dtype=np.uint8)
        lib.ref_unpack(_ptr(raw.copy()), w, h, bd, int(packed), _ptr(u))
    else:
        u = raw[: w * h * (2 if bd > 8 else 1)].copy()
    if black_level_ is not None:
        bl = np.asarray(black_level_, dtype=np.float32).ravel()
        bl = np.ascontiguousarray(np.repeat(bl, 4) if bl.size == 1 else bl, dtype=np.float32)
        e = lib.ref_black_level(_ptr(u), w, h, pattern, bd, bl.ctypes.data_as(_fp), int(per_channel))
        if e != 0:
            raise RuntimeError(f"reference black level returned {e}")
    img = u.view(np.uint16 if bd > 8 else np.uint8).reshape(h, w)
    if lens is not None:
        img = ref_lens_shading(img, pattern, bd, lens[0], lens[1])
    out = np.zeros(rgb_bytes(w, h, bd), dtype=np.uint8)
    e = lib.ref_demosaic_cpu(_ptr(np.ascontiguousarray(img).view(np.uint8).ravel().copy()), _ptr(out), w, h,
                             pattern, algo, bd, 0)
    if e != 0:
        raise RuntimeError(f"reference demosaic returned {e}")
    rgb = _as_rgb(out, w, h, bd)
    if epilogue is not None:
        flags, gains, m, gamma = epilogue
        if flags & 1:
            rgb = ref_wb(rgb, bd, gains)
        if flags & 2:
            rgb = ref_ccm(rgb, bd, m)
        if flags & 4:
            rgb = ref_gamma(rgb, bd, gamma)
    return rgb


# ------------------------------------------------------------ FP64 ISP chain

_dp = ctypes.POINTER(ctypes.c_double)


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f(a):
    return None if a is None else a.ctypes.data_as(_fp)


def bilinear_csr(w: int, h: int, pattern: int):
    """The assembled bilinear demosaic operator D (scipy CSR, FP64): rows =
    3 * w * h output samples (pixel-major), columns = w * h raw pixels."""
    import scipy.sparse as sp
    L = oracle()
    nnz = L.orc_bilinear_csr(w, h, pattern, None, None, None)
    rowptr = np.zeros(w * h * 3 + 1, dtype=np.int64)
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.float64)
    L.orc_bilinear_csr(w, h, pattern, rowptr.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                       col.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(val))
    return sp.csr_matrix((val[:nnz], col[:nnz], rowptr), shape=(w * h * 3, w * h))


def bilinear_apply64(x: np.ndarray, pattern: int) -> np.ndarray:
    """D x for one frame x[h, w] (FP64) -> [h, w, 3]."""
    h, w = x.shape
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(h * w * 3, dtype=np.float64)
    oracle().orc_bilinear_apply64(_d(x), w, h, pattern, _d(out))
    return out.reshape(h, w, 3)


def bilinear_apply_t64(v: np.ndarray, pattern: int) -> np.ndarray:
    """D^T v for one frame v[h, w, 3] (FP64) -> [h, w]."""
    h, w, _ = v.shape
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(h * w, dtype=np.float64)
    oracle().orc_bilinear_apply_t64(_d(v), w, h, pattern, _d(out))
    return out.reshape(h, w)


class chain_plant:
    """Context manager that plants a mistake in the FP64 chain forward (bits of
    orc_chain_set_plant: 1 CCM transposed, 2 gamma in place of 1 / gamma,
    4 white balance unclamped, 8 CCM before white balance)."""

    def __init__(self, mask: int):
        self.mask = mask

    def __enter__(self):
        oracle().orc_chain_set_plant(self.mask)
        return self

    def __exit__(self, *exc):
        oracle().orc_chain_set_plant(0)
        return False


def chain_fwd64(raw: np.ndarray, pattern: int, max_value: float, gains, ccm, gamma: float, target=None):
    """FP64 chain forward over raw[n, h, w] (float32, units of max_value):
    returns (out[n, h, w, 3] FP64, MSE vs target or None)."""
    n, h, w = raw.shape
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    g = np.ascontiguousarray(gains, dtype=np.float64)
    M = np.ascontiguousarray(ccm, dtype=np.float64).ravel()
    out = np.zeros((n, h, w, 3), dtype=np.float64)
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
    loss = ctypes.c_double(0.0)
    rc = oracle().orc_chain_fwd64(_f(raw), n, w, h, pattern, float(max_value), _d(g), _d(M), 1.0 / gamma,
                                  _d(out), _f(t), ctypes.byref(loss))
    assert rc == 0
    return out, (loss.value if target is not None else None)


def chain_vjp64(raw: np.ndarray, pattern: int, max_value: float, gains, ccm, gamma: float, grad_out=None,
                target=None):
    """FP64 VJP of the chain (transposed assembled operator): returns
    (grad_raw[n, h, w], grad_gains[3], grad_ccm[3, 3]); cotangent grad_out or
    the MSE gradient against target."""
    n, h, w = raw.shape
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    g = np.ascontiguousarray(gains, dtype=np.float64)
    M = np.ascontiguousarray(ccm, dtype=np.float64).ravel()
    go = None if grad_out is None else np.ascontiguousarray(grad_out, dtype=np.float32)
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
    gr = np.zeros((n, h, w), dtype=np.float64)
    gg = np.zeros(3, dtype=np.float64)
    gm = np.zeros(9, dtype=np.float64)
    rc = oracle().orc_chain_vjp64(_f(raw), n, w, h, pattern, float(max_value), _d(g), _d(M), 1.0 / gamma, _f(go),
                                  _f(t), _d(gr), _d(gg), _d(gm))
    assert rc == 0
    return gr, gg, gm.reshape(3, 3)
