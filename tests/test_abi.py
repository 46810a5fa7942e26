"""The C ABI library: it loads, exports every entry point include/imop_hip.h
declares plus the reference operator symbols it stands in for, and behaves
like the reference's no-GPU build when no device is visible.  No kernel is
launched here."""
from __future__ import annotations

import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import imop_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "include")

# Mangled names of the reference functions this library defines
# (include/imop/algorithms.hpp:65-84, include/black_level/algorithms.hpp:40-48).
CXX_SYMBOLS = [
    "_ZN4imop8has_cudaEv",
    "_ZN4imop16cuda_device_nameEv",
    "_ZN4imop16cuda_synchronizeEv",
    "_ZN4imop13demosaic_cudaEPKhPhiiNS_12BayerPatternENS_17DemosaicAlgorithmEib",
    "_ZN4imop19demosaic_cuda_batchEPKPKhPKPhiiiNS_12BayerPatternENS_17DemosaicAlgorithmEi",
    "_ZN11black_level8has_cudaEv",
    "_ZN11black_level24process_black_level_cudaEPhiiN4imop12BayerPatternENS_19BlackLevelAlgorithmEiRKNS_16BlackLevelParamsE",
    # src/gpu_dispatch.cpp:341-343, 317-319, 290-292, 455-457, 527-529
    "_ZN13white_balance8has_cudaEv",
    "_ZN13white_balance26process_white_balance_cudaEPKhPhiiiifff",
    "_ZN3ccm8has_cudaEv",
    "_ZN3ccm16process_ccm_cudaEPKhPhiiiiPKf",
    "_ZN4tone8has_cudaEv",
    "_ZN4tone17process_tone_cudaEPKhPhiiiNS_13ToneAlgorithmEiRKNS_10ToneParamsE",
    "_ZN12lens_shading8has_cudaEv",
    "_ZN12lens_shading25process_lens_shading_cudaEPhiiN4imop12BayerPatternENS_20LensShadingAlgorithmEiRKNS_17LensShadingParamsE",
    "_ZN14defect_correct8has_cudaEv",
    "_ZN14defect_correct27process_defect_correct_cudaEPhiiN4imop12BayerPatternENS_22DefectCorrectAlgorithmEiRKNS_19DefectCorrectParamsE",
]


def declared_functions():
    names = []
    for fn in sorted(os.listdir(HEADER_DIR)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(HEADER_DIR, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(imop_hip_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_symbol():
    lib = imop_amd.lib()
    names = declared_functions()
    assert len(names) >= 7
    for n in names:
        assert hasattr(lib, n), n


def test_reference_operator_symbols_exported():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", imop_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    for s in CXX_SYMBOLS:
        assert re.search(r"\sT\s" + re.escape(s) + r"$", out, flags=re.M), s


def test_no_device_behaves_like_reference_stub():
    lib = imop_amd.lib()
    if lib.imop_hip_device_count() != 0:
        pytest.skip("a device is visible")
    # cuda_dispatch.cpp:217-239: without a GPU every entry point refuses
    # with InternalError and the caller falls back to its CPU path.
    assert imop_amd.cuda_device_name() == "N/A"
    raw = np.zeros(64, dtype=np.uint8)
    e, _ = imop_amd.demosaic_host(raw, 8, 8, 0, imop_amd.DemosaicAlgorithm.MG, 8)
    assert e == imop_amd.DemosaicError.InternalError
    assert lib.imop_hip_synchronize() == 6
    assert imop_amd.demosaic_batch([], stream=0) == (6, -1)
    plan = imop_amd.BatchPlan([])
    assert plan.error == 6
    ep = imop_amd.rgb_epilogue(gains=[2.0, 1.0, 1.5], gamma=2.2)
    assert lib.imop_hip_demosaic_isp(None, 0, None, 0, 8, 8, 0, 2, 8, 0, None, 0, ctypes.byref(ep), None) == 6
    assert lib.imop_hip_rgb_epilogue_apply(None, 0, 8, 8, 8, ctypes.byref(ep), None) == 6
    ls = imop_amd.lens_shading_params([0.5, 0.0, 0.0] * 4)
    assert lib.imop_hip_lens_shading_poly(None, 0, 8, 8, 0, 8, ctypes.byref(ls), None) == 6
    assert lib.imop_hip_defect_adaptive(None, 0, 8, 8, 0, 8, ctypes.c_float(0.3), None) == 6
    info = imop_amd.DataInfo()
    assert lib.imop_hip_analyze_data(None, 0, ctypes.byref(info), None) == 6
    pat = ctypes.c_int(-1)
    assert lib.imop_hip_guess_pattern(None, 0, 8, 8, 8, 0, ctypes.byref(pat), None) == 6


def test_launch_shape_auto_rows_round_trip():
    """rows 0 is the automatic band height: get reports it, a get / set round
    trip keeps it, and -1 / out-of-range values leave a setting unchanged.
    Only host state; no kernel runs."""
    saved = imop_amd.get_launch_shape()
    try:
        imop_amd.set_launch_shape(0, 4, 3)
        assert imop_amd.get_launch_shape() == (0, 4, 3)
        imop_amd.set_launch_shape(*imop_amd.get_launch_shape())
        assert imop_amd.get_launch_shape() == (0, 4, 3)
        imop_amd.set_launch_shape(2, 4, 1)
        assert imop_amd.get_launch_shape() == (2, 4, 1)
        imop_amd.set_launch_shape(-1, 0, -1)
        assert imop_amd.get_launch_shape() == (2, 4, 1)
        # rows 2 / 4 and groups of 4 waves only
        imop_amd.set_launch_shape(8, 2, 7)
        assert imop_amd.get_launch_shape() == (2, 4, 1)
        imop_amd.set_launch_shape(4, 1, 2)
        assert imop_amd.get_launch_shape() == (4, 4, 2)
        imop_amd.set_launch_shape(0)
        assert imop_amd.get_launch_shape() == (0, 4, 2)
    finally:
        imop_amd.set_launch_shape(*saved)


@pytest.mark.parametrize("ctype,cname", [(imop_amd.FrameDesc, "imop_hip_frame_desc"),
                                         (imop_amd.RgbEpilogue, "imop_hip_rgb_epilogue"),
                                         (imop_amd.DataInfo, "imop_hip_data_info")])
def test_struct_layout_matches_header(tmp_path, ctype, cname):
    """The ctypes mirrors of the header's structs have the C compiler's layout."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    fields = [f[0] for f in ctype._fields_]
    cnames = ["in" if f == "in_" else f for f in fields]
    body = "".join(f'printf("%zu\\n", offsetof({cname}, {c}));' for c in cnames)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "imop_hip.h"\n'
                   f'int main(void) {{ printf("%zu\\n", sizeof({cname})); ' + body + ' return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-I", HEADER_DIR, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(ctype)] + [getattr(ctype, f).offset for f in fields]
    assert got == want


def test_enums_match_reference_numbering():
    A = imop_amd.DemosaicAlgorithm
    assert (A.SUPER_FAST, A.HQLI, A.MG, A.RCD, A.PRISM) == (0, 1, 2, 7, 8)
    P = imop_amd.BayerPattern
    assert (P.RGGB, P.BGGR, P.GRBG, P.GBRG) == (0, 1, 2, 3)
    E = imop_amd.DemosaicError
    assert (E.Ok, E.NullInput, E.ImageTooSmall, E.InternalError) == (0, 1, 5, 6)


def test_byte_size_formulas():
    # pixel_utils.hpp:21-34
    assert imop_amd.bayer_byte_size(4024, 3032, 12, True) == 18_301_152
    assert imop_amd.bayer_byte_size(4024, 3032, 10, True) == 15_250_960
    assert imop_amd.bayer_byte_size(1006, 758, 10, True) == 953_185
    assert imop_amd.rgb_byte_size(4024, 3032, 12) == 73_204_608
    assert imop_amd.rgb_byte_size(1006, 758, 8) == 2_287_644


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/algorithms"), reason="reference sources absent")
def test_reference_dispatch_links_against_library(tmp_path):
    """imop::demosaic (src/algorithms/dispatch.cpp) compiled from the
    reference's own source, linked with this library in place of
    cuda_dispatch.cpp: every GPU symbol it needs resolves here."""
    ref = "/root/reference"
    obj = tmp_path / "dispatch.o"
    subprocess.run(["g++", "-std=c++17", "-c", "-w", f"-I{ref}/include", f"-I{ref}/src/algorithms",
                    f"{ref}/src/algorithms/dispatch.cpp", "-o", str(obj)], check=True)
    und = subprocess.run(["nm", "-u", "-C", str(obj)], capture_output=True, text=True, check=True).stdout
    needed = [l.split(" U ")[-1].strip() for l in und.splitlines() if "imop::" in l]
    gpu = [n for n in needed if "cuda" in n]
    assert gpu, needed
    defined = subprocess.run(["nm", "-D", "-C", "--defined-only", imop_amd.LIB_PATH], capture_output=True,
                             text=True, check=True).stdout
    for n in gpu:
        assert n in defined, n
