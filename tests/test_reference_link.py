"""The drop-in recipe, checked on the reference's own sources (CPU only).

The reference's im_operator library sources (its CMakeLists.txt
add_library list) are compiled with g++ in place, without
src/algorithms/cuda_dispatch.cpp and src/gpu_dispatch.cpp and without
IM_OPERATOR_HAS_CUDA, together with the integration unit
im-operator_amd/integration/modules_without_gpu.cpp, and linked into a
shared object against libimop_hip.so with --no-undefined: every GPU hook the
dispatchers call resolves, to the library or to the unit.  A program linked
the same way then runs imop::demosaic and the stage dispatchers on this
machine, which has no GPU, and gets the CPU path's bytes.  The struct
layouts imop_compat.hpp declares equal the reference headers' layouts.

Skipped when the reference sources are absent (they never reach the GPU
box).  The reference's src/local_contrast/{bilateral,unsharp}.cpp call
std::memcpy without including <cstring>; they are compiled with
-include cstring.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import imop_amd
import oracle as O

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UNIT = os.path.join(ROOT, "im-operator_amd", "integration", "modules_without_gpu.cpp")
COMPAT_DIR = os.path.join(ROOT, "im-operator_amd", "csrc")
LIBDIR = os.path.join(ROOT, "im-operator_amd")
REPLACED = ("src/algorithms/cuda_dispatch.cpp", "src/gpu_dispatch.cpp")

pytestmark = pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "CMakeLists.txt")) or shutil.which("g++") is None,
                                reason="reference sources or g++ absent")


def _library_sources():
    text = open(os.path.join(REF, "CMakeLists.txt")).read()
    body = re.search(r"add_library\(im_operator STATIC(.*?)\)", text, flags=re.S).group(1)
    return [s for s in re.findall(r"(src/\S+\.cpp)", body) if s not in REPLACED]


def _includes():
    dirs = sorted(d for d in os.listdir(os.path.join(REF, "src")) if os.path.isdir(os.path.join(REF, "src", d)))
    return [f"-I{REF}/include"] + [f"-I{REF}/src/{d}" for d in dirs]


@pytest.fixture(scope="module")
def objects(tmp_path_factory):
    out = tmp_path_factory.mktemp("refobj")
    inc = _includes()
    srcs = [os.path.join(REF, s) for s in _library_sources()] + [UNIT]
    assert len(srcs) > 50

    def cc(src):
        obj = out / (src.replace("/", "_") + ".o")
        subprocess.run(["g++", "-std=c++17", "-O0", "-fPIC", "-w", "-fopenmp", "-mavx2", "-mfma", "-include", "cstring",
                        *inc, "-c", src, "-o", str(obj)], check=True)
        return str(obj)

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(cc, srcs))
    return out, objs


def test_reference_library_links_with_no_undefined_symbol(objects):
    out, objs = objects
    so = out / "libim_operator_hip.so"
    r = subprocess.run(["g++", "-shared", "-fopenmp", "-o", str(so), *objs, "-Wl,--no-undefined", f"-L{LIBDIR}",
                        "-limop_hip", f"-Wl,-rpath,{LIBDIR}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    # the hooks come from libimop_hip.so, not from a leftover definition
    und = subprocess.run(["nm", "-uC", str(so)], capture_output=True, text=True, check=True).stdout
    for name in ("imop::demosaic_cuda(", "black_level::process_black_level_cuda(",
                 "white_balance::process_white_balance_cuda(", "ccm::process_ccm_cuda(", "tone::process_tone_cuda(",
                 "lens_shading::process_lens_shading_cuda(", "defect_correct::process_defect_correct_cuda("):
        assert name in und, name


DRIVER = r'''
#include <cstdio>
#include <vector>
#include "imop/algorithms.hpp"
#include "black_level/algorithms.hpp"
#include "white_balance/algorithms.hpp"
#include "ccm/algorithms.hpp"
#include "tone/algorithms.hpp"
#include "lens_shading/algorithms.hpp"
#include "defect_correct/algorithms.hpp"

// argv: w h bd algo infile outfile; the RGB stages run on the demosaic output
int main(int argc, char** argv) {
    int w = atoi(argv[1]), h = atoi(argv[2]), bd = atoi(argv[3]), algo = atoi(argv[4]);
    size_t ib = (size_t)w * h * (bd > 8 ? 2 : 1), ob = ib * 3;
    std::vector<uint8_t> raw(ib), rgb(ob), tmp(ob);
    FILE* f = fopen(argv[5], "rb");
    if (fread(raw.data(), 1, ib, f) != ib) return 2;
    fclose(f);
    black_level::BlackLevelParams bp; bp.r_offset = bp.gr_offset = bp.gb_offset = bp.b_offset = 16.0f;
    if (black_level::process_black_level(raw.data(), w, h, imop::BayerPattern::RGGB,
                                         black_level::BlackLevelAlgorithm::GLOBAL, bd, bp) != black_level::BlackLevelError::Ok) return 3;
    lens_shading::LensShadingParams lp;
    lp.r_coef.a2 = 0.3f; lp.gr_coef.a2 = 0.2f; lp.gb_coef.a2 = 0.2f; lp.b_coef.a2 = 0.4f;
    if (lens_shading::process_lens_shading(raw.data(), w, h, imop::BayerPattern::RGGB,
                                           lens_shading::LensShadingAlgorithm::POLYNOMIAL, bd, lp) != lens_shading::LensShadingError::Ok) return 4;
    defect_correct::DefectCorrectParams dp; dp.threshold = 0.3f;
    if (defect_correct::process_defect_correct(raw.data(), w, h, imop::BayerPattern::RGGB,
                                               defect_correct::DefectCorrectAlgorithm::ADAPTIVE, bd, dp) != defect_correct::DefectCorrectError::Ok) return 5;
    if (imop::demosaic(raw.data(), rgb.data(), w, h, imop::BayerPattern::RGGB, (imop::DemosaicAlgorithm)algo, bd) != imop::DemosaicError::Ok) return 6;
    white_balance::WBCoefficients g; g.r_gain = 1.5f; g.g_gain = 1.0f; g.b_gain = 1.25f;
    if (white_balance::process_white_balance(rgb.data(), tmp.data(), w, h, 3, white_balance::WhiteBalanceAlgorithm::MANUAL, bd, 6.0f, g) != white_balance::WhiteBalanceError::Ok) return 7;
    ccm::CCMatrix3x3 m;
    const float mm[9] = {1.5f, -0.25f, -0.25f, -0.125f, 1.25f, -0.125f, 0.0f, -0.5f, 1.5f};
    for (int k = 0; k < 9; ++k) m.m[k] = mm[k];
    if (ccm::process_ccm(tmp.data(), rgb.data(), w, h, 3, ccm::CCMAlgorithm::LINEAR_3X3, bd, &m) != ccm::CCMError::Ok) return 8;
    tone::ToneParams tp; tp.gamma = 2.2f;
    if (tone::process_tone(rgb.data(), tmp.data(), w, h, 3, tone::ToneAlgorithm::GAMMA, bd, tp) != tone::ToneError::Ok) return 9;
    f = fopen(argv[6], "wb");
    fwrite(raw.data(), 1, ib, f);
    fwrite(tmp.data(), 1, ob, f);
    fclose(f);
    printf("has_cuda %d %d %d %d %d %d %d\n", (int)imop::has_cuda(), (int)black_level::has_cuda(),
           (int)white_balance::has_cuda(), (int)ccm::has_cuda(), (int)tone::has_cuda(), (int)lens_shading::has_cuda(),
           (int)defect_correct::has_cuda());
    return 0;
}
'''


def test_reference_dispatchers_run_through_the_library(objects, tmp_path):
    """A program on the reference's API, linked against the reference objects
    and libimop_hip.so: without a device every hook reports no GPU and the
    dispatchers' CPU path gives the oracle's bytes for black level -> lens
    shading -> defect correction -> demosaic -> white balance -> CCM ->
    gamma."""
    if imop_amd.lib().imop_hip_device_count() != 0:
        pytest.skip("a device is visible: the GPU form of this chain is tests/test_gpu_reference_api.py")
    _, objs = objects
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    subprocess.run(["g++", "-std=c++17", "-O1", "-w", "-fopenmp", *_includes(), str(src), *objs, "-o", str(exe),
                    f"-L{LIBDIR}", "-limop_hip", f"-Wl,-rpath,{LIBDIR}"], check=True)
    w, h, bd, algo = 64, 48, 12, O.MG
    rng = np.random.default_rng(7)
    raw = rng.integers(0, 4096, size=(h, w)).astype(np.uint16)
    (tmp_path / "in.raw").write_bytes(raw.tobytes())
    r = subprocess.run([str(exe), str(w), str(h), str(bd), str(algo), str(tmp_path / "in.raw"), str(tmp_path / "out")],
                       capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    assert r.stdout.split() == ["has_cuda"] + ["0"] * 7
    got = np.frombuffer((tmp_path / "out").read_bytes(), dtype=np.uint16)
    got_bayer, got_rgb = got[: w * h].reshape(h, w), got[w * h:].reshape(h, w, 3)
    b = raw.copy()
    O.oracle().orc_black_level(b.view(np.uint8).ctypes.data_as(O._u8p), w, h, 0, bd,
                               np.full(4, 16.0, np.float32).ctypes.data_as(O._fp), 0)
    b = O.lens_shading(b, 0, bd, [0.3, 0, 0, 0.2, 0, 0, 0.2, 0, 0, 0.4, 0, 0])
    e, b = O.defect_adaptive(b, bd, 0.3)
    assert e == 0
    np.testing.assert_array_equal(got_bayer, b)
    e, rgb = O.demosaic(b, w, h, 0, algo, bd)
    assert e == 0
    want = O.isp_epilogue(rgb, bd, 7, [1.5, 1.0, 1.25], [1.5, -0.25, -0.25, -0.125, 1.25, -0.125, 0.0, -0.5, 1.5], 2.2)
    np.testing.assert_array_equal(got_rgb, want)


LAYOUT = r'''
#include <cstddef>
#include <cstdio>
%s
int main() {
    printf("%%zu %%zu %%zu\n", sizeof(tone::ToneParams), offsetof(tone::ToneParams, gamma), offsetof(tone::ToneParams, mid_point));
    printf("%%zu %%zu %%zu %%zu %%zu\n", sizeof(lens_shading::LensShadingParams), offsetof(lens_shading::LensShadingParams, gb_coef),
           offsetof(lens_shading::LensShadingParams, center_y), offsetof(lens_shading::LensShadingParams, flat_field),
           offsetof(lens_shading::LensShadingParams, flat_field_height));
    printf("%%zu %%zu %%zu\n", sizeof(defect_correct::DefectCorrectParams), offsetof(defect_correct::DefectCorrectParams, map),
           offsetof(defect_correct::DefectCorrectParams, map_count));
    printf("%%zu %%zu\n", sizeof(black_level::BlackLevelParams), offsetof(black_level::BlackLevelParams, b_offset));
    printf("%%d %%d %%d %%d\n", (int)tone::ToneAlgorithm::SHADOWS_HIGHLIGHTS, (int)lens_shading::LensShadingAlgorithm::FLAT_FIELD,
           (int)defect_correct::DefectCorrectAlgorithm::MAP_BASED, (int)ccm::CCMError::InternalError);
    return 0;
}
'''


def test_compat_declarations_match_reference_headers(tmp_path):
    outs = []
    for k, (hdr, inc) in enumerate([('#include "imop_compat.hpp"', [f"-I{COMPAT_DIR}"]),
                                    ('#define IMOP_WITH_REFERENCE_HEADERS\n#include "imop_compat.hpp"',
                                     [f"-I{COMPAT_DIR}"] + _includes())]):
        src = tmp_path / f"layout{k}.cpp"
        src.write_text(LAYOUT % hdr)
        exe = tmp_path / f"layout{k}"
        subprocess.run(["g++", "-std=c++17", *inc, str(src), "-o", str(exe)], check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1], outs
