"""The C ABI driven from a plain C program (tests/capi_driver.c, built with
gcc against include/imop_hip.h and libimop_hip.so): the host-buffer entry
point equals the oracle with no Python binding in between."""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

import imop_amd
import oracle as O
from imop_amd import scene

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = tmp_path_factory.mktemp("capi") / "capi_driver"
    libdir = os.path.dirname(imop_amd.LIB_PATH)
    subprocess.run([gcc, "-O2", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "capi_driver.c"), "-o",
                    str(exe), "-L", libdir, "-limop_hip", f"-Wl,-rpath,{libdir}"], check=True)
    return str(exe)


@pytest.mark.parametrize("w,h,pat,algo,bd,packed,bl", [
    (4024, 3032, 0, 2, 12, 1, 64.0),
    (1006, 758, 1, 1, 8, 0, -1.0),
    (640, 482, 2, 8, 10, 1, 16.0),
    (131, 67, 3, 2, 14, 0, -1.0),
])
def test_c_program_matches_oracle(driver, tmp_path, w, h, pat, algo, bd, packed, bl):
    vals = scene.make_scene(h, w, bd, pat, seed=w + bd)
    raw = scene.raw_bytes(vals, bd, bool(packed))
    src, dst = tmp_path / "in.raw", tmp_path / "out.rgb"
    raw.tofile(src)
    print(f"capi {w}x{h} pat {pat} algo {algo} bd {bd} packed {packed}", flush=True)
    r = subprocess.run([driver, str(src), str(dst), str(w), str(h), str(pat), str(algo), str(bd), str(packed),
                        str(bl)], timeout=120)
    assert r.returncode == 0
    got = np.fromfile(dst, dtype=np.uint16 if bd > 8 else np.uint8).reshape(h, w, 3)
    e, want = O.demosaic(raw, w, h, pat, algo, bd, bool(packed), None if bl < 0 else [bl] * 4)
    assert e == 0
    np.testing.assert_array_equal(got, want)
