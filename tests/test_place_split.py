"""Work-item placement of the single-frame and batched register-resident
kernels: item i of a frame's strips x group rows splits as i / d and i % d
through a 40-bit fixed-point reciprocal of d (device_common.hpp,
place_reciprocal / place_split).  place_split.cpp checks every divisor
1..65535 at the largest item count the reciprocal accepts, that this count
covers every one the kernels form, and that a floor instead of a ceiling
reciprocal is caught.  frame_major.cpp checks the batched kernels'
frame-major placement (batch.hpp, frame_major_local) for every frame size up
to 20000 items and a spread of larger ones, and that dropping the remainder
items is caught."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "im-operator_amd", "csrc")
ROCM_INCLUDE = "/opt/rocm/include"


def _build(tmp_path_factory, name):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(ROCM_INCLUDE):
        pytest.skip("g++ or the HIP headers not available")
    exe = str(tmp_path_factory.mktemp("place") / name)
    subprocess.run([gxx, "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INCLUDE, "-I" + CSRC, "-o", exe,
                    os.path.join(HERE, name + ".cpp")], check=True)
    return exe


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    return _build(tmp_path_factory, "place_split")


@pytest.fixture(scope="module")
def frame_major(tmp_path_factory):
    return _build(tmp_path_factory, "frame_major")


def test_every_divisor_splits_exactly(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "2097120 cases, 0 mismatches" in r.stdout


def test_planted_floor_reciprocal_is_caught(checker):
    r = subprocess.run([checker, "plant"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and " 0 mismatches" not in r.stdout, r.stdout


def test_frame_major_places_every_item_once(frame_major):
    r = subprocess.run([frame_major], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "20014 frames, 0 mismatches" in r.stdout


def test_planted_dropped_remainder_is_caught(frame_major):
    r = subprocess.run([frame_major, "plant"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and " 0 mismatches" not in r.stdout, r.stdout
