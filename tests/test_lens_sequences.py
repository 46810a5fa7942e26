"""The register-resident lens-shading decode replaces the correctly rounded
sqrt and division by shorter sequences (sqrt_rn, div_rn in
im-operator_amd/csrc/fast_impl.hpp) on frames whose centre and coefficients
the host admits (make_lens_params, isp_kernels.hip).  lens_sequences.c
checks both sequences on the CPU over that range against the C library's
correctly rounded sqrtf and division; the GPU parity tests
(test_gpu_pipeline.py) then check the kernels themselves against the oracle."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path_factory.mktemp("lens") / "lens_sequences")
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "lens_sequences.c"), "-lm"],
                   check=True)
    return exe


def test_sequences_are_correctly_rounded(checker):
    r = subprocess.run([checker, "211"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout


def test_planted_mistake_is_caught(checker):
    r = subprocess.run([checker, "1001", "plant"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "div s=" in r.stdout, r.stdout
