"""PRISM and RCD normalise each CFA sample as (float)v / (float)max
(prism.cpp:58-78; rcd.cpp:24-41).  The GPU passes compute it as v * y plus
one fused remainder correction, y = 1 / max rounded to nearest.
cfa_norm.c checks that sequence against the correctly rounded quotient for
every sample value 0..65535 at every bit depth 1..16 -- the whole domain, so
the check is a proof for the kernels' inputs -- and that dropping the
correction is caught."""
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
    exe = str(tmp_path_factory.mktemp("cfa") / "cfa_norm")
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "cfa_norm.c"), "-lm"],
                   check=True)
    return exe


def test_every_sample_value_is_correctly_rounded(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "1048576 cases, 0 mismatches" in r.stdout


def test_planted_mistake_is_caught(checker):
    r = subprocess.run([checker, "plant"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and " 0 mismatches" not in r.stdout, r.stdout
