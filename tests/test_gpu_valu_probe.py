"""imop_hip_valu_issue_rate, the VALU issue-rate probe the bench's VALU
rooflines are calibrated with: a positive rate, within the part's range
(1,024 SIMDs at most 2.4 GHz / 4 cycles per wave64 instruction, with margin
for clocks), and the null-pointer error."""
from __future__ import annotations

import ctypes

import pytest

import imop_amd

pytestmark = pytest.mark.gpu

MODEL = 1024 * 2.4e9 / 4  # wave64 VALU instructions per second


def test_rate_is_in_the_parts_range():
    r = imop_amd.valu_issue_rate()
    assert 0.3 * MODEL < r < 1.25 * MODEL, r


def test_null_pointer_is_rejected():
    assert imop_amd.lib().imop_hip_valu_issue_rate(None, None) == 1
