"""Shared access to the golden vectors (tests/golden/golden.json) and their
inputs.  Inputs are regenerated from seeds, or read from the committed crops
of the reference's own test frames."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

import make_golden as MG  # tests/golden/make_golden.py

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "golden.json")
CROPS = os.path.join(HERE, "golden", "crops.npz")

_crops = None


def cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def case_id(c) -> str:
    tag = c.get("name", c["kind"])
    return f"{tag}-{c['w']}x{c['h']}-a{c['algo']}-bd{c['bd']}{'p' if c['packed'] else ''}-pat{c['pattern']}-{c['bl']}"


def case_input(c) -> np.ndarray:
    global _crops
    if c["kind"] == "crop":
        if _crops is None:
            _crops = dict(np.load(CROPS))
        return _crops[c["name"]]
    return MG.case_input(c)


def case_black_level(c):
    if c["kind"] == "crop":
        return (None, False) if c["bl"] == "none" else ([c["bl_value"]] * 4, False)
    return MG.case_black_level(c)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def is_big(c) -> bool:
    return c["w"] * c["h"] > 2_000_000
