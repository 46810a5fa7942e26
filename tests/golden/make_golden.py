"""Write the golden vectors for the demosaic path from the compiled reference.

Run in the build container only (needs oracle/_ref/libimop_ref.so, built by
oracle/build_ref.sh from the reference sources, and the reference's
test_data/ frames):

    python tests/golden/make_golden.py

Outputs, committed next to this script:
  * golden.json -- for each case the parameters (the input is regenerated
    from a seed, or is one of the crops below) and the SHA-256 of the
    reference's RGB output bytes (zero-initialised output buffer).
  * crops.npz   -- small crops of the reference's own RG8 / RG10 / RG12 /
    RG12_Packed frames (test_data/*_w1006_h758_*.raw), raw bytes.

The reference is composed as the fused path is defined: with a black level,
unpack (pixel::get_raw) -> black_level::process_global / _per_channel ->
demosaic_cpu of the unpacked buffer; otherwise demosaic_cpu directly.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "im-operator_amd"))

import oracle as O  # noqa: E402
from imop_amd import scene  # noqa: E402

TEST_DATA = "/root/reference/test_data"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def synthetic_cases():
    """Small seeded cases: every algorithm x pattern x layout, odd and even
    sizes, minimum sizes, black level off / global / per channel."""
    cases = []
    layouts = [(8, False), (8, True), (10, False), (12, False), (10, True), (12, True), (14, True), (16, False),
               (9, True), (6, False)]
    rng = np.random.default_rng(20260510)
    k = 0
    for algo, min_dim in ((O.HQLI, 6), (O.MG, 6), (O.PRISM, 10), (O.RCD, 10), (O.SUPER_FAST, 2), (O.L7, 8), (O.AHD, 6),
                             (O.AMAZE, 6), (O.DFPD, 12)):
        for bd, packed in layouts:
            for pattern in range(4):
                w = int(rng.integers(min_dim, 41))
                h = int(rng.integers(min_dim, 37))
                bl_kind = ["none", "global", "per_channel"][k % 3]
                k += 1
                cases.append(dict(kind="random", w=w, h=h, bd=bd, packed=packed, pattern=pattern, algo=algo,
                                  bl=bl_kind, seed=1000 + k))
    # minimum and near-minimum sizes
    for algo, min_dim in ((O.HQLI, 6), (O.MG, 6), (O.PRISM, 10), (O.RCD, 10), (O.SUPER_FAST, 2), (O.L7, 8), (O.AHD, 6),
                             (O.AMAZE, 6), (O.DFPD, 12)):
        for (w, h) in ((min_dim, min_dim), (min_dim + 1, min_dim), (min_dim, min_dim + 3)):
            for bd, packed in ((8, False), (12, True), (10, True)):
                k += 1
                cases.append(dict(kind="random", w=w, h=h, bd=bd, packed=packed, pattern=k % 4, algo=algo,
                                  bl="none", seed=1000 + k))
    # seeded scenes at moderate size, all algorithms
    for algo in (O.HQLI, O.MG, O.PRISM, O.RCD, O.SUPER_FAST, O.L7, O.AHD, O.AMAZE, O.DFPD):
        for bd, packed in ((8, False), (12, True), (10, True)):
            k += 1
            cases.append(dict(kind="scene", w=256, h=192, bd=bd, packed=packed, pattern=0, algo=algo,
                              bl="global", seed=2000 + k, lines=50 if algo in (O.PRISM, O.RCD) else 0))
    return cases


def config_cases():
    """The workload configurations at their real sizes."""
    return [
        dict(kind="scene", w=1006, h=758, bd=8, packed=False, pattern=0, algo=O.HQLI, bl="none", seed=1, lines=0),
        dict(kind="scene", w=4024, h=3032, bd=12, packed=True, pattern=0, algo=O.MG, bl="global", bl_value=64.0,
             seed=2, lines=0),
        dict(kind="scene", w=4024, h=3032, bd=10, packed=True, pattern=0, algo=O.PRISM, bl="global", bl_value=64.0,
             seed=3, lines=200),
        dict(kind="scene", w=4024, h=3032, bd=10, packed=True, pattern=0, algo=O.RCD, bl="global", bl_value=64.0,
             seed=3, lines=200),
        dict(kind="scene", w=4024, h=3032, bd=12, packed=True, pattern=0, algo=O.SUPER_FAST, bl="global",
             bl_value=64.0, seed=2, lines=0),
        dict(kind="scene", w=4024, h=3032, bd=12, packed=True, pattern=0, algo=O.L7, bl="global",
             bl_value=64.0, seed=2, lines=0),
        dict(kind="scene", w=4024, h=3032, bd=12, packed=True, pattern=0, algo=O.AHD, bl="global",
             bl_value=64.0, seed=2, lines=0),
        dict(kind="scene", w=4024, h=3032, bd=12, packed=True, pattern=0, algo=O.AMAZE, bl="global",
             bl_value=64.0, seed=2, lines=0),
        dict(kind="scene", w=4024, h=3032, bd=12, packed=True, pattern=0, algo=O.DFPD, bl="global",
             bl_value=64.0, seed=2, lines=0),
    ]


def case_input(c):
    """Raw bytes of a case (the same function the tests use)."""
    if c["kind"] == "random":
        r = np.random.default_rng(c["seed"])
        vals = r.integers(0, 1 << c["bd"], size=(c["h"], c["w"]))
        vals = vals.astype(np.uint8 if c["bd"] <= 8 else np.uint16)
    else:
        vals = scene.make_scene(c["h"], c["w"], c["bd"], c["pattern"], seed=c["seed"], n_lines=c.get("lines", 0))
    return scene.raw_bytes(vals, c["bd"], c["packed"])


def case_black_level(c):
    kind = c["bl"]
    if kind == "none":
        return None, False
    if kind == "global":
        return [c.get("bl_value", 16.0)] * 4, False
    return [16.0, 3.5, 7.0, 64.0], True


def load_crops():
    files = {
        "RG8": ("Image_20260510144302933_w1006_h758_pBayerRG8.raw", 8, False),
        "RG10": ("Image_20260510144219156_w1006_h758_pBayerRG10.raw", 10, False),
        "RG12": ("Image_20260510144315897_w1006_h758_pBayerRG12.raw", 12, False),
        "RG12_Packed": ("Image_20260510144339211_w1006_h758_pBayerRG12_Packed.raw", 12, True),
    }
    W, H = 1006, 758
    crops = {}
    for name, (fn, bd, packed) in files.items():
        data = np.fromfile(os.path.join(TEST_DATA, fn), dtype=np.uint8)
        for (x0, y0, cw, ch) in ((300, 200, 64, 48), (0, 0, 40, 30), (W - 38, H - 26, 38, 26)):
            if packed:
                rows = data[: W * H * 3 // 2].reshape(H, W * 3 // 2)
                crop = rows[y0:y0 + ch, x0 * 3 // 2:(x0 + cw) * 3 // 2].copy().ravel()
            elif bd <= 8:
                crop = data.reshape(H, W)[y0:y0 + ch, x0:x0 + cw].copy().ravel()
            else:
                crop = data.view("<u2").reshape(H, W)[y0:y0 + ch, x0:x0 + cw].copy().view(np.uint8).ravel()
            crops[f"{name}_{x0}_{y0}_{cw}x{ch}"] = (crop, cw, ch, bd, packed)
    return crops


def main():
    assert O.reference_available(), "build oracle/_ref first (oracle/build_ref.sh)"
    out = {"note": "SHA-256 of the compiled reference's RGB bytes; see make_golden.py", "cases": []}
    for c in synthetic_cases() + config_cases():
        raw = case_input(c)
        bl, pc = case_black_level(c)
        e, rgb = O.ref_demosaic(raw, c["w"], c["h"], c["pattern"], c["algo"], c["bd"], c["packed"], bl, pc)
        c = dict(c)
        c["err"] = int(e)
        c["sha256"] = sha(rgb)
        out["cases"].append(c)
        print(c["kind"], c["w"], c["h"], c["algo"], c["bd"], c["packed"], c["bl"], e, flush=True)
    crops = load_crops()
    arrays = {}
    for name, (crop, cw, ch, bd, packed) in crops.items():
        arrays[name] = crop
        for algo in (O.HQLI, O.MG, O.PRISM, O.RCD, O.SUPER_FAST, O.L7, O.AHD, O.AMAZE, O.DFPD):
            for bl in (None, [64.0] * 4):
                e, rgb = O.ref_demosaic(crop, cw, ch, 0, algo, bd, packed, bl, False)
                out["cases"].append(dict(kind="crop", name=name, w=cw, h=ch, bd=bd, packed=packed, pattern=0,
                                         algo=algo, bl="none" if bl is None else "global", bl_value=64.0,
                                         err=int(e), sha256=sha(rgb)))
    np.savez_compressed(os.path.join(HERE, "crops.npz"), **arrays)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("cases:", len(out["cases"]))


if __name__ == "__main__":
    main()
