"""Seeded inputs of the raw-data analyzer tests (imop::analyze_data and
imop::guess_pattern, src/analyzer.cpp:151-321), shared by the oracle tests
and the GPU parity tests: every branch of detect_bit_depth (known 8-bit size,
u16 storage at each depth threshold, packed streams, the generic 16-bit rule,
the 8-bit fallback), the packed pixel count's +1 rounding correction, odd and
tiny sizes, and Bayer frames of every pattern and layout."""
from __future__ import annotations

import numpy as np

from imop_amd import scene


def lsb_stream(values, bd: int) -> np.ndarray:
    """values as an LSB-first bit stream of bd bits each (the layout
    analyze_data's packed sample decodes), ceil(n * bd / 8) bytes."""
    v = np.ascontiguousarray(values, dtype=np.uint64).ravel()
    bits = ((v[:, None] >> np.arange(bd, dtype=np.uint64)[None, :]) & 1).astype(np.uint8).ravel()
    return np.packbits(bits, bitorder="little")


def _u16(v) -> np.ndarray:
    return np.ascontiguousarray(v, dtype="<u2").view(np.uint8).ravel()


def _rng(tag: str):
    return np.random.default_rng(int.from_bytes(tag.encode()[:8].ljust(8, b"."), "little") % (1 << 32))


def _threshold(mx: int, at: int) -> np.ndarray:
    v = _rng(f"t{mx}").integers(0, 200, 640 * 480)
    v[at] = mx
    return _u16(v)


def analyze_buffers(big: bool = False) -> list[tuple[str, object]]:
    """(name, thunk) pairs, the thunk returning the bytes (so that listing
    the cases builds nothing); big adds full-sensor buffers (tens of MB)."""
    ms = scene.make_scene
    out = [
        ("u8-640x480-scene", lambda: scene.raw_bytes(ms(480, 640, 8, 0, seed=1), 8, False)),
        ("u8-1006x758-noise", lambda: _rng("u8n").integers(0, 256, 1006 * 758, dtype=np.uint8)),
        ("u16-1006x758-msb12", lambda: _u16(ms(758, 1006, 12, 2, seed=3).astype(np.uint32) << 4)),
    ]
    for bd in (8, 10, 12, 14, 16):
        out.append((f"u16-640x480-bd{bd}", lambda bd=bd: scene.raw_bytes(ms(480, 640, bd, 1, seed=bd), bd, False)))
    for mx, at in ((255, 5), (256, 99999), (1023, 7), (1024, 3), (4095, 11), (4096, 12), (16383, 0),
                   (16384, 99999), (65535, 100000)):
        out.append((f"u16-threshold-{mx}-at{at}", lambda mx=mx, at=at: _threshold(mx, at)))
    for bd in (10, 12, 14, 16):
        out.append((f"packed{bd}-640x480", lambda bd=bd: scene.pack(ms(480, 640, bd, 3, seed=10 + bd), bd)))
        out.append((f"lsb{bd}-1006x758", lambda bd=bd: lsb_stream(ms(758, 1006, bd, 0, seed=bd), bd)))
    # pixel count one short of a known size: the +1 correction
    out += [
        ("lsb14-1006x758-short", lambda: lsb_stream(ms(758, 1006, 14, 0, seed=5), 14)[:-1]),
        ("lsb10-640x480-short", lambda: lsb_stream(_rng("l10").integers(300, 1024, 640 * 480), 10)[:-1]),
        ("generic16-1000x700", lambda: _u16(_rng("g16").integers(256, 65536, 1000 * 700))),
        ("generic16-odd", lambda: _u16(_rng("g16o").integers(0, 65536, 54321))[:-1]),
        ("zeros-u16-640x480", lambda: np.zeros(2 * 640 * 480, dtype=np.uint8)),
        ("hi-bytes-no-size", lambda: _rng("hib").integers(1, 256, 200000, dtype=np.uint8)),
    ]
    for n in (1, 2, 3, 4, 7, 8, 9, 10, 257, 12345):
        out.append((f"random-{n}", lambda n=n: _rng(f"r{n}").integers(0, 256, n, dtype=np.uint8)))
    if big:
        out += [
            ("packed12-4024x3032", lambda: scene.pack(ms(3032, 4024, 12, 0, seed=2), 12)),
            ("u16-4024x3032-bd12", lambda: scene.raw_bytes(ms(3032, 4024, 12, 0, seed=4), 12, False)),
            ("u16-7680x4320-noise", lambda: _u16(_rng("big").integers(0, 4096, 7680 * 4320))),
            ("lsb16-3840x2160-short", lambda: lsb_stream(_rng("l16").integers(0, 65536, 3840 * 2160), 16)[:-1]),
        ]
    return out


def pattern_frames(big: bool = False) -> list[tuple[str, object, int, int, int, bool]]:
    """(name, thunk, w, h, bd, packed) Bayer frames for guess_pattern, the
    thunk returning the bytes."""
    ms = scene.make_scene
    out = []
    for w, h in [(4, 4), (5, 7), (64, 48), (130, 66), (640, 480)]:
        for pat in range(4):
            for bd, packed in ((8, False), (12, True), (12, False), (10, True), (16, False), (14, True)):
                if (w * h) > 5000 and (pat + bd) % 3:  # thin out the larger frames
                    continue
                out.append((f"{w}x{h}-pat{pat}-bd{bd}{'p' if packed else ''}",
                            lambda w=w, h=h, pat=pat, bd=bd, packed=packed:
                            scene.raw_bytes(ms(h, w, bd, pat, seed=w + pat + bd), bd, packed), w, h, bd, packed))
    for pat in range(4):
        out.append((f"ramp16-pat{pat}", lambda pat=pat: _u16(ramp_frame(pat)), 256, 256, 16, False))
    out += [
        ("msb12-130x66", lambda: _u16(ms(66, 130, 12, 1, seed=1).astype(np.uint32) << 4), 130, 66, 12, False),
        ("u8-as-packed-6bit", lambda: _rng("p6").integers(0, 64, 96 * 64, dtype=np.uint8), 96, 64, 6, True),
        ("constant", lambda: np.full(64 * 48, 77, dtype=np.uint8), 64, 48, 8, False),
    ]
    if big:
        out.append(("4024x3032-pat2-bd12p", lambda: scene.raw_bytes(ms(3032, 4024, 12, 2, seed=9), 12, True),
                    4024, 3032, 12, True))
    return out


def ramp_frame(pattern: int, n: int = 256) -> np.ndarray:
    """A 16-bit frame whose pattern guess_pattern recovers by construction: a
    gentle diagonal ramp (0 .. ~40000) plus 5000 at the R sites of
    `pattern` (G and B carry the ramp alone).  Every 2x2 block then votes for
    the pattern's family (|R - B| on one diagonal, G - G on the other), and R's
    offset puts the R-site mean 25% above the B-site mean."""
    yy, xx = np.mgrid[0:n, 0:n]
    ry, rx = {0: (0, 0), 1: (1, 1), 2: (0, 1), 3: (1, 0)}[pattern]
    v = (xx + yy) * 78 + 5000 * (((yy & 1) == ry) & ((xx & 1) == rx))
    return v.astype(np.uint16)
