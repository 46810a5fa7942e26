"""Seeded synthetic Bayer scenes and the raw byte layouts the demosaic path
reads.

Scene recipe: RGB = sum of 3 sinusoidal gratings (frequency uniform in
0.005-0.2 cycles/px, random orientation and phase, random colour) plus 20
random axis-aligned constant-colour rectangles, optionally 1-px diagonal
lines; scaled to [0, 2^bd - 1]; sampled through the CFA; Poisson shot noise
at 1 DN per electron plus Gaussian read noise (sigma 2 DN); rounded and
clipped.

Layouts (include/imop/pixel_utils.hpp:58-118 and the reference's packer,
tools/synthetic_test.cpp:418-426):
  * 12-bit packed: 2 px / 3 B, [P0 hi8][P1 lo4 << 4 | P0 lo4][P1 hi8]
  * 10-bit packed: 4 px / 5 B, four high bytes then the 2-bit tails, P0 in
    bits 1:0 of the fifth byte
  * other depths packed: LSB-first bitstream
Packing runs over the linear pixel index with no row padding; the stream is
ceil(w*h*bd/8) bytes.
"""
from __future__ import annotations

import numpy as np

RGGB, BGGR, GRBG, GBRG = 0, 1, 2, 3
_PHASE = {RGGB: (0, 0, 1, 1), BGGR: (1, 1, 0, 0), GRBG: (0, 1, 1, 0), GBRG: (1, 0, 0, 1)}


def site_map(h: int, w: int, pattern: int) -> np.ndarray:
    """0 = R, 1 = G, 2 = B for every pixel (PatternOffsets,
    include/imop/types.hpp:88-101)."""
    rr, rc, br, bc = _PHASE[pattern]
    y = (np.arange(h) & 1)[:, None]
    x = (np.arange(w) & 1)[None, :]
    s = np.ones((h, w), dtype=np.int8)
    s[(y == rr) & (x == rc)] = 0
    s[(y == br) & (x == bc)] = 2
    return s


def scene_rgb_fn(h: int, w: int, rng: np.random.Generator, n_lines: int = 0):
    """Return (channel-evaluator, rectangles, lines): the scene is evaluated
    per CFA site so that full-frame RGB planes are never materialised."""
    gr = []
    for _ in range(3):
        f = rng.uniform(0.005, 0.2)
        th = rng.uniform(0.0, np.pi)
        ph = rng.uniform(0.0, 2 * np.pi)
        col = rng.uniform(0.2, 1.0, size=3)
        gr.append((f, th, ph, col))
    rects = []
    for _ in range(20):
        rw = int(rng.integers(max(2, w // 40), max(3, w // 5)))
        rh = int(rng.integers(max(2, h // 40), max(3, h // 5)))
        x0 = int(rng.integers(0, max(1, w - rw)))
        y0 = int(rng.integers(0, max(1, h - rh)))
        rects.append((y0, x0, rh, rw, rng.uniform(0.0, 1.0, size=3)))
    lines = []
    for _ in range(n_lines):
        ln = int(rng.integers(16, max(17, min(h, w) // 4)))
        x0 = int(rng.integers(0, w))
        y0 = int(rng.integers(0, h))
        d = int(rng.choice([-1, 1]))
        lines.append((y0, x0, ln, d, rng.uniform(0.0, 1.0, size=3)))
    return gr, rects, lines


def make_scene(h: int, w: int, bd: int, pattern: int = RGGB, seed: int = 0,
               n_lines: int = 0) -> np.ndarray:
    """Seeded Bayer frame, values in [0, 2^bd - 1], dtype u8 (bd <= 8) or
    u16, shape (h, w)."""
    rng = np.random.default_rng(seed)
    gr, rects, lines = scene_rgb_fn(h, w, rng, n_lines)
    sites = site_map(h, w, pattern)
    yy = np.arange(h, dtype=np.float32)[:, None]
    xx = np.arange(w, dtype=np.float32)[None, :]
    field = np.zeros((h, w), dtype=np.float32)
    for f, th, ph, col in gr:
        arg = (2.0 * np.pi * f) * (xx * np.float32(np.cos(th)) + yy * np.float32(np.sin(th))) + ph
        amp = np.asarray(col, dtype=np.float32)[sites]
        field += amp * (0.5 + 0.5 * np.cos(arg, dtype=np.float32))
    field /= 3.0
    for y0, x0, rh, rw, col in rects:
        sub = sites[y0:y0 + rh, x0:x0 + rw]
        field[y0:y0 + rh, x0:x0 + rw] = np.asarray(col, dtype=np.float32)[sub]
    for y0, x0, ln, d, col in lines:
        t = np.arange(ln)
        ys = y0 + t
        xs = x0 + d * t
        ok = (ys >= 0) & (ys < h) & (xs >= 0) & (xs < w)
        ys, xs = ys[ok], xs[ok]
        field[ys, xs] = np.asarray(col, dtype=np.float32)[sites[ys, xs]]
    mx = (1 << bd) - 1
    lo, hi = float(field.min()), float(field.max())
    field = (field - lo) * (mx / max(hi - lo, 1e-6))
    noisy = rng.poisson(np.maximum(field, 0.0)).astype(np.float32)
    noisy += rng.normal(0.0, 2.0, size=(h, w)).astype(np.float32)
    out = np.clip(np.rint(noisy), 0, mx)
    return out.astype(np.uint8 if bd <= 8 else np.uint16)


def pack(values: np.ndarray, bd: int) -> np.ndarray:
    """Pack a Bayer frame (any shape, row-major) into the packed byte stream
    the reference decodes; returns u8 of ceil(n*bd/8) bytes."""
    v = np.ascontiguousarray(values).astype(np.uint32).ravel()
    n = v.size
    nbytes = (n * bd + 7) // 8
    out = np.zeros(nbytes, dtype=np.uint8)
    if bd == 12:
        pairs = n // 2
        p0 = v[0:2 * pairs:2]
        p1 = v[1:2 * pairs:2]
        g = np.empty((pairs, 3), dtype=np.uint8)
        g[:, 0] = (p0 >> 4) & 0xFF
        g[:, 1] = ((p1 & 0xF) << 4) | (p0 & 0xF)
        g[:, 2] = (p1 >> 4) & 0xFF
        out[:pairs * 3] = g.ravel()
        if n & 1:
            last = v[-1]
            o = pairs * 3
            out[o] = (last >> 4) & 0xFF
            if o + 1 < nbytes:
                out[o + 1] = last & 0xF
        return out
    if bd == 10:
        quads = n // 4
        q = v[:quads * 4].reshape(quads, 4)
        g = np.empty((quads, 5), dtype=np.uint8)
        g[:, :4] = (q >> 2) & 0xFF
        g[:, 4] = (q[:, 0] & 3) | ((q[:, 1] & 3) << 2) | ((q[:, 2] & 3) << 4) | ((q[:, 3] & 3) << 6)
        out[:quads * 5] = g.ravel()
        o = quads * 5
        rest = v[quads * 4:]
        tail = 0
        for s, p in enumerate(rest):
            if o + s < nbytes:
                out[o + s] = (p >> 2) & 0xFF
            tail |= int(p & 3) << (2 * s)
        if rest.size and o + 4 < nbytes:
            out[o + 4] = tail
        return out
    if bd == 8:
        out[:] = v.astype(np.uint8)
        return out
    # LSB-first bitstream
    bits = ((v[:, None] >> np.arange(bd, dtype=np.uint32)[None, :]) & 1).astype(np.uint8).ravel()
    pad = nbytes * 8 - bits.size
    if pad:
        bits = np.concatenate([bits, np.zeros(pad, dtype=np.uint8)])
    return np.packbits(bits.reshape(-1, 8), axis=1, bitorder="little").ravel()


def pack_gige10(values: np.ndarray) -> np.ndarray:
    """10-bit samples in the GigE Vision 3-byte packing (BayerXX10Packed):
    per pixel pair [P0 9:2][P1 1:0 << 4 | P0 1:0][P1 9:2]; an odd final pixel
    gets a zero partner."""
    v = np.ascontiguousarray(values, dtype=np.uint32).ravel() & 0x3FF
    if v.size & 1:
        v = np.concatenate([v, np.zeros(1, np.uint32)])
    e, o = v[0::2], v[1::2]
    g = np.empty((e.size, 3), np.uint8)
    g[:, 0] = e >> 2
    g[:, 1] = (e & 3) | ((o & 3) << 4)
    g[:, 2] = o >> 2
    return g.ravel()


def raw_bytes(values: np.ndarray, bd: int, packed) -> np.ndarray:
    """The input buffer for demosaic: packed stream (packed = 2: the GigE
    Vision 10-bit packing), or u8 / little-endian u16 per pixel."""
    if not isinstance(packed, bool) and packed == 2:
        return pack_gige10(values)
    if packed:
        return pack(values, bd)
    if bd <= 8:
        return np.ascontiguousarray(values, dtype=np.uint8).ravel()
    return np.ascontiguousarray(values, dtype="<u2").view(np.uint8).ravel()


def mixed_batch_spec(seed: int = 4, n_small: int = 48, n_large: int = 16, small=(1006, 758),
                     large=(4024, 3032)) -> list[dict]:
    """Frames of the mixed-size batch workload: n_small frames of `small` and
    n_large of `large` (w, h), each with a bit depth drawn from {8-bit
    unpacked, 10-bit packed, 12-bit packed} and a global black level from
    {0, 16, 64}, in a seeded order.  Entries: dict(w, h, bd, packed, bl, seed)."""
    rng = np.random.default_rng(seed)
    depths = [(8, False), (10, True), (12, True)]
    frames = []
    for k in range(n_small + n_large):
        w, h = small if k < n_small else large
        bd, packed = depths[int(rng.integers(0, 3))]
        bl = [0.0, 16.0, 64.0][int(rng.integers(0, 3))]
        frames.append(dict(w=w, h=h, bd=bd, packed=packed, bl=bl, seed=seed * 1000 + k))
    order = rng.permutation(len(frames))
    return [frames[i] for i in order]
