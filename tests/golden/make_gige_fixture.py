"""Write tests/golden/rg10_gige.npz: row crops of the reference's own
GigE-packed 10-bit frame (test_data/*_w1006_h758_pBayerRG10_Packed.raw,
3 bytes per pixel pair, 1509 bytes per row), raw bytes as the file holds
them, plus a crop of its unpacked RG10 frame (little-endian u16) of the same
camera.  Run in the build container only (reads /root/reference):

    python tests/golden/make_gige_fixture.py

Keys: top (rows 0..63, 64 x 1509 bytes), bottom (rows 742..757, the end of
the stream, 16 x 1509), u16_top (rows 0..63 of the unpacked RG10 frame,
64 x 1006 u16)."""
from __future__ import annotations

import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TEST_DATA = "/root/reference/test_data"
W, H = 1006, 758


def main():
    packed = np.fromfile(glob.glob(os.path.join(TEST_DATA, "*_w1006_h758_pBayerRG10_Packed.raw"))[0], np.uint8)
    assert packed.size == W * H // 2 * 3
    rows = packed.reshape(H, W // 2 * 3)
    u16 = np.fromfile(glob.glob(os.path.join(TEST_DATA, "*_w1006_h758_pBayerRG10.raw"))[0], "<u2").reshape(H, W)
    np.savez_compressed(os.path.join(HERE, "rg10_gige.npz"), top=rows[:64].copy(), bottom=rows[H - 16:].copy(),
                        u16_top=u16[:64].copy())
    print("wrote rg10_gige.npz")


if __name__ == "__main__":
    main()
