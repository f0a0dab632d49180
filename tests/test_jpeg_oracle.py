"""The frame grab's JPEG restatement (oracle/jpeg_oracle.cpp) against the
files libjpeg-turbo writes for the same framebuffers (tests/golden/
jpeg_golden.npz, made by tests/golden/make_jpeg_golden.py), byte for byte,
and against what T.81 fixes by hand for a flat image."""
import io
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "jpeg_golden.npz")


def _golden():
    g = np.load(GOLDEN)
    n = len([k for k in g.files if k.startswith("px_")])
    return [(g["px_%d" % i], int(g["q_%d" % i]), g["jpg_%d" % i].tobytes()) for i in range(n)]


def _segments(jpg):
    """(marker, payload) of every segment up to SOS, then the scan bytes."""
    assert jpg[:2] == b"\xff\xd8" and jpg[-2:] == b"\xff\xd9"
    i, segs = 2, []
    while True:
        assert jpg[i] == 0xFF
        m, n = jpg[i + 1], jpg[i + 2] << 8 | jpg[i + 3]
        segs.append((m, jpg[i + 4:i + 2 + n]))
        i += 2 + n
        if m == 0xDA:
            return segs, jpg[i:-2]


@pytest.mark.parametrize("case", range(10))
def test_oracle_matches_libjpeg_turbo(orc, case):
    px, q, ref = _golden()[case]
    assert orc.jpeg_encode(px, quality=q) == ref


def test_oracle_pixel_order_and_row_order(orc):
    px, q, ref = _golden()[2]
    bgra = px[..., [2, 1, 0, 3]]
    assert orc.jpeg_encode(bgra, quality=q, bgra=True) == ref
    assert orc.jpeg_encode(px[::-1], quality=q, bottom_up=False) == ref


def test_marker_layout(orc):
    rng = np.random.default_rng(3)
    px = rng.integers(0, 256, (19, 30, 4), dtype=np.uint8)
    segs, scan = _segments(orc.jpeg_encode(px, quality=100))
    assert [m for m, _ in segs] == [0xE0, 0xDB, 0xDB, 0xC0, 0xC4, 0xC4, 0xC4, 0xC4, 0xDA]
    assert segs[0][1] == b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00"
    assert segs[1][1] == b"\x00" + b"\x01" * 64 and segs[2][1] == b"\x01" + b"\x01" * 64
    assert segs[3][1] == bytes([8, 0, 19, 0, 30, 3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1])
    assert [p[0] for _, p in segs[4:8]] == [0x00, 0x10, 0x01, 0x11]
    # stuffing: every 0xFF in the scan is followed by 0x00
    idx = [i for i in range(len(scan)) if scan[i] == 0xFF]
    assert idx and all(scan[i + 1] == 0 for i in idx)


def test_flat_grey_closed_form(orc):
    # Y = Cb = Cr = 128: every coefficient is 0, so each MCU is luma DC
    # category 0 ('00') + EOB ('1010'), then twice chroma DC 0 ('00') + EOB
    # ('00'): 14 bits; the first MCU's DC has no predecessor either.
    px = np.full((8, 8, 4), 128, np.uint8)
    _, scan = _segments(orc.jpeg_encode(px, quality=100))
    assert scan == bytes([0b00101000, 0b00000011])  # 14 bits + two 1 bits of padding
    px = np.full((8, 32, 4), 128, np.uint8)          # four MCUs: 56 bits, 7 bytes
    _, scan = _segments(orc.jpeg_encode(px, quality=100))
    bits = "00101000000000" * 4
    assert scan == bytes(int(bits[i:i + 8], 2) for i in range(0, 56, 8))


def test_edge_replication(orc):
    # a 13x21 frame encodes as its 16x24 extension by the last row and column
    rng = np.random.default_rng(5)
    px = rng.integers(0, 256, (13, 21, 4), dtype=np.uint8)
    top = px[::-1]
    ext = np.concatenate([top, np.repeat(top[-1:], 3, 0)], 0)
    ext = np.concatenate([ext, np.repeat(ext[:, -1:], 3, 1)], 1)
    a = _segments(orc.jpeg_encode(px, quality=90))
    b = _segments(orc.jpeg_encode(ext, quality=90, bottom_up=False))
    assert a[1] == b[1]


def test_decodes_close_to_input(orc):
    pil = pytest.importorskip("PIL.Image")
    h, w = 40, 56
    y, x = np.mgrid[0:h, 0:w]
    px = np.stack([x * 4, y * 6, (x + y) * 2, np.full_like(x, 255)], -1).astype(np.uint8)
    img = np.asarray(pil.open(io.BytesIO(orc.jpeg_encode(px, quality=100))).convert("RGB"))
    err = np.abs(img.astype(int) - px[::-1, :, :3].astype(int))
    assert err.max() <= 3, err.max()
