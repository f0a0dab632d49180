"""Writes tests/golden/jpeg_golden.npz: framebuffers and the JPEG files that
libjpeg-turbo writes for them, the library the reference's frame grabber calls
(livre/eq/FrameGrabber.cpp:49-84). libjpeg-turbo is reached through Pillow,
which links it; Pillow's encoder at subsampling 4:4:4 takes the same path as
tjCompress2 at quality >= 96 (accurate integer DCT, standard Huffman tables,
JFIF 1.01 header), and at lower qualities it still uses the accurate DCT.

Each case: `px_<i>` an (H, W, 4) uint8 framebuffer with row 0 at the bottom,
`q_<i>` the quality, `jpg_<i>` the file bytes. Run from the repository root:
    python tests/golden/make_jpeg_golden.py
"""
import io
import os

import numpy as np
from PIL import Image, features

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jpeg_golden.npz")


def cases():
    rng = np.random.default_rng(20240)
    for (h, w), q, kind in [((1, 1), 100, "noise"), ((8, 8), 100, "noise"),
                            ((13, 21), 100, "noise"), ((13, 21), 75, "noise"),
                            ((37, 91), 100, "smooth"), ((37, 91), 50, "smooth"),
                            ((64, 48), 90, "noise"), ((64, 48), 10, "smooth"),
                            ((9, 17), 1, "noise"), ((72, 96), 100, "frame")]:
        y, x = np.mgrid[0:h, 0:w].astype(np.float64)
        if kind == "noise":
            rgb = rng.integers(0, 256, (h, w, 3))
        elif kind == "smooth":
            rgb = np.stack([255 * x / max(w - 1, 1), 255 * y / max(h - 1, 1),
                            127.5 + 127.5 * np.sin((x + 2 * y) / 5.0)], -1)
        else:  # a rendered-looking frame: a soft disc on black, saturated core
            r = np.hypot(x - w / 2, y - h / 2) / (0.4 * min(w, h))
            a = np.clip(1.2 - r, 0, 1)
            rgb = np.stack([300 * a, 200 * a ** 2, 120 * a ** 3], -1)
        rgb = np.clip(np.rint(rgb), 0, 255).astype(np.uint8)
        alpha = rng.integers(0, 256, (h, w, 1)).astype(np.uint8)
        yield np.concatenate([rgb, alpha], -1)[::-1].copy(), q, rgb


def main():
    assert features.check_feature("libjpeg_turbo"), "Pillow is not linked to libjpeg-turbo"
    arrays = {"library": np.frombuffer(("libjpeg-turbo via Pillow %s"
                                        % Image.__version__).encode(), np.uint8)}
    for i, (px, q, rgb) in enumerate(cases()):
        buf = io.BytesIO()
        Image.fromarray(rgb).save(buf, format="JPEG", quality=q, subsampling=0)
        arrays["px_%d" % i] = px
        arrays["q_%d" % i] = np.array(q, np.int32)
        arrays["jpg_%d" % i] = np.frombuffer(buf.getvalue(), np.uint8)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, len([k for k in arrays if k.startswith("px_")]), "cases")


if __name__ == "__main__":
    main()
