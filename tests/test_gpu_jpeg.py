"""GPU frame grab (lvr_jpeg_encode_host, lvr_jpeg_encode, lvr_frame_grab_jpeg)
against the JPEG
restatement and libjpeg-turbo's own files for the same framebuffers: bit
exact, through the C ABI."""
import io
import os

import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import Scene

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "jpeg_golden.npz")


@pytest.fixture(scope="module")
def enc(lvr):
    e = livre_amd.JpegEncoder(0)
    yield e
    e.close()


def _gpu(enc, px, **kw):
    return enc.encode(px, **kw)


def test_golden_files(enc):
    g = np.load(GOLDEN)
    n = len([k for k in g.files if k.startswith("px_")])
    for i in range(n):
        px, q = g["px_%d" % i], int(g["q_%d" % i])
        print("golden case", i, px.shape, "quality", q, flush=True)
        assert _gpu(enc, px, quality=q) == g["jpg_%d" % i].tobytes(), i


@pytest.mark.parametrize("h,w,q,seed", [(1, 1, 100, 0), (7, 9, 100, 1), (8, 8, 95, 2),
                                        (9, 8, 100, 3), (33, 65, 100, 4), (100, 3, 60, 5),
                                        (129, 257, 100, 6), (240, 320, 30, 7)])
def test_noise_against_oracle(enc, h, w, q, seed):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    print("noise", h, w, q, flush=True)
    assert _gpu(enc, px, quality=q) == oracle.jpeg_encode(px, quality=q)


def test_formats_and_row_order(enc):
    rng = np.random.default_rng(11)
    px = rng.integers(0, 256, (45, 70, 4), dtype=np.uint8)
    ref = oracle.jpeg_encode(px, quality=100)
    assert _gpu(enc, px[..., [2, 1, 0, 3]], pixel_format=livre_amd.PIXEL_BGRA) == ref
    assert _gpu(enc, px[::-1], bottom_up=False) == ref


def test_pitch(enc):
    rng = np.random.default_rng(12)
    wide = rng.integers(0, 256, (50, 40, 4), dtype=np.uint8)
    out = enc.encode(wide, width=33)
    assert out == oracle.jpeg_encode(np.ascontiguousarray(wide[:, :33]), quality=100)


def test_large_then_small_reuses_scratch(enc):
    # the second, smaller encode runs over words and counters the first left
    rng = np.random.default_rng(13)
    big = rng.integers(0, 256, (1080, 1920, 4), dtype=np.uint8)
    print("1920x1080 noise", flush=True)
    assert _gpu(enc, big) == oracle.jpeg_encode(big)
    small = rng.integers(0, 256, (17, 23, 4), dtype=np.uint8)
    assert _gpu(enc, small, quality=80) == oracle.jpeg_encode(small, quality=80)
    assert _gpu(enc, small, quality=80) == oracle.jpeg_encode(small, quality=80)


def test_4k_smooth_frame_decodes(enc):
    pil = pytest.importorskip("PIL.Image")
    h, w = 2160, 3840
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    px = np.stack([128 + 127 * np.sin(x / 97.0), 128 + 127 * np.cos(y / 53.0),
                   (x + y) / (w + h) * 255, np.full_like(x, 255)], -1)
    px = np.clip(np.rint(px), 0, 255).astype(np.uint8)
    print("3840x2160 smooth", flush=True)
    jpg = _gpu(enc, px)
    assert jpg == oracle.jpeg_encode(px)
    img = np.asarray(pil.open(io.BytesIO(jpg)).convert("RGB"))
    # quality 100 still rounds YCbCr and DCT coefficients: a few levels at most
    err = np.abs(img.astype(np.int32) - px[::-1, :, :3].astype(np.int32))
    assert int(err.max()) <= 6 and float(err.mean()) < 0.6, (int(err.max()), float(err.mean()))
    st = enc.stats()
    assert (st.width, st.height, st.bytes) == (w, h, len(jpg)) and st.device_ms > 0


def test_invalid_arguments_fail_on_host(enc):
    px = np.zeros((8, 8, 4), np.uint8)
    for kw in [dict(quality=0), dict(quality=101), dict(pixel_format=2), dict(width=9)]:
        with pytest.raises(livre_amd.LivreError):
            enc.encode(px, **kw)
    with pytest.raises(livre_amd.LivreError):
        enc.encode(px, width=0)
    with pytest.raises(livre_amd.LivreError):
        enc.encode(np.zeros((1, 16385, 4), np.uint8))


def test_frame_grab_of_rendered_frame(lvr, orc):
    s = Scene(livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs")
    ctx = s.context()
    try:
        with pytest.raises(livre_amd.LivreError):
            ctx.grab_jpeg()  # nothing rendered yet
        tf = scenes.piecewise_tf(256, seed=2)
        ctx.set_tf(tf, s.range)
        for k, (w, h) in enumerate([(96, 80), (203, 157)]):
            mv, proj, fr = s.camera(scenes.orbit(k, 4, radius=2.0), w, h)
            vis = ctx.select_visibles(fr, h, 1.0, 0, s.vi.depth - 1)
            ctx.render(fr, w, h, livre_amd.order(s.vi, fr, vis), livre_amd.render_params(256))
            img, _ = ctx.read_frame()
            assert img[..., 3].max() > 0
            jpg = ctx.grab_jpeg()
            assert jpg == oracle.jpeg_encode(img, quality=100)
            assert ctx.grab_stats().bytes == len(jpg)
            assert ctx.grab_jpeg(quality=75) == oracle.jpeg_encode(img, quality=75)
            # the device entry point on the same framebuffer
            ptr, fw, fh = ctx.frame_device_ptr()
            with livre_amd.JpegEncoder(0) as e:
                assert e.encode_device(ptr, fw, fh) == jpg
                assert e.encode_device(ptr, fw, fh, bottom_up=False) == \
                    oracle.jpeg_encode(img, bottom_up=False)
    finally:
        ctx.close()
