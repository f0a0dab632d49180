"""The 1080p single-volume configuration at its full size against the
oracle: 2048^3 uint16 value noise (seed 3, generated on the device),
64^3 bricks + halo, 6 levels, FP32 4096-entry TF, 2048 samples per unit,
SSE 1. For frames 0, 30, 60 and 90 of the 120-frame orbit-and-zoom path
(radius 2.0, 1.5, 1.0 and 1.5: frame 60 is the closest and heaviest): the
visible set and its order are bit-exact; a band of 24 rows through the
middle of the 1920x1080 frame is within 2/255 max and 0.5/255 mean of the
FP32 oracle; against the FP64 transcription of the shader its mean error is
within 0.5/255 and at most 1e-4 of the channels are beyond 2/255 (printed
per frame beside the FP32 restatement's own error, lines starting FP64ERR).
For frame 0 the 65536-bin histogram of the rendered set is bit-exact and
sums to the scaled interior voxel count."""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import diff_stats, fp64_error

pytestmark = pytest.mark.gpu

W, H = 1920, 1080


@pytest.fixture(scope="module")
def c2(lvr, orc):
    vi = livre_amd.volume_info(2048, 64, 1, livre_amd.DT_UINT16)
    ovi = oracle.volume_info([2048] * 3, [64] * 3, [1] * 3, oracle.DT_UINT16)
    ids = scenes.all_nodes(vi)
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
    ctx.generate_noise(livre_amd.noise_params(3, octaves=5, base_cells=8, gain=1.5), ids)
    yield vi, ovi, ctx
    ctx.close()


@pytest.mark.parametrize("t", [0, 30, 60, 90])
def test_c2_frame_band(c2, t):
    vi, ovi, ctx = c2
    tf = scenes.piecewise_tf(4096, seed=4, points=8, max_alpha=0.35)
    ctx.set_tf(tf, (0.0, 65535.0))
    eye = scenes.orbit_zoom(t, 120)
    mv, proj = scenes.camera(eye, W, H, 60.0)
    fr = livre_amd.frustum(mv, proj)
    vis = ctx.select_visibles(fr, H, 1.0, 0, vi.depth - 1)
    assert np.array_equal(vis, oracle.select_visibles(ovi, mv, proj, H, 1.0, 0, ovi.depth - 1))
    ordered = livre_amd.order(vi, fr, vis)
    assert np.array_equal(ordered, oracle.order(ovi, mv, vis))
    ctx.render(fr, W, H, ordered, livre_amd.render_params(2048))
    img, _samples = ctx.read_frame()
    bricks = [ctx.download(i) for i in ordered]
    cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                        fr.far_plane, W, H)
    y0, y1 = H // 2 - 12, H // 2 + 12
    band = oracle.render_params(2048, 0.999, (0.0, 65535.0), window=(0, y0, W, y1))
    ref, ref_samples = oracle.render(ovi, cam, ordered, bricks, tf, band, "f32")
    mx, mean = diff_stats(img[y0:y1], oracle.to_rgba8(ref)[y0:y1])
    print("frame", t, "bricks", len(ordered), "band samples", ref_samples, "max", mx,
          "mean", mean, flush=True)
    assert ref_samples > 0 and mx <= 2 and mean <= 0.5
    ref64, samples64 = oracle.render(ovi, cam, ordered, bricks, tf, band, "f64")
    fp64_error(img[y0:y1], ref[y0:y1], ref64[y0:y1], "c2 frame %d rows %d-%d" % (t, y0, y1),
               (-1, ref_samples, samples64))
    if t == 0:
        got = ctx.histogram(ordered, 65536)
        want = oracle.histogram(ovi, ordered, bricks, 65536, (0.0, 65535.0))
        assert np.array_equal(got, want)
        scale = [8 ** (vi.depth - 1 - (int(i) & 0xF)) for i in ordered]
        assert int(got.sum()) == 64 ** 3 * sum(scale)


@pytest.mark.parametrize("t", [30, 60])
def test_c2_trilinear_band(c2, t):
    """The same volume and frames sampled with GL_LINEAR (trilinear, halo
    voxels at the brick faces): a 16-row band within 2/255 max and 0.5/255
    mean of the FP32 oracle's trilinear march, and its error against the FP64
    transcription within the FP64 bound (lines starting FP64ERR)."""
    vi, ovi, ctx = c2
    tf = scenes.piecewise_tf(4096, seed=4, points=8, max_alpha=0.35)
    ctx.set_tf(tf, (0.0, 65535.0))
    mv, proj = scenes.camera(scenes.orbit_zoom(t, 120), W, H, 60.0)
    fr = livre_amd.frustum(mv, proj)
    ordered = livre_amd.order(vi, fr, ctx.select_visibles(fr, H, 1.0, 0, vi.depth - 1))
    ctx.render(fr, W, H, ordered, livre_amd.render_params(2048, linear=True))
    img, _samples = ctx.read_frame()
    bricks = [ctx.download(i) for i in ordered]
    cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                        fr.far_plane, W, H)
    y0, y1 = H // 2 - 8, H // 2 + 8
    band = oracle.render_params(2048, 0.999, (0.0, 65535.0), linear=True, window=(0, y0, W, y1))
    ref, ref_samples = oracle.render(ovi, cam, ordered, bricks, tf, band, "f32")
    mx, mean = diff_stats(img[y0:y1], oracle.to_rgba8(ref)[y0:y1])
    print("trilinear frame", t, "bricks", len(ordered), "band samples", ref_samples, "max", mx,
          "mean", mean, flush=True)
    assert ref_samples > 0 and mx <= 2 and mean <= 0.5
    ref64, samples64 = oracle.render(ovi, cam, ordered, bricks, tf, band, "f64")
    fp64_error(img[y0:y1], ref[y0:y1], ref64[y0:y1],
               "c2 trilinear frame %d rows %d-%d" % (t, y0, y1), (-1, ref_samples, samples64))
