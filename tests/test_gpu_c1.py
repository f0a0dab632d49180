"""The CPU-runnable configuration end to end against the oracle.

Volume: 256^3 uint8, 64 anisotropic Gaussian blobs + 10% uniform noise,
seed 1, generated on the device by the procedural source and read back for
the oracle (the device source equals the host source bit for bit). Octree:
32^3 bricks + 1-voxel halo, 4 levels. TF: 256 entries, 6 control points,
seed 2, uploaded as the reference's uint8 LUT (TransferFunction1D::getLUT).
Render: 1024x1024, 60 degree FOV, 8 cameras on a radius-2 orbit, 256 samples
per unit (1 per voxel), early exit 0.99, SSE 1 px. Checked: visible NodeId
sets and their order (bit-exact), the 256-bin histogram of each rendered set
(bit-exact), the RGBA8 frames (within 2/255 max and 0.5/255 mean of the FP32
oracle; identical so far) and the composite-step counts. Each frame is also
compared with the FP64 transcription of the shader (mean within 0.5/255,
at most 1e-4 of the channels beyond 2/255), and the error is printed per
camera beside the FP32 restatement's own (lines starting FP64ERR)."""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import diff_stats, fp64_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(lvr, orc):
    vi = livre_amd.volume_info(256, 32, 1, livre_amd.DT_UINT8)
    ovi = oracle.volume_info([256] * 3, [32] * 3, [1] * 3, oracle.DT_UINT8)
    ids = scenes.all_nodes(vi)
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
    ctx.generate_field(livre_amd.field_params(livre_amd.FIELD_BLOBS, 64, 1, 0.1), ids)
    bricks = {int(i): ctx.download(i) for i in ids}
    yield vi, ovi, ctx, bricks
    ctx.close()


@pytest.mark.parametrize("k", range(8))
def test_c1_orbit_frame(c1, k):
    vi, ovi, ctx, bricks = c1
    w = h = 1024
    tf = scenes.piecewise_tf(256, seed=2, points=6, max_alpha=0.6)
    lut = scenes.lut_rgba8(tf)
    ctx.set_tf_rgba8(lut, (0.0, 255.0))
    eye = scenes.orbit(k, 8, radius=2.0)
    mv, proj = scenes.camera(eye, w, h, 60.0)
    fr = livre_amd.frustum(mv, proj)
    # visible set and order
    vis = ctx.select_visibles(fr, h, 1.0, 0, vi.depth - 1)
    ovis = oracle.select_visibles(ovi, mv, proj, h, 1.0, 0, ovi.depth - 1)
    assert np.array_equal(vis, ovis)
    ordered = livre_amd.order(vi, fr, vis)
    assert np.array_equal(ordered, oracle.order(ovi, mv, ovis))
    # frame
    ctx.render(fr, w, h, ordered, livre_amd.render_params(256, early_exit=0.99))
    img, samples = ctx.read_frame()
    cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                        fr.far_plane, w, h)
    ref, ref_samples = oracle.render(ovi, cam, ordered, [bricks[int(i)] for i in ordered],
                                     lut.astype(np.float32) / np.float32(255.0),
                                     oracle.render_params(256, 0.99, (0.0, 255.0)), "f32")
    mx, mean = diff_stats(img, oracle.to_rgba8(ref))
    print("camera", k, "bricks", len(ordered), "max", mx, "mean", mean, "samples", samples,
          ref_samples, flush=True)
    assert mx <= 2 and mean <= 0.5
    assert abs(int(samples) - int(ref_samples)) <= 1e-4 * ref_samples
    # the FP64 transcription of the shader on the same bricks
    ref64, samples64 = oracle.render(ovi, cam, ordered, [bricks[int(i)] for i in ordered],
                                     lut.astype(np.float32) / np.float32(255.0),
                                     oracle.render_params(256, 0.99, (0.0, 255.0)), "f64")
    fp64_error(img, ref, ref64, "c1 camera %d" % k, (samples, ref_samples, samples64))
    # histogram of the rendered set
    got = ctx.histogram(ordered, 256)
    want = oracle.histogram(ovi, ordered, [bricks[int(i)] for i in ordered], 256, (0.0, 255.0))
    assert np.array_equal(got, want)
