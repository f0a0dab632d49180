"""Frame-start rules and closed forms through the HIP path.

- The transfer-function lookup of the kernel against the closed form of a
  1D GL_LINEAR, CLAMP_TO_EDGE texture (RayCastRenderer.cpp:175-193,
  fragRayCast.glsl:199): on a constant volume every pixel's colour / alpha
  is the filtered TF colour, lerp(e0, e1, clamp(2 d - 0.5, 0, 1)) for a
  2-entry TF. Read from the unconverted RGBA32F accumulation of a partial
  render, so FP32 precision is checked, not 8-bit rounding.
- The samples-per-ray default (RayCastRenderer.cpp:216-234) and the data
  range default (RayCastRenderer.cpp:236-285): a frame rendered with the
  default equals the frame rendered with the value the rule gives.
- The classified table of an 8-bit volume stays valid when the procedural
  source first allocates its scratch after a classified frame.
"""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import Scene, diff_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def constant_u8(lvr):
    vi = livre_amd.volume_info(16, 16, 1, livre_amd.DT_UINT8)
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * 2)
    yield vi, ctx
    ctx.close()


@pytest.mark.parametrize("classify", [1, 2])
@pytest.mark.parametrize("d", [-0.4, 0.1, 0.25, 0.3, 0.5, 0.62, 0.75, 0.9, 1.35])
def test_tf_two_entries_closed_form_on_device(constant_u8, d, classify):
    vi, ctx = constant_u8
    value = 77
    ctx.upload([0], np.full((1, 18, 18, 18), value, np.uint8))
    e0 = np.array([0.9, 0.1, 0.3], np.float64)
    e1 = np.array([0.2, 0.7, 0.95], np.float64)
    tf = np.zeros((2, 4), np.float32)
    tf[0, :3], tf[1, :3] = e0, e1
    tf[:, 3] = 0.04
    r0 = np.float32(value - 100.0 * d)
    ctx.set_tf(tf, (float(r0), float(r0) + 100.0))
    mv, proj = scenes.camera((1.3, 0.6, 1.0), 24, 24, 45.0)
    fr = livre_amd.frustum(mv, proj)
    ctx.render_partial(fr, 24, 24, [0], livre_amd.render_params(32, early_exit=2.0,
                                                                classify=classify), 0)
    lay = ctx.read_layer(0).astype(np.float64)
    a = lay[..., 3]
    hit = a > 1e-3
    assert hit.sum() > 50
    got = lay[..., :3][hit] / a[hit, None]
    r1 = np.float32(r0 + np.float32(100.0))
    dd = value * (1.0 / (float(r1) - float(r0))) - float(r0) / (float(r1) - float(r0))
    w = min(max(2.0 * dd - 0.5, 0.0), 1.0)
    want = e0 * (1.0 - w) + e1 * w
    err = np.abs(got - want).max()
    print("d", d, "classify", classify, "max |colour/alpha - closed form|", err, flush=True)
    assert err < 5e-5


@pytest.fixture(scope="module")
def blobs_u8(lvr, orc):
    s = Scene(livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs")
    ctx = s.context()
    yield s, ctx
    ctx.close()


def test_default_samples_per_ray_frame(blobs_u8):
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=2)
    ctx.set_tf(tf, s.range)
    mv, proj, fr = s.camera((1.3, 0.5, 1.2), 64, 48)
    ordered = s.visible_ordered(mv, proj, 48)
    ctx.render(fr, 64, 48, ordered, livre_amd.render_params(0))
    a, sa = ctx.read_frame()
    n = livre_amd.samples_per_ray(s.vi, ordered)
    assert n == 512
    ctx.render(fr, 64, 48, ordered, livre_amd.render_params(n))
    b, sb = ctx.read_frame()
    assert sa == sb and np.array_equal(a, b)


@pytest.mark.parametrize("bad", [(0.0, 0.0), (10.0, 5.0), (-5.0, 0.0)])
def test_default_data_range_frame(blobs_u8, bad):
    """A TF range that is not valid (range1 <= 0 or range1 <= range0) maps
    the type's full range, as the reference falls back to numeric_limits."""
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=4)
    mv, proj, fr = s.camera((-1.2, 0.7, 1.1), 64, 48)
    ordered = s.visible_ordered(mv, proj, 48)
    params = livre_amd.render_params(64)
    ctx.set_tf(tf, (0.0, 255.0))
    ctx.render(fr, 64, 48, ordered, params)
    a, sa = ctx.read_frame()
    ctx.set_tf(tf, bad)
    ctx.render(fr, 64, 48, ordered, params)
    b, sb = ctx.read_frame()
    assert sa == sb and np.array_equal(a, b)


def test_classified_table_survives_source_scratch(lvr, orc):
    """A classified 8-bit frame, then the procedural source's first scratch
    allocation on the same context, then the same frame again: identical,
    and within tolerance of the oracle."""
    vi = livre_amd.volume_info(64, 16, 1, livre_amd.DT_UINT8)
    ovi = oracle.volume_info([64] * 3, [16] * 3, [1] * 3, oracle.DT_UINT8)
    ids = scenes.all_nodes(vi)
    fp = livre_amd.field_params(livre_amd.FIELD_BLOBS, 12, 5, 0.1)
    bricks = np.stack([livre_amd.field_brick_host(vi, fp, i) for i in ids])
    with livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids)) as ctx:
        ctx.upload(ids, bricks)
        tf = scenes.piecewise_tf(256, seed=7, max_alpha=0.5)
        ctx.set_tf(tf, (0.0, 255.0))
        mv, proj = scenes.camera((1.4, 0.6, -1.1), 72, 56, 60.0)
        fr = livre_amd.frustum(mv, proj)
        vis = oracle.select_visibles(ovi, mv, proj, 56, 1.0, 0, ovi.depth - 1)
        ordered = oracle.order(ovi, mv, vis)
        params = livre_amd.render_params(64, classify=1)
        ctx.render(fr, 72, 56, ordered, params)
        first, s1 = ctx.read_frame()
        ctx.generate_field(fp, ids)  # first use of the source scratch
        ctx.render(fr, 72, 56, ordered, params)
        second, s2 = ctx.read_frame()
    assert s1 == s2 and np.array_equal(first, second)
    by_id = {int(i): b for i, b in zip(ids, bricks)}
    cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                        fr.far_plane, 72, 56)
    ref, ref_s = oracle.render(ovi, cam, ordered, [by_id[int(i)] for i in ordered], tf,
                               oracle.render_params(64, 0.999, (0.0, 255.0)))
    mx, mean = diff_stats(second, oracle.to_rgba8(ref))
    assert mx <= 2 and mean <= 0.5 and s2 == ref_s


def test_composite_rejects_mismatched_layers(blobs_u8):
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=6)
    ctx.set_tf(tf, s.range)
    mv, proj, fr = s.camera((1.1, 0.4, 1.3), 40, 30)
    ordered = s.visible_ordered(mv, proj, 30)
    params = livre_amd.render_params(64)
    ctx.render_partial(fr, 40, 30, ordered, params, 2)
    p2 = ctx.layer_ptr(2)
    ctx.composite([p2], 40, 30)
    with pytest.raises(livre_amd.LivreError):
        ctx.composite([p2], 30, 40)  # same pixel count, other shape
    with pytest.raises(livre_amd.LivreError):
        ctx.composite([p2], 64, 48)  # more pixels than the layer holds
    # a larger partial render replaces the buffer: the old pointer is refused
    mv2, proj2, fr2 = s.camera((1.1, 0.4, 1.3), 400, 300)
    ctx.render_partial(fr2, 400, 300, ordered, params, 2)
    if ctx.layer_ptr(2) != p2:
        with pytest.raises(livre_amd.LivreError):
            ctx.composite([p2], 40, 30)
    ctx.sync()
    ctx.composite([ctx.layer_ptr(2)], 400, 300)
