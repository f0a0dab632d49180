"""Sort-last compositing on one GPU: the frame's ordered bricks split into
consecutive ranges (SelectVisibles.cpp:112-131), each ray-cast into an RGBA32F
layer (lvr_render_partial), the layers composited front to back
(lvr_composite_layers, Channel.cpp:464-603). Checked against the frame
rendered whole, against the oracle's render of each range, and against the
"over" of the oracle's layers. The static decomposition (lvr_sortlast_parts:
the halves of the block grid, a brick's part fixed whatever the camera) is
checked the same way for 2, 4 and 8 parts from cameras on every side, its
layers composited in lvr_sortlast_order. Tolerance: RGBA8 within 2/255 max
and 0.5/255 mean; one layer reproduces the whole frame exactly."""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import Scene, diff_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def noise_u16(lvr, orc):
    s = Scene(livre_amd.DT_UINT16, n=64, block=16, overlap=1, source="noise", seed=3)
    ctx = s.context()
    yield s, ctx
    ctx.close()


def _over(layers):
    """Front-to-back over of premultiplied RGBA32F images (front first)."""
    acc = np.zeros_like(layers[0], dtype=np.float64)
    for lay in layers:
        acc[..., :] += (1.0 - acc[..., 3:4]) * lay.astype(np.float64)
    return acc.astype(np.float32)


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_layers_composite_to_the_whole_frame(noise_u16, k):
    s, ctx = noise_u16
    w, h = 120, 90
    tf = scenes.piecewise_tf(256, seed=5, max_alpha=0.4)
    ctx.set_tf(tf, s.range)
    mv, proj, fr = s.camera((1.4, 0.7, 1.2), w, h)
    ordered = s.visible_ordered(mv, proj, h)
    params = livre_amd.render_params(64)
    ctx.render(fr, w, h, ordered, params)
    whole, _ = ctx.read_frame()
    parts = np.array_split(np.asarray(ordered), k)
    for i, part in enumerate(parts):
        ctx.render_partial(fr, w, h, part, params, i)
    ctx.composite([ctx.layer_ptr(i) for i in range(k)], w, h)
    img, _ = ctx.read_frame()
    if k == 1:
        assert np.array_equal(img, whole)
    mx, mean = diff_stats(img, whole)
    assert mx <= 2 and mean <= 0.5, (k, mx, mean)
    # the oracle's layers, composited with the over operator
    ref_layers = [s.oracle_render(fr, w, h, part, tf, 64)[0] for part in parts]
    ref = oracle.to_rgba8(_over(ref_layers).reshape(-1, 4)).reshape(h, w, 4)
    mx, mean = diff_stats(img, ref)
    assert mx <= 2 and mean <= 0.5, (k, mx, mean)
    # each layer alone (composited by itself) is the oracle's render of its range
    for i, part in enumerate(parts):
        ctx.composite([ctx.layer_ptr(i)], w, h)
        one, _ = ctx.read_frame()
        mx, mean = diff_stats(one, oracle.to_rgba8(ref_layers[i].reshape(-1, 4)).reshape(h, w, 4))
        assert mx <= 2 and mean <= 0.5, (k, i, mx, mean)


@pytest.mark.parametrize("n_parts", [2, 4, 8])
def test_static_parts_composite_to_the_whole_frame(noise_u16, n_parts):
    s, ctx = noise_u16
    w, h = 120, 90
    tf = scenes.piecewise_tf(256, seed=9, max_alpha=0.4)
    ctx.set_tf(tf, s.range)
    params = livre_amd.render_params(64)
    owner = {}
    for eye in ((1.4, 0.7, 1.2), (-1.1, -0.6, 1.5), (0.2, 1.6, -0.9), (-0.9, 0.3, -1.7)):
        mv, proj, fr = s.camera(eye, w, h)
        ordered = np.asarray(s.visible_ordered(mv, proj, h))
        ctx.render(fr, w, h, ordered, params)
        whole, whole_smp = ctx.read_frame()
        parts = livre_amd.sortlast_parts(s.vi, ordered, n_parts)
        assert (parts >= 0).all()
        for nid, p in zip(ordered, parts):  # a brick's part never changes
            assert owner.setdefault(int(nid), int(p)) == p
        order = livre_amd.sortlast_order(s.vi, list(fr.eye), n_parts)
        smp = 0
        for i, p in enumerate(order):
            ctx.render_partial(fr, w, h, ordered[parts == p], params, i)
            smp += ctx.read_frame()[1]
        ctx.composite([ctx.layer_ptr(i) for i in range(n_parts)], w, h)
        img, _ = ctx.read_frame()
        mx, mean = diff_stats(img, whole)
        print("static parts", n_parts, "eye", eye, "sizes",
              [int((parts == p).sum()) for p in range(n_parts)], "max", mx, "mean", mean,
              "samples", smp, whole_smp, flush=True)
        assert mx <= 2 and mean <= 0.5, (n_parts, eye, mx, mean)
        assert smp >= whole_smp


def test_empty_range_and_errors(noise_u16):
    s, ctx = noise_u16
    w, h = 64, 48
    tf = scenes.piecewise_tf(256, seed=6, max_alpha=0.4)
    ctx.set_tf(tf, s.range)
    mv, proj, fr = s.camera((-1.2, 0.4, 1.5), w, h)
    ordered = s.visible_ordered(mv, proj, h)
    params = livre_amd.render_params(64)
    ctx.render(fr, w, h, ordered, params)
    whole, _ = ctx.read_frame()
    ctx.render_partial(fr, w, h, ordered, params, 0)
    ctx.render_partial(fr, w, h, ordered[:0], params, 1)  # no bricks: a clear layer
    ctx.composite([ctx.layer_ptr(1), ctx.layer_ptr(0), ctx.layer_ptr(1)], w, h)
    img, _ = ctx.read_frame()
    assert np.array_equal(img, whole)
    with pytest.raises(livre_amd.LivreError):
        ctx.render_partial(fr, w, h, ordered, params, 16)
    with pytest.raises(livre_amd.LivreError):
        ctx.composite([], w, h)
    with pytest.raises(livre_amd.LivreError):
        ctx.composite([ctx.layer_ptr(0)] * 17, w, h)
