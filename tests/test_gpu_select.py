"""Device selection with clip planes and sort-last ranges, and the clipped
march (fragRayCast.glsl:152-167), through the C ABI.

- lvr_select_visibles_ex equals the oracle's SelectVisibles bit for bit
  (set and DFS order) with clip planes and with sort-last ranges
  (SelectVisibles.cpp:30-131).
- Culling by a clip plane removes only bricks whose every sample the march
  clips away: the frame rendered from the clipped selection equals the frame
  rendered from the unclipped one, both with the planes in the render
  parameters, pixel for pixel and in composite steps.
- The clipped frame is within 2/255 max and 0.5/255 mean of the oracle's
  FP32 restatement of the clipped march.
"""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import Scene, diff_stats

pytestmark = pytest.mark.gpu

PLANES = [
    [(1.0, 0.0, 0.0, 0.0137)],
    [(0.6, 0.0, 0.8, -0.2071), (0.0, 1.0, 0.0, 0.3129)],
    [(-0.3, 0.5, -0.81, 0.0517)],
]
EYES = [(1.6, 0.4, 1.2), (-0.5, 1.4, -1.3), (0.9, -1.1, 1.5)]


@pytest.fixture(scope="module")
def big(lvr, orc):
    vi = livre_amd.volume_info(256, 16, 1, livre_amd.DT_UINT8)
    ovi = oracle.volume_info([256] * 3, [16] * 3, [1] * 3, oracle.DT_UINT8)
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * 4)
    yield vi, ovi, ctx
    ctx.close()


@pytest.mark.parametrize("planes", [[]] + PLANES)
@pytest.mark.parametrize("eye", EYES)
def test_selection_with_clip_planes_matches_oracle(big, eye, planes):
    vi, ovi, ctx = big
    mv, proj = scenes.camera(eye, 320, 240, 60.0)
    fr = livre_amd.frustum(mv, proj)
    got = ctx.select_visibles_ex(fr, livre_amd.select_params(240, 1.0, 0, vi.depth - 1,
                                                              clip_planes=planes))
    want = oracle.select_visibles(ovi, mv, proj, 240, 1.0, 0, ovi.depth - 1,
                                  clip_planes=planes)
    assert np.array_equal(got, want)
    assert len(got) > 0


@pytest.mark.parametrize("k", [2, 3, 8])
def test_selection_ranges_match_oracle(big, k):
    vi, ovi, ctx = big
    mv, proj = scenes.camera((1.3, 0.8, 1.5), 320, 240, 60.0)
    fr = livre_amd.frustum(mv, proj)
    whole = ctx.select_visibles_ex(fr, livre_amd.select_params(240, 1.0, 0, vi.depth - 1))
    parts = []
    for i in range(k):
        rng = (float(np.float32(i / k)), float(np.float32((i + 1) / k)))
        got = ctx.select_visibles_ex(fr, livre_amd.select_params(240, 1.0, 0, vi.depth - 1,
                                                                  rng=rng))
        want = oracle.select_visibles(ovi, mv, proj, 240, 1.0, 0, ovi.depth - 1, rng=rng)
        assert np.array_equal(got, want), (k, i)
        parts.append(got)
    assert np.array_equal(np.concatenate(parts), whole)


@pytest.fixture(scope="module")
def blobs(lvr, orc):
    s = Scene(livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs", seed=9)
    ctx = s.context()
    yield s, ctx
    ctx.close()


@pytest.mark.parametrize("planes", PLANES)
def test_clipped_frames(blobs, planes):
    s, ctx = blobs
    w, h = 120, 96
    tf = scenes.piecewise_tf(256, seed=31, max_alpha=0.5)
    ctx.set_tf(tf, s.range)
    mv, proj, fr = s.camera((1.4, 0.6, 1.3), w, h)
    vi = s.vi
    params = livre_amd.render_params(64, clip_planes=planes)
    culled = ctx.select_visibles_ex(fr, livre_amd.select_params(h, 1.0, 0, vi.depth - 1,
                                                                 clip_planes=planes))
    full = ctx.select_visibles_ex(fr, livre_amd.select_params(h, 1.0, 0, vi.depth - 1))
    assert len(culled) < len(full)
    ctx.render(fr, w, h, livre_amd.order(vi, fr, culled), params)
    a, sa = ctx.read_frame()
    ctx.render(fr, w, h, livre_amd.order(vi, fr, full), params)
    b, sb = ctx.read_frame()
    assert sa == sb and np.array_equal(a, b)
    # against the oracle's clipped march over the culled set
    ordered = livre_amd.order(vi, fr, culled)
    cam = s.oracle_camera(fr, w, h)
    ref, ref_s = oracle.render(s.ovi, cam, ordered, [s.bricks[int(i)] for i in ordered], tf,
                               oracle.render_params(64, 0.999, s.range, clip_planes=planes))
    mx, mean = diff_stats(a, oracle.to_rgba8(ref))
    print("planes", planes, "bricks", len(culled), "of", len(full), "max", mx, "mean", mean,
          "samples", sa, ref_s, flush=True)
    assert mx <= 2 and mean <= 0.5
    assert abs(int(sa) - int(ref_s)) <= 1e-4 * ref_s
    # the planes remove part of the image
    ctx.render(fr, w, h, ordered, livre_amd.render_params(64))
    unclipped, su = ctx.read_frame()
    assert su > sa and not np.array_equal(unclipped, a)
