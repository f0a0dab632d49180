"""Autotuner: request classes, the cache (shipped table + user file), the
batch-wide samples-per-thread vote (CPU, with a stand-in context), and on the
GPU that lvr_autotune times every candidate and that every candidate renders
the same frame."""
import json

import numpy as np
import pytest

import livre_amd
from livre_amd import autotune


class FakeCtx:
    """Stands in for Context.autotune: returns a fixed choice per dtype."""

    def __init__(self, vi, choice):
        self.vi = vi
        self.choice = choice
        self.calls = 0

    def autotune(self, fr, w, h, ordered, params, reps):
        self.calls += 1
        out = livre_amd.RenderParams.from_buffer_copy(params)
        out.tile_w, out.tile_h, out.samples_per_thread, out.waves_per_simd = self.choice
        return out, [(8, 8, 2, 2, 4, 1.0), tuple(self.choice[:3]) + (2, self.choice[3], 0.5)]


def test_request_class_buckets(lvr):
    vi = lvr.volume_info(512, 64, 1, lvr.DT_UINT8)
    p = lvr.render_params(512)
    assert autotune.request_class(vi, 1024, 1024, p) == "gfx950/dt1/px20/rate+0/nearest"
    assert autotune.request_class(vi, 256, 256, lvr.render_params(256)) == \
        "gfx950/dt1/px16/rate-1/nearest"
    assert autotune.request_class(vi, 1024, 1024, lvr.render_params(1024, linear=True)) \
        .endswith("rate+1/linear")


def test_cache_hits_persist_and_vote(lvr, tmp_path):
    path = str(tmp_path / "tune.json")
    t = autotune.Tuner(path=path)
    t.table = {}  # ignore the shipped table for this test
    u8 = FakeCtx(lvr.volume_info(256, 32, 1, lvr.DT_UINT8), (16, 4, 3, 6))
    f32 = FakeCtx(lvr.volume_info(256, 32, 1, lvr.DT_FLOAT), (4, 16, 1, 8))
    p = lvr.render_params(256)
    a = t.params_for(u8, None, 512, 512, [], p)
    assert (a.tile_w, a.tile_h, a.samples_per_thread, a.waves_per_simd) == (16, 4, 3, 6)
    t.params_for(u8, None, 500, 520, [], p)  # same class: cached
    assert u8.calls == 1 and t.tuned == 1
    with open(path) as f:
        saved = json.load(f)
    assert saved[autotune.request_class(u8.vi, 512, 512, p)]["tile"] == [16, 4]
    assert saved[autotune.request_class(u8.vi, 512, 512, p)]["waves_per_simd"] == 6
    # a new tuner reads the user cache back
    t2 = autotune.Tuner(path=path)
    b = t2.params_for(u8, None, 512, 512, [], p)
    assert u8.calls == 1 and (b.tile_w, b.tile_h) == (16, 4)
    # an entry without a waves-per-SIMD choice (older cache) is tuned again
    key = autotune.request_class(u8.vi, 512, 512, p)
    del t2.table[key]["waves_per_simd"]
    t2.params_for(u8, None, 512, 512, [], p)
    assert u8.calls == 2 and t2.table[key]["waves_per_simd"] == 6
    # batch: tiles per class, samples per thread by pixel-weighted vote per
    # voxel type, waves per SIMD by vote over the batch
    u8b = FakeCtx(u8.vi, (8, 8, 2, 8))
    t2.table[autotune.request_class(u8.vi, 2048, 2048, p)] = {
        "tile": [8, 8], "samples_per_thread": 2, "classify": 0, "waves_per_simd": 8}
    reqs = [(u8, None, 512, 512, [], p), (f32, None, 2048, 2048, [], p),
            (u8, None, 512, 512, [], p), (u8b, None, 2048, 2048, [], p)]
    tuned = t2.tune_batch(reqs)
    assert [(x.tile_w, x.tile_h) for x in tuned] == [(16, 4), (4, 16), (16, 4), (8, 8)]
    # u8: 2048^2 pixels voted 2 against 2 x 512^2 for 3; f32 alone keeps 1
    assert [x.samples_per_thread for x in tuned] == [2, 1, 2, 2]
    assert {x.waves_per_simd for x in tuned} == {8}
    assert p.samples_per_thread == 0 and p.waves_per_simd == 0  # inputs untouched


@pytest.mark.gpu
def test_autotune_candidates_render_the_same_frame(lvr):
    from scene_kit import Scene
    s = Scene(lvr.DT_UINT16, n=64, block=16, overlap=1, source="noise", seed=21)
    ctx = s.context()
    try:
        mv, proj, fr = s.camera((1.6, 0.9, 1.2), 96, 80)
        ordered = s.visible_ordered(mv, proj, 80)
        from livre_amd import scenes
        ctx.set_tf(scenes.piecewise_tf(256, seed=3, max_alpha=0.5), s.range)
        base = lvr.render_params(128)
        ctx.render(fr, 96, 80, ordered, base)
        ref, ref_smp = ctx.read_frame()
        tuned, cands = ctx.autotune(fr, 96, 80, ordered, base, reps=2)
        assert len(cands) == 120
        assert {(c[0], c[1]) for c in cands} == {(8, 8), (16, 4), (4, 16), (32, 2), (2, 32)}
        assert {c[2] for c in cands} == {1, 2, 3} and {c[3] for c in cands} == {1, 2}
        assert {c[4] for c in cands} == {4, 5, 6, 8}
        best = min(cands, key=lambda c: c[5])
        assert (tuned.tile_w, tuned.tile_h, tuned.samples_per_thread, tuned.classify,
                tuned.waves_per_simd) == best[:5]
        for tw, th, d, m, w, _ms in cands:
            ctx.render(fr, 96, 80, ordered,
                       lvr.render_params(128, tile=(tw, th), samples_per_thread=d, classify=m,
                                         waves_per_simd=w))
            img, smp = ctx.read_frame()
            assert smp == ref_smp and np.array_equal(img, ref), (tw, th, d, m, w)
    finally:
        ctx.close()


def test_cached_choice_is_applied_without_tuning(lvr, tmp_path):
    """Tuner.cached applies the table's choice of the frame's class and leaves
    the parameters of a class without an entry unchanged."""
    vi = livre_amd.volume_info(256, 32, 1, livre_amd.DT_UINT8)
    t = autotune.Tuner(path=str(tmp_path / "none.json"))
    key = autotune.request_class(vi, 1024, 1024, livre_amd.render_params(256))
    t.table = {key: {"tile": [16, 4], "samples_per_thread": 3, "classify": 1,
                     "waves_per_simd": 6}}
    p = t.cached(vi, 1024, 1024, livre_amd.render_params(256, early_exit=0.99))
    assert (p.tile_w, p.tile_h, p.samples_per_thread, p.classify, p.waves_per_simd) == \
        (16, 4, 3, 1, 6)
    assert abs(p.early_exit - 0.99) < 1e-6 and p.n_samples_per_ray == 256
    q = t.cached(vi, 640, 480, livre_amd.render_params(256))
    assert (q.tile_w, q.tile_h, q.samples_per_thread) == (0, 0, 0)
    assert t.tuned == 0
