"""GPU parity of the HIP render path against the oracle, through the C ABI.

Tolerances (from the path's specification): RGBA8 frames within max 2/255
and mean 0.5/255 per channel of the FP32 restatement of fragRayCast.glsl;
executed composite steps within 1e-4 relative; visible-brick sets, their
order and integer histograms bit-exact.
"""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import Scene, diff_stats

pytestmark = pytest.mark.gpu

MAX_DIFF = 2
MEAN_DIFF = 0.5


def _check_frame(scene, ctx, eye, w, h, tf, n_samples, linear=False, exit_a=0.999,
                 missing=(), rgba8_tf=True, fovy=60.0):
    mv, proj, fr = scene.camera(eye, w, h, fovy)
    vis = oracle.select_visibles(scene.ovi, mv, proj, h, 1.0, 0, scene.ovi.depth - 1)
    ordered = oracle.order(scene.ovi, mv, vis)
    # same DFS input order on both sides: std::sort's tie handling then agrees
    assert np.array_equal(livre_amd.order(scene.vi, fr, vis), ordered)
    if rgba8_tf:
        lut = scenes.lut_rgba8(tf)
        tf_used = lut.astype(np.float32) / np.float32(255.0)
        ctx.set_tf_rgba8(lut, scene.range)
    else:
        tf_used = tf
        ctx.set_tf(tf, scene.range)
    for m in missing:
        if ctx.resident(m):
            ctx.evict(m)
    ref, ref_samples = scene.oracle_render(fr, w, h, ordered, tf_used, n_samples,
                                           early_exit=exit_a, linear=linear,
                                           missing=missing)
    params = livre_amd.render_params(n_samples, early_exit=exit_a, linear=linear)
    ctx.render(fr, w, h, ordered, params)
    img, samples = ctx.read_frame()
    ref8 = oracle.to_rgba8(ref)
    mx, mean = diff_stats(img, ref8)
    print("eye", eye, "bricks", len(ordered), "max", mx, "mean", mean,
          "samples", samples, ref_samples, flush=True)
    assert mx <= MAX_DIFF and mean <= MEAN_DIFF
    assert abs(int(samples) - int(ref_samples)) <= 1e-4 * max(1, ref_samples)
    assert ref_samples > 0 or not ordered.size
    return img


@pytest.fixture(scope="module")
def blobs_u8(lvr, orc):
    s = Scene(livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs")
    ctx = s.context()
    yield s, ctx
    ctx.close()


@pytest.mark.parametrize("k", range(4))
def test_orbit_u8_nearest(blobs_u8, k):
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=2)
    _check_frame(s, ctx, scenes.orbit(k, 4, radius=2.0, height=0.3 * k - 0.4), 96, 80,
                 tf, 64)


def test_tile_shapes_agree(blobs_u8):
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=3)
    lut = scenes.lut_rgba8(tf)
    ctx.set_tf_rgba8(lut, s.range)
    mv, proj, fr = s.camera((1.2, 0.9, 1.4), 100, 70)
    ordered = s.visible_ordered(mv, proj, 70)
    imgs = []
    for tile in ((8, 8), (16, 4), (4, 16), (32, 2), (64, 1)):
        ctx.render(fr, 100, 70, ordered, livre_amd.render_params(64, tile=tile))
        imgs.append(ctx.read_frame())
    for img, smp in imgs[1:]:
        assert np.array_equal(img, imgs[0][0]) and smp == imgs[0][1]


def test_degenerate_inputs(blobs_u8):
    """An empty brick list renders the background (all zero, no composite
    steps); one-pixel-wide and one-pixel-high viewports (a partial ray tile
    at every shape) match the oracle; invalid viewports, launch shapes and
    more than 2^20 samples per ray are rejected before anything is queued,
    and the context renders on afterwards."""
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=11)
    ctx.set_tf(tf, s.range)
    mv, proj, fr = s.camera((1.3, 0.7, 1.2), 48, 40)
    ctx.render(fr, 48, 40, [], livre_amd.render_params(64))
    img, smp = ctx.read_frame()
    assert smp == 0 and img.shape == (40, 48, 4) and not img.any()
    for w, h in ((1, 1), (1, 45), (45, 1)):
        _check_frame(s, ctx, (1.3, 0.7, 1.2), w, h, tf, 64)
    ordered = s.visible_ordered(mv, proj, 40)
    bad = [((0, 40), {}), ((48, 0), {}), ((16385, 40), {}),
           ((48, 40), dict(tile=(7, 9))), ((48, 40), dict(samples_per_thread=4)),
           ((48, 40), dict(waves_per_simd=9)), ((48, 40), dict(classify=3))]
    for (w, h), kw in bad:
        with pytest.raises(livre_amd.LivreError):
            ctx.render(fr, w, h, ordered, livre_amd.render_params(64, **kw))
    with pytest.raises(livre_amd.LivreError):
        ctx.render(fr, 48, 40, ordered, livre_amd.render_params((1 << 20) + 1))
    ctx.render(fr, 48, 40, ordered, livre_amd.render_params(1 << 20))  # the bound itself
    assert ctx.read_frame()[1] > 0
    _check_frame(s, ctx, (1.3, 0.7, 1.2), 48, 40, tf, 64)


def test_frame_stats_of_frames_in_flight(blobs_u8):
    """Frames queued back to back keep their own statistics: the stats read
    `back` frames later equal those of the same frame rendered alone."""
    s, ctx = blobs_u8
    ctx.set_tf(scenes.piecewise_tf(256, seed=6), s.range)
    ctx.set_timing(True)
    eyes = [(1.2, 0.9, 1.4), (-1.3, 0.5, 1.1), (0.8, -1.2, 1.3)]
    frames = []
    for eye in eyes:
        mv, proj, fr = s.camera(eye, 80, 64)
        frames.append((fr, s.visible_ordered(mv, proj, 64)))
    alone = []
    for fr, ordered in frames:
        ctx.render(fr, 80, 64, ordered, livre_amd.render_params(64))
        st = ctx.frame_stats()
        alone.append((st.samples, st.bricks, st.tile_entries))
    for fr, ordered in frames:
        ctx.render(fr, 80, 64, ordered, livre_amd.render_params(64))
    for back, want in zip((2, 1, 0), alone):
        st = ctx.frame_stats(back=back)
        assert (st.samples, st.bricks, st.tile_entries) == want and st.raycast_ms > 0
    with pytest.raises(livre_amd.LivreError):
        ctx.frame_stats(back=4)
    ctx.set_timing(False)


def test_queued_frames_of_different_sizes(blobs_u8):
    """Frames of different viewports queued back to back: their records,
    tile lists and counters alternate between two sets and are binned beside
    the previous frame's ray cast, and each keeps the samples (and the last
    its image) of the same frame rendered alone."""
    s, ctx = blobs_u8
    ctx.set_tf(scenes.piecewise_tf(256, seed=7), s.range)
    ctx.set_timing(True)
    shots = [((1.2, 0.9, 1.4), 96, 80), ((-1.3, 0.5, 1.1), 40, 24),
             ((0.8, -1.2, 1.3), 130, 97), ((1.5, 0.2, -1.0), 64, 64),
             ((0.3, 1.6, 0.9), 17, 33)]
    prepared = []
    for eye, w, h in shots:
        mv, proj, fr = s.camera(eye, w, h)
        prepared.append((fr, w, h, s.visible_ordered(mv, proj, h)))
    alone = []
    for fr, w, h, ordered in prepared:
        ctx.render(fr, w, h, ordered, livre_amd.render_params(64))
        alone.append(ctx.read_frame())
    for fr, w, h, ordered in prepared:  # queued without waiting
        ctx.render(fr, w, h, ordered, livre_amd.render_params(64))
    for back in range(4):
        assert ctx.frame_stats(back=back).samples == alone[len(shots) - 1 - back][1]
    img, smp = ctx.read_frame()
    assert smp == alone[-1][1] and np.array_equal(img, alone[-1][0])
    ctx.set_timing(False)


def test_camera_inside_volume(blobs_u8):
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=4, max_alpha=0.3)
    mv, proj, fr = s.camera((0.1, 0.05, 0.2), 64, 64, 70.0, target=(0.4, 0.1, -0.3))
    ordered = s.visible_ordered(mv, proj, 64)
    lut = scenes.lut_rgba8(tf)
    ctx.set_tf_rgba8(lut, s.range)
    ref, ref_samples = s.oracle_render(fr, 64, 64, ordered,
                                       lut.astype(np.float32) / np.float32(255), 64)
    ctx.render(fr, 64, 64, ordered, livre_amd.render_params(64))
    img, samples = ctx.read_frame()
    mx, mean = diff_stats(img, oracle.to_rgba8(ref))
    assert mx <= MAX_DIFF and mean <= MEAN_DIFF
    assert abs(int(samples) - int(ref_samples)) <= 1e-4 * max(1, ref_samples)


def test_exit_threshold_and_sampling_rate(blobs_u8):
    s, ctx = blobs_u8
    tf = scenes.piecewise_tf(256, seed=5, max_alpha=0.9)
    _check_frame(s, ctx, (0.3, 1.6, 1.3), 64, 64, tf, 128, exit_a=0.99)
    _check_frame(s, ctx, (0.3, 1.6, 1.3), 64, 64, tf, 32)


def test_missing_bricks_are_skipped(lvr, orc):
    s = Scene(livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs", seed=7)
    ctx = s.context()
    try:
        mv, proj, fr = s.camera((1.7, 0.4, 1.0), 64, 64)
        ordered = s.visible_ordered(mv, proj, 64)
        missing = tuple(int(x) for x in ordered[::3])
        _check_frame(s, ctx, (1.7, 0.4, 1.0), 64, 64, scenes.piecewise_tf(256, seed=8),
                     64, missing=missing)
    finally:
        ctx.close()


@pytest.mark.parametrize("dt", [livre_amd.DT_UINT16, livre_amd.DT_FLOAT])
def test_noise_volumes_fp32_tf(lvr, orc, dt):
    s = Scene(dt, n=64, block=16, overlap=1, source="noise", seed=11)
    ctx = s.context()
    try:
        tf = scenes.piecewise_tf(4096 if dt == livre_amd.DT_UINT16 else 256, seed=12,
                                 max_alpha=0.4)
        _check_frame(s, ctx, (-1.4, 0.8, 1.5), 80, 64, tf, 64, rgba8_tf=False)
        _check_frame(s, ctx, (0.2, -2.1, 0.6), 64, 64, tf, 128, rgba8_tf=False)
    finally:
        ctx.close()


@pytest.mark.parametrize("dt,linear", [(livre_amd.DT_UINT16, False),
                                       (livre_amd.DT_UINT8, True),
                                       (livre_amd.DT_FLOAT, False)])
def test_data_range_narrower_than_the_volume(lvr, orc, dt, linear):
    """Densities below 0 and above 1 take the first and the last transfer
    function entry (clamp to edge, fragRayCast.glsl:184-204). Both edge
    entries differ sharply from their neighbours, so reading a neighbour
    instead of the edge entry shows far above the tolerance."""
    s = Scene(dt, n=64, block=16, overlap=1, source="noise", seed=13)
    lo, hi = s.range
    s.range = (lo + 0.35 * (hi - lo), lo + 0.65 * (hi - lo))
    tf = scenes.piecewise_tf(256, seed=14, max_alpha=0.4).copy()
    tf[0] = (1.0, 0.0, 0.0, 0.05)
    tf[-1] = (0.0, 0.0, 1.0, 0.3)
    ctx = s.context()
    try:
        _check_frame(s, ctx, (-1.4, 0.8, 1.5), 80, 64, tf, 64, linear=linear,
                     rgba8_tf=False)
    finally:
        ctx.close()


def test_non_finite_voxels_stay_inside_the_table(lvr):
    """Float voxels that are NaN or infinite make the TF coordinate NaN or
    infinite. The TF index is a saturating floor-convert clamped in integers
    to [-1, nTF - 1], so every march reads inside the table: the brick-by-brick
    marches at 1 and 3 fetches in flight and the two-brick window at 2, with
    two ray-tile shapes, all render the same frame and composite the same
    number of samples. (The reference shader leaves non-finite texels to the
    GL implementation, so there is no oracle frame to compare with.)"""
    s = Scene(livre_amd.DT_FLOAT, n=64, block=16, overlap=1, source="noise", seed=15)
    rng = np.random.default_rng(15)
    for i in list(s.bricks):
        b = np.array(s.bricks[i], dtype=np.float32, copy=True)
        flat = b.reshape(-1)
        idx = rng.choice(flat.size, 60, replace=False)
        flat[idx[:20]] = np.nan
        flat[idx[20:40]] = np.inf
        flat[idx[40:]] = -np.inf
        s.bricks[i] = b
    ctx = s.context()
    try:
        ctx.set_tf(scenes.piecewise_tf(256, seed=16, max_alpha=0.4), s.range)
        mv, proj, fr = s.camera((1.5, 0.6, 1.2), 64, 48)
        ordered = s.visible_ordered(mv, proj, 48)
        frames = []
        for depth in (1, 2, 3):
            for tile in ((8, 8), (16, 4)):
                ctx.render(fr, 64, 48, ordered,
                           livre_amd.render_params(64, samples_per_thread=depth, tile=tile))
                frames.append(ctx.read_frame())
        img0, smp0 = frames[0]
        assert smp0 > 0
        for img, smp in frames[1:]:
            assert smp == smp0 and np.array_equal(img, img0)
    finally:
        ctx.close()


def test_trilinear_against_fp32_oracle(lvr, orc):
    s = Scene(livre_amd.DT_UINT16, n=64, block=16, overlap=1, source="noise", seed=13)
    ctx = s.context()
    try:
        tf = scenes.piecewise_tf(1024, seed=14, max_alpha=0.5)
        _check_frame(s, ctx, (1.3, 1.1, -1.2), 72, 72, tf, 64, linear=True,
                     rgba8_tf=False)
    finally:
        ctx.close()


@pytest.mark.parametrize("dt", [livre_amd.DT_UINT8, livre_amd.DT_UINT16, livre_amd.DT_FLOAT])
def test_trilinear_non_cubic_volume(lvr, orc, dt):
    """A 72 x 56 x 40 volume: brick extents along y and z are not powers of two,
    so the trilinear march takes the corrected texture coordinate
    (texturePosition<false>) beside the power-of-two one along x, and texel
    indices clamp at the short axes' edges."""
    s = Scene(dt, n=(72, 56, 40), block=16, overlap=1, source="noise", seed=21)
    assert any(float(e) != 2.0 ** round(np.log2(float(e))) for e in s.vi.world_size)
    ctx = s.context()
    try:
        tf = scenes.piecewise_tf(256, seed=22, max_alpha=0.5)
        _check_frame(s, ctx, (1.2, 0.9, 1.3), 80, 64, tf, 96, linear=True,
                     rgba8_tf=dt == livre_amd.DT_UINT8)
        _check_frame(s, ctx, (-0.9, -1.1, 1.0), 64, 72, tf, 160, linear=True,
                     rgba8_tf=dt == livre_amd.DT_UINT8)
    finally:
        ctx.close()


def test_device_noise_equals_host_noise(lvr):
    vi = livre_amd.volume_info(128, 32, 1, livre_amd.DT_UINT16)
    p = livre_amd.noise_params(3, octaves=5, base_cells=8)
    ids = scenes.all_nodes(vi)
    with livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids)) as ctx:
        assert ctx.stats()[0] == len(ids)
        ctx.generate_noise(p, ids)
        for nid in ids[:: max(1, len(ids) // 20)]:
            assert np.array_equal(ctx.download(nid), livre_amd.noise_brick_host(vi, p, nid))


@pytest.mark.parametrize("dt", [livre_amd.DT_UINT8, livre_amd.DT_UINT16, livre_amd.DT_FLOAT])
def test_device_blob_noise_equals_host(lvr, dt):
    vi = livre_amd.volume_info(128, 32, 1, dt)
    src = livre_amd.BlobNoise(vi, livre_amd.noise_params(4, octaves=5, base_cells=8), 9,
                              count=24, cells=16, weight=144)
    ids = scenes.all_nodes(vi)
    try:
        with livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids)) as ctx:
            ctx.generate_blobnoise(src, ids)
            for nid in ids:
                assert np.array_equal(ctx.download(nid), src.brick_host(nid)), nid
            other = livre_amd.volume_info(64, 32, 1, dt)
            with pytest.raises(livre_amd.LivreError):
                ctx.generate_blobnoise(livre_amd.BlobNoise(other, src.noise, 9), ids[:1])
    finally:
        src.close()


@pytest.mark.parametrize("kind,dt", [(livre_amd.FIELD_BLOBS, livre_amd.DT_UINT8),
                                     (livre_amd.FIELD_BLOBS, livre_amd.DT_UINT16),
                                     (livre_amd.FIELD_SINES, livre_amd.DT_FLOAT)])
def test_device_field_equals_host_field(lvr, kind, dt):
    vi = livre_amd.volume_info(96, 32, 1, dt)
    p = livre_amd.field_params(kind, 24, 8, noise=0.1 if kind == livre_amd.FIELD_BLOBS else 0)
    ids = scenes.all_nodes(vi)
    with livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids)) as ctx:
        ctx.generate_field(p, ids)
        for nid in ids:
            assert np.array_equal(ctx.download(nid), livre_amd.field_brick_host(vi, p, nid)), nid


@pytest.mark.parametrize("dt,bins", [(livre_amd.DT_UINT8, 256),
                                     (livre_amd.DT_UINT16, 65536),
                                     (livre_amd.DT_FLOAT, 256)])
def test_histogram_bit_exact(lvr, orc, dt, bins):
    s = Scene(dt, n=64, block=16, overlap=1, source="noise" if dt != livre_amd.DT_UINT8
              else "blobs", seed=21)
    ctx = s.context()
    try:
        if dt == livre_amd.DT_FLOAT:
            ctx.set_tf(scenes.piecewise_tf(256, seed=1), s.range)
        mv, proj, fr = s.camera((1.5, 0.5, 1.2), 128, 128)
        ids = s.visible_ordered(mv, proj, 128)
        got = ctx.histogram(ids, bins)
        ref = oracle.histogram(s.ovi, ids, [s.bricks[int(i)] for i in ids], bins,
                               s.range)
        assert np.array_equal(got, ref)
        # every brick of the tree: sum = interior voxels * scale
        got = ctx.histogram(s.ids, bins)
        ref = oracle.histogram(s.ovi, s.ids, [s.bricks[int(i)] for i in s.ids], bins,
                               s.range)
        assert np.array_equal(got, ref)
        assert int(got.sum()) == s.vi.depth * 64 ** 3
    finally:
        ctx.close()


def test_histogram_incremental_sequence(lvr, orc):
    """Consecutive frames update the running histogram by the bricks that
    entered and left the set; every frame equals a full recount."""
    s = Scene(livre_amd.DT_UINT16, n=64, block=16, overlap=1, source="noise", seed=31)
    ctx = s.context()
    try:
        for k in range(8):
            if k == 5:  # new voxels for one brick: the running sum is rebuilt
                nid = int(s.ids[-1])
                s.bricks[nid] = (s.bricks[nid] // 2).astype(np.uint16)
                ctx.upload([nid], s.bricks[nid][None])
            eye = scenes.orbit(k, 16, radius=1.6 + 0.1 * k, height=0.2)
            mv, proj, fr = s.camera(eye, 96, 96)
            ids = s.visible_ordered(mv, proj, 96, sse=0.7)
            got = ctx.histogram(ids, 65536)
            ref = oracle.histogram(s.ovi, ids, [s.bricks[int(i)] for i in ids], 65536)
            assert np.array_equal(got, ref), k
    finally:
        ctx.close()


@pytest.mark.parametrize("dt,n,block,bins", [(livre_amd.DT_UINT8, 64, 16, 256),
                                              (livre_amd.DT_UINT16, 128, 64, 65536)])
def test_histogram_cache_sequence(lvr, orc, dt, n, block, bins):
    """The per-slot cache of brick histograms (HistogramObject per brick):
    frames over changing sets, a brick rewritten in place, a brick evicted and
    reloaded into another slot, a brick of the previous set evicted (its row
    is gone: the sum starts afresh); every frame equals the oracle's count."""
    s = Scene(dt, n=n, block=block, overlap=1,
              source="blobs" if dt == livre_amd.DT_UINT8 else "noise", seed=37)
    slot = livre_amd.slot_bytes(s.vi)
    assert slot * 2 >= bins * 4  # the cache's rows fit: its path is the one taken
    ctx = s.context(pool_bytes=slot * (len(s.ids) + 2))
    try:
        prev = None
        for k in range(9):
            eye = scenes.orbit(k, 12, radius=1.5 + 0.12 * k, height=0.3 - 0.05 * k)
            mv, proj, fr = s.camera(eye, 96, 96)
            ids = s.visible_ordered(mv, proj, 96, sse=0.8)
            if k == 3:  # new voxels for a brick of the set, in its slot
                nid = int(ids[len(ids) // 2])
                s.bricks[nid] = (s.bricks[nid] // 2).astype(s.bricks[nid].dtype)
                ctx.upload([nid], s.bricks[nid][None])
            if k == 5:  # a brick of the set moves to another slot
                nid = int(ids[0])
                other = int(next(i for i in s.ids if int(i) not in set(int(x) for x in ids)))
                ctx.evict(nid)
                ctx.evict(other)
                ctx.upload([nid], s.bricks[nid][None])
                ctx.upload([other], s.bricks[other][None])
            if k == 7 and prev is not None:  # a brick of the previous set leaves the pool
                gone = int(prev[-1])
                ctx.evict(gone)
                ids = np.array([i for i in ids if int(i) != gone], np.uint64)
            got = ctx.histogram(ids, bins)
            # bricks not resident are skipped (the histogram cache cannot
            # load them)
            here = [int(i) for i in ids if ctx.resident(int(i))]
            ref = oracle.histogram(s.ovi, here, [s.bricks[i] for i in here], bins, s.range)
            assert np.array_equal(got, ref), k
            prev = ids
    finally:
        ctx.close()


def test_selection_bit_exact_against_goldens(lvr, orc):
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "lod_selection.json")))
    vi = livre_amd.volume_info(fx["voxels"], fx["block"], fx["overlap"], livre_amd.DT_UINT8)
    fr = livre_amd.frustum(fx["mv"], fx["proj"])
    with livre_amd.Context(vi, 1 << 26) as ctx:  # a few slots: selection needs none
        for case in fx["cases"]:
            got = ctx.select_visibles(fr, case["window_height"], case["sse"],
                                      case["min_lod"], case["max_lod"])
            assert sorted(int(x) for x in got) == case["expected"]


def test_selection_dfs_order_matches_oracle(lvr, orc):
    vi = livre_amd.volume_info(2048, 64, 1, livre_amd.DT_UINT16)
    ovi = orc.volume_info([2048] * 3, [64] * 3, [1] * 3, orc.DT_UINT16)
    rng = np.random.default_rng(5)
    with livre_amd.Context(vi, 1 << 26) as ctx:  # a few slots: selection needs none
        for k in range(12):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            h = int(rng.integers(256, 2160))
            w = int(h * 16 / 9)
            mv, proj = scenes.camera(tuple(d * rng.uniform(0.8, 3.0)), w, h)
            fr = livre_amd.frustum(mv, proj)
            sse = float(rng.choice([0.5, 1.0, 2.0, 4.0]))
            got = ctx.select_visibles(fr, h, sse, 0, vi.depth - 1)
            ref = orc.select_visibles(ovi, mv, proj, h, sse, 0, ovi.depth - 1)
            assert np.array_equal(got, ref), k


def test_async_selection_matches_blocking_selection(lvr, orc):
    """A selection queued while a frame renders returns what the blocking
    call returns; one selection may be pending at a time."""
    vi = livre_amd.volume_info(2048, 64, 1, livre_amd.DT_UINT16)
    ovi = orc.volume_info([2048] * 3, [64] * 3, [1] * 3, orc.DT_UINT16)
    with livre_amd.Context(vi, 1 << 26) as ctx:
        for k, eye in enumerate(((1.5, 0.4, 1.1), (-0.9, 1.6, 0.7), (0.3, -0.2, 2.4))):
            mv, proj = scenes.camera(eye, 640, 360)
            fr = livre_amd.frustum(mv, proj)
            ctx.select_visibles_async(fr, 360, 1.0, 0, vi.depth - 1)
            with pytest.raises(livre_amd.LivreError):
                ctx.select_visibles_async(fr, 360, 1.0, 0, vi.depth - 1)
            got = ctx.select_visibles_wait()
            assert np.array_equal(got, ctx.select_visibles(fr, 360, 1.0, 0, vi.depth - 1)), k
            assert np.array_equal(got, orc.select_visibles(ovi, mv, proj, 360, 1.0, 0,
                                                           ovi.depth - 1)), k
        with pytest.raises(livre_amd.LivreError):
            ctx.select_visibles_wait()


def test_pool_policy_and_eviction(lvr):
    """A pool of 2 slots evicts in policy order (Cache.cpp:111-123) and the
    load-order policy reproduces the reference cache test
    (tests/core/cache.cpp:29-78)."""
    vi = livre_amd.volume_info(64, 16, 1, livre_amd.DT_UINT8)
    slot = livre_amd.slot_bytes(vi)
    brick = np.zeros(vi.brick_shape, np.uint8)
    a, b, c = (scenes.node_id(2, i, 0, 0) for i in range(3))
    with livre_amd.Context(vi, 2 * slot + 100) as ctx:
        assert ctx.stats()[0] == 2
        ctx.set_policy(1)
        ctx.upload([a], brick[None])
        ctx.upload([b], brick[None])
        ctx.upload([a], brick[None])  # already resident: no reload, no refresh
        ctx.upload([c], brick[None])  # evicts a (first loaded)
        assert not ctx.resident(a) and ctx.resident(b) and ctx.resident(c)
        ctx.set_policy(0)
        ctx.upload([b], (brick + 1)[None])  # LRU touch
        ctx.upload([a], brick[None])        # evicts c
        assert ctx.resident(a) and ctx.resident(b) and not ctx.resident(c)
        assert np.array_equal(ctx.download(b), brick + 1)
        ctx.evict(a)
        assert ctx.stats()[1] == 1
        with pytest.raises(livre_amd.LivreError):
            ctx.evict(a)
