"""Batched rendering: every request of one persistent launch (mixed volumes,
data types, viewports, cameras and transfer functions) equals the same frame
rendered alone by lvr_render, bit for bit, with the same composite-step
count; graph replays reproduce the first launch; and the frames stay within
the image tolerance of the oracle."""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import Scene, diff_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def volumes(lvr, orc):
    out = [Scene(livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs", seed=3),
           Scene(livre_amd.DT_UINT16, n=64, block=16, overlap=1, source="noise", seed=4),
           Scene(livre_amd.DT_FLOAT, n=48, block=16, overlap=1, source="noise", seed=5)]
    ctxs = [s.context() for s in out]
    yield list(zip(out, ctxs))
    for c in ctxs:
        c.close()


def _requests(volumes, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(7):
        s, ctx = volumes[k % 3]
        w, h = int(rng.integers(24, 120)), int(rng.integers(24, 120))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = tuple(d * rng.uniform(1.4, 2.6))
        mv, proj, fr = s.camera(eye, w, h)
        ordered = s.visible_ordered(mv, proj, h)
        n_tf = 1024 if k == 4 else 256  # k == 4 reads its TF from global memory
        tf = scenes.piecewise_tf(n_tf, seed=100 + k, max_alpha=0.5)
        nsamp = int(rng.choice([32, 64, 128]))
        reqs.append((s, ctx, fr, w, h, ordered, tf, nsamp))
    return reqs


def _alone(req):
    s, ctx, fr, w, h, ordered, tf, nsamp = req
    ctx.set_tf(tf, s.range)
    ctx.render(fr, w, h, ordered, livre_amd.render_params(nsamp))
    return ctx.read_frame()


def test_batch_equals_single_frames(volumes):
    reqs = _requests(volumes, 1)
    with livre_amd.Batch() as b:
        for s, ctx, fr, w, h, ordered, tf, nsamp in reqs:
            b.add(ctx, fr, w, h, ordered, livre_amd.render_params(nsamp), tf, s.range)
        b.run()
        st = b.wait()
        assert st.requests == len(reqs) and st.graph_captures == 1
        total = 0
        for i, req in enumerate(reqs):
            img, smp = b.read(i)
            ref, ref_smp = _alone(req)
            print("request", i, req[3], req[4], "samples", smp, ref_smp, flush=True)
            assert np.array_equal(img, ref), i
            assert smp == ref_smp
            total += smp
        assert st.samples == total
        # replay of the captured graph: same shape, same frames
        first = [b.read(i)[0] for i in range(len(reqs))]
        for i, (s, ctx, fr, w, h, ordered, tf, nsamp) in enumerate(reqs):
            b.add(ctx, fr, w, h, ordered, livre_amd.render_params(nsamp), tf, s.range)
        b.run()
        st = b.wait()
        assert st.graph_captures == 1 and st.graph_launches == 2
        for i in range(len(reqs)):
            assert np.array_equal(b.read(i)[0], first[i])


def test_batch_without_graph_and_oracle_tolerance(volumes, orc):
    reqs = _requests(volumes, 2)
    with livre_amd.Batch() as b:
        b.set_graph(False)
        for s, ctx, fr, w, h, ordered, tf, nsamp in reqs:
            b.add(ctx, fr, w, h, ordered, livre_amd.render_params(nsamp), tf, s.range)
        b.run()
        st = b.wait()
        assert st.graph_launches == 0
        for i, (s, ctx, fr, w, h, ordered, tf, nsamp) in enumerate(reqs):
            img, smp = b.read(i)
            ref, ref_smp = s.oracle_render(fr, w, h, ordered, tf, nsamp)
            mx, mean = diff_stats(img, oracle.to_rgba8(ref))
            assert mx <= 2 and mean <= 0.5, (i, mx, mean)
            assert smp == ref_smp


@pytest.mark.parametrize("min_tiles", [110, 1, -1])
def test_solo_requests_in_a_mixed_batch(volumes, min_tiles):
    """Requests the single-frame window kernel covers (the 16-bit and float
    ones; the 8-bit ones classify through a table) with at least min_tiles
    8x8 ray tiles get launches of their own after the shared one: with 110
    some do and some do not, with 1 all of them, with -1 none (the batch is
    then captured as a graph). Every frame and sample count equals the frame
    rendered alone, and the batch total is their sum."""
    reqs = _requests(volumes, 3)
    tiles = [((w + 7) // 8) * ((h + 7) // 8) for _s, _c, _f, w, h, _o, _t, _n in reqs]
    print("tiles", tiles, flush=True)
    with livre_amd.Batch() as b:
        b.set_solo_tiles(min_tiles)
        for s, ctx, fr, w, h, ordered, tf, nsamp in reqs:
            b.add(ctx, fr, w, h, ordered, livre_amd.render_params(nsamp), tf, s.range)
        b.run()
        st = b.wait()
        assert st.graph_captures == (1 if min_tiles < 0 else 0)
        eligible = [k for k in range(len(reqs)) if k % 3 != 0]  # 16-bit and float
        assert st.solo_requests == {110: sum(tiles[k] >= 110 for k in eligible),
                                    1: len(eligible), -1: 0}[min_tiles]
        total = 0
        for i, req in enumerate(reqs):
            img, smp = b.read(i)
            ref, ref_smp = _alone(req)
            assert smp == ref_smp and np.array_equal(img, ref), i
            total += smp
        assert st.samples == total


@pytest.mark.parametrize("vol,classify", [(1, 0), (2, 0), (0, 2)])
def test_single_request_batch(volumes, vol, classify):
    """A batch of one request with per-sample classification runs the
    single-frame kernel (no graph capture); run again (a replay of the same
    request) it gives the same frame; both equal lvr_render's frame."""
    s, ctx = volumes[vol]
    w, h = 88, 64
    mv, proj, fr = s.camera((1.2, -0.8, 1.5), w, h)
    ordered = s.visible_ordered(mv, proj, h)
    tf = scenes.piecewise_tf(256, seed=40 + vol, max_alpha=0.5)
    params = livre_amd.render_params(96, classify=classify)
    ctx.set_tf(tf, s.range)
    ctx.render(fr, w, h, ordered, params)
    ref, ref_smp = ctx.read_frame()
    with livre_amd.Batch() as b:
        b.add(ctx, fr, w, h, ordered, params, tf, s.range)
        b.run()
        st = b.wait()
        img, smp = b.read(0)
        assert st.graph_captures == 0 and st.samples == smp == ref_smp
        assert np.array_equal(img, ref)
        b.run()
        st = b.wait()
        again, smp2 = b.read(0)
        assert st.graph_captures == 0 and smp2 == ref_smp and np.array_equal(again, ref)


def test_batch_rejects_what_it_does_not_render(volumes):
    s, ctx = volumes[0]
    mv, proj, fr = s.camera((1.5, 0.2, 1.1), 32, 32)
    with livre_amd.Batch() as b:
        with pytest.raises(livre_amd.LivreError):
            b.add(ctx, fr, 32, 32, s.ids[:4], livre_amd.render_params(32, linear=True),
                  scenes.piecewise_tf(256, 1), s.range)
        with pytest.raises(livre_amd.LivreError):
            b.add(ctx, fr, 32, 32, s.ids[:4], livre_amd.render_params(32, tile=(5, 5)),
                  scenes.piecewise_tf(256, 1), s.range)
        # one samples-per-thread value per voxel type in a batch (it selects
        # the kernel)
        b.add(ctx, fr, 32, 32, s.ids[:4], livre_amd.render_params(32, samples_per_thread=1),
              scenes.piecewise_tf(256, 1), s.range)
        with pytest.raises(livre_amd.LivreError):
            b.add(ctx, fr, 32, 32, s.ids[:4],
                  livre_amd.render_params(32, samples_per_thread=3),
                  scenes.piecewise_tf(256, 1), s.range)


def test_replay_reads_slots_from_the_page_table(lvr, orc):
    """The batch's brick records carry page-table indices, not slots: the
    kernel reads each brick's slot from the device page table when it draws
    it. A replay after one of the request's bricks was evicted skips that
    brick (the oracle's frame without it); after the brick came back into
    another slot, the replay reads it from there (the first frame again)."""
    s = Scene(livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs", seed=8)
    slot = livre_amd.slot_bytes(s.vi)
    ctx = s.context(pool_bytes=slot * (len(s.ids) + 4))
    try:
        w, h = 96, 72
        tf = scenes.piecewise_tf(256, seed=12, max_alpha=0.5)
        mv, proj, fr = s.camera((1.3, 0.7, 1.2), w, h)
        ordered = s.visible_ordered(mv, proj, h)
        params = livre_amd.render_params(64)
        with livre_amd.Batch() as b:
            b.add(ctx, fr, w, h, ordered, params, tf, s.range)
            b.run()
            b.wait()
            first, s1 = b.read(0)
            ctx.set_tf(tf, s.range)
            ctx.render(fr, w, h, ordered, params)
            assert np.array_equal(ctx.read_frame()[0], first)
            # a brick near the front of the order that covers many pixels
            m = int(ordered[len(ordered) // 4])
            q = int(next(i for i in s.ids if int(i) not in set(int(x) for x in ordered)))
            ctx.evict(m)
            b.run()  # replay: same keys and ranges, m now reads -1
            b.wait()
            without, s2 = b.read(0)
            ref, ref_s = s.oracle_render(fr, w, h, ordered, tf, 64, missing=(m,))
            mx, mean = diff_stats(without, oracle.to_rgba8(ref))
            print("evicted", m, "samples", s1, s2, ref_s, "max", mx, "mean", mean, flush=True)
            assert s2 == ref_s and s2 < s1
            assert mx <= 2 and mean <= 0.5
            assert not np.array_equal(without, first)
            # m comes back in q's slot, q in m's old one
            ctx.evict(q)
            ctx.upload([m], s.bricks[m][None])
            ctx.upload([q], s.bricks[q][None])
            b.run()
            b.wait()
            again, s3 = b.read(0)
            assert s3 == s1 and np.array_equal(again, first)
    finally:
        ctx.close()
