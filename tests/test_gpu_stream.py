"""Out-of-core streaming (lvr_stream_*): bricks streamed from the host
DataSource through pinned memory equal the source bit for bit; frames
rendered while bricks are still arriving use the coarser-LOD rendering set
and equal the same set rendered from a fully resident pool; transfers run
while a frame renders; a pool smaller than the volume evicts least recently
used bricks but never a pinned one."""
import numpy as np
import pytest

import livre_amd
from livre_amd import scenes

pytestmark = pytest.mark.gpu


def _setup(pool_bricks=None):
    vi = livre_amd.volume_info(128, 16, 1, livre_amd.DT_UINT16)
    ids = scenes.all_nodes(vi)
    noise = livre_amd.noise_params(41, octaves=5, base_cells=4)
    n = pool_bricks or len(ids)
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * n)
    full = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
    full.generate_noise(noise, ids)
    return vi, ids, noise, ctx, full


def _frame(ctx, fr, w, h, ordered, tf, rng):
    ctx.set_tf(tf, rng)
    ctx.render(fr, w, h, ordered, livre_amd.render_params(128))
    return ctx.read_frame()


def test_streamed_bricks_equal_the_source():
    vi, ids, noise, ctx, full = _setup()
    with livre_amd.Stream(ctx, noise, threads=4, staging_bricks=24) as st:
        st.request(ids)
        st.wait()
        s = st.stats()
        assert s.bricks_loaded == len(ids) and s.source_errors == 0
        for nid in ids[::7]:
            assert np.array_equal(ctx.download(nid), livre_amd.noise_brick_host(vi, noise, nid))
    ctx.close()
    full.close()


def test_fallback_frames_while_streaming_with_a_small_pool():
    vi, ids, noise, ctx, full = _setup(pool_bricks=260)
    tf = scenes.piecewise_tf(256, seed=12, max_alpha=0.5)
    rng = scenes.type_range(vi.data_type)
    w = h = 96
    rs = np.random.default_rng(3)
    with livre_amd.Stream(ctx, noise, threads=3, staging_bricks=16) as st:
        # the two coarsest levels first, so every frame has a fallback
        coarse = [i for i in ids if scenes.decode_node(i)[0] <= 1]
        st.request(coarse)
        st.wait()
        fallbacks = 0
        for k in range(6):
            d = rs.normal(size=3)
            d /= np.linalg.norm(d)
            mv, proj = scenes.camera(tuple(d * 1.6), w, h, 60.0,
                                     up=(0, 1, 0) if abs(d[1]) < 0.95 else (0, 0, 1))
            fr = livre_amd.frustum(mv, proj)
            vis = ctx.select_visibles(fr, h, 1.0, 0, vi.depth - 1)
            st.request(vis)
            st.pump()
            rset, avail = st.rendering_set(vis)
            fallbacks += avail < len(vis)
            ordered = livre_amd.order(vi, fr, rset)
            ctx.pin(ordered, 1)
            ctx.set_tf(tf, rng)
            ctx.render(fr, w, h, ordered, livre_amd.render_params(128))
            for _ in range(20):  # transfers while the frame renders
                st.pump()
            img, smp = ctx.read_frame()
            ctx.pin(ordered, -1)
            ref, ref_smp = _frame(full, fr, w, h, ordered, tf, rng)
            print("frame", k, "visible", len(vis), "available", avail, "set", len(rset),
                  "samples", smp, ref_smp, flush=True)
            assert smp == ref_smp and np.array_equal(img, ref), k
            # once everything arrived the frame is the full-LOD frame
            st.wait()
            rset2, avail2 = st.rendering_set(vis)
            assert avail2 == len(vis) and sorted(rset2) == sorted(vis)
        s = st.stats()
        print("loaded", s.bricks_loaded, "pool", ctx.stats(), "fallback frames", fallbacks,
              flush=True)
        assert fallbacks > 0 and ctx.stats()[1] <= 260
        assert s.bricks_loaded > 260  # more than the pool holds: LRU eviction ran
    ctx.close()
    full.close()
