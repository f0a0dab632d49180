"""The out-of-core configuration at its real size: an 8192^3 uint8
blobs-plus-value-noise volume (512 GB logical, 64^3 bricks, 8 levels) streamed from the
host DataSource through pinned memory into a brick pool capped at 6000 slots
(1.7 GB), rendered at 3840x2160 along the fly-through path. For three
frames the renderer draws the coarser-LOD rendering set of the bricks that
have arrived; each such frame equals, bit for bit, the same rendering set
rendered from a pool whose bricks were generated on the device (the host and
device sources are bit-identical), and at least one of them is a fallback
frame (some visible bricks drawn at a coarser level)."""
import numpy as np
import pytest

import livre_amd
from livre_amd import scenes, workloads

pytestmark = pytest.mark.gpu

W4, H4 = 3840, 2160


def test_c4_fallback_frames_at_4k(lvr):
    vi = workloads.c4_volume()
    slot = livre_amd.slot_bytes(vi)
    ctx = livre_amd.Context(vi, slot * 6000)
    source = workloads.c4_source(vi)
    tf = scenes.piecewise_tf(256, seed=8, points=6, max_alpha=0.3)
    rng = (0.0, 255.0)
    params = livre_amd.render_params(0)  # samples per ray from the finest rendered level
    fallbacks = 0
    try:
        with livre_amd.Stream(ctx, source, threads=8, staging_bricks=128) as st:
            coarse = []
            for level in range(4):
                n = 1 << level
                coarse += [scenes.node_id(level, x, y, z) for z in range(n) for y in range(n)
                           for x in range(n)]
            st.request(coarse)
            st.wait()
            for t in (0, 60, 120):
                pos, ahead = workloads.flythrough(t, 300)
                mv = scenes.look_at(tuple(pos), tuple(ahead), (0.0, 1.0, 0.0))
                proj = scenes.perspective(60.0, W4 / H4, 0.01, 6.0)
                fr = livre_amd.frustum(mv, proj)
                vis = ctx.select_visibles(fr, H4, 1.0, 0, vi.depth - 1)
                st.request(vis, ancestors=True)
                for _ in range(40):
                    st.pump()
                rset, avail = st.rendering_set(vis)
                ordered = livre_amd.order(vi, fr, rset)
                ctx.pin(ordered, 1)
                ctx.set_tf(tf, rng)
                ctx.render(fr, W4, H4, ordered, params)
                img, smp = ctx.read_frame()
                ctx.pin(ordered, -1)
                ref_ctx = livre_amd.Context(vi, slot * (len(ordered) + 16))
                try:
                    ref_ctx.generate_blobnoise(source, ordered)
                    ref_ctx.set_tf(tf, rng)
                    ref_ctx.render(fr, W4, H4, ordered, params)
                    ref, ref_smp = ref_ctx.read_frame()
                finally:
                    ref_ctx.close()
                fallbacks += avail < len(vis)
                print("c4 frame", t, "visible", len(vis), "at full LOD", avail, "rendered",
                      len(rset), "samples", smp, flush=True)
                assert smp > 0 and smp == ref_smp and np.array_equal(img, ref), t
    finally:
        ctx.close()
        source.close()
    assert fallbacks >= 1
