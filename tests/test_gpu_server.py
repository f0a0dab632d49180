"""Serving loop (lvr_server_*): every frame rendered in the mix of sessions
equals the same frame rendered alone (pixels and composite steps), under
overlapping launches, brick-pool pressure with misses produced on the device,
admission by budget and coalescing of superseded requests."""
import numpy as np
import pytest

import livre_amd
from livre_amd import scenes

pytestmark = pytest.mark.gpu


def _volume(kind, n=64, block=16):
    if kind == "noise":
        vi = livre_amd.volume_info(n, block, 1, livre_amd.DT_UINT16)
        src = dict(noise=livre_amd.noise_params(31, octaves=4, base_cells=4))
    else:
        vi = livre_amd.volume_info(n, block, 1, livre_amd.DT_UINT8)
        src = dict(field=livre_amd.field_params(livre_amd.FIELD_BLOBS, 24, 32, 0.1))
    return vi, src


def _generate(ctx, ids, src):
    if "noise" in src:
        ctx.generate_noise(src["noise"], ids)
    else:
        ctx.generate_field(src["field"], ids)


def _camera(seed, w, h):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    mv, proj = scenes.camera(tuple(d * rng.uniform(1.4, 2.4)), w, h, 60.0,
                             up=(0, 1, 0) if abs(d[1]) < 0.95 else (0, 0, 1))
    return livre_amd.frustum(mv, proj)


def _alone(ctx, fr, w, h, tf, rng, params, vi):
    vis = ctx.select_visibles(fr, h, 1.0, 0, vi.depth - 1)
    ordered = livre_amd.order(vi, fr, vis)
    ctx.set_tf(tf, rng)
    ctx.render(fr, w, h, ordered, params)
    return ctx.read_frame()


def test_frames_in_the_mix_equal_frames_alone():
    vols = [_volume("noise"), _volume("blobs")]
    ctxs = []
    with livre_amd.Server(budget_bytes=1 << 34, max_batch=8) as srv:
        vids = []
        for vi, src in vols:
            ids = scenes.all_nodes(vi)
            ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
            ctxs.append(ctx)
            vids.append(srv.add_volume(ctx, **src))
        sessions = []
        for k in range(6):
            v = k % 2
            w, h = 48 + 16 * k, 40 + 12 * k
            tf = scenes.piecewise_tf(256, seed=50 + k, max_alpha=0.5)
            params = livre_amd.render_params([64, 96, 128][k % 3])
            s, adm = srv.open(vids[v], w, h, tf, scenes.type_range(vols[v][0].data_type),
                              params)
            assert adm
            sessions.append((s, v, w, h, tf, params))
        for rnd in range(3):
            frs = {}
            for s, v, w, h, tf, params in sessions:
                frs[s] = _camera(100 * rnd + s, w, h)
                srv.submit(s, frs[s])
                srv.step()  # launches overlap the next submissions
            srv.drain()
            done = srv.poll()
            assert sorted(d["session"] for d in done if d["status"] == 0) == \
                sorted(s for s, *_ in sessions)
            for s, v, w, h, tf, params in sessions:
                img, smp = srv.read(s)
                vi = vols[v][0]
                ref, ref_smp = _alone(ctxs[v], frs[s], w, h, tf, scenes.type_range(vi.data_type),
                                      params, vi)
                print("round", rnd, "session", s, w, h, "samples", smp, ref_smp, flush=True)
                assert smp == ref_smp and np.array_equal(img, ref), (rnd, s)
        st = srv.stats()
        assert st.frames == 18 and st.failed == 0 and st.bricks_generated > 0
    for c in ctxs:
        c.close()


def test_pool_pressure_lru_and_device_sources():
    vi, src = _volume("noise", n=128, block=16)
    ids = scenes.all_nodes(vi)
    ref_ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
    _generate(ref_ctx, ids, src)
    small = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * (len(ids) // 3))
    tf = scenes.piecewise_tf(256, seed=7, max_alpha=0.5)
    rng = scenes.type_range(vi.data_type)
    params = livre_amd.render_params(128)
    with livre_amd.Server(budget_bytes=1 << 34, max_batch=4) as srv:
        v = srv.add_volume(small, **src)
        sessions = [srv.open(v, 64, 64, tf, rng, params)[0] for _ in range(3)]
        for rnd in range(4):
            frs = {s: _camera(7 * rnd + s, 64, 64) for s in sessions}
            for s in sessions:
                srv.submit(s, frs[s])
            srv.drain()
            for s in sessions:
                img, smp = srv.read(s)
                ref, ref_smp = _alone(ref_ctx, frs[s], 64, 64, tf, rng, params, vi)
                assert smp == ref_smp and np.array_equal(img, ref), (rnd, s)
        st = srv.stats()
        print("generated", st.bricks_generated, "frames", st.frames, flush=True)
        assert st.bricks_generated > len(ids) // 3  # the pool evicted and regenerated
        assert small.stats()[1] <= len(ids) // 3
    small.close()
    ref_ctx.close()


def test_admission_by_budget_and_coalescing():
    vi, src = _volume("blobs")
    ids = scenes.all_nodes(vi)
    pool = livre_amd.slot_bytes(vi) * len(ids)
    ctx = livre_amd.Context(vi, pool)
    tf = scenes.piecewise_tf(256, seed=9, max_alpha=0.5)
    rng = scenes.type_range(vi.data_type)
    params = livre_amd.render_params(64)
    per = 64 * 64 * 4 * 5  # five framebuffers per session
    with livre_amd.Server(budget_bytes=pool + 2 * per, max_batch=8) as srv:
        v = srv.add_volume(ctx, **src)
        a, adm_a = srv.open(v, 64, 64, tf, rng, params)
        b, adm_b = srv.open(v, 64, 64, tf, rng, params)
        c, adm_c = srv.open(v, 64, 64, tf, rng, params)
        assert adm_a and adm_b and not adm_c
        assert srv.stats().waiting_sessions == 1
        # three requests of one session before a step: the first two are superseded
        frs = [_camera(40 + i, 64, 64) for i in range(3)]
        for fr in frs:
            srv.submit(a, fr)
        srv.submit(c, frs[0])  # queued until c is admitted
        srv.drain()
        done = srv.poll()
        assert [d["status"] for d in done if d["session"] == a] == [1, 1, 0]
        assert not any(d["session"] == c for d in done)
        img, smp = srv.read(a)
        ref, ref_smp = _alone(ctx, frs[2], 64, 64, tf, rng, params, vi)
        assert smp == ref_smp and np.array_equal(img, ref)
        srv.close(b)  # frees budget: c is admitted and its queued request runs
        assert srv.stats().waiting_sessions == 0
        srv.drain()
        done = srv.poll()
        assert [d["status"] for d in done if d["session"] == c] == [0]
        img, smp = srv.read(c)
        ref, ref_smp = _alone(ctx, frs[0], 64, 64, tf, rng, params, vi)
        assert smp == ref_smp and np.array_equal(img, ref)
    ctx.close()
