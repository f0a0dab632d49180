"""Benchmark of the render hot path on one MI355X.

Workload (one step = one frame): 2048^3 uint16 procedural value-noise volume
(5 octaves, seed 3, log-normal remap), 64^3 bricks with a 1-voxel halo,
6 octree levels, all 37,449 bricks resident in the HBM brick pool (21.5 GB);
1920x1080, 120-frame orbit-and-zoom camera path, 2048 samples per unit
length, FP32 4096-entry transfer function, screen-space error 1 px, 65536-bin
histogram. A frame is: device SSE selection, front-to-back order, binning +
ray casting, histogram of the rendered set.

Prints one JSON line. `value` is executed composite steps per second over
the timed frames (Gsamples/s); the pool is filled on the device before the
timed region.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Gsamples/s & fps @1080p 2048³ u16; p50/p99 frame latency (serving)"
HBM_PEAK = 8.0e12
W, H = 1920, 1080
FRAMES = 120
N_SAMPLES = 2048
TF_ENTRIES = 4096
BINS = 65536


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_value(name, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/,
    or None when the summary or the line is absent."""
    path = os.path.join(HERE, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1].split()[0])
    except (OSError, ValueError, IndexError):
        pass
    return None


HBM_PROFILE = "raycast_2048u16_1080p_hbm.txt"
COUNTER_PROFILE = "raycast_2048u16_1080p_counters.txt"


def roofline_limiter():
    """What bounds the headline ray cast, from the committed counter summary."""
    vals = [profile_value(COUNTER_PROFILE, k) for k in (
        "valu_busy", "active_lanes_per_valu", "valu_issue_ceiling_gsamples_per_s",
        "l1_lines_per_voxel_load", "texture_addresser_busy")]
    txt = ["?" if x is None else ("%.2f" % x if x < 1 else "%.1f" % x) for x in vals]
    return ("VALU issue with the vector memory path beside it: VALU busy %s of cycles at %s "
            "of 64 lanes active (a VALU-issue ceiling of %s Gsamples/s at the measured "
            "clock); each voxel load touches %s L1 lines, texture addresser busy %s "
            "(profiles/%s)" % tuple(txt + [COUNTER_PROFILE]))


def valu_ceiling_entry(kernel_gsamples_per_s):
    """The kernel's rate against its VALU-issue ceiling (1024 SIMDs x the
    measured clock / 4 cycles per wave64 instruction / VALU instructions per
    voxel fetch x active lanes), as profiles/ measured it."""
    ceil = profile_value(COUNTER_PROFILE, "valu_issue_ceiling_gsamples_per_s")
    return {"bound": "valu", "achieved_gsamples_per_s": kernel_gsamples_per_s,
            "ceiling_gsamples_per_s": ceil,
            "frac": kernel_gsamples_per_s / ceil if ceil else None,
            "source": "profiles/" + COUNTER_PROFILE}


def schedule(steps):
    """Frame indices of the timed steps: the 120-frame orbit-zoom path
    sampled evenly (stride 120 / steps), so any --steps covers the whole
    path; above 120 steps the path repeats."""
    if steps >= FRAMES:
        return [i % FRAMES for i in range(steps)]
    return [(i * FRAMES) // steps for i in range(steps)]


def build_scene(livre_amd, scenes):
    vi = livre_amd.volume_info(2048, 64, 1, livre_amd.DT_UINT16)
    ids = scenes.all_nodes(vi)
    per = int(np.prod(vi.brick_shape)) * 2
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * (len(ids) + 64))
    noise = livre_amd.noise_params(3, octaves=5, base_cells=8, gain=1.5)
    t0 = time.perf_counter()
    ctx.generate_noise(noise, ids)
    ctx.sync()
    log("generated %d bricks (%.1f GB) on the device in %.1f s"
        % (len(ids), len(ids) * per / 1e9, time.perf_counter() - t0))
    tf = scenes.piecewise_tf(TF_ENTRIES, seed=4, points=8, max_alpha=0.35)
    ctx.set_tf(tf, (0.0, 65535.0))
    return vi, ctx, ids, tf


def camera(livre_amd, scenes, t):
    eye = scenes.orbit_zoom(t % FRAMES, FRAMES)
    mv, proj = scenes.camera(eye, W, H, 60.0)
    return mv, proj, livre_amd.frustum(mv, proj)


def frame(livre_amd, vi, ctx, fr, params):
    vis = ctx.select_visibles(fr, H, 1.0, 0, vi.depth - 1)
    ordered = livre_amd.order(vi, fr, vis)
    ctx.render(fr, W, H, ordered, params)
    ctx.histogram(ordered, BINS, read=False)
    return ordered, ctx.frame_stats()


def cpu_baseline(livre_amd, scenes, vi, ctx, tf, target_s):
    """The FP32 restatement (oracle) on the host cores over frames 0, 30, 60
    and 90 of the orbit path: per frame, selection, order and the per-pixel
    march over a band of rows around the centre sized to about target_s / 4
    seconds, plus the frame's histogram as a sum of cached per-brick ones."""
    import oracle
    ovi = oracle.volume_info([2048] * 3, [64] * 3, [1] * 3, oracle.DT_UINT16)
    threads = os.cpu_count() or 1
    total_samples, total_s, parts = 0, 0.0, []
    for t in (0, 30, 60, 90):
        mv, proj, fr = camera(livre_amd, scenes, t)
        cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye),
                            fr.near_plane, fr.far_plane, W, H)
        t0 = time.perf_counter()
        vis = oracle.select_visibles(ovi, mv, proj, H, 1.0, 0, ovi.depth - 1)
        ordered = oracle.order(ovi, mv, vis)
        t_sel = time.perf_counter() - t0
        bricks = [ctx.download(i) for i in ordered]
        rows, elapsed, samples = 8, 0.0, 0
        while True:
            y0 = H // 2 - rows // 2
            p = oracle.render_params(N_SAMPLES, 0.999, (0.0, 65535.0),
                                     window=(0, y0, W, y0 + rows))
            t0 = time.perf_counter()
            _img, samples = oracle.render(ovi, cam, ordered, bricks, tf, p, "f32", threads)
            elapsed = time.perf_counter() - t0
            if elapsed >= target_s / 16 or rows >= H:
                break
            rows = min(H, int(rows * max(2.0, target_s / 4 / max(elapsed, 1e-3))))
        frac = rows / H
        # the frame's histogram with warm caches (HistogramFilter.cpp:78-149):
        # each rendered brick's HistogramObject is binned once, when the brick
        # is loaded, and a frame sums the cached ones; timed here as the sum
        # of one 65536-bin histogram per rendered brick (32 distinct ones
        # binned by the oracle, reused)
        nh = min(len(ordered), 32)
        per = [oracle.histogram(ovi, [ordered[i]], [bricks[i]], BINS, (0.0, 65535.0))
               for i in range(nh)]
        t0 = time.perf_counter()
        acc = np.zeros(BINS, np.uint64)
        for i in range(len(ordered)):
            acc += per[i % max(nh, 1)]
        t_hist = time.perf_counter() - t0
        del bricks
        spent = elapsed + (t_sel + t_hist) * frac
        total_samples += samples
        total_s += spent
        parts.append("frame %d rows %d..%d (%d samples, %.1f s)"
                     % (t, H // 2 - rows // 2, H // 2 - rows // 2 + rows, samples, spent))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f
                          if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": total_samples / total_s / 1e9, "unit": "Gsamples/s", "cores": threads,
            "kind": "port", "cpu_model": model,
            "sample": "frames 0, 30, 60, 90 of the orbit path, a band of rows around the "
                      "centre of each: oracle selection + order + FP32 march + the frame's "
                      "65536-bin histogram as the sum of cached per-brick histograms, "
                      "selection and histogram charged by the band's share of the frame; "
                      + "; ".join(parts)}


def batching_1080p(livre_amd, scenes, vi, ctx, tf, reps):
    """16 concurrent 1080p sessions of the 2048^3 volume (cameras 7.5 frames
    apart on the orbit path), rendered the plain way (lvr_render per session,
    queued back to back, no wait per request) and as one lvr_batch launch,
    both with the default launch shape. Selection and ordering are done
    before either timed region."""
    sess = []
    for k in range(16):
        fr = camera(livre_amd, scenes, (k * FRAMES) // 16)[2]
        vis = ctx.select_visibles(fr, H, 1.0, 0, vi.depth - 1)
        sess.append((fr, livre_amd.order(vi, fr, vis)))
    params = livre_amd.render_params(N_SAMPLES)
    samples = 0
    for fr, ordered in sess:  # warm-up, and the samples of one pass
        ctx.render(fr, W, H, ordered, params)
        samples += ctx.samples()
    t0 = time.perf_counter()
    for _ in range(reps):
        for fr, ordered in sess:
            ctx.render(fr, W, H, ordered, params)
        ctx.sync()
    t_plain = time.perf_counter() - t0
    with livre_amd.Batch() as b0, livre_amd.Batch() as b1:
        def queue(b):
            for fr, ordered in sess:
                b.add(ctx, fr, W, H, ordered, params, tf, (0.0, 65535.0))
            b.run()
        for b in (b0, b1):  # warm-up
            queue(b)
            solo = b.wait().solo_requests
        # two batches in turn, as the plain pass queues without waiting: the
        # host assembles pass k+1 while pass k is on the GPU
        t0 = time.perf_counter()
        batch_samples, prev = 0, None
        for r in range(reps):
            b = (b0, b1)[r % 2]
            queue(b)
            if prev is not None:
                batch_samples += prev.wait().samples
            prev = b
        batch_samples += prev.wait().samples
        t_batch = time.perf_counter() - t0
        b = prev
        fr, ordered = sess[5]
        ctx.render(fr, W, H, ordered, params)
        img, smp = ctx.read_frame()
        bimg, bsmp = b.read(5)
        same = smp == bsmp and bool(np.array_equal(img, bimg))
    plain = samples * reps / t_plain / 1e9
    batched_rate = batch_samples / t_batch / 1e9
    log("16 sessions at 1080p: plain %.1f Gsamples/s, batched %.1f Gsamples/s (x%.2f)"
        % (plain, batched_rate, batched_rate / plain))
    return {"sessions": 16, "reps": reps, "plain_gsamples_per_s": plain,
            "batched_gsamples_per_s": batched_rate, "ratio": batched_rate / plain,
            "batched_equals_plain": same,
            "solo_requests": solo,
            "plain": "lvr_render per session, queued back to back, one sync per pass",
            "batched": "one lvr_batch per pass, two batches in turn; requests of at least "
                       "4 tiles per resident wavefront run the single-frame windowed kernel "
                       "on their own (solo_requests of 16), the rest one shared launch",
            "shape": "default (8x8 tiles, 2 fetches in flight, 5 waves/SIMD for the "
                     "windowed kernel) for both"}


def batch_workload(args):
    """64 uneven requests over four volumes: the plain renderer (lvr_render
    per request, queued without a wait per request) against one persistent
    batched launch (graph-replayed), both with the autotuned launch shape of
    each request, then 16 concurrent 1080p sessions of the 2048^3 volume the
    same two ways. Selection and ordering are done before either timed
    region."""
    import livre_amd
    from livre_amd import autotune, scenes, workloads
    livre_amd.lib()
    vols = {}
    for name in workloads.BATCH_VOLUMES:
        vols[name] = workloads.Volume(name)
        log("%s: %d bricks generated in %.1f s" % (name, len(vols[name].ids),
                                                    vols[name].gen_s))

    def plain_count(items, tuned):
        total = 0
        for (ctx, fr, w, h, ordered, _p, tf, rng), tp in zip(items, tuned):
            ctx.set_tf(tf, rng)
            ctx.render(fr, w, h, ordered, tp)
            total += ctx.samples()  # waits for the frame
        return total

    def plain_pass(items, tuned):
        # queued: no wait per request; lvr_tf_set waits for the frames of its
        # own context in flight, so a context's TF changes only when it must
        last_tf = {}
        for (ctx, fr, w, h, ordered, _p, tf, rng), tp in zip(items, tuned):
            if last_tf.get(id(ctx)) is not tf:
                ctx.set_tf(tf, rng)
                last_tf[id(ctx)] = tf
            ctx.render(fr, w, h, ordered, tp)
        for ctx in {id(it[0]): it[0] for it in items}.values():
            ctx.sync()

    def batch_pass(batch, items, tuned):
        for (ctx, fr, w, h, ordered, _p, tf, rng), tp in zip(items, tuned):
            batch.add(ctx, fr, w, h, ordered, tp, tf, rng)
        batch.run()
        return batch.wait()

    def compare(items, reps, label):
        for ctx, _f, _w, _h, _o, _p, tf, rng in items:  # a class the table lacks is timed
            ctx.set_tf(tf, rng)
        tuner = autotune.Tuner()
        t0 = time.perf_counter()
        tuned = tuner.tune_batch([(c, f, w, h, o, p) for c, f, w, h, o, p, _t, _r in items])
        t_tune = time.perf_counter() - t0
        plain_samples = plain_count(items, tuned) * reps  # also the warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            plain_pass(items, tuned)
        t_plain = time.perf_counter() - t0
        with livre_amd.Batch() as b0, livre_amd.Batch() as b1:
            pair = (b0, b1)
            for batch in pair:
                batch.set_solo_tiles(args.solo_tiles)
            for batch in pair:  # warm-up and graph capture of both
                batch_pass(batch, items, tuned)
            # two batches in turn: the host assembles pass k+1 while pass k
            # is on the GPU
            gpu_ms, batch_samples, prev = [], 0, None
            t0 = time.perf_counter()
            for r in range(reps):
                batch = pair[r % 2]
                for (ctx, fr, w, h, ordered, _p, tf, rng), tp in zip(items, tuned):
                    batch.add(ctx, fr, w, h, ordered, tp, tf, rng)
                batch.run()
                if prev is not None:
                    st = prev.wait()
                    gpu_ms.append(st.gpu_ms)
                    batch_samples += st.samples
                prev = batch
            st = prev.wait()
            gpu_ms.append(st.gpu_ms)
            batch_samples += st.samples
            t_batch = time.perf_counter() - t0
            batch = prev
            # every batched frame equals the plain frame of the same request
            same = True
            for i in range(0, len(items), max(1, len(items) // 8)):
                ctx, fr, w, h, ordered, params, tf, rng = items[i]
                ctx.set_tf(tf, rng)
                ctx.render(fr, w, h, ordered, tuned[i])
                img, smp = ctx.read_frame()
                bimg, bsmp = batch.read(i)
                same = same and smp == bsmp and bool(np.array_equal(img, bimg))
            graph = (st.graph_captures, st.graph_launches)
            solo = st.solo_requests
        plain = plain_samples / t_plain / 1e9
        batched = batch_samples / t_batch / 1e9
        log("%s: plain %.1f Gsamples/s, batched %.1f Gsamples/s (x%.2f), tuning %.1f s"
            % (label, plain, batched, batched / plain, t_tune))
        return {"requests": len(items), "reps": reps,
                "samples_per_pass": batch_samples // reps,
                "plain_gsamples_per_s": plain, "plain_ms_per_pass": t_plain / reps * 1e3,
                "batched_gsamples_per_s": batched, "batched_ms_per_pass": t_batch / reps * 1e3,
                "batch_gpu_ms_mean": float(np.mean(gpu_ms)), "speedup": batched / plain,
                "frames_identical": same, "graph_captures_launches": list(graph),
                "solo_requests": solo,
                "tuned_classes": tuner.tuned, "tuning_s": t_tune,
                "launch_shapes": sorted({(t.tile_w, t.tile_h, t.samples_per_thread)
                                         for t in tuned})}

    reqs = workloads.batch_requests(64, seed=11)
    items = []
    t0 = time.perf_counter()
    for r in reqs:
        vol = vols[r["volume"]]
        fr, ordered = workloads.prepare(r, vol)
        items.append((vol.ctx, fr, r["width"], r["height"], ordered,
                      livre_amd.render_params(workloads.samples_per_ray(r, vol)), r["tf"],
                      vol.range))
    t_prep = time.perf_counter() - t0
    uneven = compare(items, args.steps, "64 uneven requests")
    uneven["selection_order_ms_per_pass"] = t_prep * 1e3
    uneven["pixels"] = int(sum(r["width"] * r["height"] for r in reqs))

    big = vols["noise2048_u16"]
    sess = []
    tf = scenes.piecewise_tf(TF_ENTRIES, seed=4, points=8, max_alpha=0.35)
    for k in range(16):
        mv, proj, fr = camera(livre_amd, scenes, k * (FRAMES // 16))
        vis = big.ctx.select_visibles(fr, H, 1.0, 0, big.vi.depth - 1)
        sess.append((big.ctx, fr, W, H, livre_amd.order(big.vi, fr, vis),
                     livre_amd.render_params(N_SAMPLES), tf, (0.0, 65535.0)))
    sessions = compare(sess, max(2, args.steps // 2), "16 sessions at 1080p")
    for v in vols.values():
        v.close()
    result = {
        "metric": "Gsamples/s, batched uneven requests in one persistent launch",
        "value": uneven["batched_gsamples_per_s"], "unit": "Gsamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "higher_is_better": True, "vs_baseline": None,
        "dtype": "u8/u16/f32", "data": "synthetic: procedural volumes generated on the device",
        "config": {"workload": "64 requests, square viewports log-uniform in [256^2, 2048^2], "
                               "volumes 512^3 u8 blobs / 1024^3 u16 noise / 1536^3 f32 sines "
                               "(48^3 bricks) / 2048^3 u16 noise, 256-entry TF each, rates "
                               "{0.5,1,2} samples/voxel, sse 1; plus 16 x 1080p sessions of "
                               "the 2048^3 volume (4096-entry TF)",
                   "plain": "lvr_render per request, queued (no wait per request; a "
                            "context's TF upload waits for its frames in flight), autotuned "
                            "launch shape per request",
                   "batched": "one lvr_batch launch per pass (graph replay), two batches "
                              "in turn so the host assembles the next pass during the "
                              "current one, autotuned tiles and samples per thread"},
        "uneven_requests": uneven,
        "sessions_1080p": sessions,
    }
    print(json.dumps(result), flush=True)


def serve_workload(args):
    print(json.dumps(serve_measure(args, args.rates, args.duration)), flush=True)


def serve_measure(args, rates, duration):
    """32 concurrent sessions over the five volumes (720p..4K viewports,
    random-walk cameras, own TFs, 1 sample per voxel) served by the
    continuous-batching loop; Poisson arrivals at each rate for `duration`
    seconds. Latency = completion - arrival of every rendered frame."""
    import livre_amd
    from livre_amd import workloads
    livre_amd.lib()
    # pools of pool_frac of each pyramid, filled on demand by the server's
    # device sources (LRU eviction when a pool is smaller than the working
    # sets of its sessions)
    vols = {}
    for name in workloads.SERVE_VOLUMES:
        spec = workloads.VOLUMES[name]
        vi = livre_amd.volume_info(spec["voxels"], spec["block"], 1, spec["dtype"])
        n_ids = len(livre_amd.scenes.all_nodes(vi))
        vols[name] = workloads.Volume(name, pool_bricks=int(n_ids * args.pool_frac) + 64,
                                      generate=False)
    sessions = workloads.serve_sessions(32, seed=9)
    pools = sum(v.pool_bytes for v in vols.values())
    frames_bytes = sum(5 * 4 * s["width"] * s["height"] for s in sessions)  # 5 frames each
    budget = pools + frames_bytes
    result_rates = []
    with livre_amd.Server(budget, max_batch=args.max_batch) as srv:
        vid = {n: srv.add_volume(v.ctx, noise=v.noise, field=v.field) for n, v in vols.items()}
        sid, last_fr, req_fr, shown = [], {}, {}, {}
        for s in sessions:
            v = vols[s["volume"]]
            params = livre_amd.render_params(v.vmax)
            k, admitted = srv.open(vid[s["volume"]], s["width"], s["height"], s["tf"], v.range,
                                   params)
            sid.append(k)
        # warm-up: one frame per session (fills the pools)
        t0 = time.perf_counter()
        for i, s in enumerate(sessions):
            last_fr[i] = s["walk"].next()
            req_fr[srv.submit(sid[i], last_fr[i])] = (i, last_fr[i])
        srv.drain()

        def collect(frames):
            # the frustum of the newest rendered frame of every session
            for d in frames:
                k, fr = req_fr.pop(d["request"])
                if d["status"] == 0:
                    shown[k] = fr
            return frames

        collect(srv.poll())
        log("warm-up: 32 frames in %.2f s, %d bricks generated"
            % (time.perf_counter() - t0, srv.stats().bricks_generated))
        for rate in rates:
            arrivals = workloads.poisson_arrivals(rate, duration, len(sessions), seed=rate)
            before = srv.stats()
            done = []
            t0 = srv.now_ms()
            i = 0
            while i < len(arrivals):
                if srv.now_ms() - t0 >= arrivals[i][0]:
                    k = arrivals[i][1]
                    last_fr[k] = sessions[k]["walk"].next()
                    req_fr[srv.submit(sid[k], last_fr[k], t0 + arrivals[i][0])] = \
                        (k, last_fr[k])
                    i += 1
                    continue
                if srv.step()[1]:
                    done += collect(srv.poll())
            srv.drain()
            done += collect(srv.poll())
            t_end = max([d["done_ms"] for d in done] + [srv.now_ms()])
            after = srv.stats()
            ok = [d for d in done if d["status"] == 0]
            lat = np.array([d["done_ms"] - d["arrive_ms"] for d in ok])
            wall = (t_end - t0) / 1e3
            samples = sum(d["samples"] for d in ok)
            row = {"rate_req_per_s": rate, "requests": len(arrivals), "rendered": len(ok),
                   "superseded": sum(1 for d in done if d["status"] == 1),
                   "failed": sum(1 for d in done if d["status"] < 0),
                   "failure_codes": sorted({d["status"] for d in done if d["status"] < 0}),
                   "fps": len(ok) / wall, "gsamples_per_s": samples / wall / 1e9,
                   "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
                   "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
                   "max_ms": float(lat.max()) if len(lat) else None,
                   "batches": after.batches - before.batches,
                   "host_ms": {k: round(getattr(after, k) - getattr(before, k), 1) for k in (
                       "select_ms", "order_ms", "residency_ms", "add_ms", "run_ms",
                       "retire_ms", "gpu_ms")},
                   "bricks_generated": after.bricks_generated - before.bricks_generated}
            log("rate %d/s: %d rendered, %d superseded, %.1f fps, %.1f Gsamples/s, p50 %.1f ms, "
                "p99 %.1f ms, host %s" % (rate, row["rendered"], row["superseded"], row["fps"],
                                          row["gsamples_per_s"], row["p50_ms"] or -1,
                                          row["p99_ms"] or -1, row["host_ms"]))
            result_rates.append(row)
        # the last frame of a few sessions equals the same frame rendered alone
        same = True
        for k in (0, 7, 15, 31):
            s = sessions[k]
            v = vols[s["volume"]]
            img, smp = srv.read(sid[k])
            fr = shown[k]
            vis = v.ctx.select_visibles(fr, s["height"], 1.0, 0, v.vi.depth - 1)
            ordered = livre_amd.order(v.vi, fr, vis)
            missing = [x for x in ordered if not v.ctx.resident(x)]
            if missing:
                if v.noise is not None:
                    v.ctx.generate_noise(v.noise, missing)
                else:
                    v.ctx.generate_field(v.field, missing)
            v.ctx.set_tf(s["tf"], v.range)
            v.ctx.render(fr, s["width"], s["height"], ordered,
                         livre_amd.render_params(v.vmax))
            ref, ref_smp = v.ctx.read_frame()
            same = same and smp == ref_smp and bool(np.array_equal(img, ref))
        st = srv.stats()
    for v in vols.values():
        v.close()
    top = result_rates[-1]
    result = {
        "metric": "served Gsamples/s and p50/p99 frame latency, 32 concurrent sessions",
        "value": top["gsamples_per_s"], "unit": "Gsamples/s", "n_gpus": 1,
        "higher_is_better": True, "vs_baseline": None, "dtype": "u8/u16/f32",
        "data": "synthetic: procedural volumes, bricks generated on the device on demand",
        "config": {"workload": "32 sessions, viewports {720p,1080p,1440p,4K}, volumes "
                               "256^3 u8 / 512^3 u8 / 1024^3 u16 / 1536^3 f32 / 2048^3 u16, "
                               "random-walk cameras, 256-entry TFs, 1 sample/voxel, sse 1; "
                               "Poisson arrivals, %.0f s per rate" % duration,
                   "server": "continuous batching, 4 launches in flight, max %d requests "
                             "per launch, newest request per session wins" % args.max_batch,
                   "budget_bytes": budget, "pool_bytes": pools,
                   "pool_fraction_of_pyramid": args.pool_frac},
        "rates": result_rates, "served_equals_alone": same, "admitted_sessions": st.admitted,
    }
    return result


def ooc_workload(args):
    """8192^3 uint8 (512 GB logical) from the host blobs-plus-value-noise
    DataSource, 64^3 bricks, depth 8, pool capped at --pool-gb; bricks stream through
    pinned memory on the upload stream while frames render the coarser-LOD
    rendering set; 3840x2160, 300-frame fly-through."""
    import livre_amd
    from livre_amd import scenes, workloads
    livre_amd.lib()
    W4, H4 = 3840, 2160
    vi = workloads.c4_volume()
    slot = livre_amd.slot_bytes(vi)
    n_slots = int(args.pool_gb * 1e9) // slot
    ctx = livre_amd.Context(vi, slot * n_slots)
    source = workloads.c4_source(vi)
    tf = scenes.piecewise_tf(256, seed=8, points=6, max_alpha=0.3)
    ctx.set_tf(tf, (0.0, 255.0))
    ctx.set_timing(True)
    threads = max(1, min(14, (os.cpu_count() or 2) - 2))
    rows = []
    with livre_amd.Stream(ctx, source, threads=threads, staging_bricks=256) as st:
        t0 = time.perf_counter()
        # levels 0..3 (585 bricks) before the first frame: every region has a
        # fallback from then on
        lv = []
        for level in range(4):
            n = 1 << level
            lv += [scenes.node_id(level, x, y, z) for z in range(n) for y in range(n)
                   for x in range(n)]
        st.request(lv)
        st.wait()
        log("preloaded %d coarse bricks in %.2f s (%d source threads)"
            % (len(lv), time.perf_counter() - t0, threads))
        start = st.stats()
        host_ms = []

        def camera4k(t):
            pos, ahead = workloads.flythrough(t, args.frames)
            mv = scenes.look_at(tuple(pos), tuple(ahead), (0.0, 1.0, 0.0))
            proj = scenes.perspective(60.0, W4 / H4, 0.01, 6.0)
            return livre_amd.frustum(mv, proj)

        def select(t):
            """Queue frame t's selection; it runs between two ray casts."""
            fr = camera4k(t)
            ctx.select_visibles_async(fr, H4, 1.0, 0, vi.depth - 1)
            return time.perf_counter(), fr

        def prepare(t, pending):
            """Frame t's visible set (queued one frame earlier), stream
            requests, rendering set and order; its bricks are pinned at once,
            so uploads for later frames cannot evict them before it renders.
            Frame t+1's selection is queued."""
            f0, fr = pending
            h0 = time.perf_counter()
            vis = ctx.select_visibles_wait()
            nxt = select(t + 1) if t + 1 < args.frames else None
            h1 = time.perf_counter()
            # wishes in selection (depth-first, spatially coherent) order, with
            # their ancestors; the stream loads coarse levels first
            st.request(vis, ancestors=True)
            h2 = time.perf_counter()
            st.pump()
            rset, avail = st.rendering_set(vis)
            h3 = time.perf_counter()
            ordered = livre_amd.order(vi, fr, rset)
            host_ms.append(((h1 - h0) * 1e3, (h2 - h1) * 1e3, (h3 - h2) * 1e3))
            ctx.pin(ordered, 1)
            # samples per ray from the finest rendered level, as
            # RayCastRenderer::onFrameStart does when none is given
            # (RayCastRenderer.cpp:216-234)
            max_lod = max(int(i) & 0xF for i in ordered) if len(ordered) else 0
            n_samples = max(8192 >> (vi.depth - max_lod - 1), 512)
            return (f0, fr, ordered, len(vis), avail, len(rset), max_lod, n_samples), nxt

        # pipelined as RenderPipeline::renderAsync: frame t+1 is prepared on
        # the host while frame t ray-casts; transfers are started throughout
        t_start = time.perf_counter()
        cur, pending = prepare(0, select(0))
        ctx.render(cur[1], W4, H4, cur[2], livre_amd.render_params(cur[7]))
        for t in range(args.frames):
            nxt = None
            if pending is not None:
                nxt, pending = prepare(t + 1, pending)
            while not ctx.query():  # transfers overlap the ray casting
                st.pump()
                time.sleep(0.0002)
            fs = ctx.frame_stats()
            f0, fr, ordered, n_vis, avail, n_rset, max_lod, _ = cur
            ctx.pin(ordered, -1)
            if nxt is not None:
                ctx.render(nxt[1], W4, H4, nxt[2], livre_amd.render_params(nxt[7]))
            rows.append((time.perf_counter() - f0, fs.raycast_ms, fs.samples, n_vis, avail,
                         n_rset, fs.bricks, max_lod))
            if t % 30 == 0:
                s = st.stats()
                log("frame %d: visible %d, at full LOD %d, rendered %d (finest level %d), "
                    "raycast %.1f ms, frame %.1f ms, loaded %d"
                    % (t, n_vis, avail, n_rset, max_lod, fs.raycast_ms,
                       rows[-1][0] * 1e3, s.bricks_loaded))
            cur = nxt
        total = time.perf_counter() - t_start
        end = st.stats()
    pool_used = ctx.stats()[1]
    ctx.close()
    a = np.array(rows, dtype=np.float64)
    loaded = end.bricks_loaded - start.bricks_loaded
    result = {
        "metric": "out-of-core fly-through: fps, Gsamples/s, streamed GB/s, full-LOD fraction",
        "value": float(a[:, 2].sum() / total / 1e9), "unit": "Gsamples/s", "n_gpus": 1,
        "higher_is_better": True, "vs_baseline": None, "dtype": "u8",
        "data": "synthetic: 8192^3 u8 blobs plus value noise (64 blobs, seed 8) from the "
                "host DataSource",
        "config": {"workload": "8192^3 uint8 (512 GB logical), 64^3 bricks + 1 halo, depth 8, "
                               "pool %.0f GB (%d slots) with LRU, 3840x2160, %d-frame "
                               "fly-through, samples/unit from the finest rendered level "
                               "(RayCastRenderer rule), sse 1"
                               % (args.pool_gb, n_slots, args.frames),
                   "streaming": "%d host source threads, 256 pinned staging bricks, "
                                "hipMemcpyAsync + tiling on the upload stream" % threads},
        "frames": len(rows), "fps": len(rows) / total,
        "p50_ms": float(np.percentile(a[:, 0], 50) * 1e3),
        "p99_ms": float(np.percentile(a[:, 0], 99) * 1e3),
        "raycast_ms_mean": float(a[:, 1].mean()),
        "visible_mean": float(a[:, 3].mean()),
        "full_lod_fraction_mean": float(np.mean(a[:, 4] / np.maximum(a[:, 3], 1))),
        "full_lod_fraction_last": float(a[-1, 4] / max(a[-1, 3], 1)),
        "finest_rendered_level_mean": float(a[:, 7].mean()),
        "host_ms_mean": dict(zip(("select", "order_and_request", "pump_and_rendering_set"),
                                 np.mean(np.array(host_ms), axis=0).round(2).tolist())),
        "bricks_streamed": int(loaded),
        "streamed_gb_per_s": loaded * vi.max_block[0] ** 3 / total / 1e9,
        "pool_bricks_resident": int(pool_used),
    }
    print(json.dumps(result), flush=True)


def raw_workload(args):
    """A raw file on disk (1024^3 uint16, written first from the value-noise
    field) opened as RawDataSource does, its LOD pyramid built in HBM (64^3
    bricks + 1 halo, depth 5) and every brick cut into the pool; then the
    1080p orbit-zoom rendered from it."""
    import tempfile
    import livre_amd
    from livre_amd import scenes
    livre_amd.lib()
    n = args.raw_size
    tmpdir = tempfile.mkdtemp(prefix="livre_raw_")
    path = os.path.join(tmpdir, "volume.raw")
    # the file: the value-noise field at n^3, generated on the device one
    # z-slab of bricks at a time and written slab by slab
    vi_src = livre_amd.volume_info(n, 64, 0, livre_amd.DT_UINT16)
    noise = livre_amd.noise_params(3, octaves=5, base_cells=8, gain=1.5)
    nb = n // 64
    t0 = time.perf_counter()
    with livre_amd.Context(vi_src, livre_amd.slot_bytes(vi_src) * nb * nb) as gen, \
            open(path, "wb") as f:
        for z in range(nb):
            slab_ids = [scenes.node_id(vi_src.depth - 1, x, y, z)
                        for y in range(nb) for x in range(nb)]
            for nid in slab_ids:
                if gen.resident(nid):
                    gen.evict(nid)
            gen.generate_noise(noise, slab_ids)
            slab = np.empty((64, n, n), np.uint16)
            for nid in slab_ids:
                _lv, x, y, _z = scenes.decode_node(nid)
                slab[:, y * 64:(y + 1) * 64, x * 64:(x + 1) * 64] = gen.download(nid)
            slab.tofile(f)
            for nid in slab_ids:
                gen.evict(nid)
    t_write = time.perf_counter() - t0
    log("wrote %s (%.1f GB) in %.1f s" % (path, n ** 3 * 2 / 1e9, t_write))
    try:
        with livre_amd.RawSource("raw://%s#%d,%d,%d,uint16" % (path, n, n, n)) as src:
            vi = src.pyramid_info(64, 1)
            ids = scenes.all_nodes(vi)
            ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
            # the first load builds the levels: file -> HBM, downsampling,
            # bricks cut into slots
            t0 = time.perf_counter()
            ctx.load_pyramid(src, ids)
            ctx.sync()
            t_build = time.perf_counter() - t0
            # again with the levels resident: the cut alone
            t0 = time.perf_counter()
            ctx.load_pyramid(src, ids)
            ctx.sync()
            t_cut = time.perf_counter() - t0
            ctx.release_pyramid()
            brick_bytes = len(ids) * vi.max_block[0] ** 3 * 2
            log("pyramid: %d bricks, build %.2f s, cut %.3f s" % (len(ids), t_build, t_cut))
            tf = scenes.piecewise_tf(TF_ENTRIES, seed=4, points=8, max_alpha=0.35)
            ctx.set_tf(tf, (0.0, 65535.0))
            ctx.set_timing(True)
            params = livre_amd.render_params(n, early_exit=0.999)
            samples, rc_ms, lat = 0, [], []
            frames = args.steps or 60
            t0 = time.perf_counter()
            for t in range(frames):
                f1 = time.perf_counter()
                fr = camera(livre_amd, scenes, t * FRAMES // frames)[2]
                vis = ctx.select_visibles(fr, H, 1.0, 0, vi.depth - 1)
                ctx.render(fr, W, H, livre_amd.order(vi, fr, vis), params)
                st = ctx.frame_stats()
                samples += st.samples
                rc_ms.append(st.raycast_ms)
                lat.append((time.perf_counter() - f1) * 1e3)
            t_render = time.perf_counter() - t0
            ctx.close()
    finally:
        os.remove(path)
        os.rmdir(tmpdir)
    result = {
        "metric": "raw volume: LOD pyramid build GB/s; render Gsamples/s",
        "value": n ** 3 * 2 / t_build / 1e9, "unit": "GB/s (file bytes into a resident "
                                                     "pyramid)",
        "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "u16",
        "data": "synthetic: %d^3 uint16 value noise written to a raw file first" % n,
        "config": {"workload": "raw://volume.raw#%d,%d,%d,uint16 -> pyramid of 64^3 "
                               "bricks + 1 halo, depth %d, %d bricks; 1920x1080 "
                               "orbit-zoom, %d frames" % (n, n, n, vi.depth, len(ids),
                                                          frames)},
        "file_write_s": t_write,
        "build_s": t_build, "cut_s": t_cut,
        "cut_gb_per_s": brick_bytes / t_cut / 1e9,
        "bricks": len(ids),
        "render_gsamples_per_s": samples / t_render / 1e9,
        "render_fps": frames / t_render,
        "raycast_ms_mean": float(np.mean(rc_ms)),
        "frame_p50_ms": float(np.percentile(lat, 50)),
    }
    print(json.dumps(result), flush=True)


def c1_workload(args):
    """The CPU-runnable configuration on the GPU: 256^3 uint8 blobs (64
    Gaussians + 10% noise, seed 1), 32^3 bricks + 1 halo, 4 levels, a
    256-entry uint8 LUT (seed 2), 1024x1024, 8 cameras on a radius-2 orbit,
    256 samples per unit, early exit 0.99, sse 1, 256-bin histogram. Each
    frame: selection + order + ray cast + histogram, one frame at a time."""
    import livre_amd
    from livre_amd import scenes
    livre_amd.lib()
    w = h = 1024
    vi = livre_amd.volume_info(256, 32, 1, livre_amd.DT_UINT8)
    ids = scenes.all_nodes(vi)
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
    ctx.generate_field(livre_amd.field_params(livre_amd.FIELD_BLOBS, 64, 1, 0.1), ids)
    tf = scenes.piecewise_tf(256, seed=2, points=6, max_alpha=0.6)
    lut = scenes.lut_rgba8(tf)
    ctx.set_tf_rgba8(lut, (0.0, 255.0))
    ctx.set_timing(True)
    # the launch shape the shipped tuning table holds for this class
    from livre_amd import autotune
    tuner = autotune.Tuner()
    params = tuner.cached(vi, w, h, livre_amd.render_params(256, early_exit=0.99))
    cams = []
    for k in range(8):
        mv, proj = scenes.camera(scenes.orbit(k, 8, radius=2.0), w, h, 60.0)
        cams.append((mv, proj, livre_amd.frustum(mv, proj)))

    def one(k):
        fr = cams[k % 8][2]
        vis = ctx.select_visibles(fr, h, 1.0, 0, vi.depth - 1)
        ordered = livre_amd.order(vi, fr, vis)
        ctx.render(fr, w, h, ordered, params)
        ctx.histogram(ordered, 256, read=True)
        return ordered, ctx.frame_stats()

    for k in range(args.warmup):
        one(k)
    frames = args.steps or 80
    samples, rc_ms, lat, alg = 0, [], [], []
    t0 = time.perf_counter()
    for k in range(frames):
        f1 = time.perf_counter()
        ordered, st = one(k)
        lat.append((time.perf_counter() - f1) * 1e3)
        samples += st.samples
        rc_ms.append(st.raycast_ms)
        alg.append(st.bricks * 34 ** 3 + w * h * 4 + 256 * 16)
    total = time.perf_counter() - t0
    rc = np.array(rc_ms)
    achieved = float(np.mean(np.array(alg) / (rc * 1e-3))) / 1e9
    # the reference-parity variant (early exit 0.999) over the 8 cameras
    params = tuner.cached(vi, w, h, livre_amd.render_params(256, early_exit=0.999))
    one(0)
    p_samples, p_rc, t1 = 0, [], time.perf_counter()
    for k in range(8):
        _o, st = one(k)
        p_samples += st.samples
        p_rc.append(st.raycast_ms)
    p_total = time.perf_counter() - t1
    result = {
        "metric": "C1: Gsamples/s & fps, 256^3 u8 at 1024x1024, 8-camera orbit",
        "value": samples / total / 1e9, "unit": "Gsamples/s", "n_gpus": 1, "steps": frames,
        "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None, "dtype": "u8",
        "data": "synthetic: 64 seeded Gaussian blobs + 10% noise, generated on the device",
        "config": {"workload": "256^3 uint8, 32^3 bricks + 1 halo, 4 levels, 256-entry uint8 "
                               "LUT, 1024x1024, 8 orbit cameras (radius 2), 256 samples/unit, "
                               "exit 0.99, sse 1, 256-bin histogram; one frame at a time "
                               "(select + order + ray cast + histogram, synchronised)",
                   "launch": "tile %dx%d, %d fetches in flight, classify %d, %d waves/SIMD "
                             "(livre_amd/tune_gfx950.json)"
                             % (params.tile_w, params.tile_h, params.samples_per_thread,
                                params.classify, params.waves_per_simd)},
        "fps": frames / total, "frame_p50_ms": float(np.percentile(lat, 50)),
        "frame_p99_ms": float(np.percentile(lat, 99)), "raycast_ms_mean": float(rc.mean()),
        "kernel_gsamples_per_s": samples / (rc.sum() * 1e-3) / 1e9,
        "exit_0999": {"gsamples_per_s": p_samples / p_total / 1e9, "fps": 8 / p_total,
                      "raycast_ms_mean": float(np.mean(p_rc)),
                      "samples_per_frame": p_samples / 8,
                      "note": "the reference-parity run (early exit 0.999), the 8 cameras "
                              "once each after one warm-up frame"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / (HBM_PEAK / 1e9), "traffic": None,
                     "kernel": "raycastKernel<unsigned char> (classified table)",
                     "bytes_per_launch": "rendered bricks x 39304 B + 1024x1024x4 + 256x16",
                     "note": "the 23 MB pool fits in the 256 MB Infinity Cache: HBM does "
                             "not bind here"},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        import oracle
        ovi = oracle.volume_info([256] * 3, [32] * 3, [1] * 3, oracle.DT_UINT8)
        mv, proj, fr = cams[0]
        cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                            fr.far_plane, w, h)
        t1 = time.perf_counter()
        vis = oracle.select_visibles(ovi, mv, proj, h, 1.0, 0, ovi.depth - 1)
        ordered = oracle.order(ovi, mv, vis)
        bricks = [ctx.download(i) for i in ordered]
        t_sel = time.perf_counter() - t1
        threads = os.cpu_count() or 1
        t1 = time.perf_counter()
        _img, smp = oracle.render(ovi, cam, ordered, bricks,
                                  lut.astype(np.float32) / np.float32(255.0),
                                  oracle.render_params(256, 0.99, (0.0, 255.0)), "f32", threads)
        t_march = time.perf_counter() - t1
        # warm caches: a frame sums the cached per-brick histograms
        # (HistogramFilter.cpp:78-149); each brick is binned once, on load
        per = [oracle.histogram(ovi, [i], [b], 256) for i, b in zip(ordered, bricks)]
        t1 = time.perf_counter()
        acc = np.zeros(256, np.uint64)
        for hb in per:
            acc += hb
        t_hist = time.perf_counter() - t1
        result["cpu_baseline"] = {
            "value": smp / (t_sel + t_march + t_hist) / 1e9, "unit": "Gsamples/s",
            "cores": threads, "kind": "port",
            "sample": "camera 0: oracle selection + order (%.2f s, brick reads included), FP32 "
                      "march of the full 1024x1024 frame (%.2f s, %d samples) and the 256-bin "
                      "histogram as the sum of the cached per-brick ones (%.4f s)"
                      % (t_sel, t_march, smp, t_hist)}
    ctx.close()
    print(json.dumps(result), flush=True)


def sortlast_workload(args):
    """One GPU's part of sort-last rendering at the headline shape: the
    bricks of each 1080p orbit-zoom frame of the 2048^3 volume are split by
    the static decomposition into the 8 octants of the block grid (the 8
    GPUs' shares: a brick's owner never changes with the camera), each part
    ray-cast into an RGBA32F layer, and the 8 layers composited in the
    octants' front-to-back order for the eye into RGBA8 (the kernel every GPU
    runs on its strip after the exchange). The exchange itself (an
    all-to-all over xGMI) is not run on one GPU."""
    import livre_amd
    from livre_amd import scenes
    livre_amd.lib()
    vi, ctx, _ids, _tf = build_scene(livre_amd, scenes)
    params = livre_amd.render_params(N_SAMPLES, early_exit=0.999)
    ctx.set_timing(True)
    n_layers = 8
    frames = args.steps or 30
    comp_ms, part_max, part_sum, whole_ms, diffs = [], [], [], [], []
    owner, sizes = {}, []
    for t in range(args.warmup + frames):
        fr = camera(livre_amd, scenes, t * FRAMES // frames)[2]
        vis = ctx.select_visibles(fr, H, 1.0, 0, vi.depth - 1)
        ordered = np.asarray(livre_amd.order(vi, fr, vis))
        ctx.render(fr, W, H, ordered, params)
        whole_st = ctx.frame_stats()
        whole, _ = ctx.read_frame()
        part_of = livre_amd.sortlast_parts(vi, ordered, n_layers)
        for nid, p in zip(ordered, part_of):
            if owner.setdefault(int(nid), int(p)) != p:
                raise RuntimeError("a brick changed owner")
        order = livre_amd.sortlast_order(vi, list(fr.eye), n_layers)
        rc = []
        for i, p in enumerate(order):
            ctx.render_partial(fr, W, H, ordered[part_of == p], params, i)
            rc.append(ctx.frame_stats().raycast_ms)
        sizes.append([int((part_of == p).sum()) for p in range(n_layers)])
        ms = ctx.composite([ctx.layer_ptr(i) for i in range(n_layers)], W, H, timed=True,
                           reps=10)
        img, _ = ctx.read_frame()
        if t < args.warmup:
            continue
        d = np.abs(img.astype(np.int32) - whole.astype(np.int32))
        diffs.append((int(d.max()), float(d.mean())))
        comp_ms.append(ms)
        part_max.append(max(rc))
        part_sum.append(sum(rc))
        whole_ms.append(whole_st.raycast_ms)
    comp = float(np.mean(comp_ms))
    bytes_per = n_layers * W * H * 16 + W * H * 4
    achieved = bytes_per / (comp * 1e-3) / 1e9
    result = {
        "metric": "sort-last: 8-layer composite GB/s at 1920x1080",
        "value": achieved, "unit": "GB/s", "n_gpus": 1, "steps": frames,
        "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: frames of the 2048^3 uint16 value-noise orbit-zoom",
        "config": {"workload": "each frame's ordered bricks split into the 8 octants of the "
                               "block grid (static decomposition: owners fixed across the "
                               "orbit), each part ray-cast into an RGBA32F layer, the 8 "
                               "layers composited in the octants' front-to-back order for "
                               "the eye (1920x1080)"},
        "bricks_per_part_mean": np.mean(np.array(sizes), axis=0).round(1).tolist(),
        "distinct_bricks_owned": len(owner),
        "composite_ms_mean": comp,
        "composite_timing": "HIP events around 10 back-to-back launches per frame",
        "partial_raycast_ms_max_mean": float(np.mean(part_max)),
        "partial_raycast_ms_sum_mean": float(np.mean(part_sum)),
        "whole_raycast_ms_mean": float(np.mean(whole_ms)),
        "vs_whole_frame_max_diff": max(x[0] for x in diffs),
        "vs_whole_frame_mean_diff": float(np.mean([x[1] for x in diffs])),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / (HBM_PEAK / 1e9),
                     "traffic": profile_value("composite_8layer_1080p_hbm.txt",
                                              "traffic_bytes_per_launch"),
                     "traffic_source": "profiles/composite_8layer_1080p_hbm.txt (rocprofv3 "
                                       "FETCH_SIZE and WRITE_SIZE per dispatch, each scaled "
                                       "by the same counter on a 1 GiB device copy)",
                     "kernel": "compositeKernel",
                     "bytes_per_launch": "8 x 1920x1080x16 + 1920x1080x4"},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # the over of the same 8 layers on the host (numpy, float32)
        rng = np.random.default_rng(0)
        lay = [(rng.random((H, W, 4), np.float32) * 0.1) for _ in range(n_layers)]
        t0, n = time.perf_counter(), 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 5.0) or n == 0:
            acc = np.zeros((H, W, 4), np.float32)
            for l in lay:
                acc += (np.float32(1) - acc[..., 3:4]) * l
            out = np.rint(np.clip(acc, 0, 1) * 255).astype(np.uint8)
            n += 1
        el = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": bytes_per * n / el / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "numpy float32 over of 8 1920x1080 RGBA32F layers into RGBA8, "
                      "%d composites in %.1f s" % (n, el)}
    ctx.close()
    print(json.dumps(result), flush=True)


def jpeg_workload(args):
    """Frame grab (FrameGrabber::_encodeJpeg: JPEG quality 100, 4:4:4, rows
    bottom-up) of the 2048^3 orbit-zoom frames at 1920x1080 and 3840x2160:
    each frame is rendered, then grabbed into host memory. A grab = encode on
    the device + copy of the file to the host."""
    import livre_amd
    from livre_amd import scenes
    livre_amd.lib()
    vi, ctx, _ids, _tf = build_scene(livre_amd, scenes)
    params = livre_amd.render_params(N_SAMPLES, early_exit=0.999)
    frames = args.steps or 60
    res, frames_host = {}, {}
    for (w, h) in [(W, H), (3840, 2160)]:
        dev, stage, e2e, size = [], [], [], []
        for t in range(args.warmup + frames):
            eye = scenes.orbit_zoom((t * FRAMES // frames) % FRAMES, FRAMES)
            mv, proj = scenes.camera(eye, w, h, 60.0)
            fr = livre_amd.frustum(mv, proj)
            vis = ctx.select_visibles(fr, h, 1.0, 0, vi.depth - 1)
            ctx.render(fr, w, h, livre_amd.order(vi, fr, vis), params)
            ctx.sync()
            t0 = time.perf_counter()
            jpg = ctx.grab_jpeg(100)
            t1 = time.perf_counter()
            if t < args.warmup:
                continue
            st = ctx.grab_stats()
            dev.append(st.device_ms)
            stage.append((st.dct_ms, st.offsets_ms, st.emit_ms, st.stuff_ms))
            e2e.append((t1 - t0) * 1e3)
            size.append(len(jpg))
            if t == args.warmup:
                frames_host[(w, h)] = ctx.read_frame()[0]
        sm = np.mean(np.array(stage), 0)
        res["%dx%d" % (w, h)] = {
            "grab_ms_mean": float(np.mean(e2e)), "grab_ms_p50": float(np.percentile(e2e, 50)),
            "grab_ms_p99": float(np.percentile(e2e, 99)),
            "device_ms_mean": float(np.mean(dev)),
            "stage_ms_mean": {"dct": float(sm[0]), "offsets": float(sm[1]),
                              "emit": float(sm[2]), "stuff": float(sm[3])},
            "jpeg_bytes_mean": float(np.mean(size)),
            "grabs_per_s": 1e3 / float(np.mean(e2e)),
            "device_grabs_per_s": 1e3 / float(np.mean(dev)),
            "input_gb_per_s": w * h * 4 / (float(np.mean(dev)) * 1e-3) / 1e9}
        log("%dx%d: grab %.2f ms (device %.3f ms: dct %.3f, offsets %.3f, emit %.3f, "
            "stuff %.3f), %.2f MB per file" % (w, h, np.mean(e2e), np.mean(dev), sm[0], sm[1],
                                              sm[2], sm[3], np.mean(size) / 1e6))
    ctx.close()
    hd = res["%dx%d" % (W, H)]
    # the dominant kernel, the Huffman emit stage: it reads the quantised
    # coefficients (2 B each) and each block's bit offset, and writes the
    # entropy-coded scan (before byte stuffing, about the file's size)
    blocks = 3 * ((W + 7) // 8) * ((H + 7) // 8)
    emit_bytes = blocks * (64 * 2 + 4) + hd["jpeg_bytes_mean"]
    achieved = emit_bytes / (hd["stage_ms_mean"]["emit"] * 1e-3) / 1e9
    result = {
        "metric": "frame grab: JPEG frames/s at 1920x1080 (quality 100, 4:4:4)",
        "value": hd["grabs_per_s"], "unit": "frames/s", "n_gpus": 1, "steps": frames,
        "warmup": args.warmup, "higher_is_better": True, "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: frames of the 2048^3 uint16 value-noise orbit-zoom",
        "config": {"workload": "render then grab each of %d frames of the orbit-zoom at "
                               "1920x1080 and at 3840x2160; grab = device encode + file to "
                               "host memory" % frames},
        "sizes": res,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / (HBM_PEAK / 1e9), "traffic": None,
                     "kernel": "jpegEmitKernel",
                     "bytes_per_launch": "97200 blocks x (64 x 2 + 4) + the scan's bytes "
                                         "(~0.6 MB)"},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        import oracle
        px = frames_host[(W, H)]
        t0, n = time.perf_counter(), 0
        while time.perf_counter() - t0 < args.cpu_seconds or n == 0:
            oracle.jpeg_encode(px, quality=100)
            n += 1
        el = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": n / el, "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": "oracle JPEG restatement (scalar C++, one core) on the first timed "
                      "1920x1080 frame, %d encodes in %.1f s" % (n, el)}
    print(json.dumps(result), flush=True)


def hbmcopy_workload(args):
    """A plain device-to-device copy of 1 GiB, four times: the reference for
    the HBM traffic counters (FETCH_SIZE / WRITE_SIZE of a dispatch that reads
    and writes exactly 1 GiB) and the copy bandwidth the part reaches."""
    import torch
    n = 1 << 30
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": 4 * 2 * n / el / 1e9,
                      "unit": "GB/s", "n_gpus": 1, "bytes_per_copy": n, "copies": 4,
                      "higher_is_better": True}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra-legs", action="store_true",
                    help="render leg only: no 1080p batching ratio, no serving leg")
    ap.add_argument("--workload", choices=("render", "batch", "serve", "ooc", "raw", "jpeg",
                                           "c1", "sortlast", "hbmcopy"),
                    default="render",
                    help="render: the 1080p single-volume frame loop (default); "
                         "batch: uneven requests, plain vs one batched launch; "
                         "serve: 32 sessions under Poisson arrivals; "
                         "ooc: out-of-core 8192^3 fly-through; "
                         "raw: a raw file's LOD pyramid built in HBM, then rendered; "
                         "jpeg: frame grab of rendered frames as JPEG; "
                         "c1: the CPU-runnable 256^3 configuration at 1024x1024; "
                         "sortlast: 8 partial layers per frame and their composite; "
                         "hbmcopy: a 1 GiB device copy (HBM counter reference)")
    ap.add_argument("--rates", type=int, nargs="+", default=[50, 100, 200, 400])
    ap.add_argument("--duration", type=float, default=3.0)
    ap.add_argument("--max-batch", type=int, default=1,
                    help="requests per serving launch; one request per launch runs the "
                         "single-frame kernel (two-brick window)")
    ap.add_argument("--pool-frac", type=float, default=1.0)
    ap.add_argument("--solo-tiles", type=int, default=0,
                    help="batch workload: requests of at least this many ray tiles run "
                         "single-frame launches of their own (0: the library's default, "
                         "negative: only a lone request)")
    ap.add_argument("--linear", action="store_true",
                    help="render workload with GL_LINEAR (trilinear) sampling instead of "
                         "the GL_NEAREST parity filter; no serving or batching legs, no "
                         "CPU baseline")
    ap.add_argument("--pool-gb", type=float, default=200.0)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--raw-size", type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the brick pool, kernels and measurements are "
              "single-device; of multi-GPU sort-last rendering only one GPU's part is "
              "built (--workload sortlast: partial layers and their composite), not the "
              "exchange between GPUs", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        if args.steps is None:
            args.steps = 4
        batch_workload(args)
        return
    if args.workload == "serve":
        serve_workload(args)
        return
    if args.workload == "ooc":
        ooc_workload(args)
        return
    if args.workload == "raw":
        raw_workload(args)
        return
    if args.workload == "jpeg":
        jpeg_workload(args)
        return
    if args.workload == "c1":
        c1_workload(args)
        return
    if args.workload == "sortlast":
        sortlast_workload(args)
        return
    if args.workload == "hbmcopy":
        hbmcopy_workload(args)
        return
    if args.steps is None:
        args.steps = FRAMES

    import livre_amd
    from livre_amd import scenes
    livre_amd.lib()
    vi, ctx, _ids, tf = build_scene(livre_amd, scenes)
    params = livre_amd.render_params(N_SAMPLES, early_exit=0.999, linear=args.linear)
    if args.linear:
        args.no_extra_legs = True
        args.no_cpu_baseline = True
    ctx.set_timing(True)

    frames = schedule(args.steps)
    for i in range(args.warmup):
        frame(livre_amd, vi, ctx, camera(livre_amd, scenes, frames[i % len(frames)])[2],
              params)
    ctx.sync()

    # Frames are pipelined as RenderPipeline::renderAsync does: while frame t
    # ray-casts, frame t+1 (whose selection was queued one frame earlier, so
    # it ran between two ray casts) is ordered, prepared and queued behind it,
    # and frame t+2's selection is queued; then frame t's statistics are read
    # (the context keeps the last four frames'). A frame's latency runs from
    # the queueing of its selection to the moment its completion is observed.
    lat, samples, bricks, entries, rc_ms, bin_ms, hist_ms = [], 0, [], [], [], [], []
    host_ms = []
    alg_bytes = []
    last = args.steps - 1

    def select(i):
        fr = camera(livre_amd, scenes, frames[i])[2]
        ctx.select_visibles_async(fr, H, 1.0, 0, vi.depth - 1)
        return time.perf_counter(), fr

    def submit(pending, i):
        f0, fr = pending
        vis = ctx.select_visibles_wait()
        nxt = select(i + 1) if i < last else None
        ordered = livre_amd.order(vi, fr, vis)
        ctx.render(fr, W, H, ordered, params)
        ctx.histogram(ordered, BINS, read=False)
        return f0, nxt

    t_start = time.perf_counter()
    f0, pending = submit(select(0), 0)
    starts = [f0]
    for i in range(args.steps):
        t = frames[i]
        if i < last:
            f0, pending = submit(pending, i + 1)
            starts.append(f0)
            st = ctx.frame_stats(back=1)  # waits for step i only
        else:
            st = ctx.frame_stats()
        lat.append((time.perf_counter() - starts[i]) * 1e3)
        samples += st.samples
        bricks.append(st.bricks)
        entries.append(st.tile_entries)
        rc_ms.append(st.raycast_ms)
        bin_ms.append(st.bin_ms)
        hist_ms.append(st.histogram_ms)
        host_ms.append(st.host_ms)
        alg_bytes.append(st.bricks * st.brick_bytes + W * H * 4 + TF_ENTRIES * 16)
        if i % 20 == 0:
            log("frame %d: %d bricks, %.2f Gsamples, raycast %.2f ms, frame %.2f ms"
                % (t, st.bricks, st.samples / 1e9, st.raycast_ms, lat[-1]))
    ctx.sync()
    total = time.perf_counter() - t_start

    # the same frames once more, one at a time (nothing else in flight): the
    # latency of a frame on its own
    sync_lat = []
    for i in range(args.steps):
        f1 = time.perf_counter()
        frame(livre_amd, vi, ctx, camera(livre_amd, scenes, frames[i])[2], params)
        sync_lat.append((time.perf_counter() - f1) * 1e3)

    rc = np.array(rc_ms)
    # algorithmic bytes of a launch over its duration, averaged over launches
    achieved = float(np.mean(np.array(alg_bytes) / (rc * 1e-3))) / 1e9
    # HBM bytes per algorithmic byte, measured over the same sampled path
    ratio = profile_value(HBM_PROFILE, "traffic_per_algorithmic_byte")
    traffic = ratio * float(np.mean(alg_bytes)) if ratio is not None else None
    # samples per second if every algorithmic byte moved at the HBM peak
    ceiling = samples / (float(np.sum(alg_bytes)) / HBM_PEAK) / 1e9
    result = {
        "metric": METRIC,
        "value": samples / total / 1e9,
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic: procedural 5-octave value noise (seed 3), generated on the "
                "device into the brick pool before timing",
        "config": {"workload": "2048^3 uint16, 64^3 bricks + 1 halo, depth 6, all 37449 "
                               "bricks resident; 1920x1080, %d frames sampled evenly from "
                               "the 120-frame orbit-zoom (frames %s), 2048 samples/unit, "
                               "FP32 4096-entry TF, sse 1, 65536-bin histogram%s"
                               % (args.steps, "0..119" if args.steps >= FRAMES else
                                  ", ".join(map(str, frames[:3])) +
                                  (", ..., %d" % frames[-1] if len(frames) > 3 else ""),
                                  "; GL_LINEAR (trilinear) sampling" if args.linear else ""),
                   "frame": "select (device) + order + bin + raycast + histogram; the "
                            "next frame is ordered and queued while the current one "
                            "ray-casts, its selection queued one frame ahead"},
        "fps": args.steps / total,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "latency": "p50/p99_ms: pipelined, selection queued to frame end; "
                   "sync_p50/p99_ms: the frame alone",
        "sync_p50_ms": float(np.percentile(sync_lat, 50)),
        "sync_p99_ms": float(np.percentile(sync_lat, 99)),
        "raycast_ms_mean": float(rc.mean()),
        "bin_ms_mean": float(np.mean(bin_ms)),
        "histogram_ms_mean": float(np.mean(hist_ms)),
        "render_call_host_ms_mean": float(np.mean(host_ms)),
        "kernel_gsamples_per_s": samples / (rc.sum() * 1e-3) / 1e9,
        "bandwidth_ceiling_gsamples_per_s": ceiling,
        # the ray cast's binding ceiling: VALU issue at the measured clock,
        # from the committed counter summary (same kernel, same sampled path)
        "valu_ceiling": valu_ceiling_entry(samples / (rc.sum() * 1e-3) / 1e9),
        # HBM read + write bytes per second of the ray cast: the measured
        # traffic per launch over the mean launch time
        "hbm_gb_per_s": traffic / (float(rc.mean()) * 1e-3) / 1e9 if traffic else None,
        "l2_hit_rate": profile_value(COUNTER_PROFILE, "l2_hit_rate"),
        "bricks_mean": float(np.mean(bricks)),
        "tile_entries_mean": float(np.mean(entries)),
        # priced against HBM (the path has no matrix work); the counters say
        # what limits it instead: VALU issue, not bandwidth
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / (HBM_PEAK / 1e9),
                     "traffic": traffic,
                     "traffic_source": "mean algorithmic bytes of the timed launches x "
                                       "HBM bytes per algorithmic byte from profiles/"
                                       + HBM_PROFILE + " (rocprofv3 FETCH_SIZE and "
                                       "WRITE_SIZE per dispatch over the same sampled path, "
                                       "each scaled by the same counter on a 1 GiB copy)",
                     "kernel": ("raycastKernel<unsigned short, LINEAR, the next sample's 8 "
                                "texels in flight>" if args.linear else
                                "raycastKernel<unsigned short, NEAREST, 2 fetches in flight>"),
                     "bytes_per_launch": "rendered bricks x 574992 B + 1920x1080x4 + "
                                         "4096x16",
                     "limiter": roofline_limiter()},
        "cpu_baseline": None,
    }
    if args.linear:
        # the committed counter and traffic summaries are of the GL_NEAREST
        # kernel; none is claimed for this variant
        result["roofline"].update(traffic=None, traffic_source=None,
                                  limiter="not profiled (the counter summaries under "
                                          "profiles/ are of the GL_NEAREST kernel)")
        result["hbm_gb_per_s"] = None
        result["l2_hit_rate"] = None
        result["valu_ceiling"] = None
    if not args.no_cpu_baseline:
        try:
            result["cpu_baseline"] = cpu_baseline(livre_amd, scenes, vi, ctx, tf,
                                                  args.cpu_seconds)
        except Exception as e:  # reported, never silently replaced
            result["cpu_baseline"] = {"error": repr(e)}
    if not args.no_extra_legs:
        result["batching_1080p"] = batching_1080p(livre_amd, scenes, vi, ctx, tf, reps=3)
    ctx.close()
    if not args.no_extra_legs:
        # the metric's serving latency: 32 sessions at 200 and 400 requests/s
        srv = serve_measure(args, (200, 400), 2.0)
        result["serving"] = {
            "config": srv["config"]["workload"] + "; " + srv["config"]["server"],
            "served_equals_alone": srv["served_equals_alone"],
            "rates": [{k: r[k] for k in ("rate_req_per_s", "requests", "rendered",
                                         "superseded", "failed", "p50_ms", "p99_ms",
                                         "max_ms", "fps", "gsamples_per_s")}
                      for r in srv["rates"]]}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
