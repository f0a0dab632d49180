"""The batched uneven-request configuration on its real volumes: 512^3 uint8
blobs, 1024^3 uint16 noise, 1536^3 float32 sines with 48^3 bricks and
2048^3 uint16 noise, all generated on the device. Eight requests cover
sampling rates 0.5, 1 and 2 samples per voxel and square viewports from
256^2 to 2048^2, each with its own seeded 256-entry TF and camera, in the
launch shapes the autotuner picks for their classes. The batched frames
equal the frames lvr_render gives for the same requests bit for bit (pixels
and composite steps); two of them (the 1536^3 float volume at 2 samples per
voxel and the 512^3 uint8 volume at 0.5) are checked against the FP32 oracle
on a band of rows (2/255 max, 0.5/255 mean) and against the FP64
transcription (lines starting FP64ERR)."""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import autotune, scenes, workloads

from scene_kit import diff_stats, fp64_error

pytestmark = pytest.mark.gpu

# (volume, viewport side, samples per voxel)
REQUESTS = [("sines1536_f32", 2048, 0.5), ("sines1536_f32", 256, 2.0),
            ("noise2048_u16", 1024, 1.0), ("blobs512_u8", 512, 2.0),
            ("noise1024_u16", 768, 0.5), ("sines1536_f32", 640, 1.0),
            ("noise2048_u16", 256, 2.0), ("blobs512_u8", 1536, 0.5)]
ORACLE_CHECKED = (1, 7)


@pytest.fixture(scope="module")
def c3(lvr, orc):
    vols = {name: workloads.Volume(name) for name in workloads.BATCH_VOLUMES}
    yield vols
    for v in vols.values():
        v.close()


def _items(vols):
    rng = np.random.default_rng(33)
    items = []
    for k, (name, side, rate) in enumerate(REQUESTS):
        vol = vols[name]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        req = dict(volume=name, width=side, height=side, eye=tuple(d * rng.uniform(1.5, 3.0)),
                   tf=scenes.piecewise_tf(256, seed=3000 + k, points=6, max_alpha=0.5),
                   rate=rate)
        fr, ordered = workloads.prepare(req, vol)
        params = livre_amd.render_params(workloads.samples_per_ray(req, vol))
        items.append((vol, fr, side, ordered, params, req["tf"]))
    return items


def test_c3_batched_equals_plain_and_oracle(c3):
    items = _items(c3)
    for vol, _fr, _s, _o, _p, tf in items:  # a class the table lacks is timed with a TF set
        vol.ctx.set_tf(tf, vol.range)
    tuned = autotune.Tuner().tune_batch([(v.ctx, fr, s, s, o, p)
                                         for v, fr, s, o, p, _tf in items])
    with livre_amd.Batch() as b:
        for (vol, fr, side, ordered, _p, tf), tp in zip(items, tuned):
            b.add(vol.ctx, fr, side, side, ordered, tp, tf, vol.range)
        b.run()
        st = b.wait()
        assert st.samples > 0
        frames = [b.read(i) for i in range(len(items))]
    for k, ((vol, fr, side, ordered, _p, tf), tp) in enumerate(zip(items, tuned)):
        vol.ctx.set_tf(tf, vol.range)
        vol.ctx.render(fr, side, side, ordered, tp)
        img, smp = vol.ctx.read_frame()
        bimg, bsmp = frames[k]
        print("request", k, vol.name, side, "rate", REQUESTS[k][2], "bricks", len(ordered),
              "tile", (tp.tile_w, tp.tile_h), "fetches", tp.samples_per_thread, "samples",
              smp, flush=True)
        assert smp > 0 and smp == bsmp and np.array_equal(img, bimg), k
        if k not in ORACLE_CHECKED:
            continue
        vi = vol.vi
        ovi = oracle.volume_info(list(vi.voxels), [vi.max_block[0] - 2] * 3, [1] * 3,
                                 int(vi.data_type))
        cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                            fr.far_plane, side, side)
        bricks = [vol.ctx.download(i) for i in ordered]
        y0, y1 = side // 2 - 8, side // 2 + 8
        band = oracle.render_params(tp.n_samples_per_ray, 0.999, vol.range,
                                    window=(0, y0, side, y1))
        ref, ref_samples = oracle.render(ovi, cam, ordered, bricks, tf, band, "f32")
        mx, mean = diff_stats(img[y0:y1], oracle.to_rgba8(ref)[y0:y1])
        print("request", k, "oracle band samples", ref_samples, "max", mx, "mean", mean,
              flush=True)
        assert ref_samples > 0 and mx <= 2 and mean <= 0.5, k
        ref64, samples64 = oracle.render(ovi, cam, ordered, bricks, tf, band, "f64")
        fp64_error(img[y0:y1], ref[y0:y1], ref64[y0:y1],
                   "c3 request %d %s %d^2 rate %.1f rows %d-%d"
                   % (k, vol.name, side, REQUESTS[k][2], y0, y1),
                   (-1, ref_samples, samples64))
