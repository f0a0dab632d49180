"""Small seeded scenes shared by the parity tests: the same voxels, camera and
transfer function go to the oracle and to the HIP library."""
import numpy as np

import oracle
import livre_amd
from livre_amd import scenes

ORC_DT = {livre_amd.DT_UINT8: oracle.DT_UINT8, livre_amd.DT_UINT16: oracle.DT_UINT16,
          livre_amd.DT_FLOAT: oracle.DT_FLOAT}


class Scene:
    def __init__(self, dt=livre_amd.DT_UINT8, n=64, block=16, overlap=1, source="blobs",
                 seed=1):
        # n: voxels per axis, or one count for a cube (the blob source is cubic)
        self.dt = dt
        self.vi = livre_amd.volume_info(n, block, overlap, dt)
        self.ovi = oracle.volume_info(list(n) if not np.isscalar(n) else [n] * 3, [block] * 3,
                                      [overlap] * 3, ORC_DT[dt])
        self.ids = scenes.all_nodes(self.vi)
        if source == "blobs":
            vol = scenes.blob_volume(n, seed=seed, blobs=16, dtype=livre_amd.NP_DTYPE[dt])
            bricks = scenes.blob_bricks(self.vi, vol, self.ids)
        else:
            self.noise = livre_amd.noise_params(seed, octaves=4, base_cells=4)
            bricks = scenes.noise_bricks(self.vi, self.noise, self.ids)
        self.bricks = {int(i): b for i, b in zip(self.ids, bricks)}
        self.range = scenes.type_range(dt)

    def camera(self, eye, w, h, fovy=60.0, target=(0.0, 0.0, 0.0)):
        mv, proj = scenes.camera(eye, w, h, fovy, target)
        return mv, proj, livre_amd.frustum(mv, proj)

    def oracle_camera(self, fr, w, h):
        return oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye),
                             fr.near_plane, fr.far_plane, w, h)

    def visible_ordered(self, mv, proj, h, sse=1.0):
        vis = oracle.select_visibles(self.ovi, mv, proj, h, sse, 0, self.ovi.depth - 1)
        return oracle.order(self.ovi, mv, vis)

    def oracle_render(self, fr, w, h, ordered, tf, n_samples, precision="f32",
                      early_exit=0.999, linear=False, missing=()):
        cam = self.oracle_camera(fr, w, h)
        params = oracle.render_params(n_samples, early_exit, self.range, linear)
        bricks = [None if int(i) in missing else self.bricks[int(i)] for i in ordered]
        img, samples = oracle.render(self.ovi, cam, ordered, bricks, tf, params,
                                     precision)
        return img, samples

    def context(self, pool_bytes=None):
        slot = livre_amd.slot_bytes(self.vi)
        ctx = livre_amd.Context(self.vi, pool_bytes or slot * len(self.ids))
        assert ctx.stats()[0] >= len(self.ids)
        ctx.upload(self.ids, np.stack([self.bricks[int(i)] for i in self.ids]))
        return ctx


def diff_stats(gpu_rgba8, ref_rgba8):
    d = np.abs(gpu_rgba8.astype(np.int32) - ref_rgba8.astype(np.int32))
    return int(d.max()), float(d.mean())


def fp64_error(gpu_rgba8, ref32_rgba, ref64_rgba, label, samples=None):
    """Error of a frame against the FP64 transcription of the shader, printed
    beside the FP32 restatement's own error against FP64 (the precision the
    reference shader computes in). Returns (max, mean, channels above 2/255)
    of the frame. The FP64 bound: mean within 0.5/255, and at most 1e-4 of
    the channels beyond 2/255 (FP32 rounding flips a texel or a sample at a
    few pixels, in the FP32 restatement as much as in the kernel)."""
    import oracle
    ref64 = oracle.to_rgba8(ref64_rgba).astype(np.int32)
    d = np.abs(gpu_rgba8.astype(np.int32) - ref64)
    d32 = np.abs(oracle.to_rgba8(ref32_rgba).astype(np.int32) - ref64)
    over, over32 = int((d > 2).sum()), int((d32 > 2).sum())
    extra = ""
    if samples is not None:
        extra = " | samples kernel %d fp32 %d fp64 %d" % samples
    print("FP64ERR %s max %d mean %.6f over2 %d of %d | fp32 oracle vs fp64: max %d mean %.6f "
          "over2 %d%s" % (label, d.max(), d.mean(), over, d.size, d32.max(), d32.mean(), over32,
                          extra), flush=True)
    assert d.mean() <= 0.5 and over <= 1e-4 * d.size, (label, d.max(), d.mean(), over)
    return int(d.max()), float(d.mean()), over

