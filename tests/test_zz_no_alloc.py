"""lvr_render_noalloc against lvr_render through the C ABI: two calls in a
row on one context (the second with a smaller viewport and fewer bricks),
each frame and its composite-step count identical to lvr_render's, for the
tabulated 8-bit and 16-bit classification and the per-sample one; and the
capacity limits are refused before anything runs. The output buffers are
allocated by the test, as a caller that cannot allocate in its frame loop
would have done beforehand, through the HIP runtime the library uses."""
import ctypes

import numpy as np
import pytest

import livre_amd
from livre_amd import scenes

from scene_kit import Scene

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        livre_amd.lib()  # loads the runtime the library links
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
    return _hip


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptr = p.value

    def read(self, dtype, shape):
        out = np.zeros(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # D2H
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None


@pytest.mark.parametrize("dt,classify", [(livre_amd.DT_UINT8, 0), (livre_amd.DT_UINT16, 0),
                                         (livre_amd.DT_UINT16, 1)])
def test_noalloc_frames_equal_render(lvr, orc, dt, classify):
    s = Scene(dt, n=64, block=16, overlap=1, source="blobs" if dt == livre_amd.DT_UINT8 else
              "noise", seed=6)
    ctx = s.context()
    frame = DeviceBuffer(160 * 120 * 4)
    samples = DeviceBuffer(8)
    try:
        ctx.set_tf(scenes.piecewise_tf(256, seed=21, max_alpha=0.5), s.range)
        params = livre_amd.render_params(64, classify=classify)
        counts = []
        for eye, w, h in (((1.3, 0.6, 1.1), 160, 120), ((-1.6, 0.9, 1.9), 64, 40)):
            mv, proj, fr = s.camera(eye, w, h)
            ordered = s.visible_ordered(mv, proj, h)
            counts.append(len(ordered))
            ctx.render(fr, w, h, ordered, params)
            ref, ref_samples = ctx.read_frame()
            ctx.render_noalloc(fr, w, h, ordered, params, frame.ptr, samples.ptr)
            ctx.sync()
            img = frame.read(np.uint8, (h, w, 4))
            got = int(samples.read(np.uint64, 1)[0])
            print("noalloc", dt, classify, w, h, "bricks", len(ordered), "samples", got,
                  ref_samples, flush=True)
            assert got == ref_samples and ref_samples > 0
            assert np.array_equal(img, ref)
        assert counts[1] < counts[0]
    finally:
        frame.free()
        samples.free()
        ctx.close()


def test_noalloc_refuses_what_exceeds_its_scratch(lvr):
    vi = livre_amd.volume_info(512, 16, 1, livre_amd.DT_UINT8)
    ids = scenes.all_nodes(vi)
    assert len(ids) > 32768
    with livre_amd.Context(vi, livre_amd.slot_bytes(vi) * 8) as ctx:
        ctx.set_tf(scenes.piecewise_tf(256, seed=1), (0.0, 255.0))
        mv, proj = scenes.camera((1.5, 0.5, 1.2), 64, 64)
        fr = livre_amd.frustum(mv, proj)
        out = DeviceBuffer(8192 * 8192 * 4 // 64)
        try:
            with pytest.raises(livre_amd.LivreError, match="32768"):
                ctx.render_noalloc(fr, 64, 64, ids, livre_amd.render_params(64), out.ptr)
            with pytest.raises(livre_amd.LivreError, match="262144"):
                ctx.render_noalloc(fr, 8192, 8192, ids[:0], livre_amd.render_params(64),
                                   out.ptr)
        finally:
            out.free()
