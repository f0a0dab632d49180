"""The session front end on the GPU: a session driven by the messages of
Livre's ZeroEQ communicator. A LookOut camera (converted with
lvr_lookout_to_modelview) renders the frame of the equivalent modelview; a
TransferFunction1D message and a ClipPlanes / VR-parameters message change
the following frames exactly as lvr_render with the new TF or parameters
renders them; the ImageJPEG of the session (encoded on the device from its
framebuffer) is byte-identical to the encoder's file for the frame's pixels
read back to the host."""
import numpy as np
import pytest

import livre_amd
from livre_amd import scenes

pytestmark = pytest.mark.gpu


def test_session_messages(lvr):
    vi = livre_amd.volume_info(64, 16, 1, livre_amd.DT_UINT16)
    noise = livre_amd.noise_params(23, octaves=4, base_cells=4)
    ids = scenes.all_nodes(vi)
    ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
    ref_ctx = livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids))
    ref_ctx.generate_noise(noise, ids)
    rng = scenes.type_range(vi.data_type)
    w, h = 96, 72
    tf0 = scenes.piecewise_tf(256, seed=61, max_alpha=0.5)
    tf1 = scenes.piecewise_tf(256, seed=62, max_alpha=0.6)
    p0 = livre_amd.render_params(96)
    p1 = livre_amd.render_params(128, clip_planes=[(0.6, -0.3, 0.74, 0.05)])
    proj = scenes.perspective(60.0, w / h, 0.3, 6.0)

    def alone(mv, tf, params):
        fr = livre_amd.frustum(mv, proj)
        vis = ref_ctx.select_visibles(fr, h, 1.0, 0, vi.depth - 1)
        ordered = livre_amd.order(vi, fr, vis)
        ref_ctx.set_tf(tf, rng)
        ref_ctx.render(fr, w, h, ordered, params)
        return ref_ctx.read_frame()

    try:
        with livre_amd.Server(budget_bytes=1 << 33, max_batch=1) as srv, \
                livre_amd.JpegEncoder() as enc:
            v = srv.add_volume(ctx, noise=noise)
            s, adm = srv.open(v, w, h, tf0, rng, p0)
            assert adm
            # LookOut in meters with 1000 data units per meter: the client's
            # camera 1.7 mm away
            lookout = scenes.look_at((1.1e-3, 0.6e-3, 1.1e-3))
            mv = livre_amd.lookout_to_modelview(lookout, None, 1000.0)
            eye = np.linalg.inv(np.asarray(mv, np.float64).reshape(4, 4).T)[:3, 3]
            assert np.allclose(eye, (1.1, 0.6, 1.1), atol=1e-5)
            for tf, params, msg in ((tf0, p0, "camera"), (tf1, p0, "tf"), (tf1, p1, "params")):
                if msg == "tf":
                    srv.set_tf(s, tf, rng)
                if msg == "params":
                    srv.set_params(s, params)
                srv.submit(s, livre_amd.frustum(mv, proj))
                srv.drain()
                assert [d["status"] for d in srv.poll()] == [0]
                img, smp = srv.read(s)
                ref, ref_smp = alone(mv, tf, params)
                print("message", msg, "samples", smp, ref_smp, flush=True)
                assert smp == ref_smp and np.array_equal(img, ref), msg
                jpg = srv.grab_jpeg(s, enc)
                assert jpg == enc.encode(img, bottom_up=True) and jpg[:2] == b"\xff\xd8"
            # the camera published back to the client is its own LookOut
            back = livre_amd.modelview_to_lookout(mv, None, 1000.0)
            assert np.allclose(back, np.asarray(lookout, np.float32).reshape(16), atol=1e-6)
    finally:
        ctx.close()
        ref_ctx.close()
