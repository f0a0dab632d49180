"""Python binding of liblivre_amd.so, the MI355X render hot path of Livre.

The library is C++/HIP behind a C ABI (include/livre_amd.h); this module
binds it with ctypes for tests, benchmarks and scripting. There is no
fallback: if the shared library is missing or no GPU is present, the calls
that need it raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblivre_amd.so")

DT_FLOAT, DT_UINT8, DT_UINT16, DT_UINT32, DT_INT8, DT_INT16, DT_INT32 = range(7)
NP_DTYPE = {DT_FLOAT: np.float32, DT_UINT8: np.uint8, DT_UINT16: np.uint16,
            DT_UINT32: np.uint32, DT_INT8: np.int8, DT_INT16: np.int16,
            DT_INT32: np.int32}

LVR_OK, LVR_EINVAL, LVR_EHIP, LVR_ENOMEM, LVR_ENOENT, LVR_ERANGE = 0, -1, -2, -3, -4, -5

# every entry point declared in include/livre_amd.h
EXPORTS = (
    "lvr_fill_regular_volume_info", "lvr_node_world_box", "lvr_frustum_init",
    "lvr_ctx_create", "lvr_ctx_destroy", "lvr_last_error", "lvr_ctx_sync", "lvr_ctx_query",
    "lvr_pool_set_policy", "lvr_pool_stats", "lvr_brick_upload",
    "lvr_brick_upload_many", "lvr_brick_evict", "lvr_brick_resident",
    "lvr_brick_download", "lvr_noise_params_init", "lvr_noise_brick_host",
    "lvr_noise_bricks_device", "lvr_field_brick_host",
    "lvr_field_bricks_device", "lvr_blobnoise_create", "lvr_blobnoise_destroy",
    "lvr_blobnoise_brick_host", "lvr_blobnoise_bricks_device", "lvr_blobnoise_source",
    "lvr_tf_set", "lvr_tf_set_rgba8",
    "lvr_select_visibles", "lvr_order", "lvr_render", "lvr_frame_read",
    "lvr_frame_device_ptr", "lvr_histogram", "lvr_histogram_read",
    "lvr_select_visibles_async", "lvr_select_visibles_wait", "lvr_select_visibles_ex",
    "lvr_select_visibles_ex_async",
    "lvr_raw_open", "lvr_raw_close", "lvr_raw_get_info", "lvr_raw_brick",
    "lvr_pyramid_volume_info", "lvr_pyramid_load", "lvr_pyramid_release",
    "lvr_jpeg_encoder_create", "lvr_jpeg_encoder_destroy", "lvr_jpeg_encode",
    "lvr_jpeg_encode_host",
    "lvr_render_partial", "lvr_partial_device_ptr", "lvr_partial_read", "lvr_composite_layers",
    "lvr_samples_per_ray", "lvr_render_noalloc", "lvr_sortlast_parts", "lvr_sortlast_order",
    "lvr_jpeg_max_size", "lvr_jpeg_error", "lvr_jpeg_get_stats", "lvr_frame_grab_jpeg",
    "lvr_frame_grab_stats",
    "lvr_set_timing", "lvr_frame_stats_get", "lvr_frame_stats_back", "lvr_slot_bytes",
    "lvr_autotune",
    "lvr_batch_create", "lvr_batch_destroy", "lvr_batch_last_error",
    "lvr_batch_set_graph", "lvr_batch_set_solo_tiles", "lvr_batch_add", "lvr_batch_clear", "lvr_batch_run",
    "lvr_batch_wait", "lvr_batch_read", "lvr_batch_query", "lvr_batch_set_output",
    "lvr_batch_reserve_cus",
    "lvr_server_now_ms", "lvr_server_create", "lvr_server_destroy", "lvr_server_last_error",
    "lvr_server_set_coalesce", "lvr_server_add_volume", "lvr_server_open", "lvr_server_close",
    "lvr_server_submit", "lvr_server_step", "lvr_server_drain", "lvr_server_poll",
    "lvr_server_read", "lvr_server_get_stats", "lvr_server_set_tf", "lvr_server_set_params",
    "lvr_server_grab_jpeg", "lvr_lookout_to_modelview", "lvr_modelview_to_lookout",
    "lvr_noise_source", "lvr_stream_create", "lvr_stream_destroy", "lvr_stream_last_error",
    "lvr_stream_request", "lvr_stream_request_ex", "lvr_stream_pump", "lvr_stream_wait", "lvr_stream_rendering_set",
    "lvr_rendering_set", "lvr_stream_get_stats", "lvr_brick_pin",
)


class BatchStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("samples", ctypes.c_uint64),
                ("bricks", ctypes.c_uint64), ("tile_entries", ctypes.c_uint64),
                ("gpu_ms", ctypes.c_float), ("graph_captures", ctypes.c_uint64),
                ("graph_launches", ctypes.c_uint64), ("solo_requests", ctypes.c_uint64)]


class ServerFrame(ctypes.Structure):
    _fields_ = [("request", ctypes.c_int64), ("session", ctypes.c_int32),
                ("status", ctypes.c_int32), ("samples", ctypes.c_uint64),
                ("arrive_ms", ctypes.c_double), ("launch_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "admitted", "frames", "batches", "samples", "superseded", "failed",
        "bricks_generated", "budget_bytes", "used_bytes", "pending", "waiting_sessions",
        "in_flight")] + [(n, ctypes.c_double) for n in (
        "select_ms", "order_ms", "residency_ms", "add_ms", "run_ms", "retire_ms", "gpu_ms")]


class FrameStats(ctypes.Structure):
    _fields_ = [("samples", ctypes.c_uint64), ("bricks", ctypes.c_uint64),
                ("tile_entries", ctypes.c_uint64), ("brick_bytes", ctypes.c_uint64),
                ("bin_ms", ctypes.c_float), ("raycast_ms", ctypes.c_float),
                ("histogram_ms", ctypes.c_float), ("host_ms", ctypes.c_float)]


class LivreError(RuntimeError):
    pass


class VolumeInfo(ctypes.Structure):
    _fields_ = [("voxels", ctypes.c_uint32 * 3),
                ("max_block", ctypes.c_uint32 * 3),
                ("overlap", ctypes.c_uint32 * 3),
                ("data_type", ctypes.c_int32),
                ("world_size", ctypes.c_float * 3),
                ("world_space_per_voxel", ctypes.c_float),
                ("depth", ctypes.c_uint32),
                ("root_blocks", ctypes.c_uint32 * 3)]

    @property
    def block(self):
        return [self.max_block[i] - 2 * self.overlap[i] for i in range(3)]

    @property
    def dtype(self):
        return NP_DTYPE[self.data_type]

    @property
    def brick_shape(self):
        # numpy (z, y, x) shape of one brick, x fastest
        return (self.max_block[2], self.max_block[1], self.max_block[0])


class Frustum(ctypes.Structure):
    _fields_ = [("mv", ctypes.c_float * 16), ("proj", ctypes.c_float * 16),
                ("inv_mv", ctypes.c_float * 16), ("inv_proj", ctypes.c_float * 16),
                ("mvp", ctypes.c_float * 16), ("eye", ctypes.c_float * 3),
                ("view_dir", ctypes.c_float * 3), ("near_plane", ctypes.c_float),
                ("far_plane", ctypes.c_float), ("left", ctypes.c_float),
                ("right", ctypes.c_float), ("bottom", ctypes.c_float),
                ("top", ctypes.c_float), ("planes", (ctypes.c_float * 4) * 6)]


class RenderParams(ctypes.Structure):
    _fields_ = [("n_samples_per_ray", ctypes.c_int32),
                ("max_samples_per_ray", ctypes.c_int32),
                ("early_exit", ctypes.c_float),
                ("linear_filter", ctypes.c_int32),
                ("n_clip_planes", ctypes.c_int32),
                ("clip_planes", ctypes.c_float * 24),
                ("tile_w", ctypes.c_int32), ("tile_h", ctypes.c_int32),
                ("samples_per_thread", ctypes.c_int32), ("classify", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32)]


class SelectParams(ctypes.Structure):
    _fields_ = [("window_height", ctypes.c_uint32), ("sse", ctypes.c_float),
                ("min_lod", ctypes.c_uint32), ("max_lod", ctypes.c_uint32),
                ("n_clip_planes", ctypes.c_int32), ("clip_planes", ctypes.c_float * 24),
                ("range0", ctypes.c_float), ("range1", ctypes.c_float)]


def select_params(window_height, sse, min_lod, max_lod, clip_planes=(), rng=(0.0, 1.0)):
    p = SelectParams()
    p.window_height, p.sse, p.min_lod, p.max_lod = window_height, sse, min_lod, max_lod
    p.n_clip_planes = len(clip_planes)
    for i, pl in enumerate(clip_planes):
        for j in range(4):
            p.clip_planes[4 * i + j] = pl[j]
    p.range0, p.range1 = rng
    return p


class TuneEntry(ctypes.Structure):
    _fields_ = [("tile_w", ctypes.c_int32), ("tile_h", ctypes.c_int32),
                ("samples_per_thread", ctypes.c_int32), ("classify", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32), ("raycast_ms", ctypes.c_float)]


class NoiseParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("octaves", ctypes.c_int32),
                ("base_cells", ctypes.c_int32), ("remap", ctypes.c_uint16 * 4097)]


class StreamStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "requested", "bricks_loaded", "bytes_loaded", "transfers", "source_errors",
        "pool_full", "queued", "generating", "ready", "in_flight")]


class FieldParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("count", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("noise", ctypes.c_float),
                ("pad", ctypes.c_int32)]


class NoiseSourceData(ctypes.Structure):
    _fields_ = [("info", VolumeInfo), ("params", NoiseParams)]


class RawInfo(ctypes.Structure):
    _fields_ = [("info", VolumeInfo), ("input_type", ctypes.c_int32),
                ("big_endian", ctypes.c_int32), ("data_bytes", ctypes.c_uint64)]


_lib = None


def lib():
    """Load liblivre_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LivreError("liblivre_amd.so is not built: run python -m livre_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    c = ctypes
    f32p, u64p, u8p = P(c.c_float), P(c.c_uint64), P(c.c_uint8)
    L.lvr_fill_regular_volume_info.argtypes = [P(VolumeInfo)]
    L.lvr_node_world_box.argtypes = [P(VolumeInfo), c.c_uint64, f32p]
    L.lvr_frustum_init.argtypes = [P(Frustum), f32p, f32p]
    L.lvr_ctx_create.argtypes = [P(VolumeInfo), c.c_size_t, c.c_int, P(c.c_void_p)]
    L.lvr_ctx_destroy.argtypes = [c.c_void_p]
    L.lvr_ctx_destroy.restype = None
    L.lvr_last_error.argtypes = [c.c_void_p]
    L.lvr_last_error.restype = c.c_char_p
    L.lvr_ctx_sync.argtypes = [c.c_void_p]
    L.lvr_ctx_query.argtypes = [c.c_void_p]
    L.lvr_pool_set_policy.argtypes = [c.c_void_p, c.c_int]
    L.lvr_pool_stats.argtypes = [c.c_void_p, u64p, u64p, u64p]
    L.lvr_brick_upload.argtypes = [c.c_void_p, c.c_uint64, c.c_void_p, c.c_size_t]
    L.lvr_brick_upload_many.argtypes = [c.c_void_p, u64p, c.c_int, c.c_void_p,
                                        c.c_size_t]
    L.lvr_brick_evict.argtypes = [c.c_void_p, c.c_uint64]
    L.lvr_brick_resident.argtypes = [c.c_void_p, c.c_uint64]
    L.lvr_brick_download.argtypes = [c.c_void_p, c.c_uint64, c.c_void_p, c.c_size_t]
    L.lvr_autotune.argtypes = [c.c_void_p, P(Frustum), P(c.c_int), u64p, c.c_int, c.c_int,
                               P(RenderParams), P(TuneEntry), c.c_int, P(c.c_int)]
    L.lvr_noise_params_init.argtypes = [P(NoiseParams), c.c_uint64, c.c_int, c.c_int,
                                        c.c_float]
    L.lvr_noise_brick_host.argtypes = [P(VolumeInfo), P(NoiseParams), c.c_uint64,
                                       c.c_void_p, c.c_size_t]
    L.lvr_noise_bricks_device.argtypes = [c.c_void_p, P(NoiseParams), u64p, c.c_int]
    L.lvr_field_brick_host.argtypes = [P(VolumeInfo), P(FieldParams), c.c_uint64,
                                       c.c_void_p, c.c_size_t]
    L.lvr_field_bricks_device.argtypes = [c.c_void_p, P(FieldParams), u64p, c.c_int]
    L.lvr_blobnoise_create.argtypes = [P(VolumeInfo), P(NoiseParams), c.c_uint64, c.c_int,
                                       c.c_int, c.c_int, P(c.c_void_p)]
    L.lvr_blobnoise_destroy.argtypes = [c.c_void_p]
    L.lvr_blobnoise_destroy.restype = None
    L.lvr_blobnoise_brick_host.argtypes = [c.c_void_p, c.c_uint64, c.c_void_p, c.c_size_t]
    L.lvr_blobnoise_bricks_device.argtypes = [c.c_void_p, c.c_void_p, u64p, c.c_int]
    L.lvr_raw_open.argtypes = [c.c_char_p, P(c.c_void_p), c.c_char_p, c.c_size_t]
    L.lvr_raw_close.argtypes = [c.c_void_p]
    L.lvr_raw_close.restype = None
    L.lvr_raw_get_info.argtypes = [c.c_void_p, P(RawInfo)]
    L.lvr_raw_brick.argtypes = [c.c_void_p, c.c_uint64, c.c_void_p, c.c_size_t]
    L.lvr_pyramid_volume_info.argtypes = [c.c_void_p, c.c_uint32, c.c_uint32, P(VolumeInfo)]
    L.lvr_pyramid_load.argtypes = [c.c_void_p, c.c_void_p, u64p, c.c_int]
    L.lvr_pyramid_release.argtypes = [c.c_void_p]
    L.lvr_tf_set.argtypes = [c.c_void_p, f32p, c.c_int, c.c_float, c.c_float]
    L.lvr_tf_set_rgba8.argtypes = [c.c_void_p, u8p, c.c_int, c.c_float, c.c_float]
    L.lvr_select_visibles.argtypes = [c.c_void_p, P(Frustum), c.c_uint32, c.c_float,
                                      c.c_uint32, c.c_uint32, u64p, c.c_int,
                                      P(c.c_int)]
    L.lvr_select_visibles_async.argtypes = [c.c_void_p, P(Frustum), c.c_uint32, c.c_float,
                                            c.c_uint32, c.c_uint32]
    L.lvr_select_visibles_wait.argtypes = [c.c_void_p, u64p, c.c_int, P(c.c_int)]
    L.lvr_select_visibles_ex.argtypes = [c.c_void_p, P(Frustum), P(SelectParams), u64p, c.c_int,
                                         P(c.c_int)]
    L.lvr_select_visibles_ex_async.argtypes = [c.c_void_p, P(Frustum), P(SelectParams)]
    L.lvr_order.argtypes = [P(VolumeInfo), P(Frustum), u64p, c.c_int, u64p]
    L.lvr_samples_per_ray.argtypes = [P(VolumeInfo), u64p, c.c_int, c.c_int32,
                                      P(c.c_int32)]
    L.lvr_sortlast_parts.argtypes = [P(VolumeInfo), u64p, c.c_int, c.c_int, P(c.c_int32)]
    L.lvr_sortlast_order.argtypes = [P(VolumeInfo), P(c.c_float), c.c_int, P(c.c_int32)]
    L.lvr_render.argtypes = [c.c_void_p, P(Frustum), P(c.c_int), u64p, c.c_int,
                             P(RenderParams)]
    L.lvr_render_noalloc.argtypes = [c.c_void_p, P(Frustum), P(c.c_int), u64p, c.c_int,
                                     P(RenderParams), c.c_void_p, c.c_void_p]
    L.lvr_frame_read.argtypes = [c.c_void_p, u8p, c.c_size_t, u64p]
    L.lvr_frame_device_ptr.argtypes = [c.c_void_p, P(c.c_void_p), P(c.c_int),
                                       P(c.c_int)]
    L.lvr_histogram.argtypes = [c.c_void_p, u64p, c.c_int, c.c_int]
    L.lvr_histogram_read.argtypes = [c.c_void_p, u64p, c.c_int]
    L.lvr_slot_bytes.argtypes = [P(VolumeInfo), u64p]
    L.lvr_batch_create.argtypes = [c.c_int, P(c.c_void_p)]
    L.lvr_batch_destroy.argtypes = [c.c_void_p]
    L.lvr_batch_destroy.restype = None
    L.lvr_batch_last_error.argtypes = [c.c_void_p]
    L.lvr_batch_last_error.restype = c.c_char_p
    L.lvr_batch_set_graph.argtypes = [c.c_void_p, c.c_int]
    L.lvr_batch_set_solo_tiles.argtypes = [c.c_void_p, c.c_int]
    L.lvr_batch_add.argtypes = [c.c_void_p, c.c_void_p, P(Frustum), P(c.c_int), u64p,
                                c.c_int, P(RenderParams), f32p, c.c_int, c.c_float,
                                c.c_float, P(c.c_int)]
    L.lvr_batch_clear.argtypes = [c.c_void_p]
    L.lvr_batch_run.argtypes = [c.c_void_p]
    L.lvr_batch_wait.argtypes = [c.c_void_p, P(BatchStats)]
    L.lvr_batch_read.argtypes = [c.c_void_p, c.c_int, u8p, c.c_size_t, u64p]
    L.lvr_batch_query.argtypes = [c.c_void_p]
    L.lvr_batch_set_output.argtypes = [c.c_void_p, c.c_int, c.c_void_p]
    L.lvr_batch_reserve_cus.argtypes = [c.c_void_p, c.c_int]
    L.lvr_server_now_ms.argtypes = []
    L.lvr_server_now_ms.restype = c.c_double
    L.lvr_server_create.argtypes = [c.c_int, c.c_size_t, c.c_int, P(c.c_void_p)]
    L.lvr_server_destroy.argtypes = [c.c_void_p]
    L.lvr_server_destroy.restype = None
    L.lvr_server_last_error.argtypes = [c.c_void_p]
    L.lvr_server_last_error.restype = c.c_char_p
    L.lvr_server_set_coalesce.argtypes = [c.c_void_p, c.c_int]
    L.lvr_server_add_volume.argtypes = [c.c_void_p, c.c_void_p, P(NoiseParams),
                                        P(FieldParams), P(c.c_int)]
    L.lvr_server_open.argtypes = [c.c_void_p, c.c_int, c.c_int, c.c_int, f32p, c.c_int,
                                  c.c_float, c.c_float, P(RenderParams), P(c.c_int),
                                  P(c.c_int)]
    L.lvr_server_close.argtypes = [c.c_void_p, c.c_int]
    L.lvr_server_submit.argtypes = [c.c_void_p, c.c_int, P(Frustum), c.c_double,
                                    P(c.c_int64)]
    L.lvr_server_step.argtypes = [c.c_void_p, P(c.c_int), P(c.c_int)]
    L.lvr_server_drain.argtypes = [c.c_void_p]
    L.lvr_server_poll.argtypes = [c.c_void_p, P(ServerFrame), c.c_int, P(c.c_int)]
    L.lvr_server_read.argtypes = [c.c_void_p, c.c_int, u8p, c.c_size_t, u64p]
    L.lvr_server_get_stats.argtypes = [c.c_void_p, P(ServerStats)]
    L.lvr_server_set_tf.argtypes = [c.c_void_p, c.c_int, P(c.c_float), c.c_int, c.c_float,
                                    c.c_float]
    L.lvr_server_set_params.argtypes = [c.c_void_p, c.c_int, P(RenderParams)]
    L.lvr_server_grab_jpeg.argtypes = [c.c_void_p, c.c_int, c.c_void_p, c.c_int,
                                       P(c.c_uint8), c.c_size_t, P(c.c_size_t)]
    L.lvr_lookout_to_modelview.argtypes = [P(c.c_float), P(c.c_float), c.c_float,
                                           P(c.c_float)]
    L.lvr_modelview_to_lookout.argtypes = [P(c.c_float), P(c.c_float), c.c_float,
                                           P(c.c_float)]
    L.lvr_stream_create.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_int, c.c_int,
                                    P(c.c_void_p)]
    L.lvr_stream_destroy.argtypes = [c.c_void_p]
    L.lvr_stream_destroy.restype = None
    L.lvr_stream_last_error.argtypes = [c.c_void_p]
    L.lvr_stream_last_error.restype = c.c_char_p
    L.lvr_stream_request.argtypes = [c.c_void_p, u64p, c.c_int]
    L.lvr_stream_request_ex.argtypes = [c.c_void_p, u64p, c.c_int, c.c_int]
    L.lvr_stream_pump.argtypes = [c.c_void_p, P(c.c_int)]
    L.lvr_stream_wait.argtypes = [c.c_void_p]
    L.lvr_stream_rendering_set.argtypes = [c.c_void_p, u64p, c.c_int, u64p, c.c_int,
                                           P(c.c_int), P(c.c_int)]
    L.lvr_rendering_set.argtypes = [u64p, c.c_int, u64p, c.c_int, u64p, c.c_int, P(c.c_int),
                                    P(c.c_int)]
    L.lvr_stream_get_stats.argtypes = [c.c_void_p, P(StreamStats)]
    L.lvr_brick_pin.argtypes = [c.c_void_p, u64p, c.c_int, c.c_int]
    L.lvr_set_timing.argtypes = [c.c_void_p, c.c_int]
    L.lvr_render_partial.argtypes = [c.c_void_p, P(Frustum), P(c.c_int), u64p, c.c_int,
                                     P(RenderParams), c.c_int]
    L.lvr_partial_device_ptr.argtypes = [c.c_void_p, c.c_int, P(c.c_void_p)]
    L.lvr_partial_read.argtypes = [c.c_void_p, c.c_int, f32p, c.c_size_t, P(c.c_int),
                                   P(c.c_int)]
    L.lvr_composite_layers.argtypes = [c.c_void_p, P(c.c_void_p), c.c_int, c.c_int, c.c_int,
                                       c.c_int, P(c.c_float)]
    L.lvr_jpeg_encoder_create.argtypes = [c.c_int, P(c.c_void_p)]
    L.lvr_jpeg_encoder_destroy.argtypes = [c.c_void_p]
    L.lvr_jpeg_encoder_destroy.restype = None
    L.lvr_jpeg_encode.argtypes = [c.c_void_p, c.c_void_p, c.c_int, c.c_int, c.c_int, c.c_int,
                                  c.c_int, c.c_int, u8p, c.c_size_t, P(c.c_size_t)]
    L.lvr_jpeg_encode_host.argtypes = [c.c_void_p, c.c_void_p, c.c_int, c.c_int, c.c_int,
                                       c.c_int, c.c_int, c.c_int, u8p, c.c_size_t,
                                       P(c.c_size_t)]
    L.lvr_jpeg_max_size.argtypes = [c.c_int, c.c_int]
    L.lvr_jpeg_max_size.restype = c.c_size_t
    L.lvr_jpeg_error.argtypes = [c.c_void_p]
    L.lvr_jpeg_error.restype = c.c_char_p
    L.lvr_jpeg_get_stats.argtypes = [c.c_void_p, P(JpegStats)]
    L.lvr_frame_grab_jpeg.argtypes = [c.c_void_p, c.c_int, u8p, c.c_size_t, P(c.c_size_t)]
    L.lvr_frame_grab_stats.argtypes = [c.c_void_p, P(JpegStats)]
    L.lvr_frame_stats_get.argtypes = [c.c_void_p, P(FrameStats)]
    L.lvr_frame_stats_back.argtypes = [c.c_void_p, c.c_int, P(FrameStats)]
    _lib = L
    return L


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def volume_info(voxels, block, overlap, data_type):
    """VolumeInformation filled as fillRegularVolumeInfo does."""
    vi = VolumeInfo()
    if np.isscalar(block):
        block = (block,) * 3
    if np.isscalar(overlap):
        overlap = (overlap,) * 3
    if np.isscalar(voxels):
        voxels = (voxels,) * 3
    for i in range(3):
        vi.voxels[i] = voxels[i]
        vi.max_block[i] = block[i] + 2 * overlap[i]
        vi.overlap[i] = overlap[i]
    vi.data_type = data_type
    rc = lib().lvr_fill_regular_volume_info(ctypes.byref(vi))
    if rc != LVR_OK:
        raise LivreError("invalid volume information (%d)" % rc)
    return vi


def slot_bytes(vi):
    """HBM bytes per brick in the pool: a pool of n * slot_bytes holds n bricks."""
    out = ctypes.c_uint64()
    if lib().lvr_slot_bytes(ctypes.byref(vi), ctypes.byref(out)) != LVR_OK:
        raise LivreError("invalid volume information")
    return out.value


def world_box(vi, node_id):
    out = np.zeros(6, np.float32)
    rc = lib().lvr_node_world_box(ctypes.byref(vi), int(node_id),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc != LVR_OK:
        raise LivreError("invalid node id")
    return out


def frustum(mv, proj):
    f = Frustum()
    mv, pmv = _f32(mv)
    proj, pproj = _f32(proj)
    rc = lib().lvr_frustum_init(ctypes.byref(f), pmv, pproj)
    if rc != LVR_OK:
        raise LivreError("singular matrix")
    return f


def order(vi, fr, ids):
    ids, pids = _u64(ids)
    out = np.zeros(len(ids), np.uint64)
    rc = lib().lvr_order(ctypes.byref(vi), ctypes.byref(fr), pids, len(ids),
                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if rc != LVR_OK:
        raise LivreError("order failed")
    return out


def noise_params(seed, octaves=5, base_cells=8, gain=1.5):
    p = NoiseParams()
    rc = lib().lvr_noise_params_init(ctypes.byref(p), seed, octaves, base_cells, gain)
    if rc != LVR_OK:
        raise LivreError("invalid noise parameters")
    return p


def noise_brick_host(vi, params, node_id):
    out = np.zeros(vi.brick_shape, vi.dtype)
    rc = lib().lvr_noise_brick_host(ctypes.byref(vi), ctypes.byref(params),
                                    int(node_id), out.ctypes.data, out.nbytes)
    if rc != LVR_OK:
        raise LivreError("noise brick failed (%d)" % rc)
    return out


FIELD_BLOBS, FIELD_SINES = 0, 1


def field_params(kind, count, seed, noise=0.0):
    """Procedural blob (kind FIELD_BLOBS) or sum-of-sines (FIELD_SINES) source."""
    return FieldParams(kind, count, seed, noise, 0)


def field_brick_host(vi, params, node_id):
    out = np.zeros(vi.brick_shape, vi.dtype)
    rc = lib().lvr_field_brick_host(ctypes.byref(vi), ctypes.byref(params),
                                    int(node_id), out.ctypes.data, out.nbytes)
    if rc != LVR_OK:
        raise LivreError("field brick failed (%d)" % rc)
    return out


class BlobNoise:
    """Blobs plus value noise (lvr_blobnoise_*): the value-noise source
    blended with a smooth field of Gaussian blobs tabulated on a lattice of
    `cells` cells across the largest axis; `weight` of 256 goes to the blobs.
    Host and device bricks are bit-identical."""

    def __init__(self, vi, noise, seed, count=64, cells=64, weight=128):
        self.vi = vi
        self.noise = noise
        self._h = ctypes.c_void_p()
        rc = lib().lvr_blobnoise_create(ctypes.byref(vi), ctypes.byref(noise), seed, count,
                                        cells, weight, ctypes.byref(self._h))
        if rc != LVR_OK:
            raise LivreError("invalid blob-noise parameters (%d)" % rc)

    def brick_host(self, node_id):
        out = np.zeros(self.vi.brick_shape, self.vi.dtype)
        rc = lib().lvr_blobnoise_brick_host(self._h, int(node_id), out.ctypes.data, out.nbytes)
        if rc != LVR_OK:
            raise LivreError("blob-noise brick failed (%d)" % rc)
        return out

    def close(self):
        if self._h:
            lib().lvr_blobnoise_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def samples_per_ray(vi, ids, requested=0):
    """nSamplesPerRay of a frame of these bricks: `requested`, or for 0 the
    reference's rule from the finest rendered level."""
    ids, pids = _u64(ids)
    out = ctypes.c_int32()
    if lib().lvr_samples_per_ray(ctypes.byref(vi), pids, len(ids), requested,
                                 ctypes.byref(out)) != 0:
        raise LivreError("samples_per_ray: invalid arguments")
    return out.value


def sortlast_parts(vi, ids, n_parts):
    """Static sort-last part (0..n_parts-1, or -1 for a node straddling a
    split) of every id: the halves of the block grid along x, y, z."""
    ids, pids = _u64(ids)
    out = np.zeros(len(ids), np.int32)
    if lib().lvr_sortlast_parts(ctypes.byref(vi), pids, len(ids), n_parts,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))) != 0:
        raise LivreError("sortlast_parts: invalid arguments")
    return out


def sortlast_order(vi, eye, n_parts):
    """Front-to-back order of the static parts for an eye position."""
    e = (ctypes.c_float * 3)(*[float(x) for x in eye])
    out = np.zeros(n_parts, np.int32)
    if lib().lvr_sortlast_order(ctypes.byref(vi), e, n_parts,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))) != 0:
        raise LivreError("sortlast_order: invalid arguments")
    return out


def render_params(n_samples_per_ray, early_exit=0.999, linear=False,
                  max_samples_per_ray=32, clip_planes=(), tile=(0, 0),
                  samples_per_thread=0, classify=0, waves_per_simd=0):
    p = RenderParams()
    p.n_samples_per_ray = n_samples_per_ray
    p.max_samples_per_ray = max_samples_per_ray
    p.early_exit = early_exit
    p.linear_filter = 1 if linear else 0
    p.n_clip_planes = len(clip_planes)
    for i, pl in enumerate(clip_planes):
        for j in range(4):
            p.clip_planes[4 * i + j] = pl[j]
    p.tile_w, p.tile_h = tile
    p.samples_per_thread = samples_per_thread
    p.classify = classify
    p.waves_per_simd = waves_per_simd
    return p


class JpegStats(ctypes.Structure):
    _fields_ = [("device_ms", ctypes.c_float), ("dct_ms", ctypes.c_float),
                ("offsets_ms", ctypes.c_float), ("emit_ms", ctypes.c_float),
                ("stuff_ms", ctypes.c_float), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("bytes", ctypes.c_uint64)]


PIXEL_RGBA, PIXEL_BGRA = 0, 1


class JpegEncoder:
    """The frame grabber's JPEG encoder on the device (FrameGrabber::_encodeJpeg)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        rc = lib().lvr_jpeg_encoder_create(device, ctypes.byref(self._h))
        if rc != LVR_OK:
            raise LivreError("lvr_jpeg_encoder_create failed: %s (%d)"
                             % (lib().lvr_jpeg_error(None).decode(), rc))
        self._buf = None

    def close(self):
        if getattr(self, "_h", None):
            lib().lvr_jpeg_encoder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _run(self, fn, src, width, height, pitch, pixel_format, bottom_up, quality):
        need = lib().lvr_jpeg_max_size(max(width, 1), max(height, 1))
        if self._buf is None or self._buf.size < need:
            self._buf = np.empty(need, np.uint8)
        size = ctypes.c_size_t()
        rc = fn(self._h, src, width, height, 4 * width if pitch is None else pitch,
                pixel_format, int(bool(bottom_up)), quality,
                self._buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), self._buf.size,
                ctypes.byref(size))
        if rc != LVR_OK:
            raise LivreError("jpeg encode: %s (%d)" % (lib().lvr_jpeg_error(self._h).decode(), rc))
        return self._buf[:size.value].tobytes()

    def encode(self, pixels, pixel_format=PIXEL_RGBA, bottom_up=True, quality=100, width=None):
        """JPEG file bytes of an (H, W, 4) uint8 host image (FrameGrabber's
        input); `width` < W encodes the left columns with the image's pitch."""
        px = np.ascontiguousarray(pixels, dtype=np.uint8)
        assert px.ndim == 3 and px.shape[2] == 4
        h, w = px.shape[:2]
        return self._run(lib().lvr_jpeg_encode_host, px.ctypes.data_as(ctypes.c_void_p),
                         w if width is None else width, h, 4 * w, pixel_format, bottom_up,
                         quality)

    def encode_device(self, device_ptr, width, height, pitch=None, pixel_format=PIXEL_RGBA,
                      bottom_up=True, quality=100):
        """JPEG file bytes of width x height 4-byte pixels at device_ptr."""
        return self._run(lib().lvr_jpeg_encode, ctypes.c_void_p(int(device_ptr)), width, height,
                         pitch, pixel_format, bottom_up, quality)

    def stats(self):
        st = JpegStats()
        lib().lvr_jpeg_get_stats(self._h, ctypes.byref(st))
        return st


class RawSource:
    """A raw or NRRD volume (RawDataSource): `uri` as lvr_raw_open takes it."""

    def __init__(self, uri):
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib().lvr_raw_open(uri.encode(), ctypes.byref(h), err, len(err))
        if rc != LVR_OK:
            raise LivreError(err.value.decode() or "raw source failed (%d)" % rc)
        self._h = h
        ri = RawInfo()
        lib().lvr_raw_get_info(self._h, ctypes.byref(ri))
        self.raw_info = ri
        self.vi = ri.info

    def close(self):
        if getattr(self, "_h", None):
            lib().lvr_raw_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def brick(self, node_id=0):
        """The one brick of the reference's depth-1 volume (getData)."""
        vi = self.vi
        out = np.empty(vi.brick_shape, vi.dtype)
        rc = lib().lvr_raw_brick(self._h, int(node_id), out.ctypes.data, out.nbytes)
        if rc != LVR_OK:
            raise LivreError("raw brick failed (%d)" % rc)
        return out

    def pyramid_info(self, block, overlap):
        vi = VolumeInfo()
        rc = lib().lvr_pyramid_volume_info(self._h, block, overlap, ctypes.byref(vi))
        if rc != LVR_OK:
            raise LivreError("invalid pyramid shape (%d)" % rc)
        return vi


class Context:
    """One volume's HBM brick pool and render state on one device."""

    def __init__(self, vi, pool_bytes, device=0):
        self.vi = vi
        self._h = ctypes.c_void_p()
        rc = lib().lvr_ctx_create(ctypes.byref(vi), int(pool_bytes), device,
                                  ctypes.byref(self._h))
        if rc != LVR_OK:
            raise LivreError("lvr_ctx_create failed (%d)" % rc)
        self.width = self.height = 0

    def close(self):
        if self._h:
            lib().lvr_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != LVR_OK:
            msg = lib().lvr_last_error(self._h).decode()
            raise LivreError("%s failed (%d): %s" % (what, rc, msg))

    def sync(self):
        self._check(lib().lvr_ctx_sync(self._h), "sync")

    def query(self):
        """True once the queued work of the context has finished."""
        rc = lib().lvr_ctx_query(self._h)
        if rc < 0:
            self._check(rc, "query")
        return rc == 1

    def set_policy(self, policy):
        self._check(lib().lvr_pool_set_policy(self._h, policy), "set_policy")

    def stats(self):
        s, u, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().lvr_pool_stats(self._h, ctypes.byref(s), ctypes.byref(u),
                                         ctypes.byref(b)), "stats")
        return s.value, u.value, b.value

    def upload(self, ids, bricks):
        ids, pids = _u64(ids)
        bricks = np.ascontiguousarray(bricks, dtype=self.vi.dtype)
        per = int(np.prod(self.vi.brick_shape)) * bricks.itemsize
        if bricks.nbytes != per * len(ids):
            raise ValueError("brick array does not match ids")
        self._check(lib().lvr_brick_upload_many(self._h, pids, len(ids),
                                                bricks.ctypes.data, per), "upload")

    def evict(self, node_id):
        self._check(lib().lvr_brick_evict(self._h, int(node_id)), "evict")

    def resident(self, node_id):
        rc = lib().lvr_brick_resident(self._h, int(node_id))
        if rc < 0:
            self._check(rc, "resident")
        return rc == 1

    def download(self, node_id):
        out = np.zeros(self.vi.brick_shape, self.vi.dtype)
        self._check(lib().lvr_brick_download(self._h, int(node_id), out.ctypes.data,
                                             out.nbytes), "download")
        return out

    def pin(self, ids, delta=1):
        """Pin (+1) or unpin (-1) resident bricks (never evicted while pinned)."""
        ids, pids = _u64(ids)
        self._check(lib().lvr_brick_pin(self._h, pids, len(ids), delta), "pin")

    def load_pyramid(self, source, ids):
        """Bricks of a RawSource's LOD pyramid into the pool (the context's
        volume information must be source.pyramid_info(block, overlap))."""
        ids, pids = _u64(ids)
        self._check(lib().lvr_pyramid_load(self._h, source._h, pids, len(ids)),
                    "load_pyramid")

    def release_pyramid(self):
        self._check(lib().lvr_pyramid_release(self._h), "release_pyramid")

    def generate_noise(self, params, ids):
        ids, pids = _u64(ids)
        self._check(lib().lvr_noise_bricks_device(self._h, ctypes.byref(params), pids,
                                                  len(ids)), "generate_noise")

    def generate_blobnoise(self, source, ids):
        ids, pids = _u64(ids)
        self._check(lib().lvr_blobnoise_bricks_device(self._h, source._h, pids, len(ids)),
                    "generate_blobnoise")

    def generate_field(self, params, ids):
        ids, pids = _u64(ids)
        self._check(lib().lvr_field_bricks_device(self._h, ctypes.byref(params), pids,
                                                  len(ids)), "generate_field")

    def set_tf(self, rgba, data_range):
        rgba, p = _f32(rgba)
        self._check(lib().lvr_tf_set(self._h, p, rgba.size // 4, data_range[0],
                                     data_range[1]), "set_tf")

    def set_tf_rgba8(self, lut, data_range):
        lut = np.ascontiguousarray(lut, dtype=np.uint8).reshape(-1)
        self._check(lib().lvr_tf_set_rgba8(
            self._h, lut.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), lut.size // 4,
            data_range[0], data_range[1]), "set_tf_rgba8")

    def node_count(self):
        """Nodes of the octree: the largest possible visible set."""
        if getattr(self, "_nodes", None) is None:
            # every level is the root block grid refined 2x per axis
            vi = self.vi
            self._nodes = sum((vi.root_blocks[0] << lv) * (vi.root_blocks[1] << lv) *
                              (vi.root_blocks[2] << lv) for lv in range(vi.depth))
        return self._nodes

    def select_visibles(self, fr, window_height, sse, min_lod, max_lod, cap=None):
        cap = cap or self.node_count()
        out = np.empty(cap, np.uint64)
        n = ctypes.c_int(0)
        self._check(lib().lvr_select_visibles(
            self._h, ctypes.byref(fr), window_height, sse, min_lod, max_lod,
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap,
            ctypes.byref(n)), "select_visibles")
        return out[:n.value].copy()

    def select_visibles_ex(self, fr, params, cap=None):
        """Selection with clip planes and a sort-last range (select_params)."""
        cap = cap or self.node_count()
        out = np.zeros(cap, np.uint64)
        n = ctypes.c_int()
        self._check(lib().lvr_select_visibles_ex(
            self._h, ctypes.byref(fr), ctypes.byref(params),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap, ctypes.byref(n)),
            "select_visibles_ex")
        return out[:n.value].copy()

    def select_visibles_async(self, fr, window_height, sse, min_lod, max_lod):
        """Queue a selection; select_visibles_wait() returns its result."""
        self._check(lib().lvr_select_visibles_async(
            self._h, ctypes.byref(fr), window_height, sse, min_lod, max_lod),
            "select_visibles_async")

    def select_visibles_wait(self, cap=None):
        cap = cap or self.node_count()
        out = np.empty(cap, np.uint64)
        n = ctypes.c_int(0)
        self._check(lib().lvr_select_visibles_wait(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap,
            ctypes.byref(n)), "select_visibles_wait")
        return out[:n.value].copy()

    def render(self, fr, width, height, ordered, params):
        ordered, pord = _u64(ordered)
        vp = (ctypes.c_int * 4)(0, 0, width, height)
        self._check(lib().lvr_render(self._h, ctypes.byref(fr), vp, pord, len(ordered),
                                     ctypes.byref(params)), "render")
        self.width, self.height = width, height

    def render_noalloc(self, fr, width, height, ordered, params, out_ptr, samples_ptr=None):
        """lvr_render into caller device memory (out_ptr: width * height * 4
        bytes; samples_ptr: 8 bytes or None), allocating nothing."""
        ordered, pord = _u64(ordered)
        vp = (ctypes.c_int * 4)(0, 0, width, height)
        self._check(lib().lvr_render_noalloc(self._h, ctypes.byref(fr), vp, pord, len(ordered),
                                             ctypes.byref(params), ctypes.c_void_p(out_ptr),
                                             ctypes.c_void_p(samples_ptr)), "render_noalloc")

    def render_partial(self, fr, width, height, ordered, params, layer):
        """Sort-last partial: ordered (a contiguous range of the frame's
        order) into RGBA32F layer `layer`."""
        ordered, pord = _u64(ordered)
        vp = (ctypes.c_int * 4)(0, 0, width, height)
        self._check(lib().lvr_render_partial(self._h, ctypes.byref(fr), vp, pord, len(ordered),
                                             ctypes.byref(params), layer), "render_partial")

    def layer_ptr(self, layer):
        p = ctypes.c_void_p()
        self._check(lib().lvr_partial_device_ptr(self._h, layer, ctypes.byref(p)), "layer_ptr")
        return p.value

    def read_layer(self, layer):
        """A sort-last layer as an (h, w, 4) float32 array (premultiplied
        RGBA32F, row 0 = bottom)."""
        w, h = ctypes.c_int(), ctypes.c_int()
        self._check(lib().lvr_partial_read(self._h, layer, None, 0, ctypes.byref(w),
                                           ctypes.byref(h)), "partial_read")
        out = np.zeros((h.value, w.value, 4), np.float32)
        self._check(lib().lvr_partial_read(
            self._h, layer, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.nbytes,
            ctypes.byref(w), ctypes.byref(h)), "partial_read")
        return out

    def composite(self, layer_ptrs, width, height, timed=False, reps=1):
        """Front-to-back over of RGBA32F device images into the framebuffer;
        returns the kernel's mean milliseconds per launch (over reps
        back-to-back launches) when timed."""
        arr = (ctypes.c_void_p * max(1, len(layer_ptrs)))(
            *[ctypes.c_void_p(p) for p in layer_ptrs])
        ms = ctypes.c_float(0.0)
        self._check(lib().lvr_composite_layers(self._h, arr, len(layer_ptrs), width, height,
                                               reps, ctypes.byref(ms) if timed else None),
                    "composite")
        self.width, self.height = width, height
        return ms.value if timed else None

    def autotune(self, fr, w, h, ordered, params, reps=3):
        """Fastest ray tile, samples per thread, classification mode and
        wavefronts per SIMD for this frame: returns a copy of params with them
        set, and the list of (tile_w, tile_h, samples_per_thread, classify,
        waves_per_simd, raycast_ms) candidates."""
        ordered, pids = _u64(ordered)
        vp = (ctypes.c_int * 4)(0, 0, w, h)
        out = RenderParams.from_buffer_copy(params)
        cap = 256
        res = (TuneEntry * cap)()
        n = ctypes.c_int()
        self._check(lib().lvr_autotune(self._h, ctypes.byref(fr), vp, pids, len(ordered), reps,
                                       ctypes.byref(out), res, cap, ctypes.byref(n)), "autotune")
        self.width, self.height = w, h
        cands = [(r.tile_w, r.tile_h, r.samples_per_thread, r.classify, r.waves_per_simd,
                  r.raycast_ms) for r in res[:min(n.value, cap)]]
        return out, cands

    def read_frame(self):
        img = np.zeros((self.height, self.width, 4), np.uint8)
        s = ctypes.c_uint64()
        self._check(lib().lvr_frame_read(
            self._h, img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), img.nbytes,
            ctypes.byref(s)), "read_frame")
        return img, s.value

    def samples(self):
        s = ctypes.c_uint64()
        self._check(lib().lvr_frame_read(self._h, None, 0, ctypes.byref(s)), "samples")
        return s.value

    def set_timing(self, enable=True):
        self._check(lib().lvr_set_timing(self._h, 1 if enable else 0), "set_timing")

    def frame_stats(self, back=0):
        """Statistics of the last rendered frame, or of the one `back` frames
        before it (up to 3) while later frames are in flight."""
        s = FrameStats()
        self._check(lib().lvr_frame_stats_back(self._h, back, ctypes.byref(s)), "frame_stats")
        return s

    def grab_jpeg(self, quality=100):
        """The last frame as the reference's frame grabber encodes it."""
        need = lib().lvr_jpeg_max_size(max(self.width, 1), max(self.height, 1))
        buf = getattr(self, "_jpeg_buf", None)
        if buf is None or buf.size < need:  # kept: its pages stay mapped
            buf = self._jpeg_buf = np.empty(need, np.uint8)
        size = ctypes.c_size_t()
        self._check(lib().lvr_frame_grab_jpeg(
            self._h, quality, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), buf.size,
            ctypes.byref(size)), "frame_grab_jpeg")
        return buf[:size.value].tobytes()

    def frame_device_ptr(self):
        """(device pointer, width, height) of the framebuffer."""
        p, w, h = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().lvr_frame_device_ptr(self._h, ctypes.byref(p), ctypes.byref(w),
                                               ctypes.byref(h)), "frame_device_ptr")
        return p.value, w.value, h.value

    def grab_stats(self):
        st = JpegStats()
        self._check(lib().lvr_frame_grab_stats(self._h, ctypes.byref(st)), "frame_grab_stats")
        return st

    def histogram(self, ids, bins, read=True):
        ids, pids = _u64(ids)
        self._check(lib().lvr_histogram(self._h, pids, len(ids), bins), "histogram")
        if not read:
            return None
        out = np.zeros(bins, np.uint64)
        self._check(lib().lvr_histogram_read(
            self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), bins),
            "histogram_read")
        return out


class Batch:
    """Many frames of different volumes / viewports / TFs in one persistent
    launch (lvr_batch_*)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        if lib().lvr_batch_create(device, ctypes.byref(self._h)) != LVR_OK:
            raise LivreError("lvr_batch_create failed")
        self.sizes = []

    def close(self):
        if self._h:
            lib().lvr_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != LVR_OK:
            msg = lib().lvr_batch_last_error(self._h).decode()
            raise LivreError("%s failed (%d): %s" % (what, rc, msg))

    def set_graph(self, enable):
        self._check(lib().lvr_batch_set_graph(self._h, 1 if enable else 0), "set_graph")

    def set_solo_tiles(self, min_tiles):
        """Requests the single-frame window kernel covers with at least
        min_tiles ray tiles get launches of their own (0 = default, negative
        = only a lone request)."""
        self._check(lib().lvr_batch_set_solo_tiles(self._h, int(min_tiles)), "set_solo_tiles")

    def reserve_cus(self, n):
        self._check(lib().lvr_batch_reserve_cus(self._h, n), "reserve_cus")

    def query(self):
        rc = lib().lvr_batch_query(self._h)
        if rc < 0:
            self._check(rc, "query")
        return rc == 1

    def clear(self):
        self._check(lib().lvr_batch_clear(self._h), "clear")
        self.sizes = []

    def add(self, ctx, fr, width, height, ordered, params, tf, data_range):
        ordered, pord = _u64(ordered)
        tf, ptf = _f32(tf)
        vp = (ctypes.c_int * 4)(0, 0, width, height)
        req = ctypes.c_int()
        self._check(lib().lvr_batch_add(self._h, ctx._h, ctypes.byref(fr), vp, pord,
                                        len(ordered), ctypes.byref(params), ptf,
                                        tf.size // 4, data_range[0], data_range[1],
                                        ctypes.byref(req)), "add")
        if req.value == 0:
            self.sizes = []
        self.sizes.append((width, height))
        return req.value

    def run(self):
        self._check(lib().lvr_batch_run(self._h), "run")

    def wait(self):
        s = BatchStats()
        self._check(lib().lvr_batch_wait(self._h, ctypes.byref(s)), "wait")
        return s

    def read(self, request):
        w, h = self.sizes[request]
        img = np.zeros((h, w, 4), np.uint8)
        s = ctypes.c_uint64()
        self._check(lib().lvr_batch_read(
            self._h, request, img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            img.nbytes, ctypes.byref(s)), "read")
        return img, s.value


class Server:
    """Continuous-batching request loop over concurrent sessions
    (lvr_server_*)."""

    def __init__(self, budget_bytes, max_batch=64, device=0):
        self._h = ctypes.c_void_p()
        if lib().lvr_server_create(device, int(budget_bytes), max_batch,
                                   ctypes.byref(self._h)) != LVR_OK:
            raise LivreError("lvr_server_create failed")
        self.sizes = {}
        self._keep = []  # contexts and sources must outlive the server

    @staticmethod
    def now_ms():
        return lib().lvr_server_now_ms()

    def close_server(self):
        if self._h:
            lib().lvr_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close_server()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close_server()

    def _check(self, rc, what):
        if rc != LVR_OK:
            msg = lib().lvr_server_last_error(self._h).decode()
            raise LivreError("%s failed (%d): %s" % (what, rc, msg))

    def set_coalesce(self, enable):
        self._check(lib().lvr_server_set_coalesce(self._h, 1 if enable else 0), "set_coalesce")

    def add_volume(self, ctx, noise=None, field=None):
        v = ctypes.c_int()
        self._check(lib().lvr_server_add_volume(
            self._h, ctx._h, ctypes.byref(noise) if noise is not None else None,
            ctypes.byref(field) if field is not None else None, ctypes.byref(v)), "add_volume")
        self._keep.append((ctx, noise, field))
        return v.value

    def open(self, volume, width, height, tf, data_range, params):
        tf, ptf = _f32(tf)
        sess, adm = ctypes.c_int(), ctypes.c_int()
        self._check(lib().lvr_server_open(self._h, volume, width, height, ptf, tf.size // 4,
                                          data_range[0], data_range[1], ctypes.byref(params),
                                          ctypes.byref(sess), ctypes.byref(adm)), "open")
        self.sizes[sess.value] = (width, height)
        return sess.value, bool(adm.value)

    def close(self, session):
        self._check(lib().lvr_server_close(self._h, session), "close")

    def submit(self, session, fr, arrive_ms=None):
        if arrive_ms is None:
            arrive_ms = self.now_ms()
        rid = ctypes.c_int64()
        self._check(lib().lvr_server_submit(self._h, session, ctypes.byref(fr), arrive_ms,
                                            ctypes.byref(rid)), "submit")
        return rid.value

    def step(self):
        launched, retired = ctypes.c_int(), ctypes.c_int()
        self._check(lib().lvr_server_step(self._h, ctypes.byref(launched),
                                          ctypes.byref(retired)), "step")
        return launched.value, retired.value

    def drain(self):
        self._check(lib().lvr_server_drain(self._h), "drain")

    def poll(self, cap=256):
        buf = self._poll_buf if getattr(self, "_poll_cap", 0) == cap else None
        if buf is None:
            buf = self._poll_buf = (ServerFrame * cap)()
            self._poll_cap = cap
        n = ctypes.c_int()
        out = []
        while True:
            self._check(lib().lvr_server_poll(self._h, buf, cap, ctypes.byref(n)), "poll")
            out += self._frames(buf, n.value)
            if n.value < cap:
                return out

    @staticmethod
    def _frames(buf, n):
        return [dict(request=f.request, session=f.session, status=f.status,
                     samples=f.samples, arrive_ms=f.arrive_ms, launch_ms=f.launch_ms,
                     done_ms=f.done_ms) for f in buf[:n]]

    def read(self, session):
        w, h = self.sizes[session]
        img = np.zeros((h, w, 4), np.uint8)
        s = ctypes.c_uint64()
        self._check(lib().lvr_server_read(
            self._h, session, img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), img.nbytes,
            ctypes.byref(s)), "read")
        return img, s.value

    def stats(self):
        st = ServerStats()
        self._check(lib().lvr_server_get_stats(self._h, ctypes.byref(st)), "stats")
        return st

    def set_tf(self, session, tf, data_range):
        """A session's transfer function message: requests staged from now
        on use it."""
        tf, ptf = _f32(tf)
        self._check(lib().lvr_server_set_tf(self._h, session, ptf, tf.size // 4,
                                            data_range[0], data_range[1]), "set_tf")

    def set_params(self, session, params):
        """A session's render parameters / clip planes message."""
        self._check(lib().lvr_server_set_params(self._h, session, ctypes.byref(params)),
                    "set_params")

    def grab_jpeg(self, session, encoder, quality=100):
        """ImageJPEG of the session's last finished frame, encoded on the
        device from its framebuffer."""
        w, h = self.sizes[session]
        cap = lib().lvr_jpeg_max_size(w, h)
        buf = np.empty(cap, np.uint8)
        size = ctypes.c_size_t()
        self._check(lib().lvr_server_grab_jpeg(
            self._h, session, encoder._h, quality,
            buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), cap, ctypes.byref(size)),
            "grab_jpeg")
        return buf[:size.value].tobytes()


def lookout_to_modelview(lookout, data_to_livre=None, meter_to_data_unit=1.0):
    """Livre's modelview (column-major 16 floats) from a client LookOut
    matrix (the communicator's onCamera)."""
    src, psrc = _f32(np.asarray(lookout, np.float32).reshape(16))
    d, pd = (None, None) if data_to_livre is None else _f32(
        np.asarray(data_to_livre, np.float32).reshape(16))
    out = np.zeros(16, np.float32)
    if lib().lvr_lookout_to_modelview(psrc, pd, meter_to_data_unit,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
        raise LivreError("lookout_to_modelview: invalid matrix")
    return out


def modelview_to_lookout(modelview, data_to_livre=None, meter_to_data_unit=1.0):
    """The client LookOut (column-major 16 floats) of a Livre modelview (the
    communicator's _createLookOut)."""
    src, psrc = _f32(np.asarray(modelview, np.float32).reshape(16))
    d, pd = (None, None) if data_to_livre is None else _f32(
        np.asarray(data_to_livre, np.float32).reshape(16))
    out = np.zeros(16, np.float32)
    if lib().lvr_modelview_to_lookout(psrc, pd, meter_to_data_unit,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) != 0:
        raise LivreError("modelview_to_lookout: invalid matrix")
    return out


def rendering_set(visible, resident):
    """RenderingSetGenerator rule against an explicit resident set:
    (ids, n_available)."""
    visible, pv = _u64(visible)
    resident, pr = _u64(resident)
    out = np.zeros(max(1, len(visible)), np.uint64)
    n, avail = ctypes.c_int(), ctypes.c_int()
    rc = lib().lvr_rendering_set(pv, len(visible), pr, len(resident),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(out),
                                 ctypes.byref(n), ctypes.byref(avail))
    if rc != LVR_OK:
        raise LivreError("rendering set failed (%d)" % rc)
    return out[:n.value].copy(), avail.value


class Stream:
    """Out-of-core streaming of a context's bricks from a host DataSource
    (lvr_stream_*): the value noise (NoiseParams) or blobs plus value noise
    (a BlobNoise, which must outlive the stream)."""

    def __init__(self, ctx, source, threads=4, staging_bricks=64):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        if isinstance(source, BlobNoise):
            self._data = source
            fn = ctypes.cast(lib().lvr_blobnoise_source, ctypes.c_void_p)
            user = source._h
        else:
            self._data = NoiseSourceData(ctx.vi, source)
            fn = ctypes.cast(lib().lvr_noise_source, ctypes.c_void_p)
            user = ctypes.cast(ctypes.pointer(self._data), ctypes.c_void_p)
        rc = lib().lvr_stream_create(ctx._h, fn, user, threads, staging_bricks,
                                     ctypes.byref(self._h))
        if rc != LVR_OK:
            raise LivreError("lvr_stream_create failed (%d)" % rc)

    def close(self):
        if self._h:
            lib().lvr_stream_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != LVR_OK:
            msg = lib().lvr_stream_last_error(self._h).decode()
            raise LivreError("%s failed (%d): %s" % (what, rc, msg))

    def request(self, ids, ancestors=False):
        ids, pids = _u64(ids)
        self._check(lib().lvr_stream_request_ex(self._h, pids, len(ids), 1 if ancestors else 0),
                    "request")

    def pump(self):
        n = ctypes.c_int()
        self._check(lib().lvr_stream_pump(self._h, ctypes.byref(n)), "pump")
        return n.value

    def wait(self):
        self._check(lib().lvr_stream_wait(self._h), "wait")

    def rendering_set(self, visible):
        visible, pv = _u64(visible)
        out = np.zeros(max(1, len(visible)), np.uint64)
        n, avail = ctypes.c_int(), ctypes.c_int()
        self._check(lib().lvr_stream_rendering_set(
            self._h, pv, len(visible), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            len(out), ctypes.byref(n), ctypes.byref(avail)), "rendering_set")
        return out[:n.value].copy(), avail.value

    def stats(self):
        st = StreamStats()
        self._check(lib().lvr_stream_get_stats(self._h, ctypes.byref(st)), "stats")
        return st
