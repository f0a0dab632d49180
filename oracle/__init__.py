"""ctypes binding of the CPU oracle (liblivre_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, where it is the checker and the timed CPU
restatement. The product package livre_amd never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblivre_oracle.so")

DT_FLOAT, DT_UINT8, DT_UINT16, DT_UINT32, DT_INT8, DT_INT16, DT_INT32 = range(7)
_NP_TYPES = {DT_FLOAT: np.float32, DT_UINT8: np.uint8, DT_UINT16: np.uint16,
             DT_UINT32: np.uint32, DT_INT8: np.int8, DT_INT16: np.int16,
             DT_INT32: np.int32}


class VolumeInfo(ctypes.Structure):
    _fields_ = [("voxels", ctypes.c_uint32 * 3),
                ("max_block", ctypes.c_uint32 * 3),
                ("overlap", ctypes.c_uint32 * 3),
                ("data_type", ctypes.c_int32),
                ("world_size", ctypes.c_float * 3),
                ("world_space_per_voxel", ctypes.c_float),
                ("depth", ctypes.c_uint32),
                ("root_blocks", ctypes.c_uint32 * 3)]


class Camera(ctypes.Structure):
    _fields_ = [("inv_proj", ctypes.c_float * 16),
                ("inv_mv", ctypes.c_float * 16),
                ("eye", ctypes.c_float * 3),
                ("near_plane", ctypes.c_float),
                ("far_plane", ctypes.c_float),
                ("width", ctypes.c_int32),
                ("height", ctypes.c_int32)]


class RenderParams(ctypes.Structure):
    _fields_ = [("n_samples_per_ray", ctypes.c_int32),
                ("max_samples_per_ray", ctypes.c_int32),
                ("early_exit", ctypes.c_float),
                ("range0", ctypes.c_float),
                ("range1", ctypes.c_float),
                ("linear_filter", ctypes.c_int32),
                ("n_clip_planes", ctypes.c_int32),
                ("clip_planes", ctypes.c_float * 24),
                ("window", ctypes.c_int32 * 4)]


class FrustumInfo(ctypes.Structure):
    _fields_ = [("near_plane", ctypes.c_float), ("far_plane", ctypes.c_float),
                ("left", ctypes.c_float), ("right", ctypes.c_float),
                ("bottom", ctypes.c_float), ("top", ctypes.c_float),
                ("near_plane_eq", ctypes.c_float * 4),
                ("eye", ctypes.c_float * 3), ("view_dir", ctypes.c_float * 3)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        f32p = P(ctypes.c_float)
        u64p = P(ctypes.c_uint64)
        _lib.orc_fill_volume_info.argtypes = [P(VolumeInfo)]
        _lib.orc_node_world_box.argtypes = [P(VolumeInfo), ctypes.c_uint64, f32p]
        _lib.orc_frustum_info_get.argtypes = [f32p, f32p, P(FrustumInfo)]
        _lib.orc_select_visibles.argtypes = [
            P(VolumeInfo), f32p, f32p, ctypes.c_uint32, ctypes.c_float,
            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
            ctypes.c_int, f32p, u64p, ctypes.c_int, P(ctypes.c_int)]
        _lib.orc_order.argtypes = [P(VolumeInfo), f32p, u64p, ctypes.c_int, u64p]
        for name in ("orc_render_f32", "orc_render_f64"):
            getattr(_lib, name).argtypes = [
                P(VolumeInfo), P(Camera), u64p, ctypes.c_int,
                P(ctypes.c_void_p), f32p, ctypes.c_int, P(RenderParams), f32p,
                u64p, ctypes.c_int]
        _lib.orc_to_rgba8.argtypes = [f32p, ctypes.c_int, P(ctypes.c_uint8)]
        _lib.orc_histogram.argtypes = [
            P(VolumeInfo), u64p, ctypes.c_int, P(ctypes.c_void_p), ctypes.c_int,
            ctypes.c_float, ctypes.c_float, u64p]
        _lib.orc_default_tf.argtypes = [f32p]
        _lib.orc_tf_lut.argtypes = [f32p, ctypes.c_int, P(ctypes.c_uint8)]
        _lib.orc_jpeg_encode.argtypes = [
            P(ctypes.c_uint8), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, P(ctypes.c_uint8), ctypes.c_size_t,
            P(ctypes.c_size_t)]
    return _lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def volume_info(voxels, block, overlap, data_type):
    vi = VolumeInfo()
    for i in range(3):
        vi.voxels[i] = voxels[i]
        vi.max_block[i] = block[i] + 2 * overlap[i]
        vi.overlap[i] = overlap[i]
    vi.data_type = data_type
    if lib().orc_fill_volume_info(ctypes.byref(vi)) != 0:
        raise ValueError("invalid volume information")
    return vi


def world_box(vi, node_id):
    out = np.zeros(6, np.float32)
    lib().orc_node_world_box(ctypes.byref(vi), node_id,
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def frustum_info(mv, proj):
    fi = FrustumInfo()
    mv, pmv = _f32(mv)
    proj, pproj = _f32(proj)
    lib().orc_frustum_info_get(pmv, pproj, ctypes.byref(fi))
    return fi


def select_visibles(vi, mv, proj, height, sse, min_lod, max_lod, rng=(0.0, 1.0),
                    cap=1 << 22, clip_planes=()):
    out = np.zeros(cap, np.uint64)
    n = ctypes.c_int(0)
    mv, pmv = _f32(mv)
    proj, pproj = _f32(proj)
    planes = np.zeros((max(1, len(clip_planes)), 4), np.float32)
    for i, pl in enumerate(clip_planes):
        planes[i] = pl
    planes, pplanes = _f32(planes)
    rc = lib().orc_select_visibles(
        ctypes.byref(vi), pmv, pproj, height, sse, min_lod, max_lod, rng[0],
        rng[1], len(clip_planes), pplanes,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap,
        ctypes.byref(n))
    if rc != 0:
        raise RuntimeError("selection overflow")
    return out[:n.value].copy()


def order(vi, mv, ids):
    ids, pids = _u64(ids)
    out = np.zeros(len(ids), np.uint64)
    mv, pmv = _f32(mv)
    lib().orc_order(ctypes.byref(vi), pmv, pids, len(ids),
                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def camera(inv_proj, inv_mv, eye, near, far, width, height):
    cam = Camera()
    for i in range(16):
        cam.inv_proj[i] = float(np.float32(inv_proj[i]))
        cam.inv_mv[i] = float(np.float32(inv_mv[i]))
    for i in range(3):
        cam.eye[i] = float(np.float32(eye[i]))
    cam.near_plane = near
    cam.far_plane = far
    cam.width = width
    cam.height = height
    return cam


def render_params(n_samples_per_ray, early_exit=0.999, data_range=(0.0, 255.0),
                  linear=False, max_samples_per_ray=32, clip_planes=(), window=None):
    p = RenderParams()
    if window is not None:
        for i in range(4):
            p.window[i] = window[i]
    p.n_samples_per_ray = n_samples_per_ray
    p.max_samples_per_ray = max_samples_per_ray
    p.early_exit = early_exit
    p.range0, p.range1 = data_range
    p.linear_filter = 1 if linear else 0
    p.n_clip_planes = len(clip_planes)
    for i, pl in enumerate(clip_planes):
        for j in range(4):
            p.clip_planes[4 * i + j] = pl[j]
    return p


def render(vi, cam, ordered, bricks, tf, params, precision="f32", threads=0):
    """bricks: list of numpy arrays (or None) aligned with ordered.
    Returns (float32 H x W x 4 image, executed composite steps)."""
    ordered, pord = _u64(ordered)
    ptrs = (ctypes.c_void_p * max(1, len(ordered)))()
    keep = []
    for i, b in enumerate(bricks):
        if b is None:
            ptrs[i] = None
        else:
            b = np.ascontiguousarray(b)
            keep.append(b)
            ptrs[i] = b.ctypes.data
    tf, ptf = _f32(tf)
    out = np.zeros((cam.height, cam.width, 4), np.float32)
    samples = ctypes.c_uint64(0)
    fn = lib().orc_render_f32 if precision == "f32" else lib().orc_render_f64
    rc = fn(ctypes.byref(vi), ctypes.byref(cam), pord, len(ordered), ptrs, ptf,
            len(tf) // 4 if tf.ndim == 1 else tf.shape[0], ctypes.byref(params),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            ctypes.byref(samples), threads)
    if rc != 0:
        raise RuntimeError("oracle render failed")
    return out, samples.value


def to_rgba8(img):
    img, pimg = _f32(img)
    out = np.zeros(img.shape, np.uint8)
    lib().orc_to_rgba8(pimg, img.size // 4,
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


def histogram(vi, ids, bricks, bins, data_range=(0.0, 255.0)):
    ids, pids = _u64(ids)
    ptrs = (ctypes.c_void_p * max(1, len(ids)))()
    keep = []
    for i, b in enumerate(bricks):
        if b is None:
            ptrs[i] = None
        else:
            b = np.ascontiguousarray(b)
            keep.append(b)
            ptrs[i] = b.ctypes.data
    out = np.zeros(bins, np.uint64)
    rc = lib().orc_histogram(ctypes.byref(vi), pids, len(ids), ptrs, bins,
                             data_range[0], data_range[1],
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if rc != 0:
        raise RuntimeError("oracle histogram failed (%d)" % rc)
    return out


def default_tf():
    out = np.zeros((256, 4), np.float32)
    lib().orc_default_tf(out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return out


def tf_lut(rgba):
    rgba, p = _f32(rgba)
    n = rgba.size // 4
    out = np.zeros((n, 4), np.uint8)
    lib().orc_tf_lut(p, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


# ------------------------------------------------------------ LOD pyramid
# Restatement of the raw-volume pyramid (no reference counterpart: the
# reference serves a raw volume as one brick, RawDataSource.cpp:62-219).
# Pinned by closed forms in tests/test_oracle_pyramid.py: a constant volume
# stays constant at every level, the finest level is the file padded with
# zeros, a 2x2x2 block of known values averages to its rounded mean, and a
# brick's halo repeats its neighbours' voxels.

def pyramid_levels(volume, depth, root_blocks, block):
    """volume: numpy (z, y, x) array of the output type. Returns the level
    grids, level 0 first; level L is root_blocks * 2**L * block per axis."""
    P = root_blocks * (1 << (depth - 1)) * block
    finest = np.zeros((P, P, P), volume.dtype)
    nz, ny, nx = volume.shape
    finest[:nz, :ny, :nx] = volume
    levels = [finest]
    for _ in range(depth - 1):
        f = levels[0]
        # children in the order c = dz*4 + dy*2 + dx
        v = [f[dz::2, dy::2, dx::2] for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)]
        if np.issubdtype(f.dtype, np.floating):
            s = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]))
            c = (s * np.float32(0.125)).astype(np.float32)
        else:
            s = sum(x.astype(np.int64) for x in v)
            c = ((s + 4) >> 3).astype(f.dtype)
        levels.insert(0, c)
    return levels


def pyramid_brick(levels, max_block, overlap, node_id):
    """The brick of node_id: its level's voxels [b*block - overlap,
    (b+1)*block + overlap) per axis, clamped to the level's extent."""
    level = node_id & 0xF
    pos = [(node_id >> (4 + 14 * a)) & 0x3FFF for a in range(3)]
    block = max_block - 2 * overlap
    g = levels[level]
    P = g.shape[0]
    idx = [np.clip(np.arange(max_block) + p * block - overlap, 0, P - 1) for p in pos]
    return g[np.ix_(idx[2], idx[1], idx[0])]


def jpeg_encode(pixels, quality=100, bgra=False, bottom_up=True):
    """JPEG file bytes of an (H, W, 4) uint8 image, as the frame grab writes
    it (FrameGrabber.cpp:49-84): 4:4:4, bottom_up = row H-1 is the top."""
    px = np.ascontiguousarray(pixels, dtype=np.uint8)
    assert px.ndim == 3 and px.shape[2] == 4
    h, w = px.shape[:2]
    u8p = ctypes.POINTER(ctypes.c_uint8)
    size = ctypes.c_size_t(0)
    rc = lib().orc_jpeg_encode(px.ctypes.data_as(u8p), w, h, 4 * w, int(bgra), int(bottom_up),
                               quality, None, 0, ctypes.byref(size))
    assert rc == 0, "orc_jpeg_encode failed (%d)" % rc
    out = np.empty(size.value, np.uint8)
    rc = lib().orc_jpeg_encode(px.ctypes.data_as(u8p), w, h, 4 * w, int(bgra), int(bottom_up),
                               quality, out.ctypes.data_as(u8p), out.size, ctypes.byref(size))
    assert rc == 0, "orc_jpeg_encode failed (%d)" % rc
    return out.tobytes()
