"""Synthetic workloads of the measured configurations, shared by bench.py and
the tests: the volumes (generated on the device straight into the brick
pool by the procedural data sources), seeded request lists and camera paths.

Volumes
  blobs256_u8    256^3 uint8, 64 anisotropic Gaussian blobs + 10% noise, seed 1,
                 32^3 bricks (the CPU-runnable configuration)
  blobs512_u8    512^3 uint8, 64 anisotropic Gaussian blobs + 10% noise, seed 5
  noise1024_u16  1024^3 uint16, 5-octave value noise, log-normal remap, seed 6
  sines1536_f32  1536^3 float32, mean of 8 low-frequency sines in [-1, 1],
                 seed 7, 48^3 bricks
  noise2048_u16  2048^3 uint16, 5-octave value noise, seed 3 (the 1080p
                 single-volume workload)
All use a 1-voxel halo and 64^3 bricks unless stated.

Out-of-core volume
  ooc8192_u8     8192^3 uint8 (512 GB logical), depth 8: blobs plus value
                 noise (c4_source), streamed from the host source
"""
import math
import time

import numpy as np

from . import (DT_FLOAT, DT_UINT8, DT_UINT16, FIELD_BLOBS, FIELD_SINES, BlobNoise, Context,
               field_params, frustum, noise_params, slot_bytes, volume_info)
from . import scenes

VOLUMES = {
    "blobs256_u8": dict(voxels=256, block=32, dtype=DT_UINT8, source=("blobs", 64, 1, 0.1)),
    "blobs512_u8": dict(voxels=512, block=64, dtype=DT_UINT8, source=("blobs", 64, 5, 0.1)),
    "noise1024_u16": dict(voxels=1024, block=64, dtype=DT_UINT16, source=("noise", 6)),
    "sines1536_f32": dict(voxels=1536, block=48, dtype=DT_FLOAT, source=("sines", 8, 7)),
    "noise2048_u16": dict(voxels=2048, block=64, dtype=DT_UINT16, source=("noise", 3)),
}
BATCH_VOLUMES = ("blobs512_u8", "noise1024_u16", "sines1536_f32", "noise2048_u16")


SERVE_VOLUMES = ("blobs256_u8", "noise2048_u16", "blobs512_u8", "noise1024_u16",
                 "sines1536_f32")
VIEWPORTS = ((1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))


def source_params(name):
    """(noise, field) parameters of a volume's procedural source."""
    src = VOLUMES[name]["source"]
    if src[0] == "noise":
        return noise_params(src[1], octaves=5, base_cells=8, gain=1.5), None
    if src[0] == "blobs":
        return None, field_params(FIELD_BLOBS, src[1], src[2], src[3])
    return None, field_params(FIELD_SINES, src[1], src[2])


class Volume:
    """A volume in its own brick pool: every brick of every level generated
    on the device (generate=True), or an empty pool of pool_bricks slots that
    a server fills on demand."""

    def __init__(self, name, extra_slots=64, pool_bricks=None, generate=True):
        spec = VOLUMES[name]
        self.name = name
        self.vi = volume_info(spec["voxels"], spec["block"], 1, spec["dtype"])
        self.ids = scenes.all_nodes(self.vi)
        self.range = scenes.type_range(spec["dtype"])
        self.vmax = spec["voxels"]
        n = pool_bricks if pool_bricks is not None else len(self.ids) + extra_slots
        self.pool_bricks = n
        self.pool_bytes = slot_bytes(self.vi) * n
        self.ctx = Context(self.vi, self.pool_bytes)
        self.noise, self.field = source_params(name)
        t0 = time.perf_counter()
        if generate:
            if self.noise is not None:
                self.ctx.generate_noise(self.noise, self.ids)
            else:
                self.ctx.generate_field(self.field, self.ids)
            self.ctx.sync()
        self.gen_s = time.perf_counter() - t0

    def close(self):
        self.ctx.close()


def look_from(eye, width, height, fovy=60.0):
    """Camera at eye looking at the volume centre (up = +y unless the view is
    nearly vertical)."""
    d = np.asarray(eye, np.float64)
    up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.98 * np.linalg.norm(d) else (0.0, 0.0, 1.0)
    mv, proj = scenes.camera(tuple(d), width, height, fovy, up=up)
    return mv, proj, frustum(mv, proj)


def batch_requests(n=64, seed=11, volumes=BATCH_VOLUMES):
    """Uneven requests: square viewports with log-uniform area in
    [256^2, 2048^2], a volume drawn uniformly, a random view direction at
    radius U[1.5, 3], a seeded 256-entry transfer function and a sampling rate
    from {0.5, 1, 2} samples per voxel of the finest level."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        area = math.exp(rng.uniform(math.log(256.0 ** 2), math.log(2048.0 ** 2)))
        side = int(round(math.sqrt(area)))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        out.append(dict(volume=volumes[int(rng.integers(len(volumes)))], width=side,
                        height=side, eye=tuple(d * rng.uniform(1.5, 3.0)),
                        tf=scenes.piecewise_tf(256, seed=1000 + i, points=6, max_alpha=0.5),
                        rate=float(rng.choice([0.5, 1.0, 2.0]))))
    return out


def prepare(request, vol, sse=1.0):
    """Camera, visible set (device SSE selection) and front-to-back order of
    one request."""
    from . import order
    mv, proj, fr = look_from(request["eye"], request["width"], request["height"])
    vis = vol.ctx.select_visibles(fr, request["height"], sse, 0, vol.vi.depth - 1)
    return fr, order(vol.vi, fr, vis)


def samples_per_ray(request, vol):
    return max(1, int(round(request["rate"] * vol.vmax)))


class RandomWalk:
    """Seeded random-walk camera around the volume: the view direction turns
    by a few degrees and the distance drifts within [1.5, 3] every frame."""

    def __init__(self, seed, width, height):
        self.rng = np.random.default_rng(seed)
        d = self.rng.normal(size=3)
        self.dir = d / np.linalg.norm(d)
        self.r = self.rng.uniform(1.5, 3.0)
        self.width, self.height = width, height

    def next(self):
        step = self.rng.normal(size=3) * 0.04
        d = self.dir + step - np.dot(step, self.dir) * self.dir
        self.dir = d / np.linalg.norm(d)
        self.r = float(np.clip(self.r + self.rng.normal() * 0.03, 1.5, 3.0))
        return look_from(tuple(self.dir * self.r), self.width, self.height)[2]


def serve_sessions(n=32, seed=9, volumes=SERVE_VOLUMES):
    """Sessions: a volume, a viewport between 720p and 4K, a 256-entry TF and
    a random-walk camera each (seed + session)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        w, h = VIEWPORTS[int(rng.integers(len(VIEWPORTS)))]
        out.append(dict(volume=volumes[int(rng.integers(len(volumes)))], width=w, height=h,
                        tf=scenes.piecewise_tf(256, seed=2000 + k, points=6, max_alpha=0.5),
                        walk=RandomWalk(seed + k, w, h)))
    return out


def poisson_arrivals(rate, duration_s, n_sessions, seed):
    """(time_ms, session) of a Poisson process of `rate` requests per second,
    each for a session drawn uniformly."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1000.0 / rate)
        if t > duration_s * 1000.0:
            return out
        out.append((t, int(rng.integers(n_sessions))))


def c4_volume():
    """The out-of-core volume: 8192^3 uint8, 64^3 bricks + 1 halo, depth 8."""
    return volume_info(8192, 64, 1, DT_UINT8)


def c4_source(vi):
    """Its procedural source, blobs plus value noise: 5-octave value noise
    (seed 8, log-normal remap) blended half and half with 64 anisotropic
    Gaussian blobs (seed 8) tabulated on a 64^3-cell lattice (a cell is 128
    finest voxels)."""
    return BlobNoise(vi, noise_params(8, octaves=5, base_cells=8, gain=1.5), 8, count=64,
                     cells=64, weight=128)


def flythrough(t, frames=300, seed=8):
    """Camera of frame t on a seeded Catmull-Rom path that starts outside the
    unit cube, crosses its inside and leaves again; it looks along the path."""
    rng = np.random.default_rng(seed)
    pts = [np.array([-2.2, 0.3, 1.6])]
    for _ in range(4):
        pts.append(rng.uniform(-0.35, 0.35, 3))
    pts.append(np.array([2.0, -0.4, -1.8]))
    pts = [2 * pts[0] - pts[1]] + pts + [2 * pts[-1] - pts[-2]]
    u = (len(pts) - 3) * t / max(1, frames - 1)
    i = min(int(u), len(pts) - 4)
    f = u - i
    p0, p1, p2, p3 = pts[i:i + 4]

    pos = 0.5 * ((2 * p1) + (-p0 + p2) * f + (2 * p0 - 5 * p1 + 4 * p2 - p3) * f * f +
                 (-p0 + 3 * p1 - 3 * p2 + p3) * f * f * f)
    tangent = 0.5 * ((-p0 + p2) + 2 * (2 * p0 - 5 * p1 + 4 * p2 - p3) * f +
                     3 * (-p0 + 3 * p1 - 3 * p2 + p3) * f * f)
    return pos, pos + tangent / max(np.linalg.norm(tangent), 1e-9)
