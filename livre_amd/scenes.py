"""Synthetic scenes for the render path: cameras, transfer functions and
procedural volumes cut into Livre bricks.

Cameras follow the conventions of the reference: a look-at modelview and an
OpenGL perspective projection, near/far chosen as Channel::Impl::setupFrustum
does from the model centre (livre/eq/Channel.cpp:185-213): near =
max(minNear, d - 1), far = max(2 near, d + 1) for a model centre at eye
distance d.
"""
import math

import numpy as np

from . import (DT_FLOAT, DT_UINT8, DT_UINT16, NP_DTYPE, noise_brick_host,
               noise_params, volume_info)


# ------------------------------------------------------------------ cameras
def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Column-major float32 modelview."""
    eye = np.asarray(eye, np.float64)
    target = np.asarray(target, np.float64)
    f = target - eye
    f /= np.linalg.norm(f)
    up = np.asarray(up, np.float64)
    if abs(np.dot(f, up / np.linalg.norm(up))) > 0.9999:
        up = np.array([0.0, 0.0, 1.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3] = s
    m[1, :3] = u
    m[2, :3] = -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m.T.reshape(-1).astype(np.float32)  # column-major


def perspective(fovy_deg, aspect, near, far):
    f = 1.0 / math.tan(math.radians(fovy_deg) / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = (far + near) / (near - far)
    m[2, 3] = 2.0 * far * near / (near - far)
    m[3, 2] = -1.0
    return m.T.reshape(-1).astype(np.float32)


def near_far(eye, centre=(0.0, 0.0, 0.0), min_near=1e-4):
    d = float(np.linalg.norm(np.asarray(eye, np.float64) - np.asarray(centre)))
    near = max(min_near, d - 1.0)
    far = max(2.0 * near, d + 1.0)
    return near, far


def camera(eye, width, height, fovy=60.0, target=(0.0, 0.0, 0.0), up=(0, 1, 0)):
    """(mv, proj) for a camera at eye looking at target."""
    near, far = near_far(eye, target)
    return look_at(eye, target, up), perspective(fovy, width / height, near, far)


def orbit(k, n, radius=2.0, height=0.0):
    """Eye position k of n evenly spaced on a horizontal orbit."""
    t = 2.0 * math.pi * k / n
    return (radius * math.cos(t), height, radius * math.sin(t))


def orbit_zoom(t, frames=120):
    """Orbit-and-zoom path of the 2048^3 config: theta = 2 pi t / frames,
    r = 2 - sin^2(pi t / frames)."""
    th = 2.0 * math.pi * t / frames
    r = 2.0 - math.sin(math.pi * t / frames) ** 2
    return (r * math.cos(th), 0.35 * r * math.sin(2 * th) * 0.5, r * math.sin(th))


# ------------------------------------------------------- transfer functions
def piecewise_tf(n, seed, points=6, max_alpha=0.6):
    """Seeded piecewise-linear RGBA transfer function of n float entries."""
    rng = np.random.default_rng(seed)
    xs = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, points - 2)), [1.0]])
    cols = rng.uniform(0, 1, (points, 3))
    alpha = rng.uniform(0, max_alpha, points)
    alpha[0] = 0.0
    grid = np.linspace(0.0, 1.0, n)
    out = np.zeros((n, 4), np.float32)
    for c in range(3):
        out[:, c] = np.interp(grid, xs, cols[:, c])
    out[:, 3] = np.interp(grid, xs, alpha)
    return out


def lut_rgba8(tf):
    """TransferFunction1D::getLUT truncation uint8(v * 255)."""
    return (np.asarray(tf, np.float32) * np.float32(255.0)).astype(np.uint8)


# ------------------------------------------------------------------ volumes
def blob_volume(n, seed=1, blobs=64, noise=0.1, dtype=np.uint8):
    """Sum of seeded anisotropic rotated Gaussian blobs plus uniform noise,
    normalised so that the maximum is the type's maximum. Returns (z, y, x)."""
    rng = np.random.default_rng(seed)
    grid = (np.arange(n, dtype=np.float32) + 0.5) / n
    vol = np.zeros((n, n, n), np.float32)
    for _ in range(blobs):
        c = rng.uniform(0.1, 0.9, 3)
        sig = rng.uniform(0.02, 0.12, 3)
        q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
        amp = rng.uniform(0.3, 1.0)
        ext = 3.5 * sig.max()
        lo = np.clip(((c - ext) * n).astype(int), 0, n)
        hi = np.clip(((c + ext) * n).astype(int) + 1, 0, n)
        z = grid[lo[2]:hi[2], None, None] - c[2]
        y = grid[None, lo[1]:hi[1], None] - c[1]
        x = grid[None, None, lo[0]:hi[0]] - c[0]
        acc = 0.0
        for i in range(3):
            comp = (q[i, 0] * x + q[i, 1] * y + q[i, 2] * z) / sig[i]
            acc = acc + comp * comp
        vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += amp * np.exp(-0.5 * acc)
    vol += noise * rng.uniform(0, 1, vol.shape).astype(np.float32)
    top = np.iinfo(dtype).max if np.issubdtype(dtype, np.integer) else 1.0
    vol *= top / vol.max()
    if np.issubdtype(dtype, np.integer):
        return np.clip(np.rint(vol), 0, top).astype(dtype)
    return vol.astype(dtype)


def lod_pyramid(vol, depth):
    """Levels depth-1 (full resolution) down to 0 by 2x2x2 box mean."""
    levels = [None] * depth
    levels[depth - 1] = vol
    cur = vol
    for lv in range(depth - 2, -1, -1):
        z, y, x = cur.shape
        c = cur[: z - z % 2, : y - y % 2, : x - x % 2].astype(np.float64)
        c = c.reshape(z // 2, 2, y // 2, 2, x // 2, 2).mean(axis=(1, 3, 5))
        if np.issubdtype(vol.dtype, np.integer):
            c = np.floor(c + 0.5).astype(vol.dtype)
        else:
            c = c.astype(vol.dtype)
        levels[lv] = c
        cur = c
    return levels


def node_id(level, x, y, z, t=0):
    return (level & 0xF) | ((x & 0x3FFF) << 4) | ((y & 0x3FFF) << 18) | \
        ((z & 0x3FFF) << 32) | ((t & 0x3FFFF) << 46)


def decode_node(nid):
    nid = int(nid)
    return (nid & 0xF, (nid >> 4) & 0x3FFF, (nid >> 18) & 0x3FFF, (nid >> 32) & 0x3FFF)


def all_nodes(vi):
    """Every NodeId of the octree, level by level."""
    out = []
    for lv in range(vi.depth):
        nx, ny, nz = (vi.root_blocks[i] << lv for i in range(3))
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    out.append(node_id(lv, x, y, z))
    return np.array(out, np.uint64)


def cut_brick(level_vol, vi, nid):
    """Brick of (block + 2 overlap)^3 voxels with halo, clamped at the
    volume edge, (z, y, x) with x fastest."""
    lv, x, y, z = decode_node(nid)
    b = vi.block
    ov = [vi.overlap[i] for i in range(3)]
    pos = (x, y, z)
    idx = []
    shape = level_vol.shape[::-1]  # x, y, z sizes
    for a in range(3):
        r = np.arange(pos[a] * b[a] - ov[a], pos[a] * b[a] + b[a] + ov[a])
        idx.append(np.clip(r, 0, shape[a] - 1))
    return level_vol[np.ix_(idx[2], idx[1], idx[0])]


def blob_bricks(vi, vol, ids):
    """Bricks of ids from a full-resolution volume (box-filtered LODs)."""
    levels = lod_pyramid(vol, vi.depth)
    return np.stack([cut_brick(levels[decode_node(i)[0]], vi, i) for i in ids])


def noise_bricks(vi, params, ids):
    return np.stack([noise_brick_host(vi, params, i) for i in ids])


def type_range(dt):
    return {DT_UINT8: (0.0, 255.0), DT_UINT16: (0.0, 65535.0),
            DT_FLOAT: (-1.0, 1.0)}[dt]


__all__ = ["look_at", "perspective", "near_far", "camera", "orbit", "orbit_zoom",
           "piecewise_tf", "lut_rgba8", "blob_volume", "lod_pyramid", "node_id",
           "decode_node", "all_nodes", "cut_brick", "blob_bricks", "noise_bricks",
           "type_range", "volume_info", "noise_params", "NP_DTYPE"]
