"""Closed forms that pin the oracle's transfer-function lookup and histogram
binning independently of any code written for this project.

- TF lookup: texture(transferFnTex, density) on a 1D GL_LINEAR,
  CLAMP_TO_EDGE texture of N entries (RayCastRenderer.cpp:175-193,
  fragRayCast.glsl:199) reads u = density * N - 0.5, blends entries floor(u)
  and floor(u) + 1 by the fraction, and clamps both indices to [0, N - 1]
  (GL 4.2 core 3.9.11). On a constant volume the image is src.rgb times the
  accumulated alpha, so colour / alpha is the filtered TF colour itself: a
  2-entry TF gives lerp(e0, e1, clamp(2 d - 0.5, 0, 1)); a 256-entry ramp
  with entry i = i / 255 gives clamp(256 d - 0.5, 0, 255) / 255. A missing
  half-texel offset, an index clamped at the wrong end or a transposed blend
  weight each fail one of these.
- Float histogram (HistogramObject.cpp:30-72, binDataSlow): 256 bins over
  [range0, range1] of width (range1 - range0) / 255 (the bin count minus
  one), bin = (value - min) / width truncated. Values placed at the centre
  of hand-computed bins land in those bins; a width of range / 256 moves
  every value above the middle one bin up.
- 8- and 16-bit histograms: np.bincount of the interior voxels of each brick
  times 8^(depth - level - 1) (HistogramObject.cpp:157-161), brute force.
"""
import numpy as np
import pytest

from livre_amd import scenes


def _cam(orc, mv, proj, w, h):
    from livre_amd import frustum
    f = frustum(mv, proj)
    return orc.camera(list(f.inv_proj), list(f.inv_mv), list(f.eye), f.near_plane,
                      f.far_plane, w, h)


def _constant_render(orc, value, tf, data_range, precision):
    """A 16^3 uint8 volume of one value, one 16^3 brick + 1 halo, 24 x 24
    pixels; returns (image, samples)."""
    vi = orc.volume_info([16] * 3, [16] * 3, [1] * 3, orc.DT_UINT8)
    mv, proj = scenes.camera((1.3, 0.6, 1.0), 24, 24, 45.0)
    cam = _cam(orc, mv, proj, 24, 24)
    brick = np.full((18, 18, 18), value, np.uint8)
    ids = np.array([0], np.uint64)
    p = orc.render_params(32, 2.0, data_range)
    return orc.render(vi, cam, ids, [brick], np.asarray(tf, np.float32), p, precision)


def _filtered_colour(img):
    a = img[..., 3].astype(np.float64)
    hit = a > 1e-3
    assert hit.sum() > 50
    return img[..., :3][hit].astype(np.float64) / a[hit, None]


# densities below 0, below the first texel centre, inside, past the last
# texel centre and above 1
DENSITIES = [-0.4, 0.0, 0.1, 0.25, 0.3, 0.45, 0.5, 0.62, 0.75, 0.9, 1.0, 1.35]


@pytest.mark.parametrize("d", DENSITIES)
def test_tf_two_entries_closed_form(orc, d):
    e0 = np.array([0.9, 0.1, 0.3], np.float64)
    e1 = np.array([0.2, 0.7, 0.95], np.float64)
    tf = np.zeros((2, 4), np.float32)
    tf[0, :3], tf[1, :3] = e0, e1
    tf[:, 3] = 0.04  # constant alpha: every ray keeps a recoverable colour
    value = 77.0
    # data range placing the constant value at density d exactly in f64
    r0 = value - 100.0 * d
    img, _ = _constant_render(orc, int(value), tf, (r0, r0 + 100.0), "f64")
    # the density as the shader forms it: value * (1 / (r1 - r0)) - r0 / (r1 - r0)
    dd = value * (1.0 / 100.0) + (-np.float64(np.float32(r0)) / 100.0)
    w = min(max(2.0 * dd - 0.5, 0.0), 1.0)
    want = e0 * (1.0 - w) + e1 * w
    got = _filtered_colour(img)
    assert np.abs(got - want).max() < 1e-6, (d, got[0], want)


@pytest.mark.parametrize("value", [0, 1, 2, 37, 128, 200, 254, 255])
def test_tf_ramp_256_closed_form(orc, value):
    tf = np.zeros((256, 4), np.float32)
    ramp = np.arange(256, dtype=np.float64) / 255.0
    tf[:, 0] = ramp
    tf[:, 1] = 1.0 - ramp
    tf[:, 2] = 0.5
    tf[:, 3] = 0.05
    img, _ = _constant_render(orc, value, tf, (0.0, 255.0), "f64")
    d = value / 255.0
    u = min(max(256.0 * d - 0.5, 0.0), 255.0)
    got = _filtered_colour(img)
    # entry values are float32, so the expected colour is good to ~1e-7
    assert np.abs(got[:, 0] - u / 255.0).max() < 1e-6
    assert np.abs(got[:, 1] - (1.0 - u / 255.0)).max() < 1e-6


def test_tf_lookup_alpha_follows_the_filter(orc):
    """Alpha is filtered like colour before the opacity correction: the
    image of a 2-entry TF whose alphas differ equals the image of the
    constant TF holding the blended entry."""
    tf = np.array([[0.3, 0.6, 0.9, 0.02], [0.8, 0.2, 0.1, 0.3]], np.float32)
    r0, r1 = -23.0, 177.0  # value 77 -> density 0.5 -> w = 0.5
    img, s = _constant_render(orc, 77, tf, (r0, r1), "f64")
    mid = (tf[0].astype(np.float64) + tf[1]) / 2
    flat = np.tile(mid.astype(np.float32), (2, 1))
    ref, s_ref = _constant_render(orc, 77, flat, (r0, r1), "f64")
    assert s == s_ref
    assert np.abs(img - ref).max() < 1e-6


def _float_brick_volume(orc, values, block=8):
    vi = orc.volume_info([block] * 3, [block] * 3, [1] * 3, orc.DT_FLOAT)
    m = block + 2
    brick = np.full((m, m, m), np.float32(values[0]), np.float32)
    interior = brick[1:-1, 1:-1, 1:-1].reshape(-1)
    interior[:len(values)] = values
    brick[1:-1, 1:-1, 1:-1] = interior.reshape((block,) * 3)
    # halo voxels carry a value outside any bin tested, which must not count
    brick[0, :, :] = 1e6
    return vi, brick


@pytest.mark.parametrize("rng", [(0.0, 1.0), (-1.0, 1.0), (10.0, 522.0)])
def test_float_histogram_bin_width(orc, rng):
    lo, hi = rng
    width = (hi - lo) / 255.0
    bins = [0, 1, 63, 127, 128, 129, 200, 254]
    # centres of the chosen bins; the first value fills the rest of the brick
    values = np.array([lo + (b + 0.5) * width for b in bins], np.float32)
    vi, brick = _float_brick_volume(orc, values)
    h = orc.histogram(vi, [0], [brick], 256, rng)
    want = np.zeros(256, np.uint64)
    for b in bins:
        want[b] += 1
    want[bins[0]] += 8 ** 3 - len(bins)
    assert np.array_equal(h, want), np.nonzero(h)[0]


def _bincount_histogram(vi, ids, bricks, bins):
    ov = vi.overlap[0]
    out = np.zeros(bins, np.uint64)
    for nid, b in zip(ids, bricks):
        level = int(nid) & 0xF
        scale = (1 << (vi.depth - level - 1)) ** 3
        inner = b[ov:-ov, ov:-ov, ov:-ov] if ov else b
        out += np.bincount(inner.reshape(-1).astype(np.int64),
                           minlength=bins).astype(np.uint64) * np.uint64(scale)
    return out


@pytest.mark.parametrize("dt,bins", [("u8", 256), ("u16", 65536)])
def test_integer_histogram_brute_force(orc, dt, bins):
    rng = np.random.default_rng(11)
    np_t = np.uint8 if dt == "u8" else np.uint16
    ot = orc.DT_UINT8 if dt == "u8" else orc.DT_UINT16
    vi = orc.volume_info([32] * 3, [8] * 3, [1] * 3, ot)
    assert vi.depth == 3
    ids = [scenes.node_id(0, 0, 0, 0), scenes.node_id(1, 1, 0, 1),
           scenes.node_id(2, 3, 2, 0), scenes.node_id(2, 0, 3, 3)]
    bricks = []
    for k in range(len(ids)):
        b = rng.integers(0, bins, (10, 10, 10)).astype(np_t)
        if k == 1:
            b[:] = bins - 1  # the last bin
        bricks.append(b)
    got = orc.histogram(vi, ids, bricks, bins, (0.0, float(bins - 1)))
    want = _bincount_histogram(vi, ids, bricks, bins)
    assert np.array_equal(got, want)
    # every interior voxel counted once, scaled by its level
    assert int(got.sum()) == sum(8 ** 3 * (1 << (3 - (int(i) & 0xF) - 1)) ** 3 for i in ids)
