"""Host-side logic of liblivre_amd (no GPU needed): octree shape, node boxes,
frustum, brick order and the procedural data source, against the oracle; and
the C ABI exports every declared symbol."""
import os
import re

import numpy as np
import pytest

import livre_amd
from livre_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(lvr):
    with open(os.path.join(ROOT, "include", "livre_amd.h")) as f:
        header = f.read()
    declared = set(re.findall(r"^\s*(?:int|void|double|size_t|const char\*)\s+(lvr_\w+)\(",
                              header, re.M))
    assert declared == set(lvr.EXPORTS)
    L = lvr.lib()
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("voxels,block,overlap", [
    ((256, 256, 256), 32, 1), ((2048, 2048, 2048), 64, 1), ((1536,) * 3, 48, 1),
    ((2048, 2099, 2048), 64, 0), ((1024, 1024, 512), 32, 4), ((8192,) * 3, 64, 1),
    ((100, 37, 250), 16, 2)])
def test_volume_info_matches_oracle(lvr, orc, voxels, block, overlap):
    a = lvr.volume_info(voxels, block, overlap, lvr.DT_UINT16)
    b = orc.volume_info(voxels, [block] * 3, [overlap] * 3, orc.DT_UINT16)
    assert a.depth == b.depth
    assert list(a.root_blocks) == list(b.root_blocks)
    assert list(a.world_size) == list(b.world_size)
    assert a.world_space_per_voxel == b.world_space_per_voxel


def test_world_boxes_match_oracle(lvr, orc):
    rng = np.random.default_rng(0)
    a = lvr.volume_info((2048, 2099, 1500), 64, 1, lvr.DT_UINT16)
    b = orc.volume_info((2048, 2099, 1500), [64] * 3, [1] * 3, orc.DT_UINT16)
    for _ in range(500):
        lv = int(rng.integers(0, a.depth))
        pos = [int(rng.integers(0, a.root_blocks[i] << lv)) for i in range(3)]
        nid = scenes.node_id(lv, *pos)
        assert np.array_equal(lvr.world_box(a, nid), orc.world_box(b, nid))


def _random_cameras(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = rng.uniform(0.9, 3.0)
        w, h = int(rng.integers(64, 2048)), int(rng.integers(64, 2048))
        yield scenes.camera(tuple(d * r), w, h, float(rng.uniform(30, 90))), h


def test_frustum_matches_oracle(lvr, orc):
    for (mv, proj), _h in _random_cameras(50, 1):
        f = lvr.frustum(mv, proj)
        g = orc.frustum_info(mv, proj)
        assert f.near_plane == g.near_plane and f.far_plane == g.far_plane
        assert f.top == g.top and f.bottom == g.bottom
        assert list(f.planes[4]) == list(g.near_plane_eq)
        assert list(f.eye) == pytest.approx(list(g.eye), abs=1e-6)
        assert list(f.view_dir) == pytest.approx(list(g.view_dir), abs=1e-6)


def test_order_matches_oracle(lvr, orc):
    a = lvr.volume_info(512, 32, 1, lvr.DT_UINT8)
    b = orc.volume_info([512] * 3, [32] * 3, [1] * 3, orc.DT_UINT8)
    for (mv, proj), h in _random_cameras(10, 2):
        vis = orc.select_visibles(b, mv, proj, h, 1.0, 0, b.depth - 1)
        f = lvr.frustum(mv, proj)
        assert np.array_equal(lvr.order(a, f, vis), orc.order(b, mv, vis))


def test_order_is_front_to_back(lvr):
    a = lvr.volume_info(256, 32, 1, lvr.DT_UINT8)
    mv, proj = scenes.camera((1.5, 0.7, -1.1), 512, 512)
    f = lvr.frustum(mv, proj)
    ids = scenes.all_nodes(a)
    o = lvr.order(a, f, ids)
    m = np.asarray(mv, np.float64).reshape(4, 4).T
    d = []
    for nid in o:
        box = lvr.world_box(a, nid).astype(np.float64)
        c = np.append((box[:3] + box[3:]) / 2, 1.0)
        d.append(np.linalg.norm((m @ c)[:3]))
    assert np.all(np.diff(d) >= -1e-6)


def test_noise_bricks_are_deterministic_and_consistent(lvr):
    vi = lvr.volume_info(128, 32, 1, lvr.DT_UINT16)
    p = lvr.noise_params(3, octaves=5, base_cells=8)
    a = lvr.noise_brick_host(vi, p, scenes.node_id(2, 1, 2, 3))
    b = lvr.noise_brick_host(vi, p, scenes.node_id(2, 1, 2, 3))
    assert np.array_equal(a, b)
    assert a.min() < a.max()
    # the halo of a brick equals the interior of its neighbour at that level
    c = lvr.noise_brick_host(vi, p, scenes.node_id(2, 2, 2, 3))
    assert np.array_equal(a[:, :, -1], c[:, :, 1])
    assert np.array_equal(a[:, :, -2], c[:, :, 0])
    # the remap spans the type's range and is monotone
    r = np.array(p.remap)
    assert r[0] == 0 and r[-1] == 65535 and np.all(np.diff(r.astype(np.int64)) >= 0)


def test_noise_types(lvr):
    for dt, np_t in ((lvr.DT_UINT8, np.uint8), (lvr.DT_FLOAT, np.float32)):
        vi = lvr.volume_info(64, 16, 1, dt)
        p = lvr.noise_params(5, octaves=3, base_cells=4)
        b = lvr.noise_brick_host(vi, p, scenes.node_id(1, 1, 0, 1))
        assert b.dtype == np_t and b.shape == (18, 18, 18)
        if dt == lvr.DT_FLOAT:
            assert b.min() >= -1.0 and b.max() <= 1.0


def test_field_sources_are_deterministic_and_consistent(lvr):
    for kind, dt, count in ((lvr.FIELD_BLOBS, lvr.DT_UINT8, 64),
                            (lvr.FIELD_BLOBS, lvr.DT_UINT16, 16),
                            (lvr.FIELD_SINES, lvr.DT_FLOAT, 8)):
        vi = lvr.volume_info(128, 32, 1, dt)
        p = lvr.field_params(kind, count, 21, noise=0.1 if kind == lvr.FIELD_BLOBS else 0.0)
        a = lvr.field_brick_host(vi, p, scenes.node_id(2, 1, 2, 1))
        assert np.array_equal(a, lvr.field_brick_host(vi, p, scenes.node_id(2, 1, 2, 1)))
        assert a.min() < a.max()
        c = lvr.field_brick_host(vi, p, scenes.node_id(2, 2, 2, 1))
        assert np.array_equal(a[:, :, -1], c[:, :, 1])
        assert np.array_equal(a[:, :, -2], c[:, :, 0])
        if kind == lvr.FIELD_SINES:
            assert a.min() >= -1.0 and a.max() <= 1.0
    # another seed, another volume
    vi = lvr.volume_info(64, 16, 1, lvr.DT_UINT8)
    x = lvr.field_brick_host(vi, lvr.field_params(lvr.FIELD_BLOBS, 32, 1, 0.1), 0)
    y = lvr.field_brick_host(vi, lvr.field_params(lvr.FIELD_BLOBS, 32, 2, 0.1), 0)
    assert not np.array_equal(x, y)
    with pytest.raises(lvr.LivreError):
        lvr.field_brick_host(vi, lvr.field_params(7, 32, 1), 0)


def test_blob_noise_source(lvr):
    """Blobs plus value noise: weight 0 is the value-noise source exactly; a
    weight w brick is the integer blend (blob * w + noise * (256 - w) + 128)
    >> 8 of the weight-256 and weight-0 bricks; the blob part is the
    trilinear interpolation of a lattice of 32-voxel cells, so along x it is
    affine inside each cell (second differences within the rounding of three
    fixed-point lerps) and bends only where a voxel centre crosses a cell
    boundary; 8-bit and float volumes convert the same 16-bit values."""
    vi = lvr.volume_info(256, 32, 1, lvr.DT_UINT16)
    noise = lvr.noise_params(5)
    src = {w: lvr.BlobNoise(vi, noise, 11, count=32, cells=8, weight=w) for w in (0, 96, 256)}
    for nid in (0, scenes.node_id(1, 1, 0, 1), scenes.node_id(3, 1, 2, 3)):
        n0, n96, blob = (src[w].brick_host(nid) for w in (0, 96, 256))
        assert np.array_equal(n0, lvr.noise_brick_host(vi, noise, nid))
        mix = (blob.astype(np.uint32) * 96 + n0.astype(np.uint32) * 160 + 128) >> 8
        assert np.array_equal(n96, mix)
    # finest level: voxel x = 31 .. 64 of the brick, centres (x + 0.5) / 32 cells
    blob = src[256].brick_host(scenes.node_id(3, 1, 2, 3)).astype(np.int64)
    x = np.arange(31, 65)
    cell = (x * 2 + 1) // 64
    d2 = blob[:, :, 2:] - 2 * blob[:, :, 1:-1] + blob[:, :, :-2]
    inside = (cell[2:] == cell[:-2])
    assert np.abs(d2[:, :, inside]).max() <= 8
    assert np.abs(np.diff(blob, axis=2)).max() > 100  # the field is not flat here
    # the slope changes from cell to cell (a smoothstep fade of the fraction
    # would bend inside the cells instead)
    assert np.abs(d2[:, :, ~inside]).max() > 8
    for dt in (lvr.DT_UINT8, lvr.DT_FLOAT):
        v2 = lvr.volume_info(256, 32, 1, dt)
        b = lvr.BlobNoise(v2, noise, 11, count=32, cells=8, weight=96).brick_host(
            scenes.node_id(3, 1, 2, 3))
        ref = src[96].brick_host(scenes.node_id(3, 1, 2, 3))
        if dt == lvr.DT_UINT8:
            assert np.array_equal(b, (ref >> 8).astype(np.uint8))
        else:
            assert np.array_equal(b, ref.astype(np.float32) * np.float32(2.0 / 65535.0)
                                  - np.float32(1.0))
    for bad in (dict(count=0), dict(cells=0), dict(weight=257), dict(cells=2000)):
        args = dict(count=32, cells=8, weight=96)
        args.update(bad)
        with pytest.raises(lvr.LivreError):
            lvr.BlobNoise(vi, noise, 11, **args)


def test_invalid_arguments_are_rejected(lvr):
    with pytest.raises(lvr.LivreError):
        lvr.volume_info(64, 0, 0, lvr.DT_UINT8)
    vi = lvr.volume_info(64, 16, 1, lvr.DT_UINT8)
    with pytest.raises(lvr.LivreError):
        lvr.world_box(vi, scenes.node_id(7, 0, 0, 0))


def _noise_voxel_py(p, vi, level, g):
    """Per-voxel definition of the value-noise source in Python integers:
    splitmix64 lattice hash, smoothstep fade and 16-bit lerps per octave
    (x, then y, then z), band limit by level, remap table."""
    M = (1 << 64) - 1

    def splitmix(x):
        x = (x + 0x9E3779B97F4A7C15) & M
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        return x ^ (x >> 31)

    def lattice(o, x, y, z):
        key = (x & 0xFFFFF) | ((y & 0xFFFFF) << 20) | ((z & 0xFFFFF) << 40) | (o << 60)
        return splitmix(p.seed ^ splitmix(key)) >> 48

    def fade(f):
        return (f * f * (3 * 65536 - 2 * f)) >> 32

    def lerp(a, b, s):
        return a + (((b - a) * s) >> 16)  # Python >> floors like the arithmetic shift

    vmax = max(vi.voxels)
    S = 1 << (vi.depth - 1 - level)
    u = []
    for a in range(3):
        c = (2 * g[a] + 1) * S * 32768
        u.append(min(max(c, 32768), (2 * vi.voxels[a] - 1) * 32768))
    octaves = 0
    m = 0
    for o in range(p.octaves):
        m += 65535 >> o
        cells = p.base_cells << o
        if vmax // cells >= 2 * S or o == 0:
            octaves = o + 1
    total = 0
    for o in range(octaves):
        mult = ((p.base_cells << o) << 32) // vmax
        ci, fr = [], []
        for a in range(3):
            c = (u[a] * mult) >> 32
            ci.append(c >> 16)
            fr.append(fade(c & 0xFFFF))
        v = [lattice(o, ci[0] + (k & 1), ci[1] + ((k >> 1) & 1), ci[2] + ((k >> 2) & 1))
             for k in range(8)]
        x00, x10 = lerp(v[0], v[1], fr[0]), lerp(v[2], v[3], fr[0])
        x01, x11 = lerp(v[4], v[5], fr[0]), lerp(v[6], v[7], fr[0])
        y0, y1 = lerp(x00, x10, fr[1]), lerp(x01, x11, fr[1])
        total += lerp(y0, y1, fr[2]) >> o
    seg_mult = (4096 << 32) // (m + 1)
    t = (total * seg_mult) >> 16
    seg, f = min(t >> 16, 4095), t & 0xFFFF
    a, b = p.remap[seg], p.remap[seg + 1]
    return min(max(a + (((b - a) * f) >> 16), 0), 65535)


def test_host_noise_bricks_follow_the_per_voxel_definition(lvr):
    rng = np.random.default_rng(5)
    vi = lvr.volume_info((200, 160, 128), 32, 1, lvr.DT_UINT16)
    p = lvr.noise_params(77, octaves=5, base_cells=3)
    for level in range(vi.depth):
        n = [max(1, int(np.ceil(vi.voxels[a] / (32 << (vi.depth - 1 - level)))))
             for a in range(3)]
        pos = [int(rng.integers(n[a])) for a in range(3)]
        nid = scenes.node_id(level, *pos)
        brick = lvr.noise_brick_host(vi, p, nid)
        for _ in range(40):
            x, y, z = (int(rng.integers(34)) for _ in range(3))
            g = (pos[0] * 32 - 1 + x, pos[1] * 32 - 1 + y, pos[2] * 32 - 1 + z)
            assert brick[z, y, x] == _noise_voxel_py(p, vi, level, g), (level, x, y, z)


def test_jpeg_max_size_bounds_the_worst_block(lvr):
    # 1664 bits (22 DC + 63 x 26 AC, rounded to 32) per 8x8 block and
    # component, doubled for byte stuffing, plus the JFIF header and EOI
    L = lvr.lib()
    blocks = 3 * ((1920 + 7) // 8) * ((1080 + 7) // 8)
    assert L.lvr_jpeg_max_size(1920, 1080) == 1024 + blocks * 208 * 2 + 2
    assert L.lvr_jpeg_max_size(1, 1) == 1024 + 3 * 208 * 2 + 2
    assert L.lvr_jpeg_max_size(0, 5) == 0 and L.lvr_jpeg_max_size(5, -1) == 0


def test_oracle_jpeg_worst_case_fits_the_bound(lvr, orc):
    # noise at quality 100 is near the worst case per block
    import numpy as np
    rng = np.random.default_rng(9)
    px = rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)
    px[..., :3] = np.where(rng.random((64, 64, 3)) < 0.5, 0, 255)
    assert len(orc.jpeg_encode(px, quality=100)) <= lvr.lib().lvr_jpeg_max_size(64, 64)


def test_samples_per_ray_rule_host(lvr):
    """RayCastRenderer.cpp:216-234, host side: maxVoxelDim / 2^(depth - maxLOD - 1), at least 512."""
    vi = livre_amd.volume_info(2048, 64, 1, livre_amd.DT_UINT16)
    assert vi.depth == 6
    nid = scenes.node_id
    assert livre_amd.samples_per_ray(vi, [nid(5, 3, 4, 5)]) == 2048
    assert livre_amd.samples_per_ray(vi, [nid(2, 0, 0, 0), nid(4, 1, 1, 1)]) == 1024
    assert livre_amd.samples_per_ray(vi, [nid(1, 0, 0, 0), nid(2, 1, 1, 1)]) == 512
    assert livre_amd.samples_per_ray(vi, []) == 512
    assert livre_amd.samples_per_ray(vi, [nid(5, 0, 0, 0)], requested=300) == 300
    small = livre_amd.volume_info([300, 200, 100], 32, 1, livre_amd.DT_UINT8)
    top = small.depth - 1
    assert livre_amd.samples_per_ray(small, [nid(top, 0, 0, 0)]) == 512
    with pytest.raises(livre_amd.LivreError):
        livre_amd.samples_per_ray(vi, [nid(6, 0, 0, 0)])


def test_host_source_builds_agree(lvr):
    """The host brick source is built for the x86-64 baseline and for AVX2,
    picked at run time; LIVRE_AMD_HOST_BASELINE=1 forces the baseline. Both
    give the same bricks (value noise and blobs plus value noise, three
    voxel types, several levels)."""
    import subprocess
    import sys
    code = r'''
import sys, numpy as np
sys.path.insert(0, %r)
import livre_amd as lvr
from livre_amd import scenes
out = []
for dt in (lvr.DT_UINT8, lvr.DT_UINT16, lvr.DT_FLOAT):
    vi = lvr.volume_info(512, 64, 1, dt)
    p = lvr.noise_params(7, octaves=5, base_cells=8, gain=1.5)
    bn = lvr.BlobNoise(vi, p, 3, count=16, cells=8, weight=128)
    for nid in (0, scenes.node_id(2, 1, 3, 2), scenes.node_id(3, 5, 0, 7)):
        out.append(lvr.noise_brick_host(vi, p, nid).tobytes())
        out.append(bn.brick_host(nid).tobytes())
sys.stdout.buffer.write(b"".join(out))
''' % ROOT
    runs = []
    for env in ({}, {"LIVRE_AMD_HOST_BASELINE": "1"}):
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, timeout=300,
                           env=dict(os.environ, **env))
        assert r.returncode == 0, r.stderr.decode()[-2000:]
        runs.append(r.stdout)
    assert len(runs[0]) > 0 and runs[0] == runs[1]
