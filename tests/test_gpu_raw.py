"""Raw and NRRD volumes on the GPU: the reference's nucleon volume (one
brick, as RawDataSource serves it) rendered against the oracle, and the LOD
pyramid cut into pool slots against the oracle's restatement of it."""
import os

import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

from scene_kit import diff_stats

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NRRD = os.path.join(GOLDEN, "nucleon.nrrd")
ORC_DT = {livre_amd.DT_UINT8: oracle.DT_UINT8, livre_amd.DT_UINT16: oracle.DT_UINT16,
          livre_amd.DT_FLOAT: oracle.DT_FLOAT}


def _ovi(vi):
    return oracle.volume_info(list(vi.voxels), list(vi.block), list(vi.overlap),
                              ORC_DT[vi.data_type])


def _render_check(vi, ctx, bricks, eye, w, h, tf, rng, n_samples):
    mv, proj = scenes.camera(eye, w, h)
    fr = livre_amd.frustum(mv, proj)
    ovi = _ovi(vi)
    vis = ctx.select_visibles(fr, h, 1.0, 0, vi.depth - 1)
    assert np.array_equal(vis, oracle.select_visibles(ovi, mv, proj, h, 1.0, 0,
                                                      ovi.depth - 1))
    ordered = livre_amd.order(vi, fr, vis)
    ctx.render(fr, w, h, ordered, livre_amd.render_params(n_samples))
    img, samples = ctx.read_frame()
    cam = oracle.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                        fr.far_plane, w, h)
    ref, ref_samples = oracle.render(ovi, cam, ordered, [bricks(int(i)) for i in ordered],
                                     tf, oracle.render_params(n_samples, 0.999, rng), "f32")
    mx, mean = diff_stats(img, oracle.to_rgba8(ref))
    assert mx <= 2 and mean <= 0.5, (mx, mean)
    assert abs(int(samples) - int(ref_samples)) <= 1e-4 * max(1, ref_samples)
    assert ref_samples > 0
    return img


def test_nucleon_renders_as_the_oracle(lvr, orc):
    with livre_amd.RawSource("raw://" + NRRD) as src:
        vi = src.vi
        brick = src.brick()
        root = scenes.node_id(0, 0, 0, 0)
        with livre_amd.Context(vi, 2 * livre_amd.slot_bytes(vi)) as ctx:
            ctx.upload([root], brick[None])
            tf = scenes.piecewise_tf(256, seed=7, max_alpha=0.4)
            ctx.set_tf(tf, (0.0, 255.0))
            for k, eye in enumerate(((1.4, 0.8, 1.1), (-1.2, -0.6, 1.5))):
                _render_check(vi, ctx, lambda i: brick, eye, 96, 80, tf, (0.0, 255.0), 128)


@pytest.mark.parametrize("dtype, out", [("uint8", None), ("uint16", None),
                                        ("uint16", "uint8"), ("float", None)])
def test_pyramid_bricks_match_the_oracle(lvr, orc, tmp_path, dtype, out):
    rng = np.random.default_rng(11)
    shape = (37, 40, 45)  # z, y, x: padded to the octree's cubic extent
    np_in = {"uint8": np.uint8, "uint16": np.uint16, "float": np.float32}[dtype]
    if dtype == "float":
        vol = rng.standard_normal(shape).astype(np.float32)
    else:
        vol = rng.integers(0, np.iinfo(np_in).max + 1, shape, dtype=np_in)
    p = str(tmp_path / "v.raw")
    vol.tofile(p)
    uri = p + ("?output=%s" % out if out else "") + "#45,40,37," + dtype
    with livre_amd.RawSource(uri) as src:
        vi = src.pyramid_info(16, 1)
        expect = vol if out is None else (vol >> 8).astype(np.uint8)
        levels = oracle.pyramid_levels(expect, vi.depth, vi.root_blocks[0], 16)
        ids = scenes.all_nodes(vi)
        with livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids)) as ctx:
            ctx.load_pyramid(src, ids)
            for nid in ids:
                want = oracle.pyramid_brick(levels, 18, 1, int(nid))
                got = ctx.download(nid)
                assert np.array_equal(got, want), (dtype, out, scenes.decode_node(nid))
            # loading again reuses the levels; releasing frees them
            ctx.load_pyramid(src, ids[:3])
            ctx.release_pyramid()


def test_pyramid_renders_as_the_oracle(lvr, orc, tmp_path):
    with livre_amd.RawSource("raw://" + NRRD) as src:
        vi = src.pyramid_info(16, 1)
        expect = src.brick()
        levels = oracle.pyramid_levels(expect, vi.depth, vi.root_blocks[0], 16)
        ids = scenes.all_nodes(vi)
        with livre_amd.Context(vi, livre_amd.slot_bytes(vi) * len(ids)) as ctx:
            ctx.load_pyramid(src, ids)
            tf = scenes.piecewise_tf(256, seed=9, max_alpha=0.4)
            ctx.set_tf(tf, (0.0, 255.0))
            _render_check(vi, ctx, lambda i: oracle.pyramid_brick(levels, 18, 1, i),
                          (1.3, 0.9, 1.2), 96, 80, tf, (0.0, 255.0), 128)
