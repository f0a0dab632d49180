"""Closed-form pins of the oracle's LOD pyramid restatement (the raw-volume
pyramid has no reference counterpart)."""
import numpy as np

import oracle


def test_constant_volume_is_constant_at_every_level():
    vol = np.full((20, 20, 20), 77, np.uint8)
    levels = oracle.pyramid_levels(vol, depth=3, root_blocks=1, block=8)
    assert [g.shape[0] for g in levels] == [8, 16, 32]
    # inside the volume every level keeps 77 (means of equal values)
    assert np.all(levels[2][:20, :20, :20] == 77) and np.all(levels[2][20:, :, :] == 0)
    assert np.all(levels[1][:10, :10, :10] == 77)
    assert np.all(levels[0][:5, :5, :5] == 77)


def test_mean_rounds_half_up_and_float_order():
    vol = np.zeros((2, 2, 2), np.uint16)
    vol.flat[:] = [0, 1, 1, 1, 1, 1, 1, 0]          # sum 6: 6/8 = 0.75 -> 1
    lv = oracle.pyramid_levels(vol, depth=2, root_blocks=1, block=1)
    assert lv[0][0, 0, 0] == 1
    vol.flat[:] = [0, 0, 0, 0, 1, 1, 1, 1]          # sum 4: 0.5 -> 1 (half up)
    assert oracle.pyramid_levels(vol, 2, 1, 1)[0][0, 0, 0] == 1
    vol.flat[:] = [0, 0, 0, 0, 0, 1, 1, 1]          # sum 3: 0.375 -> 0
    assert oracle.pyramid_levels(vol, 2, 1, 1)[0][0, 0, 0] == 0
    s8 = np.array([-8, -1, 0, 0, 0, 0, 0, 0], np.int8).reshape(2, 2, 2)  # -9/8 -> -1
    assert oracle.pyramid_levels(s8, 2, 1, 1)[0][0, 0, 0] == -1
    f = np.array([1e8, 1, -1e8, 1, 3, 5, 7, 9], np.float32).reshape(2, 2, 2)
    want = np.float32((np.float32(np.float32(1e8) + 1) + np.float32(-1e8 + 1)) +
                      np.float32(np.float32(3 + 5) + np.float32(7 + 9))) * np.float32(0.125)
    assert oracle.pyramid_levels(f, 2, 1, 1)[0][0, 0, 0] == want


def test_finest_bricks_are_the_file_with_halo_from_neighbours():
    rng = np.random.default_rng(1)
    vol = rng.integers(0, 256, (24, 24, 24), dtype=np.uint8)
    levels = oracle.pyramid_levels(vol, depth=3, root_blocks=1, block=8)
    from livre_amd import scenes
    b = oracle.pyramid_brick(levels, 10, 1, scenes.node_id(2, 1, 1, 1))
    assert np.array_equal(b, vol[7:17, 7:17, 7:17])
    # a corner brick clamps its outer halo to the level's edge
    c = oracle.pyramid_brick(levels, 10, 1, scenes.node_id(2, 0, 0, 0))
    assert np.array_equal(c[1:, 1:, 1:], vol[:9, :9, :9])
    assert np.array_equal(c[0, 1:, 1:], vol[0, :9, :9])
    # the root brick of a depth-3 octree is the whole coarsest level
    r = oracle.pyramid_brick(levels, 10, 1, scenes.node_id(0, 0, 0, 0))
    assert np.array_equal(r[1:9, 1:9, 1:9], levels[0])
