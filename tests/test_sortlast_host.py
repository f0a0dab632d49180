"""The static sort-last decomposition on the host (lvr_sortlast_parts and
lvr_sortlast_order): every node lies inside the half-spaces of its part, the
nodes that straddle a split are exactly those whose box crosses it, a node's
part does not depend on the frame, and every ray from the eye meets the
parts in the returned order. World boxes come from the oracle's
internalNodeToLODNode (DataSourcePlugin.cpp:40-71)."""
import numpy as np
import pytest

import oracle
import livre_amd
from livre_amd import scenes

VOLUMES = [((256, 256, 256), 32), ((96, 80, 72), 16), ((300, 200, 120), 20)]


def _planes(ovi):
    """World coordinates of the three splits: the minimum corner of the
    level-1 node at the middle block position."""
    mid = scenes.node_id(1, ovi.root_blocks[0], ovi.root_blocks[1], ovi.root_blocks[2])
    return oracle.world_box(ovi, mid)[:3].astype(np.float64)


@pytest.mark.parametrize("voxels,block", VOLUMES)
@pytest.mark.parametrize("n_parts", [1, 2, 4, 8])
def test_parts_are_half_spaces(voxels, block, n_parts):
    vi = livre_amd.volume_info(voxels, block, 1, livre_amd.DT_UINT8)
    ovi = oracle.volume_info(list(voxels), [block] * 3, [1] * 3, oracle.DT_UINT8)
    ids = scenes.all_nodes(vi)
    parts = livre_amd.sortlast_parts(vi, ids, n_parts)
    axes = {1: 0, 2: 1, 4: 2, 8: 3}[n_parts]
    plane = _planes(ovi)
    straddlers = 0
    for nid, p in zip(ids, parts):
        box = oracle.world_box(ovi, nid).astype(np.float64)
        crosses = [box[a] < plane[a] < box[3 + a] for a in range(axes)]
        if p < 0:
            straddlers += 1
            assert any(crosses), nid
            continue
        assert not any(crosses) and 0 <= p < n_parts, nid
        for a in range(axes):
            bit = (p >> (axes - 1 - a)) & 1
            if bit:
                assert box[a] >= plane[a], (nid, a)
            else:
                assert box[3 + a] <= plane[a], (nid, a)
    if axes == 0:
        assert straddlers == 0
    # same parts for the same ids in another order (no dependence on the frame)
    perm = np.random.default_rng(0).permutation(len(ids))
    assert np.array_equal(livre_amd.sortlast_parts(vi, np.asarray(ids)[perm], n_parts),
                          parts[perm])


def _slab_entry(o, d, lo, hi):
    """Entry parameter of a ray into the box [lo, hi] (inf if missed)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - o) / d
        t1 = (hi - o) / d
    tn = np.nanmax(np.minimum(t0, t1))
    tf = np.nanmin(np.maximum(t0, t1))
    return max(tn, 0.0) if tf >= max(tn, 0.0) else np.inf


@pytest.mark.parametrize("n_parts", [2, 4, 8])
def test_order_is_front_to_back_for_every_ray(n_parts):
    voxels, block = (300, 200, 120), 20
    vi = livre_amd.volume_info(voxels, block, 1, livre_amd.DT_UINT8)
    ovi = oracle.volume_info(list(voxels), [block] * 3, [1] * 3, oracle.DT_UINT8)
    axes = {2: 1, 4: 2, 8: 3}[n_parts]
    plane = _planes(ovi)
    half = np.array(vi.world_size, np.float64) / 2
    rng = np.random.default_rng(5)
    for _ in range(60):
        eye = rng.uniform(-2.0, 2.0, size=3)
        order = list(livre_amd.sortlast_order(vi, eye, n_parts))
        assert sorted(order) == list(range(n_parts))
        boxes = []
        for p in range(n_parts):
            lo, hi = -half.copy() * 3, half.copy() * 3
            for a in range(axes):
                if (p >> (axes - 1 - a)) & 1:
                    lo[a] = plane[a]
                else:
                    hi[a] = plane[a]
            boxes.append((lo, hi))
        for _ in range(40):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            entries = [(_slab_entry(eye, d, lo, hi), p) for p, (lo, hi) in enumerate(boxes)]
            met = [p for t, p in sorted(entries) if np.isfinite(t)]
            pos = [order.index(p) for p in met]
            assert pos == sorted(pos), (eye, d, met, order)


def test_invalid_arguments():
    vi = livre_amd.volume_info(64, 16, 1, livre_amd.DT_UINT8)
    with pytest.raises(livre_amd.LivreError):
        livre_amd.sortlast_parts(vi, [scenes.node_id(0, 0, 0, 0)], 3)
    with pytest.raises(livre_amd.LivreError):
        livre_amd.sortlast_order(vi, (0.0, 0.0, 3.0), 16)
    with pytest.raises(livre_amd.LivreError):
        livre_amd.sortlast_parts(vi, [scenes.node_id(9, 0, 0, 0)], 2)
