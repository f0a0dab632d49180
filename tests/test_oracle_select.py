"""The oracle's SelectVisibles with clip planes and the sort-last range.

- Clip planes (SelectVisibles.cpp:69, _clipPlanes.isOutside; Lexis is not
  vendored, so the rule is the half-space the shader clips away,
  fragRayCast.glsl:152-167): the set with planes equals the set without
  them minus the nodes whose eight corners all satisfy n . x + d < 0,
  evaluated by brute force in double precision. Culling never changes the
  decision of a node it keeps, and a node outside a plane has every
  descendant outside it too, so this is the exact set.
- The range (visitPost, SelectVisibles.cpp:109-131): [r0 * N, r1 * N) of
  the DFS-ordered set with the products in float, truncated; consecutive
  ranges partition the set in order.
"""
import itertools

import numpy as np
import pytest

from livre_amd import scenes

# offsets off the brick grid (multiples of 1/16 and 1/32 would put box
# corners exactly on a plane, where float32 and float64 may round apart)
PLANES = [
    [(1.0, 0.0, 0.0, 0.0137)],                   # keeps x >= -0.0137
    [(0.0, -1.0, 0.0, 0.1093)],                  # keeps y <= 0.1093
    [(0.6, 0.0, 0.8, -0.2071), (0.0, 1.0, 0.0, 0.3129)],
    [(-0.3, 0.5, -0.81, 0.0517)],
]


def _vi(orc):
    return orc.volume_info([256] * 3, [16] * 3, [1] * 3, orc.DT_UINT8)


def _outside(orc, vi, nid, planes):
    """1 when every corner of the node's box is in the clipped half-space of
    some plane, 0 when not, None when a corner lies within 1e-6 of a plane
    that decides it (float32 rounding may go either way there)."""
    b = orc.world_box(vi, int(nid)).astype(np.float64)
    corners = np.array(list(itertools.product(*[(b[a], b[3 + a]) for a in range(3)])))
    near = False
    for p in planes:
        p32 = np.array(p, np.float32).astype(np.float64)
        v = corners @ p32[:3] + p32[3]
        if np.all(v < -1e-6):
            return 1
        near = near or (np.all(v < 1e-6) and np.max(v) > -1e-6)
    return None if near else 0


@pytest.mark.parametrize("planes", PLANES)
@pytest.mark.parametrize("eye", [(1.6, 0.4, 1.2), (-0.5, 1.4, -1.3)])
def test_clip_planes_cull_exactly_the_boxes_outside(orc, planes, eye):
    vi = _vi(orc)
    mv, proj = scenes.camera(eye, 320, 240, 60.0)
    full = orc.select_visibles(vi, mv, proj, 240, 1.0, 0, vi.depth - 1)
    clipped = orc.select_visibles(vi, mv, proj, 240, 1.0, 0, vi.depth - 1, clip_planes=planes)
    verdict = {int(i): _outside(orc, vi, i, planes) for i in full}
    ambiguous = [i for i, v in verdict.items() if v is None]
    assert len(ambiguous) <= 2
    kept = [int(i) for i in clipped]
    assert set(kept) <= set(int(i) for i in full)
    # same DFS order, and exactly the nodes the corners say are not outside
    assert kept == [i for i in (int(x) for x in full) if verdict[i] != 1 and
                    (verdict[i] == 0 or i in set(kept))]
    assert 0 < len(clipped) < len(full)


def test_sort_last_ranges_partition_the_set(orc):
    vi = _vi(orc)
    mv, proj = scenes.camera((1.3, 0.8, 1.5), 320, 240, 60.0)
    full = [int(i) for i in orc.select_visibles(vi, mv, proj, 240, 1.0, 0, vi.depth - 1)]
    n = len(full)
    for k in (2, 3, 7):
        parts = []
        for i in range(k):
            r0, r1 = np.float32(i / k), np.float32((i + 1) / k)
            got = [int(x) for x in orc.select_visibles(vi, mv, proj, 240, 1.0, 0, vi.depth - 1,
                                                       rng=(float(r0), float(r1)))]
            s, e = int(r0 * np.float32(n)), int(r1 * np.float32(n))
            assert got == full[s:e]
            parts += got
        assert parts == full
