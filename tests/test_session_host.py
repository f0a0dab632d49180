"""Camera messages of the session front end (communicator.cpp:106-133
onCamera, 257-283 _createLookOut), checked on the host against closed forms:
with an identity data-to-Livre transform and a ratio of 1 the LookOut is
the modelview; the camera position (the inverse view applied to the origin)
moves from meters to data units and through the data-to-Livre transform
while the rotation stays; the two conversions invert each other."""
import numpy as np
import pytest

import livre_amd
from livre_amd import scenes


def _cm(m):
    """Column-major 16 floats of a 4x4 row-indexed array."""
    return np.asarray(m, np.float64).T.reshape(16)


def _mat(v):
    return np.asarray(v, np.float64).reshape(4, 4).T


def _eye_of(view):
    return np.linalg.inv(view)[:3, 3]


def _views(n=12, seed=0):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        eye = rng.normal(size=3) * rng.uniform(0.5, 3.0)
        target = rng.normal(size=3) * 0.2
        up = (0, 1, 0) if abs(eye[1]) < 0.9 * np.linalg.norm(eye) else (0, 0, 1)
        yield eye, _mat(scenes.look_at(tuple(eye), tuple(target), up))


def _affine(seed):
    rng = np.random.default_rng(seed)
    d = np.diag(list(rng.uniform(0.5, 2.0, size=3)) + [1.0])
    d[:3, 3] = rng.normal(size=3)
    return d


def test_identity_transform_keeps_the_matrix():
    for _eye, view in _views():
        mv = _mat(livre_amd.lookout_to_modelview(_cm(view)))
        assert np.allclose(mv, view, atol=2e-5)
        back = _mat(livre_amd.modelview_to_lookout(_cm(view)))
        assert np.allclose(back, view, atol=2e-5)


@pytest.mark.parametrize("ratio", [1.0, 1e3, 1e-2])
@pytest.mark.parametrize("seed", [1, 2])
def test_camera_position_through_units_and_transform(ratio, seed):
    d = _affine(seed)
    for eye, view in _views(6, seed):
        mv = _mat(livre_amd.lookout_to_modelview(_cm(view), _cm(d), ratio))
        # same rotation, camera at D (ratio * eye)
        assert np.allclose(mv[:3, :3], view[:3, :3], atol=1e-6)
        want = (d @ np.append(eye * ratio, 1.0))[:3]
        assert np.allclose(_eye_of(mv), want, rtol=2e-5, atol=2e-5 * max(1.0, ratio))
        assert np.allclose(mv[3], [0, 0, 0, 1])


@pytest.mark.parametrize("ratio", [1.0, 1e3])
def test_round_trip(ratio):
    d = _affine(5)
    for _eye, view in _views(8, 3):
        mv = livre_amd.lookout_to_modelview(_cm(view), _cm(d), ratio)
        back = _mat(livre_amd.modelview_to_lookout(mv, _cm(d), ratio))
        assert np.allclose(back, view, atol=1e-4)


def test_invalid_matrices():
    with pytest.raises(livre_amd.LivreError):
        livre_amd.lookout_to_modelview(np.zeros(16))
    with pytest.raises(livre_amd.LivreError):
        livre_amd.lookout_to_modelview(np.eye(4).reshape(16), meter_to_data_unit=0.0)
    bad = np.eye(4)
    bad[0, 0] = np.nan
    with pytest.raises(livre_amd.LivreError):
        livre_amd.modelview_to_lookout(bad.reshape(16))
