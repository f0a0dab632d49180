"""Raw and NRRD volumes (RawDataSource): the reference's own test data
(tests/data/nucleon.{raw,nrrd}, copied under tests/golden/) and the
expectations of tests/data/rawDatasource.cpp:37-78, plus the URI rules of
RawDataSource.cpp:62-219. Host only."""
import os

import numpy as np
import pytest

import livre_amd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NRRD = os.path.join(GOLDEN, "nucleon.nrrd")
RAW = os.path.join(GOLDEN, "nucleon.raw")


def _check_nucleon(src):
    # rawDatasource.cpp:41-55: depth 1, one component, uint8, 41^3 voxels,
    # no overlap, the block is the whole volume
    vi = src.vi
    assert vi.depth == 1 and list(vi.root_blocks) == [1, 1, 1]
    assert vi.data_type == livre_amd.DT_UINT8
    assert list(vi.voxels) == [41, 41, 41] and list(vi.overlap) == [0, 0, 0]
    assert list(vi.max_block) == [41, 41, 41]
    assert vi.world_space_per_voxel == np.float32(1 / 41)
    brick = src.brick()
    assert np.array_equal(brick.ravel(), np.fromfile(RAW, np.uint8))


def test_nucleon_nrrd(lvr):
    with livre_amd.RawSource("raw://" + NRRD) as src:
        _check_nucleon(src)
        assert src.raw_info.big_endian == 0 and src.raw_info.data_bytes == 41 ** 3


def test_nucleon_raw(lvr):
    for uri in ("raw://" + RAW + "#41,41,41,uint8", RAW + "#41,41,41",
                RAW + "#41,41,41,unsigned char"):
        with livre_amd.RawSource(uri) as src:
            _check_nucleon(src)


@pytest.mark.parametrize("uri, why", [
    ("http://x/nucleon.raw#41,41,41", "scheme"),
    (os.path.join(GOLDEN, "nucleon.dat") + "#41,41,41", "extension"),
    (RAW + "#41,41", "not enough parameters"),
    (RAW + "#41,41,41,double", "unsupported data format"),
    (RAW + "?output=uint16#41,41,41,uint8", "unsupported data conversion"),
    (RAW + "#41,41,42", "shorter"),
    (os.path.join(GOLDEN, "missing.raw") + "#4,4,4", "cannot open"),
])
def test_uri_errors(lvr, uri, why):
    with pytest.raises(livre_amd.LivreError, match=why):
        livre_amd.RawSource(uri)


def test_output_type_keeps_high_bits(lvr, tmp_path):
    rng = np.random.default_rng(3)
    v16 = rng.integers(0, 65536, (6, 5, 7), dtype=np.uint16)
    p16 = str(tmp_path / "v16.raw")
    v16.tofile(p16)
    with livre_amd.RawSource(p16 + "?output=uint8#7,5,6,uint16") as src:
        assert src.vi.data_type == livre_amd.DT_UINT8
        assert src.raw_info.input_type == livre_amd.DT_UINT16
        assert np.array_equal(src.brick(), (v16 >> 8).astype(np.uint8))
    v32 = rng.integers(0, 2 ** 32, (3, 4, 5), dtype=np.uint32)
    p32 = str(tmp_path / "v32.raw")
    v32.tofile(p32)
    with livre_amd.RawSource(p32 + "?output=uint16#5,4,3,uint32") as src:
        assert np.array_equal(src.brick(), (v32 >> 16).astype(np.uint16))
    with livre_amd.RawSource(p32 + "?output=uint8#5,4,3,uint32") as src:
        assert np.array_equal(src.brick(), (v32 >> 24).astype(np.uint8))


def test_nrrd_with_attached_data_and_big_endian(lvr, tmp_path):
    v = np.arange(4 * 3 * 2, dtype=">u2").reshape(2, 3, 4)
    p = str(tmp_path / "v.nrrd")
    with open(p, "wb") as f:
        f.write(b"NRRD0004\n# a comment\ntype: ushort\ndimension: 3\nsizes: 4 3 2\n"
                b"endian: big\nencoding: raw\nkey:=value\n\n")
        f.write(v.tobytes())
    with livre_amd.RawSource(p) as src:
        assert list(src.vi.voxels) == [4, 3, 2] and src.raw_info.big_endian == 1
        assert src.vi.data_type == livre_amd.DT_UINT16
        # bytes as stored, as the reference passes them on
        assert src.brick().tobytes() == v.tobytes()


def test_pyramid_volume_info(lvr):
    with livre_amd.RawSource(NRRD) as src:
        vi = src.pyramid_info(16, 1)
        assert vi.depth == 3 and list(vi.root_blocks) == [1, 1, 1]
        assert list(vi.max_block) == [18, 18, 18] and list(vi.voxels) == [41, 41, 41]
        ref = livre_amd.volume_info(41, 16, 1, livre_amd.DT_UINT8)
        assert bytes(vi) == bytes(ref)
