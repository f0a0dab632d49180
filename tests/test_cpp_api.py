"""The C++ drop-in classes (livre_amd/cpp/livre/hip/Renderer.h) driven the way
Livre's Channel drives RayCastRenderer: selection, order, sync upload from a
DataSourcePlugin, render, histogram. Compiled by the host compiler against
the C ABI (livre_amd/render_frame)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import livre_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "livre_amd", "render_frame")


@pytest.fixture(scope="module")
def driver(lvr):
    if not os.path.exists(DRIVER):
        from livre_amd import build
        build.build_cpp_driver()
    return DRIVER


def test_host_side_api(driver):
    nucleon = os.path.join(ROOT, "tests", "golden", "nucleon.raw")
    res = subprocess.run([driver, "--host-check", nucleon], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr


@pytest.mark.gpu
def test_cpp_pipeline_frame_matches_oracle(driver, orc, tmp_path):
    prefix = str(tmp_path / "frame")
    n, block, w, h = 64, 16, 96, 72
    res = subprocess.run([driver, prefix, str(n), str(block), str(w), str(h)],
                         capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    meta = {}
    for line in open(prefix + ".txt"):
        k, *v = line.split()
        meta[k] = v
    mv = np.array(meta["mv"], np.float32)
    proj = np.array(meta["proj"], np.float32)
    visible = np.array([int(x) for x in meta["visible"]], np.uint64)
    img = np.fromfile(prefix + ".rgba", np.uint8).reshape(h, w, 4)
    with open(prefix + ".jpg", "rb") as f:
        assert f.read() == orc.jpeg_encode(img, quality=100)
    hist = np.fromfile(prefix + ".hist", np.uint64)

    vi = livre_amd.volume_info(n, block, 1, livre_amd.DT_UINT16)
    ovi = orc.volume_info([n] * 3, [block] * 3, [1] * 3, orc.DT_UINT16)
    sel = orc.select_visibles(ovi, mv, proj, h, 1.0, 0, ovi.depth - 1)
    ordered = orc.order(ovi, mv, sel)
    assert np.array_equal(ordered, visible)
    p = livre_amd.noise_params(7, octaves=4, base_cells=4, gain=1.5)
    bricks = [livre_amd.noise_brick_host(vi, p, i) for i in ordered]
    x = np.arange(256, dtype=np.float32) / np.float32(255.0)
    tf = np.stack([x, np.float32(1) - x, np.full(256, 0.5, np.float32),
                   np.float32(0.4) * x * x], axis=1).astype(np.float32)
    fr = livre_amd.frustum(mv, proj)
    cam = orc.camera(list(fr.inv_proj), list(fr.inv_mv), list(fr.eye), fr.near_plane,
                     fr.far_plane, w, h)
    ref, ref_samples = orc.render(ovi, cam, ordered, bricks, tf,
                                  orc.render_params(n, 0.999, (0.0, 65535.0)))
    d = np.abs(img.astype(int) - orc.to_rgba8(ref).astype(int))
    assert d.max() <= 2 and d.mean() <= 0.5
    assert int(meta["samples"][0]) == ref_samples
    assert int(meta["sortlast_maxdiff"][0]) <= 2  # two layers composited vs whole
    assert np.array_equal(hist, orc.histogram(ovi, sel, [
        livre_amd.noise_brick_host(vi, p, i) for i in sel], 65536))


def _compile_only(src, extra_include=None):
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "livre_amd", "cpp"),
           "-I", os.path.join(ROOT, "tests", "cpp")]
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror"]
                          + inc + [src], capture_output=True, text=True, timeout=120)


def test_livre_integration_compiles_beside_livre_declarations():
    """livre/hip/LivreIntegration.h and livre/hip/Renderer.h in one
    translation unit with the declarations of Livre's own classes: no name is
    redefined, livre::hip::RayCastRenderer is a livre::Renderer with the
    reference's constructor and update(), ProceduralDataSource is a
    livre::DataSourcePlugin with the plugin statics."""
    res = _compile_only(os.path.join(ROOT, "tests", "cpp", "livre_adapter_check.cpp"))
    assert res.returncode == 0, res.stderr


def test_integration_snippet_compiles(tmp_path):
    """The C++ snippet of INTEGRATION.md marked as compiled."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    marker = "<!-- compiled by tests/test_cpp_api.py -->"
    assert marker in text
    block = text.split(marker, 1)[1]
    code = block.split("```cpp", 1)[1].split("```", 1)[0]
    src = tmp_path / "snippet.cpp"
    src.write_text('#include "livre_signatures.h"\n' + code)
    res = _compile_only(str(src))
    assert res.returncode == 0, res.stderr
