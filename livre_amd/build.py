"""Build liblivre_amd.so in-tree for gfx950 with hipcc.

Usage: python -m livre_amd.build  (or livre_amd.build.build() from Python).
Objects go to livre_amd/_build/, the library to livre_amd/liblivre_amd.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblivre_amd.so")
SOURCES = ["octree.cpp", "pool.cpp", "noise.hip", "render.hip", "histogram.hip",
           "select.hip", "fields.hip", "autotune.cpp",
           "server.cpp", "stream.cpp", "raw.hip", "jpeg.hip", "session.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LIVRE_AMD_ARCH", "gfx950")

# -ffp-contract=off: every FP32 expression of the shader is rounded as written
#   (no fused multiply-add), which keeps sample positions and texel choices
#   identical to the FP32 restatement.
# -fhip-fp32-correctly-rounded-divide-sqrt: IEEE division and square root.
# -packed-fp32-ops off: no v_pk_*_f32 generated.
FLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
         "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, res.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=6) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cpp_driver()
    return LIB


CPP_DRIVER = os.path.join(HERE, "render_frame")


def build_cpp_driver():
    """The C++ drop-in classes (livre_amd/cpp/livre/hip/Renderer.h) compiled
    by the host compiler against the C ABI, as a Livre build would."""
    root = os.path.dirname(HERE)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra",
           "-I", os.path.join(root, "include"), "-I", os.path.join(HERE, "cpp"),
           os.path.join(HERE, "cpp", "render_frame.cpp"), "-o", CPP_DRIVER + ".tmp",
           "-L", HERE, "-llivre_amd", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed on render_frame.cpp:\n" + res.stderr)
    os.replace(CPP_DRIVER + ".tmp", CPP_DRIVER)
    return CPP_DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
