"""Build the gfx950 C-ABI library libfftlib_amd.so in-tree with hipcc.

Every source under csrc/ is compiled for gfx950 only, with packed-fp32
instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops), and
linked into c7x_fftlib_amd/libfftlib_amd.so next to this file.

Usage: python -m c7x_fftlib_amd.build [-v]
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfftlib_amd.so")
OBJ_DIR = os.path.join(HERE, "_build")

SOURCES = [
    "fft_kernels.hip",
    "fft_fixed.hip",
    "fft_fixed_class0.hip",
    "fft_fixed_class1.hip",
    "fft_fixed_class2.hip",
    "fft_fixed_class3.hip",
    "fft_ragged_class.hip",
    "fft_large.hip",
    "fft_mfma.hip",
    "fft_mma.hip",
    "fftlib_runtime.cpp",
    "fftlib_fft_c7x.cpp",
    "fftlib_ragged.cpp",
    "fftlib_mma.cpp",
    "fftlib_async.cpp",
    "fftlib_noalloc.hip",
]

ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-result",
    "-I" + os.path.join(HERE, "..", "include"),
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the GPU library cannot be built")
    return exe


def _compile(src, verbose):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), *COMMON, "-x", "hip", "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def clean_stale():
    """Removes objects (and interrupted temporaries) in _build/ whose source is
    no longer in SOURCES, so they are neither linked nor shipped."""
    if not os.path.isdir(OBJ_DIR):
        return
    keep = {s + ".o" for s in SOURCES}
    for f in os.listdir(OBJ_DIR):
        if f not in keep:
            os.remove(os.path.join(OBJ_DIR, f))


def build(verbose=False):
    """Compile all sources and link libfftlib_amd.so; returns the library path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    clean_stale()
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC",
           "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"), "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
