"""Build librrlbm.so (gfx950 HIP kernels + C ABI) and the C++ driver in-tree.

Usage: python build_lib.py            (from any directory)

Objects go to _build/ next to this file; the shared library lands at
recursively-regularized-lbm_amd/librrlbm.so, where the rrlbm package loads it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librrlbm.so")
DRIVER = os.path.join(HERE, "rrlbm_channel")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed-FP32 instructions are disabled for all device code
HIPFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"), "-I" + CSRC,
]
# (source, object, extra flags): the fused kernels build once per storage type
SOURCES = [("kernels_fused.hip", "kernels_fused.o", []), ("kernels_ref.hip", "kernels_ref.o", []),
           ("rrlbm_capi.hip", "rrlbm_capi.o", [])] + [
    ("fused_storage.hip", "fused_%s.o" % name, ["-DRRLBM_FUSED_STORAGE=%d" % k])
    for k, name in enumerate(("f64", "f32", "f16"))]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = [os.path.join(BUILD, obj) for _, obj, _ in SOURCES]

    def compile_one(item):
        src, obj, extra = item
        _run([HIPCC] + HIPFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, obj)])
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # C++ driver (the counterpart of the reference main.cu), linked against the library
    drv = os.path.join(HERE, "driver", "rrlbm_channel.cpp")
    if os.path.exists(drv):
        _run([HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(REPO, "include"), drv, "-o", DRIVER,
              "-L" + HERE, "-lrrlbm", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
