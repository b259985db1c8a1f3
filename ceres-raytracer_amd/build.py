"""Build the native parts in-tree: libcrt_hip.so (HIP kernels + C ABI, gfx950)
and the pybind11 host module _crt.  Usage: python build.py [--only hip|py]."""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CRT_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["bvh_build.hip", "trace.hip", "render.hip", "sensors.hip", "server.hip", "scene_io.hip"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             # packed FP32 instructions are disabled for every device code object
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I" + INCLUDE, "-I" + CSRC]

LIB_HIP = os.path.join(HERE, "libcrt_hip.so")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs = [os.path.join(HERE, "_build", src.replace(".hip", ".o")) for src in HIP_SOURCES]
    jobs = max(1, min(len(HIP_SOURCES), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    # translation units compile independently; the link waits for all of them
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for f in [pool.submit(_run, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
                  for src, obj in zip(HIP_SOURCES, objs)]:
            f.result()
    tmp = LIB_HIP + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-lpthread", "-o", tmp])
    os.replace(tmp, LIB_HIP)


def build_py():
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_crt" + ext)
    pyinc = sysconfig.get_paths()["include"]
    src = os.path.join(CSRC, "host", "pycrt.cpp")
    if not os.path.exists(src):
        return
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
          "-I" + pybind11.get_include(), "-I" + pyinc, "-I" + INCLUDE, "-I" + CSRC,
          src, "-L" + HERE, "-lcrt_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)


def main(argv):
    only = None
    if "--only" in argv:
        only = argv[argv.index("--only") + 1]
    if only in (None, "hip"):
        build_hip()
    if only in (None, "py"):
        build_py()


if __name__ == "__main__":
    main(sys.argv[1:])
