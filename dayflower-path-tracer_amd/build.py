#!/usr/bin/env python3
"""Builds the two product libraries in-tree (dayflower_amd/lib/).

libdayflower_scene.so - host scene compiler (g++).
libdayflower_hip.so   - HIP kernels for gfx950 + the renderer C ABI (hipcc).

Usage: python dayflower-path-tracer_amd/build.py [--jobs N]
"""

import argparse
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "dayflower_amd", "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, so every float operation rounds as the
# reference's Java float arithmetic does.  Packed FP32 instructions are
# disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]

SCENE_SOURCES = ["scene/df_scene.cpp", "scene/df_configs.cpp", "scene/df_obj.cpp", "scene/df_scene_capi.cpp"]
HIP_SOURCES = ["df_kernels.hip", "df_wavefront.hip", "df_adjoint.hip", "df_runtime.hip", "df_calibrate.hip"]


def run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass reports that it does not know the packed-fp32 feature
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (p.returncode, " ".join(cmd), out))
    return out


def build_scene_lib():
    out = os.path.join(LIB_DIR, "libdayflower_scene.so")
    srcs = [os.path.join(CSRC, s) for s in SCENE_SOURCES]
    run(["g++"] + CXX_FLAGS + ["-shared", "-o", out + ".tmp"] + srcs)
    os.replace(out + ".tmp", out)
    return out


def build_hip_lib(jobs):
    out = os.path.join(LIB_DIR, "libdayflower_hip.so")
    objs = []

    def compile_one(src):
        obj = os.path.join(LIB_DIR, os.path.basename(src) + ".o")
        run([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, src)])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


STAMP_PATH = os.path.join(LIB_DIR, "sources.stamp")


def source_hash():
    """SHA-256 over the names and contents of every source and header the two libraries are built from."""
    h = hashlib.sha256()
    roots = [CSRC, os.path.join(os.path.dirname(HERE), "include")]
    for root in roots:
        for dirpath, _, files in sorted(os.walk(root)):
            for f in sorted(files):
                if f.endswith((".hip", ".cpp", ".h", ".hpp")):
                    path = os.path.join(dirpath, f)
                    h.update(os.path.relpath(path, os.path.dirname(HERE)).encode())
                    h.update(open(path, "rb").read())
    return h.hexdigest()


def is_current():
    """True when both libraries exist and were built from the sources as they are now."""
    libs = [os.path.join(LIB_DIR, n) for n in ("libdayflower_scene.so", "libdayflower_hip.so")]
    if not all(os.path.exists(p) for p in libs) or not os.path.exists(STAMP_PATH):
        return False
    return open(STAMP_PATH).read().strip() == source_hash()


def build(jobs=4):
    os.makedirs(LIB_DIR, exist_ok=True)
    digest = source_hash()
    with ThreadPoolExecutor(max_workers=2) as ex:
        f1 = ex.submit(build_scene_lib)
        f2 = ex.submit(build_hip_lib, jobs)
        out = [f1.result(), f2.result()]
    with open(STAMP_PATH, "w") as fh:
        fh.write(digest + "\n")
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    for path in build(a.jobs):
        print("built", path)
    sys.exit(0)
