"""Builds the gfx950 library and the drop-in CLIs in-tree with hipcc.

Outputs (git-ignored, shipped with the tree):
  heterosync_amd/libheterosync_amd.so   the C-ABI library (include/heterosync.h)
  heterosync_amd/bin/allSyncPrims-1kernel, seqlocks, flags, refCounter,
                     splitCounter, uts          the drop-in command lines
  tests/device_api/libhs_device_api_test.so    user kernels on the public
                                               device API (GPU tests)

Usage:  python -m heterosync_amd.build   (or heterosync_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
BIN = os.path.join(HERE, "bin")
LIB = os.path.join(HERE, "libheterosync_amd.so")

ARCH = os.environ.get("HS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-parameter",
]

LIB_SOURCES = [
    "hs_host.hip",
    "hs_syncprims.hip",
    "hs_relaxed.hip",
    "hs_uts.hip",
    "hs_batch.hip",
]

# Test extensions: kernels written against the public device API
# (include/heterosync_device.h), loaded by the GPU tests.
ROOT = os.path.dirname(HERE)
TEST_LIBS = {
    os.path.join(ROOT, "tests", "device_api", "libhs_device_api_test.so"):
        os.path.join(ROOT, "tests", "device_api", "user_kernels.hip"),
}

CLIS = {
    "allSyncPrims-1kernel": "cli/all_sync_prims.cpp",
    "seqlocks": "cli/seqlocks.cpp",
    "flags": "cli/flags.cpp",
    "refCounter": "cli/ref_counter.cpp",
    "splitCounter": "cli/split_counter.cpp",
    "uts": "cli/uts.cpp",
    "hs_probe": "cli/hs_probe.hip",
}


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = False) -> str:
    """Compiles every HIP source for gfx950 and links the library + CLIs."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")) or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for name, src in CLIS.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        out = os.path.join(BIN, name)
        if src.endswith(".hip"):  # device code of its own
            _run([HIPCC, *HIP_FLAGS, path, "-o", out])
        else:
            _run([HIPCC, "-O2", "-std=c++17", path, "-o", out, f"-L{HERE}",
                  "-lheterosync_amd", "-Wl,-rpath,$ORIGIN/.."])
    for out, src in TEST_LIBS.items():
        if os.path.exists(src):
            _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{os.path.join(ROOT, 'include')}", src,
                  "-o", out])
    if verbose:
        print("built", LIB)
    return LIB


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
