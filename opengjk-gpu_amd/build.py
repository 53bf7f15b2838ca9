"""In-tree build of the opengjk-amd libraries.

  lib/libogjk.so      HIP kernels for gfx950 + the C ABI + the C++ drop-in API
  lib/libogjk_gen.so  host-only synthetic workload generator

Run ``python opengjk-gpu_amd/build.py`` (or ``__graft_entry__.build()``).
hipcc cross-compiles gfx950 code objects without a GPU present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("OGJK_OFFLOAD_ARCH", "gfx950")

# Device code: no FMA contraction (the FP64 path must round exactly like the
# CPU checker; FMAs that the FP32 sweep wants are written explicitly), and no
# packed-FP32 instructions.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-warning-option",
    "-Wno-unused-command-line-argument",
]

LIB_SOURCES = ["ogjk_kernels.hip", "ogjk_cluster.hip", "ogjk_api.hip", "ogjk_serve.hip",
               "ogjk_broad.hip", "ogjk_hull.hip", "ogjk_dropin.hip"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_gen() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libogjk_gen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          f"-I{INC}", "-o", out, os.path.join(CSRC, "ogjk_gen.cpp"), "-lpthread"])
    return out


def build_lib() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, src.replace(".hip", ".o")) for src in LIB_SOURCES]
    # one hipcc per source, in parallel (the kernels file dominates)
    from concurrent.futures import ThreadPoolExecutor
    jobs = min(len(LIB_SOURCES), max(1, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC, "-c", *DEVICE_FLAGS, f"-I{INC}", f"-I{CSRC}", "-o",
                                     so[1], os.path.join(CSRC, so[0])]),
                    zip(LIB_SOURCES, objs)))
    out = os.path.join(LIBDIR, "libogjk.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_tests() -> str:
    """The test driver of the reference-named C++ API (tests/dropin_api_check.hip),
    linked against the in-tree libogjk.so."""
    src = os.path.join(ROOT, "tests", "dropin_api_check.hip")
    bindir = os.path.join(ROOT, "tests", "bin")
    os.makedirs(bindir, exist_ok=True)
    out = os.path.join(bindir, "dropin_api_check")
    _run([HIPCC, *DEVICE_FLAGS, f"-I{INC}", "-o", out, src, f"-L{LIBDIR}", "-logjk",
          "-Wl,-rpath,$ORIGIN/../../opengjk-gpu_amd/lib"])
    return out


def build_oracle() -> None:
    """The checker: oracle restatement always; the reference build only where
    the reference checkout exists (never on the GPU host)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "all"])
    if os.path.isdir("/root/reference/GJK/cpu"):
        _run(["make", "-s", "-C", odir, "ref"])


def build_all() -> None:
    build_gen()
    build_oracle()
    build_lib()
    build_tests()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gen": build_gen, "lib": build_lib, "oracle": build_oracle,
         "tests": build_tests}[w]()
