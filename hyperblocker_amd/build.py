"""Builds the native parts of hyperblocker_amd in-tree.

  hyperblocker_amd/libhyperblocker_amd.so   C ABI (include/hyperblocker_amd.h)
                                            + gfx950 kernels
  hyperblocker_amd/libhb_datagen.so         seeded synthetic tables
  hyperblocker_amd/bin/run_hyperblocker_exec  drop-in CLI

Run: python -m hyperblocker_amd.build   (or __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
ARCH = os.environ.get("HB_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
DEVICE_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # One rounding per operation, in source order (Jaro-Winkler tail).
    "-ffp-contract=off",
    # No packed fp32 instructions (they are generated from plain float code).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    units = [
        ([HIPCC] + DEVICE_FLAGS + ["-c", os.path.join(CSRC, "hb_kernels.hip"),
                                   "-o", os.path.join(OBJ, "hb_kernels.o")]),
        ([HIPCC] + DEVICE_FLAGS + ["-c", os.path.join(CSRC, "hb_sort.hip"),
                                   "-o", os.path.join(OBJ, "hb_sort.o")]),
        ([HIPCC] + HOST_FLAGS + ["-c", os.path.join(CSRC, "hb_runtime.cpp"),
                                 "-o", os.path.join(OBJ, "hb_runtime.o")]),
        (["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, "hb_host.cpp"),
                                 "-o", os.path.join(OBJ, "hb_host.o")]),
        (["g++"] + HOST_FLAGS + ["-shared", os.path.join(CSRC, "datagen.cpp"),
                                 "-o", os.path.join(PKG, "libhb_datagen.so"), "-lpthread"]),
    ]
    with ThreadPoolExecutor(max_workers=len(units)) as ex:
        for out in ex.map(_run, units):
            if verbose and out:
                print(out)
    lib = os.path.join(PKG, "libhyperblocker_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
         + [os.path.join(OBJ, n) for n in ("hb_kernels.o", "hb_sort.o", "hb_runtime.o",
                                           "hb_host.o")]
         + ["-lpthread"])
    os.replace(tmp, lib)
    exe = os.path.join(PKG, "bin", "run_hyperblocker_exec")
    _run([HIPCC] + HOST_FLAGS + [os.path.join(CSRC, "run_hyperblocker.cpp"), "-o", exe,
                                 "-L" + PKG, "-lhyperblocker_amd",
                                 "-Wl,-rpath,$ORIGIN/.."])
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
