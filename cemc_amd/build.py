"""Builds libcemcgpu.so (HIP, gfx950) in-tree: cemc_amd/lib/libcemcgpu.so."""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libcemcgpu.so")
SOURCES = ["ce_kernels.hip", "ce_host.hip"]
ARCH = "gfx950"


def hipcc_command(out=LIB_PATH):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return [hipcc, "--offload-arch=%s" % ARCH, "-O3", "-fPIC", "-shared", "-std=c++17", "-fvisibility=hidden",
            # no packed FP32 instructions in any HIP code object
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]\
        + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", out]


def build(verbose=False):
    """Compile every HIP source of the package into one shared library."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp%d" % os.getpid()
    cmd = hipcc_command(tmp)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(" ".join(cmd))
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
