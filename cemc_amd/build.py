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
PROBE_PATH = os.path.join(LIB_DIR, "libcemcprobe.so")
PROBE_SOURCES = ["roofline_probe.hip"]
ARCH = "gfx950"


def hipcc_command(out=LIB_PATH, sources=SOURCES):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return [hipcc, "--offload-arch=%s" % ARCH, "-O3", "-fPIC", "-shared", "-std=c++17", "-fvisibility=hidden",
            # no packed FP32 instructions in any HIP code object
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]\
        + [os.path.join(CSRC, s) for s in sources] + ["-o", out]


def _compile(out, sources, verbose):
    tmp = out + ".tmp%d" % os.getpid()
    cmd = hipcc_command(tmp, sources)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, out)
    if verbose:
        print(" ".join(cmd))


def build(verbose=False):
    """Compile the engine (libcemcgpu.so) and the roofline probes
    (libcemcprobe.so) for gfx950."""
    os.makedirs(LIB_DIR, exist_ok=True)
    _compile(LIB_PATH, SOURCES, verbose)
    _compile(PROBE_PATH, PROBE_SOURCES, verbose)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
