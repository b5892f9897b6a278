"""Builds the package's HIP libraries in-tree for gfx950:

* cemc_amd/lib/libcemcgpu.so   — the engine (C ABI of include/cemc_gpu.h)
* cemc_amd/lib/libcemcprobe.so — roofline probes (HBM copy, LDS read)

The sweep kernel's instantiations are compiled as separate objects, one per
(waves per chain, lattice path), and the Wang-Landau kernel's one per
lattice path, in parallel, then linked with the rest.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
OBJ_DIR = os.path.join(PKG, "_build")
LIB_PATH = os.path.join(LIB_DIR, "libcemcgpu.so")
PROBE_PATH = os.path.join(LIB_DIR, "libcemcprobe.so")
ARCH = "gfx950"
SWEEP_WAVES = (1, 2, 4, 8)


def _hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _flags():
    return ["--offload-arch=%s" % ARCH, "-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden",
            # no packed FP32 instructions in any HIP code object
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _units():
    """(object name, source, extra defines) of every translation unit of the engine."""
    units = [("ce_kernels", "ce_kernels.hip", []), ("ce_host", "ce_host.hip", []),
             ("ce_field", "ce_field_inst.hip", [])]
    for w in SWEEP_WAVES:
        for p in (0, 1):
            units.append(("ce_sweep_w%d_p%d" % (w, p), "ce_sweep_inst.hip",
                          ["-DCEMC_SWEEP_WAVES=%d" % w, "-DCEMC_SWEEP_PERIODIC=%d" % p]))
    for p in (0, 1):
        for sgc in (0, 1):
            units.append(("ce_wl_p%d_s%d" % (p, sgc), "ce_wl_inst.hip",
                          ["-DCEMC_WL_PERIODIC=%d" % p, "-DCEMC_WL_SGC=%d" % sgc]))
    return units


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)


def _jobs():
    env = os.environ.get("MAX_JOBS")
    n = int(env) if env and env.isdigit() else (os.cpu_count() or 4)
    return max(1, min(n, 16))


def build(verbose=False):
    """Compile the engine and the probes for gfx950 (always recompiles)."""
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    hipcc = _hipcc()
    cmds, objs = [], []
    for name, src, defs in _units():
        obj = os.path.join(OBJ_DIR, name + ".o")
        objs.append(obj)
        cmds.append([hipcc] + _flags() + defs + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB_PATH + ".tmp%d" % os.getpid()
    _run([hipcc] + _flags() + ["-shared"] + objs + ["-o", tmp])
    os.replace(tmp, LIB_PATH)
    tmp = PROBE_PATH + ".tmp%d" % os.getpid()
    _run([hipcc] + _flags() + ["-shared", os.path.join(CSRC, "roofline_probe.hip"), "-o", tmp])
    os.replace(tmp, PROBE_PATH)
    if verbose:
        for c in cmds:
            print(" ".join(c))
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=False))
