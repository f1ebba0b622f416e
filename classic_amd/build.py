"""Build libclawgpu.so in-tree with hipcc for gfx950.

``python -m classic_amd.build`` (or ``build()``) compiles every HIP source
under ``classic_amd/csrc`` into ``classic_amd/libclawgpu.so``.  FMA
contraction is switched off so that the device arithmetic follows the
reference's rounding step by step, and packed FP32 instructions are disabled.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libclawgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CLAWGPU_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]
# per-source additions: the network's MFMA results stay in VGPRs (no
# accumulator-register reads after every matrix instruction)
FILE_FLAGS = {"learned.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
OBJ_DIR = os.path.join(HERE, "_build")


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile each source to an object (in parallel) and link the library."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in srcs]
    cmds = [[HIPCC, *FLAGS, *FILE_FLAGS.get(os.path.basename(s), []), "-c", "-o", o, s]
            for s, o in zip(srcs, objs)]
    jobs = jobs or min(len(cmds), int(os.environ.get("MAX_JOBS", "8")), os.cpu_count() or 1)
    running = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append(subprocess.Popen(cmd))
        if len(running) >= jobs:
            _wait(running.pop(0), running)
    while running:
        _wait(running.pop(0), running)
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    return LIB


def build_code_object(src, out, verbose=False, extra=()):
    """Compile a HIP source of device-side user routines (include/clawgpu.h,
    clawgpu_dev_grid) or user solvers (include/clawgpu_solver.h, which needs
    the kernel headers of csrc/) into a gfx950 code object for
    clawgpu_dev_hooks_load / clawgpu_solver2_load, with the library's
    arithmetic flags (no contraction unless the source asks for it) and any
    extra flags."""
    cmd = [HIPCC, "--genco", *[f for f in FLAGS if f not in ("-fPIC",)],
           "-I", os.path.join(HERE, "csrc"), *extra, "-o", out, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


def _wait(proc, others):
    if proc.wait() != 0:
        for p in others:
            p.wait()
        raise subprocess.CalledProcessError(proc.returncode, proc.args)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
