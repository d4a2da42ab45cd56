"""Build recipe for the native library (libqsr_amd.so) and the CLI (QSR_hip).

Compiles the HIP sources for gfx950 with hipcc, in-tree, into
``lcdrsimulator_amd/lib/``.  Packed FP32 instructions are disabled
(``-packed-fp32-ops``); the host sources are compiled by hipcc as well so one
toolchain links everything.  Run ``python -m lcdrsimulator_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB_DIR = os.path.join(PKG, "lib")
OBJ_DIR = os.path.join(PKG, "lib", "obj")
LIB = os.path.join(LIB_DIR, "libqsr_amd.so")
CLI = os.path.join(LIB_DIR, "QSR_hip")
ARCH = os.environ.get("QSR_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
DEV_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"--offload-arch={ARCH}", "-Wno-unused-command-line-argument"]

DEVICE_SOURCES = ["qsr_step.hip", "qsr_obs.hip", "qsr_surr.hip"]
HOST_SOURCES = ["qsr_capi.cpp", "qsr_host.cpp", "qsr_surrogate.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in proc.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if proc.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + "\n".join(lines))
    return "\n".join(lines)


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    flags = DEV_FLAGS if src.endswith(".hip") else HOST_FLAGS
    _run([HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library and the CLI."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, DEVICE_SOURCES + HOST_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    cli_obj = _compile("qsr_cli.cpp") if os.path.exists(os.path.join(CSRC, "qsr_cli.cpp")) else None
    if cli_obj:
        tmp = CLI + ".tmp"
        _run([HIPCC, cli_obj, os.path.join(OBJ_DIR, "qsr_host.o"), "-L" + LIB_DIR, "-lqsr_amd",
              "-Wl,-rpath,$ORIGIN", "-o", tmp])
        os.replace(tmp, CLI)
    if verbose:
        print(f"built {LIB}" + (f" and {CLI}" if cli_obj else ""))
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
