"""Ahead-of-time build of libmrad.so for gfx950.

Compiles the host scene builder with g++, the HIP kernels and runtime with
hipcc for gfx950 only, and links both into master-rad_amd/mrad/libmrad.so.
The SHA-256 of the gfx950 code object (the counterpart of the reference's
compiledPtxSha256, gpu/GpuRayTracer.java:814-818) is computed from a
device-only compile of the same source with the same flags and embedded in
the library.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "mrad" / "libmrad.so"
ROOT = HERE.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Numerics contract: no FMA contraction, IEEE division and square root, no
# fast-math.  Packed FP32 instructions are disabled (they are not needed here
# and some of their forms are unreliable on the target boxes).
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]

HIP_SOURCES = ["mrad_runtime.hip", "mrad_wavefront.hip", "mrad_adjoint.hip", "mrad_calibrate.hip"]
CPP_SOURCES = ["scene_builder.cpp"]


def run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def run_all(cmds: list[list[str]]) -> None:
    """Runs independent compiles side by side (one process each)."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(run, cmds))


def code_object_sha256(host_cmds: list[list[str]] = ()) -> str:
    """SHA-256 over the gfx950 code objects of every HIP source; `host_cmds`
    are compiled alongside."""
    h = hashlib.sha256()
    outs = [BUILD / (Path(src).stem + f"-{ARCH}.hsaco") for src in HIP_SOURCES]
    run_all([*host_cmds, *([HIPCC, *DEVICE_FLAGS, "--cuda-device-only", "-o", str(out), str(CSRC / src)]
                           for src, out in zip(HIP_SOURCES, outs))])
    for out in outs:
        h.update(out.read_bytes())
    return h.hexdigest()


def build(verbose: bool = True, out: Path | None = None, defines: tuple[str, ...] = ()) -> Path:
    """Builds the library (to `out`, default master-rad_amd/mrad/libmrad.so);
    `defines` are extra -D tuning macros of the kernels."""
    BUILD.mkdir(exist_ok=True)
    lib = out or LIB
    objs, host_cmds = [], []
    for src in CPP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        host_cmds.append(["g++", *HOST_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    sha = code_object_sha256(host_cmds)
    cmds = []
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        cmds.append([HIPCC, *DEVICE_FLAGS, *(f"-D{d}" for d in defines), "-fPIC",
                     f'-DMRAD_CODE_OBJECT_SHA256="{sha}"', "-I", str(ROOT / "include"),
                     "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    run_all(cmds)
    tmp = lib.with_suffix(".so.tmp")
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build()
    sys.exit(0)
