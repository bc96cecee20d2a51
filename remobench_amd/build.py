"""Builds the product library libremolod.so (HIP kernels for gfx950 + the C ABI).

The library is built in-tree (remobench_amd/_lib/) so it travels with the
repository; nothing is installed. `python -m remobench_amd.build` rebuilds it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libremolod.so"

HIP_SOURCES = ["construct.hip", "accumulate.hip", "analysis.hip", "mlp.hip",
               "refine.hip", "context.hip"]
CPP_SOURCES = ["scene.cpp"]

# gfx950 only; IEEE division and no FMA contraction (the quantisation and the
# voxel centres must round exactly as the oracle does); packed FP32 disabled.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    objs = []
    procs = []
    for src in HIP_SOURCES + CPP_SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = objdir / (src + ".o")
        flags = list(HIP_FLAGS) if src.endswith(".hip") else ["-O3", "-std=c++17", "-fPIC", "-Wall"]
        cmd = [cc, *flags, "-c", str(path), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd = ["g++", *flags, "-c", str(path), "-o", str(obj)]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(f"--- {src}\n{text}")
        elif verbose and text.strip():
            print(f"--- {src}\n{text}")
    if failed:
        raise RuntimeError("HIP build failed:\n" + "\n".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    tmp.replace(LIB)
    return LIB


HOST = PKG / "host"
HOST_SOURCES = ["PointSource.cpp", "SimlodReader.cpp", "RemolodHipPipeline.cpp", "PipelineRegistry.cpp"]
HOST_LIB = LIBDIR / "libremolod_host.so"
HEADLESS = LIBDIR / "remolod_headless"


def build_host() -> Path:
    """The C++ host side (ILodPipeline mirror, point source, registry) over the C
    ABI, plus the headless frame-loop driver. Plain g++: no HIP headers needed."""
    inc = ["-I", str(ROOT / "include"), "-I", str(HOST)]
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]
    rpath = "-Wl,-rpath,$ORIGIN"
    cmd = ["g++", *flags, *inc, "-shared", "-o", str(HOST_LIB) + ".tmp",
           *[str(HOST / s) for s in HOST_SOURCES], "-L", str(LIBDIR), "-lremolod", rpath]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("host library build failed:\n" + r.stdout.decode(errors="replace"))
    Path(str(HOST_LIB) + ".tmp").replace(HOST_LIB)
    cmd = ["g++", *flags, *inc, "-o", str(HEADLESS) + ".tmp", str(HOST / "remolod_headless.cpp"),
           "-L", str(LIBDIR), "-lremolod_host", "-lremolod", rpath]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("headless driver build failed:\n" + r.stdout.decode(errors="replace"))
    Path(str(HEADLESS) + ".tmp").replace(HEADLESS)
    return HEADLESS


def build_oracle() -> Path:
    """Compiles the CPU oracle (test infrastructure) with its own Makefile."""
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout.decode(errors="replace"))
    return ROOT / "oracle" / "_build" / "liboracle.so"


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_host())
    print(build_oracle())
