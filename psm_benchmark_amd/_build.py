"""Builds libpsm.so (HIP kernels for gfx950 + host orchestration) in-tree.

    python -m psm_benchmark_amd._build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsm.so")
SOURCES = ["device.hip", "psm_abi.cpp", "plan.cpp", "gen.cpp", "textio.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PSM_OFFLOAD_ARCH", "gfx950")


def flags() -> list[str]:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            # packed FP32 instructions are kept out of every HIP build
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value",
            "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmd = [HIPCC, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for cmd, p in procs:
        out, _ = p.communicate()
        text = "\n".join(l for l in out.decode().splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            raise RuntimeError("compile failed: " + " ".join(cmd) + "\n" + text)
        if verbose and text.strip():
            print(text)
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    build_dropin()
    return LIB


DROPIN = os.path.join(HERE, "dropin")
DROPIN_LIB = os.path.join(HERE, "libpsm_dropin.so")
DRIVER = os.path.join(HERE, "SubgraphMatching.out")


def build_dropin() -> None:
    """The reference's C++ call surfaces over the C ABI (host C++ only)."""
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + DROPIN]
    common = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wno-unused-result", *inc]
    rpath = "-Wl,-rpath,$ORIGIN"
    subprocess.check_call(["g++", *common, "-shared", "-o", DROPIN_LIB + ".tmp",
                           os.path.join(DROPIN, "psm_dropin.cpp"), "-L" + HERE, "-lpsm", rpath])
    os.replace(DROPIN_LIB + ".tmp", DROPIN_LIB)
    subprocess.check_call(["g++", *common, "-o", DRIVER + ".tmp", os.path.join(DROPIN, "subgraph_matching.cpp"),
                           "-L" + HERE, "-lpsm_dropin", "-lpsm", rpath])
    os.replace(DRIVER + ".tmp", DRIVER)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
