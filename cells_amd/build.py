"""Builds libcells_hip.so in-tree with hipcc for gfx950.

Every HIP source is compiled with -ffp-contract=off (products and sums round as written, so the
factor's pairwise sums are the defined order) and with the packed-FP32 target feature off. Each
build first proves that the feature switch reaches the device pass: a probe kernel whose float2
multiply-add compiles to a packed FP32 instruction without the switch must compile to none with
it (the host pass's warning that it does not know the feature is expected and harmless).
"""

import os
import re
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libcells_hip.so")
SOURCES = ("cells_hip.hip", "symbolic.cpp")
ARCH = "gfx950"


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

PROBE = """#include <hip/hip_runtime.h>
__global__ void probe(float2* a, const float2* b, float s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float2 x = a[i], y = b[i];
  x.x = x.x * s + y.x;
  x.y = x.y * s + y.y;
  a[i] = x;
}
"""
PACKED = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def command():
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
        "-ffp-contract=off", *NO_PACKED_FP32,
        "-I", SRC, "-o", LIB, *(os.path.join(SRC, s) for s in SOURCES),
    ]


def packed_fp32_count(flags):
    """Packed FP32 instructions in the probe kernel's gfx950 assembly compiled with `flags`."""
    with tempfile.TemporaryDirectory() as d:
        src, asm = os.path.join(d, "probe.hip"), os.path.join(d, "probe.s")
        with open(src, "w") as f:
            f.write(PROBE)
        r = subprocess.run([hipcc(), f"--offload-arch={ARCH}", "-O3", "--offload-device-only", "-S", *flags,
                            src, "-o", asm], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed on the packed-FP32 probe:\n" + r.stdout + r.stderr)
        return len(PACKED.findall(open(asm).read()))


def check_packed_fp32_guard():
    """Raise unless the switch removes the probe's packed FP32 instructions."""
    if packed_fp32_count([]) == 0:
        raise RuntimeError("the packed-FP32 probe no longer produces a packed instruction: it proves nothing")
    if packed_fp32_count(NO_PACKED_FP32) != 0:
        raise RuntimeError("the packed-FP32 switch does not reach gfx950's device pass")


def build(verbose=False):
    check_packed_fp32_guard()
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = command()
    cmd[cmd.index("-o") + 1] = tmp
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
