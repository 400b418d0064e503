"""Build the package's native libraries in-tree.

  libpga_hip.so   hipcc, gfx950 only: csrc/pga_api.hip (includes the kernels)
  libpga_host.so  gcc: host C (synthetic genomes, GIX / seed-stream writers)
  pga_drive       gcc: the C host driver (links libpga_hip.so and libpga_host.so)

Run as ``python -m distributedpanga_amd.build``.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I" + INC, "-I" + CSRC]
CC_FLAGS = ["-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + INC]

HOST_SRCS = ["pga_synth.c", "pga_gix.c", "pga_gdbfile.c"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "unknown target feature" not in l and "packed-fp32-ops" not in l
                    and "warning generated" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out:
        print(out, file=sys.stderr)


def build_hip():
    out = os.path.join(PKG, "libpga_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp, os.path.join(CSRC, "pga_api.hip")])
    os.replace(tmp, out)
    return out


def build_host():
    out = os.path.join(PKG, "libpga_host.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in HOST_SRCS]
    missing = [s for s in srcs if not os.path.exists(s)]
    if missing:
        raise FileNotFoundError("host library sources missing: " + ", ".join(missing))
    _run(["gcc", "-shared"] + CC_FLAGS + ["-o", tmp] + srcs + ["-lm"])
    os.replace(tmp, out)
    return out


def build_driver():
    src = os.path.join(CSRC, "pga_drive.c")
    if not os.path.exists(src):
        raise FileNotFoundError("C driver source missing: " + src)
    out = os.path.join(PKG, "pga_drive")
    tmp = out + ".tmp"
    _run(["gcc"] + CC_FLAGS + ["-o", tmp, src, "-L" + PKG, "-lpga_hip", "-lpga_host", "-lm",
                               "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_host(), build_hip(), build_driver()]


if __name__ == "__main__":
    for p in build_all():
        if p:
            print(p)
