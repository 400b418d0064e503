"""Locating the package's native libraries.

Both are built in-tree by ``distributedpanga_amd/build.py`` (called from
``__graft_entry__.build()``):
  * ``libpga_host.so`` -- host C: synthetic genomes, GIX/seed file writers,
    the CPU-side helpers of the C host driver;
  * ``libpga_hip.so``  -- the HIP kernels for gfx950 behind the C-ABI declared
    in ``include/libpga_hip.h``.
A missing library is an error: there is no fallback path.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HOST_SO = os.path.join(PKG_DIR, "libpga_host.so")
HIP_SO = os.path.join(PKG_DIR, "libpga_hip.so")

_host = None
_hip = None


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_SO):
            raise RuntimeError(f"{HOST_SO} is not built (run __graft_entry__.build())")
        _host = C.CDLL(HOST_SO)
    return _host


def hip_lib():
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_SO):
            raise RuntimeError(f"{HIP_SO} is not built (run __graft_entry__.build()); "
                               "the device path has no CPU fallback")
        _hip = C.CDLL(HIP_SO)
    return _hip
