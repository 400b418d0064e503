"""The C-ABI libraries load and export every function their headers declare.
(No compute calls: this runs without a GPU.)"""
import ctypes as C
import os
import re

import pytest

from distributedpanga_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(pga_[A-Za-z0-9_]+)\s*\(", re.M)


def declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
    return sorted(set(DECL.findall(text)))


@pytest.mark.parametrize("header,loader", [("libpga_hip.h", _lib.hip_lib), ("pga_synth.h", _lib.host_lib),
                                           ("pga_gix.h", _lib.host_lib), ("pga_gdbfile.h", _lib.host_lib)])
def test_exports_every_declared_symbol(header, loader):
    names = declared(header)
    assert len(names) >= 3
    L = loader()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_hip_library_declares_the_serving_api():
    names = declared("libpga_hip.h")
    for n in ("pga_create", "pga_build_index", "pga_merge", "pga_seeds_export", "pga_server_run",
              "pga_server_seeds_export"):
        assert n in names


def test_default_opts_without_a_device():
    from distributedpanga_amd import pga
    L = pga._setup(_lib.hip_lib())
    o = pga.Opts()
    L.pga_default_opts(C.byref(o))
    assert o.freq == 10 and o.tile_records in (2048, 4096, 8192) and o.build_cum == 1
    assert o.digit2_bits in (8, 9, 10)
