"""Failure handling of the C-ABI, forced from the host (no GPU needed): the
reference aborts on these (src/dsort.c:470-490, exit(1) on OOM); the library
returns a PGA_E* code instead.  Device-side cases are in test_gpu_edges.py."""
import ctypes as C

import numpy as np
import pytest

from distributedpanga_amd import pga
from distributedpanga_amd._lib import hip_lib

EINVAL, EOVERFLOW = -3, -4


def _lib():
    return pga._setup(hip_lib())


def _opts(**kw):
    L = _lib()
    o = pga.Opts()
    L.pga_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@pytest.mark.parametrize("field,value", [("freq", 0), ("freq", 65), ("tile_records", 1000),
                                         ("tile_records", 16384), ("chunk_records", 1024),
                                         ("merge_tile", 128), ("merge_tile", 2048),
                                         ("digit2_bits", 7), ("digit2_bits", 12)])
def test_create_rejects_bad_options(field, value):
    L = _lib()
    h = C.c_void_p()
    assert L.pga_create(C.byref(_opts(**{field: value})), C.byref(h)) == EINVAL
    assert not h.value


def test_calls_without_a_context():
    L = _lib()
    assert L.pga_build_index(None) == EINVAL
    assert L.pga_merge(None) == EINVAL
    m = pga.IndexMeta()
    assert L.pga_index_info(None, 0, C.byref(m)) == EINVAL
    s = pga.SeedsMeta()
    assert L.pga_seeds_info(None, 0, C.byref(s)) == EINVAL
    assert L.pga_index_poke(None, 0, 0, 0, 0) == EINVAL
    assert L.pga_last_error(None) == b"no context"


def test_destination_map_rejects_bad_arguments():
    L = _lib()
    counts = np.ones(1024, np.int64)
    ks = (C.c_int * 2)()
    sel = (C.c_int * 1024)()
    ptr = counts.ctypes.data_as(C.POINTER(C.c_int64))
    assert L.pga_destination_map(ptr, 0, ks, sel) == EINVAL
    assert L.pga_destination_map(None, 1, ks, sel) == EINVAL
    assert L.pga_destination_map(ptr, 1, ks, sel) == 0 and list(ks) == [0, 1024]


def test_noalloc_merge_rejects_bad_arguments_on_the_host():
    """Argument checks of pga_merge_gix_noalloc run before any device call."""
    L = hip_lib()
    f = L.pga_merge_gix_noalloc
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                  C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                  C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_void_p]
    n, c = C.c_int64(), C.c_int64()
    fake = C.c_void_p(16)                     # never dereferenced: each call fails its checks first
    assert f(fake, 100, 4, 0, 1, 2, 10, fake, 0, fake, 0, fake, fake, C.byref(n), C.byref(c), None) == EINVAL
    assert f(fake, 100, 4, 4, 4, 2, 10, fake, 0, fake, 0, fake, fake, C.byref(n), C.byref(c), None) == EINVAL
    assert f(fake, 100, 4, 3, 1, 2, 0, fake, 0, fake, 0, fake, fake, C.byref(n), C.byref(c), None) == EINVAL
    assert f(fake, 100, 4, 3, 1, 2, 65, fake, 0, fake, 0, fake, fake, C.byref(n), C.byref(c), None) == EINVAL
    assert f(fake, 100, 4, 3, 1, 0, 10, fake, 0, fake, 0, fake, fake, C.byref(n), C.byref(c), None) == EINVAL
    assert f(fake, 100, (1 << 23) + 1, 3, 1, 2, 10, fake, 0, fake, 0, fake, fake,
             C.byref(n), C.byref(c), None) == -5
